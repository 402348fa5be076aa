"""Build recipe for the oracle's native parts (test infrastructure only).

* ``oracle/banded.c`` -> ``oracle/_build/liboracle.so`` (gcc -O3, OpenMP).
* When the reference tree is present (this container only): the reference's
  ``shenfun/legendre/fastgl/fastgl.cpp`` compiled where it lies, together with
  ``oracle/fastgl_shim.cpp``, -> ``oracle/_ref/libfastgl_ref.so``; and the
  reference's Cython kernels ``shenfun/optimization/cython/{la,transforms,
  Matvec}.pyx``, translated where they lie by Cython (language level 3, C++)
  into ``oracle/_ref/*.cpp`` and compiled with g++ against the Python and
  NumPy headers into extension modules ``oracle/_ref/<name>*.so`` (they need
  nothing else: numpy, libc, libcpp).  Nothing is copied out of the
  reference; the outputs are git-ignored and travel with the working tree.
  The extension modules are rebuilt only when their source is newer.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
REF_FASTGL = "/root/reference/shenfun/legendre/fastgl"
REF_CYTHON = "/root/reference/shenfun/optimization/cython"
CYTHON_MODULES = ("la", "transforms", "Matvec")


def run(cmd):
    print(" ".join(cmd), flush=True)
    subprocess.check_call(cmd)


def build(verbose=True):
    bdir = os.path.join(HERE, "_build")
    os.makedirs(bdir, exist_ok=True)
    run(["gcc", "-O3", "-fopenmp", "-shared", "-fPIC", "-std=c99",
         os.path.join(HERE, "banded.c"), "-o", os.path.join(bdir, "liboracle.so"), "-lm"])
    if os.path.isdir(REF_FASTGL):
        rdir = os.path.join(HERE, "_ref")
        os.makedirs(rdir, exist_ok=True)
        run(["g++", "-O2", "-shared", "-fPIC", "-I", REF_FASTGL,
             os.path.join(REF_FASTGL, "fastgl.cpp"), os.path.join(HERE, "fastgl_shim.cpp"),
             "-o", os.path.join(rdir, "libfastgl_ref.so")])
    if os.path.isdir(REF_CYTHON):
        build_cython_ref()


def build_cython_ref():
    """oracle/_ref/<name><ext suffix> for the reference's Cython kernels."""
    import sysconfig
    import numpy
    rdir = os.path.join(HERE, "_ref")
    os.makedirs(rdir, exist_ok=True)
    suffix = sysconfig.get_config_var("EXT_SUFFIX")
    inc = [sysconfig.get_paths()["include"], numpy.get_include()]
    for name in CYTHON_MODULES:
        src = os.path.join(REF_CYTHON, name + ".pyx")
        cpp = os.path.join(rdir, name + ".cpp")
        so = os.path.join(rdir, name + suffix)
        if os.path.exists(so) and os.path.getmtime(so) >= os.path.getmtime(src):
            continue
        run([sys.executable, "-m", "cython", "-3", "--cplus", src, "-o", cpp])
        run(["g++", "-O2", "-shared", "-fPIC", "-std=c++14", "-w"] + [f"-I{d}" for d in inc]
            + [cpp, "-o", so])
        os.remove(cpp)  # the translation is an intermediate


def load_cython_ref(name):
    """The compiled reference module ``name`` (la, transforms, Matvec), or None
    where it was not built."""
    import importlib.machinery
    import importlib.util
    import sysconfig
    so = os.path.join(HERE, "_ref", name + sysconfig.get_config_var("EXT_SUFFIX"))
    if not os.path.exists(so):
        return None
    loader = importlib.machinery.ExtensionFileLoader(name, so)
    spec = importlib.util.spec_from_file_location(name, so, loader=loader)
    mod = importlib.util.module_from_spec(spec)
    loader.exec_module(mod)
    return mod


if __name__ == "__main__":
    build()
    sys.exit(0)
