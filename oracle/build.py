"""Build recipe for the oracle's native parts (test infrastructure only).

* ``oracle/banded.c`` -> ``oracle/_build/liboracle.so`` (gcc -O3, OpenMP).
* When the reference tree is present (this container only): the reference's
  ``shenfun/legendre/fastgl/fastgl.cpp`` compiled where it lies, together with
  ``oracle/fastgl_shim.cpp``, -> ``oracle/_ref/libfastgl_ref.so``.  Nothing is
  copied out of the reference; the built library is git-ignored and travels
  with the working tree.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
REF_FASTGL = "/root/reference/shenfun/legendre/fastgl"


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


if __name__ == "__main__":
    build()
    sys.exit(0)
