"""The name-based backend dispatch (shenfun/optimization/__init__.py:37-80) on
the host: resolution by __name__ then __qualname__ with '.' -> '_', the
configured mode read at decoration (optimizer) or at every call
(runtimeoptimizer), and the HIP backend exporting shenfun's Cython names."""
import types

import pytest

from shenfun_amd import optimization as opt
from shenfun_amd.config import config

# shenfun's Cython entry points on the path (la.pyx, transforms.pyx, Matvec.pyx)
CYTHON_NAMES = [
    "TDMA_LU", "TDMA_Solve", "PDMA_LU", "PDMA_Solve", "TDMA_O_LU", "TDMA_O_Solve", "FDMA_LU",
    "FDMA_Solve", "HeptaDMA_LU", "HeptaDMA_Solve", "TwoDMA_Solve", "ThreeDMA_Solve", "DiagMA_Solve",
    "TDMA_inner_solve", "PDMA_inner_solve", "TDMA_O_inner_solve", "FDMA_inner_solve",
    "HeptaDMA_inner_solve", "TwoDMA_inner_solve", "ThreeDMA_inner_solve", "DiagMA_inner_solve",
    "SolverGeneric1ND_solve_data", "LU_Helmholtz", "Solve_Helmholtz", "Helmholtz_matvec",
    "Helmholtz_Neumann_matvec", "Poisson_Solve_ADD", "scalar_product", "evaluate_expansion_all",
    "LU_Biharmonic", "Biharmonic_factor_pr", "Biharmonic_Solve", "Biharmonic_matvec"]


@pytest.fixture
def fake_backend(monkeypatch):
    mod = types.ModuleType("fake_backend")
    mod.Solve = lambda *a: "by-name"
    mod.TDMA_LU = lambda *a: "by-qualname"
    monkeypatch.setitem(opt.BACKENDS, "fake", "fake_backend")
    monkeypatch.setitem(__import__("sys").modules, "fake_backend", mod)
    return mod


def _f(name, qualname):
    f = lambda *a: "python"  # noqa: E731
    f.__name__, f.__qualname__ = name, qualname
    return f


def test_resolution_order(fake_backend, monkeypatch):
    monkeypatch.setitem(config["optimization"], "mode", "fake")
    assert opt.optimizer(_f("Solve", "TDMA.Solve"))() == "by-name"
    assert opt.optimizer(_f("LU", "TDMA.LU"))() == "by-qualname"
    assert opt.optimizer(_f("other", "X.other"))() == "python"  # missing: the function itself
    monkeypatch.setitem(config["optimization"], "mode", "python")
    assert opt.optimizer(_f("Solve", "TDMA.Solve"))() == "python"  # no backend of that name


def test_runtimeoptimizer_reads_the_mode_per_call(fake_backend, monkeypatch):
    r = opt.runtimeoptimizer(_f("LU", "TDMA.LU"))
    monkeypatch.setitem(config["optimization"], "mode", "fake")
    assert r() == "by-qualname"
    monkeypatch.setitem(config["optimization"], "mode", "none")
    assert r() == "python"


def test_hip_backend_exports_the_cython_names():
    hip = opt.backend("hip")
    missing = [n for n in CYTHON_NAMES if not callable(getattr(hip, n, None))]
    assert not missing, missing
    assert sorted(hip.__all__) == sorted(CYTHON_NAMES)


def test_solver_classes_resolve_to_the_backend():
    """la's classes carry shenfun's static methods; with mode 'hip' they are the
    backend's functions (TDMA.Solve -> TDMA_Solve, Helmholtz.LU_Helmholtz ->
    LU_Helmholtz at call time)."""
    from shenfun_amd import la
    hip = opt.backend("hip")
    for cls, names in ((la.TDMA, ("LU", "Solve")), (la.PDMA, ("LU", "Solve")), (la.FDMA, ("LU", "Solve")),
                       (la.HeptaDMA, ("LU", "Solve")), (la.TDMA_O, ("LU", "Solve")),
                       (la.TwoDMA, ("Solve",)), (la.ThreeDMA, ("Solve",)), (la.DiagMA, ("Solve",))):
        for n in names:
            f = getattr(cls, n)
            assert f.__wrapped__.__qualname__ == f"{cls.__name__}.{n}"
    assert la.Helmholtz.LU_Helmholtz.__qualname__ == "Helmholtz.LU_Helmholtz"
    assert la.SolverGeneric1ND.solve_data.__qualname__ == "SolverGeneric1ND.solve_data"
    assert callable(hip.SolverGeneric1ND_solve_data)


def test_no_host_fallback(monkeypatch):
    """With a mode that names no backend, the runtime-resolved methods raise
    (there is no CPU implementation to fall back to)."""
    from shenfun_amd import la
    monkeypatch.setitem(config["optimization"], "mode", "cython")
    with pytest.raises(NotImplementedError):
        la.Helmholtz.Solve_Helmholtz(None, None, False, None, None, None, None, 0)
