"""CPU checks of the C ABI library and the host-side logic (no GPU calls)."""
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "shenfun_amd.h")


def header_functions():
    txt = open(HEADER).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(?:int|int64_t|const char\*)\s+(sfa_\w+)\s*\(", txt)))


def test_library_exports_every_header_symbol():
    import ctypes
    import torch  # noqa: F401  (shares its HIP runtime with the library)
    from shenfun_amd import _lib
    assert os.path.exists(_lib.LIBPATH), "build the library first (__graft_entry__.build())"
    L = ctypes.CDLL(_lib.LIBPATH)
    names = header_functions()
    assert len(names) >= 20
    for n in names:
        assert hasattr(L, n), n
    # every header function is bound with a signature on the Python side
    assert set(names) == set(_lib.exported_symbols())
    assert _lib.lib().sfa_version() >= 100


def test_error_reporting_without_gpu_work():
    """Argument errors are reported before any device work."""
    from shenfun_amd import _lib
    L = _lib.lib()
    rc = L.sfa_tdma_lu(None, 1, 1, None)  # n < 2 is rejected on the host
    assert rc == -1
    assert b"TDMA" in L.sfa_last_error()
    rc = L.sfa_fft_r2c(None, None, 1, _lib.arr([7]), _lib.arr([1]), _lib.arr([1]), 0, 1.0, None)
    assert rc == -1 and b"null" in L.sfa_last_error()
    rc = L.sfa_pad_axis(None, None, 0, 1, _lib.arr([4]), _lib.arr([1]), _lib.arr([6]), _lib.arr([1]),
                        0, 0, 0, 4, 0, 0, 0, None)
    assert rc == -1
    rc = L.sfa_helmholtz3d(None, None, 4, 4, 4, 1.0, 0, 0, 0, None, None, 0, None)
    assert rc == -1


def test_band_pattern_solver_selection_and_size_checks():
    """SparseMatrix.get_solver (matrixbase.py:439-474) and the host-side
    minimum sizes of the band solvers (no device work)."""
    import shenfun_amd as sfa
    from shenfun_amd import _lib, la
    cases = {(0,): la.DiagMA, (0, 2): la.TwoDMA, (-2, 0, 2): la.TDMA, (0, 2, 4): la.ThreeDMA,
             (-2, 0, 2, 4): la.FDMA, (-4, -2, 0, 2, 4): la.PDMA,
             (-4, -2, 0, 2, 4, 6, 8): la.HeptaDMA}
    for keys, cls in cases.items():
        assert sfa.SparseMatrix({k: 1.0 + (k == 0) for k in keys}, (12, 12)).get_solver() is cls
    assert sfa.SparseMatrix({-1: 1.0, 0: -2.0, 1: 1.0}, (9, 9)).get_solver() is la.TDMA_O
    assert sfa.SparseMatrix({-1: 2.0, 0: -2.0, 1: 1.0}, (9, 9)).get_solver() is la.Solve
    assert sfa.SparseMatrix({-2: 1.0, 0: 2.0, 4: 1.0}, (9, 9)).get_solver() is la.Solve
    L = _lib.lib()
    assert L.sfa_heptadma_lu(None, 8, 1, None) == -1  # null data
    for fn, nmin in ((L.sfa_fdma_lu, 4), (L.sfa_heptadma_lu, 8), (L.sfa_tdma_o_lu, 1)):
        assert fn(None, nmin - 1, 1, None) == -1
    for name, nmin in (("sfa_twodma_solve", 2), ("sfa_threedma_solve", 4), ("sfa_fdma_solve", 4),
                       ("sfa_heptadma_solve", 8), ("sfa_tdma_o_solve", 1)):
        rc = getattr(L, name)(None, _lib.F64, 1, _lib.arr([16]), _lib.arr([1]), 0, None, nmin - 1,
                              0, 0, None)
        assert rc == -1 and name.encode() in L.sfa_last_error()


def test_matrices_match_oracle_closed_forms():
    from oracle import spectral as S
    import shenfun_amd as sfa
    for fam in ("chebyshev", "legendre"):
        for bc in ("D", "N"):
            V = sfa.FunctionSpace(14, fam, bc=bc)
            A = sfa.stiffness_matrix(V)
            B = sfa.mass_matrix(V)
            fA = {("chebyshev", "D"): S.cheb_A_SD, ("chebyshev", "N"): S.cheb_A_SN,
                  ("legendre", "D"): S.leg_A_SD, ("legendre", "N"): S.leg_A_SN}[(fam, bc)]
            fB = {("chebyshev", "D"): S.cheb_B_SD, ("chebyshev", "N"): S.cheb_B_SN,
                  ("legendre", "D"): S.leg_B_SD, ("legendre", "N"): S.leg_B_SN}[(fam, bc)]
            assert np.allclose(A.todense(), S.dense(fA(14), 12), rtol=1e-14, atol=1e-12)
            assert np.allclose(B.todense(), S.dense(fB(14), 12), rtol=1e-14, atol=1e-14)


def test_leggauss_matches_reference_fastgl():
    from shenfun_amd.bases import leggauss
    g = np.load(os.path.join(ROOT, "tests", "golden", "fastgl_leggauss.npz"))
    eps = np.finfo(float).eps
    for N in g["Ns"]:
        x, w = leggauss(int(N))
        assert np.max(np.abs(x - g[f"x{N}"])) <= 4 * eps
        assert np.max(np.abs(w - g[f"w{N}"]) / g[f"w{N}"]) <= 32 * eps


def test_solver_generic_assembly_host():
    """Per-mode DIA data of SolverGeneric1ND equals A + (alpha + k^2) B (host only)."""
    import shenfun_amd as sfa
    from oracle import spectral as S

    class FakeT:
        dimensions = 2

        def __init__(self, bases):
            self.bases = bases

        def shape(self, fo=True):
            return (self.bases[0].N, self.bases[1].N // 2 + 1)

        def nonperiodic_axis(self):
            return 0

    SD = sfa.FunctionSpace(12, "legendre", bc=(0, 0))
    K = sfa.FunctionSpace(8, "F", dtype="d")
    T = FakeT([SD, K])
    mats = sfa.helmholtz_tpmatrices(T, 3.0)
    H = sfa.la.SolverGeneric1ND(mats)
    assert H.kind == "tdma" and H._data.shape == (5, 3, 10)
    for j, k in enumerate(np.fft.rfftfreq(8, 1 / 8)):
        sc = 3.0 + k * k
        d = {0: S.leg_A_SD(12)[0] + sc * S.leg_B_SD(12)[0], -2: sc * S.leg_B_SD(12)[-2],
             2: sc * S.leg_B_SD(12)[2]}
        assert np.allclose(H._data[j], S.tdma_dia(d, 10))


def test_functionspace_factory():
    import shenfun_amd as sfa
    assert isinstance(sfa.FunctionSpace(8, "F", dtype="D"), sfa.C2C)
    assert isinstance(sfa.FunctionSpace(8, "F", dtype="d"), sfa.R2C)
    assert sfa.FunctionSpace(8, "chebyshev", bc=(0, 0)).bc_kind == "D"
    assert sfa.FunctionSpace(8, "chebyshev", bc={"left": {"N": 0}, "right": {"N": 0}}).bc_kind == "N"
    assert sfa.FunctionSpace(8, "legendre").bc_kind is None
    V = sfa.FunctionSpace(8, "chebyshev", bc=(1, -2))
    assert V.bc_kind == "D" and V.bc_values == (1.0, -2.0) and V.has_nonhomogeneous_bcs
    W = sfa.FunctionSpace(8, "legendre", bc={"left": {"N": 0.5}, "right": {"N": 0}})
    assert W.bc_kind == "N" and W.bc_values == (0.5, 0.0)
    assert not sfa.FunctionSpace(8, "chebyshev", bc=(0, 0)).has_nonhomogeneous_bcs
    M = sfa.FunctionSpace(8, "chebyshev", bc={"left": {"D": 1}, "right": {"N": 0}})
    assert M.bc_kind == "DN" and M.bc_values == (1.0, 0.0) and M.has_nonhomogeneous_bcs
    with pytest.raises(NotImplementedError):
        sfa.FunctionSpace(8, "laguerre")
