"""The oracle's restatements against the reference's own kernels: shenfun's
Cython modules la.pyx, transforms.pyx and Matvec.pyx, translated and compiled
where they lie into oracle/_ref by oracle/build.py (this container only; the
tests skip where that build is absent).  Same inputs, results equal to
round-off (the restatements follow the reference's operation order, so most
agree bit for bit)."""
import numpy as np
import pytest

from oracle import build as obuild
from oracle import spectral as S

la = obuild.load_cython_ref("la")
tr = obuild.load_cython_ref("transforms")
mv = obuild.load_cython_ref("Matvec")
pytestmark = pytest.mark.skipif(la is None or tr is None or mv is None,
                                reason="reference Cython kernels not built (oracle/_ref)")


class Mat(dict):
    """The part of shenfun's SpectralMatrix the kernels read: diagonals by
    offset and a scale."""

    def __init__(self, d, scale=1.0):
        super().__init__({k: np.ascontiguousarray(np.asarray(v, dtype=float)) for k, v in d.items()})
        self.scale = scale


def close(a, b, tol=1e-14):
    return np.max(np.abs(np.asarray(a) - np.asarray(b))) <= tol * max(np.max(np.abs(b)), 1e-300)


def _bands(rng, offs, n):
    d = {o: rng.uniform(0.2, 1.0, n - abs(o)) * (1 if o % 4 else -1) for o in offs}
    d[0] = rng.uniform(4.0, 5.0, n)
    return d


BANDS = [("TDMA", (-2, 0, 2), "tdma_lu", "tdma_solve"),
         ("PDMA", (-4, -2, 0, 2, 4), "pdma_lu", "pdma_solve"),
         ("TDMA_O", (-1, 0, 1), "tdma_o_lu", "tdma_o_solve"),
         ("FDMA", (-2, 0, 2, 4), "fdma_lu", "fdma_solve"),
         ("HeptaDMA", (-4, -2, 0, 2, 4, 6, 8), "heptadma_lu", "heptadma_solve"),
         ("TwoDMA", (0, 2), None, "twodma_solve"),
         ("ThreeDMA", (0, 2, 4), None, "threedma_solve")]


@pytest.mark.parametrize("name,offs,lu,sol", BANDS)
@pytest.mark.parametrize("n", [12, 17])
def test_banded_lu_and_solve(name, offs, lu, sol, n):
    """la.pyx:29-303 (LU) and the *_inner_solve loops along axis 0 and 1."""
    rng = np.random.default_rng(n + len(offs))
    data = S.band_dia(_bands(rng, offs, n), n, offs)
    ref = data.copy()
    ours = data.copy()
    if lu:
        getattr(la, f"{name}_LU")(ref)
        ours = getattr(S, lu)(ours)
        assert close(ours, ref, 0.0), "LU"
    kw = {"compat": True} if name == "ThreeDMA" else {}
    for axis in (0, 1):
        b = rng.standard_normal((n, 4)) + 1j * rng.standard_normal((n, 4))
        b = b if axis == 0 else np.ascontiguousarray(b.T)
        u = b.copy()
        getattr(la, f"{name}_Solve")(u, ref, axis)
        assert close(getattr(S, sol)(b, ours, axis, **kw), u, 1e-15), axis


def test_diagma_solve():
    rng = np.random.default_rng(3)
    d = rng.uniform(1, 2, (1, 9))
    b = rng.standard_normal((9, 3))
    u = b.copy()
    la.DiagMA_Solve(u, d, 0)
    assert close(S.diagma_solve(b, d[0], 0), u, 0.0)


def test_solver_generic_1nd_solve_data():
    """la.pyx:371-426: all modes but the first, per-mode TDMA factors."""
    n, m = 14, 6
    rng = np.random.default_rng(1)
    data = np.stack([S.tdma_lu(S.band_dia(_bands(rng, (-2, 0, 2), n), n, (-2, 0, 2))) for _ in range(m)])
    u0 = rng.standard_normal((n, m)) + 1j * rng.standard_normal((n, m))
    u = u0.copy()
    la.SolverGeneric1ND_solve_data(u, data, la.TDMA_inner_solve, 0, True)
    assert close(u[:, 0], u0[:, 0], 0.0)
    for i in range(1, m):
        assert close(S.tdma_solve(u0[:, i], data[i], 0), u[:, i], 1e-15)


@pytest.mark.parametrize("bc", ["D", "N"])
def test_helmholtz_lu_solve_matvec(bc):
    """LU_Helmholtz / Solve_Helmholtz (la.pyx:708-973) on a 3D array along
    axis 0 with per-mode alfa, beta, and Helmholtz_matvec (Matvec.pyx:656-720).
    The reference's Helmholtz_Neumann_matvec does not run when translated by
    Cython 3 (an array assigned to a typed memoryview slice, Matvec.pyx:788);
    the oracle's Neumann matvec rests on the dense product
    (tests/test_oracle.py)."""
    N = 18
    A = Mat(S.cheb_A_SD(N) if bc == "D" else S.cheb_A_SN(N))
    B = Mat(S.cheb_B_SD(N) if bc == "D" else S.cheb_B_SN(N))
    M = N - 2
    B[2] = np.ascontiguousarray(np.broadcast_to(B[2], (M - 2,)))
    B[-2] = np.ascontiguousarray(np.broadcast_to(B[-2], (M - 2,)))
    rng = np.random.default_rng(4)
    alfa = -np.ones((1, 3, 2))
    beta = rng.uniform(1, 20, (1, 3, 2))
    fac = [np.zeros((N, 3, 2)) for _ in range(4)]
    la.LU_Helmholtz(A, B, alfa, beta, bc == "N", *fac, 0)
    ours = S.helmholtz_lu(bc, N, alfa[0], beta[0])
    for f, o in zip(fac, ours):
        assert close(o[:M], f[:M], 1e-15)
    b = rng.standard_normal((N, 3, 2)) + 1j * rng.standard_normal((N, 3, 2))
    u = np.zeros_like(b)
    la.Solve_Helmholtz(b, u, bc == "N", *fac, 0)
    assert close(S.helmholtz_solve(b, *ours, neumann=(bc == "N"), axis=0)[:M], u[:M], 1e-14)
    if bc == "D":
        c = np.zeros_like(b)
        mv.Helmholtz_matvec(u, c, alfa, beta, A, B, 0)
        assert close(S.helmholtz_matvec(u, alfa[0], beta[0], bc, 0)[:M], c[:M], 1e-12)


def test_poisson_solve_add():
    """Poisson_Solve_ADD (la.pyx:614-706), scale != 1."""
    N = 20
    A = Mat(S.cheb_A_SD(N, nbands=1), scale=-3.0)
    b = np.random.default_rng(5).standard_normal((N, 4))
    u = b.copy()
    la.Poisson_Solve_ADD(A, b, u, 0)
    assert close(S.poisson_add_solve(b, N, -3.0, 0)[:N - 2], u[:N - 2], 1e-15)


@pytest.mark.parametrize("N", [16, 17, 30])
def test_biharmonic_lu_solve(N):
    """LU_Biharmonic + Biharmonic_factor_pr + Biharmonic_Solve (la.pyx:975-1310)
    on one line.  The reference's solve takes the even chain's length for both
    chains, so with N - 4 odd it divides by the odd chain's unset last pivot
    (la.pyx:1273, ZeroDivisionError): there the factors are compared, the
    oracle's solve rests on the dense solves (tests/test_oracle.py)."""
    a0, al, be = 1.0, -18.0, 83.0
    Sm, Am, Bm = S.cheb_SB_matrices(N)
    dg = [Sm[0], Sm[2], Sm[4], Am[-2], Am[0], Am[2], Bm[-4], Bm[-2], Bm[0], Bm[2], Bm[4]]
    dg = [np.ascontiguousarray(d, dtype=float) for d in dg]
    M = dg[0][::2].shape[0]
    fac = [np.zeros((2, M)) for _ in range(7)]
    la.LU_Biharmonic(a0, al, be, *dg, *fac[:5], 0)
    la.Biharmonic_factor_pr(fac[5], fac[6], fac[3], fac[4], 0)
    ours = S.biharmonic_lu(N, a0, al, be)
    for name, f, o in zip(("u0", "u1", "u2", "l0", "l1", "ak", "bk"), fac, ours):
        assert close(o, f, 1e-15), name
    if (N - 4) % 2:
        return
    b = np.random.default_rng(N).standard_normal(N)
    u = np.zeros_like(b)
    la.Biharmonic_Solve(b, u, *fac, a0, 0)
    assert close(S.biharmonic_solve(b, ours, a0, 0)[:N - 4], u[:N - 4], 1e-14)


@pytest.mark.parametrize("N", [8, 33, 64])
def test_recursive_legendre_transforms(N):
    """transforms.pyx:175-267 with the Legendre recursion rows vs the oracle's
    recurrence (same update order)."""
    x, w = S.leg_points_weights(N)
    a = np.ascontiguousarray(np.stack(S._leg_recursion_rows(N + 3)))
    f = np.random.default_rng(N).standard_normal((N, 3)) + 0j
    out = np.zeros_like(f)
    tr.scalar_product(f, out, x, w, 0, a)
    assert close(S.leg_scalar_product(f, axis=0, x=x, w=w), out, 1e-14)
    back = np.zeros_like(f)
    tr.evaluate_expansion_all(out, back, x, 0, a)
    assert close(S.leg_backward(out, axis=0, x=x), back, 1e-14)
