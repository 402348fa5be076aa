"""GPU parity: banded solvers and the Chebyshev Helmholtz solver vs the FP64 oracle."""
import numpy as np
import pytest
import torch

from oracle import spectral as S

pytestmark = pytest.mark.gpu
TOL = 1e-12


def rel(a, b):
    a = a.cpu().numpy() if isinstance(a, torch.Tensor) else a
    return np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300)


def dev(a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


@pytest.fixture(scope="module")
def sfa():
    import shenfun_amd
    return shenfun_amd


# tests/test_la.py:8-15 of the reference (N = 10 matrices: DiagMA, TwoDMA,
# TDMA_O, TDMA, FDMA, PDMA), plus ThreeDMA and HeptaDMA systems
N10 = 10
MATS = [
    {0: np.arange(N10) + 1.0},
    {0: -2.0, 2: 1.0},
    {-1: 1.0, 0: -2.0, 1: 1.0},
    {-2: 1.0, 0: -2.0, 2: 1.0},
    {-2: 1.0, 0: -2.0, 2: 1.0, 4: 0.1},
    {-4: 0.1, -2: 1.0, 0: -2.0, 2: 1.0, 4: 0.1},
    {0: -2.0, 2: 1.0, 4: 0.3},
    {-4: 0.1, -2: 0.5, 0: -4.0, 2: 1.0, 4: 0.3, 6: -0.2, 8: 0.1},
]

# pattern -> (offsets, oracle LU, oracle solve, C-ABI LU, C-ABI solve)
BANDS = {
    "twodma": ((0, 2), None, "twodma_solve", None, "sfa_twodma_solve"),
    "tdma_o": ((-1, 0, 1), "tdma_o_lu", "tdma_o_solve", "sfa_tdma_o_lu", "sfa_tdma_o_solve"),
    "threedma": ((0, 2, 4), None, "threedma_solve", None, "sfa_threedma_solve"),
    "fdma": ((-2, 0, 2, 4), "fdma_lu", "fdma_solve", "sfa_fdma_lu", "sfa_fdma_solve"),
    "heptadma": ((-4, -2, 0, 2, 4, 6, 8), "heptadma_lu", "heptadma_solve", "sfa_heptadma_lu",
                 "sfa_heptadma_solve"),
}


def _rand_bands(rng, offs, n):
    d = {o: rng.uniform(0.2, 1.0, n - abs(o)) * (1 if o % 4 else -1) for o in offs}
    d[0] = rng.uniform(4.0, 5.0, n)
    return d


@pytest.mark.parametrize("kind", sorted(BANDS))
@pytest.mark.parametrize("n", [8, 9, 64, 255])
def test_band_patterns_per_line_factors(sfa, kind, n):
    """Per-line factors (data_line_stride) and skip_lines through the C ABI, LU
    on the device for all lines at once, against the oracle per line; the
    space axis is strided (axis 1 of (lines, n, 3))."""
    from shenfun_amd import _lib
    print(f"[band {kind} n={n}]", flush=True)
    offs, olu, osol, clu, csol = BANDS[kind]
    rng = np.random.default_rng(n + len(offs))
    nl = 5
    mats = [_rand_bands(rng, offs, n) for _ in range(nl * 3)]
    data = np.stack([S.band_dia(d, n, offs) for d in mats])  # (lines, rows, n)
    dd = dev(data)
    if clu is not None:
        _lib.call(clu, _lib.ptr(dd), n, nl * 3, _lib.stream_ptr())
    b = rng.standard_normal((nl, n, 3)) + 1j * rng.standard_normal((nl, n, 3))
    u = dev(b)
    _lib.call(csol, _lib.ptr(u), _lib.C128, 3, _lib.arr(u.shape), _lib.arr(u.stride()), 1,
              _lib.ptr(dd), n, data.shape[1] * n, 1, _lib.stream_ptr())
    got = u.cpu().numpy()
    assert np.array_equal(got[0, :, 0], b[0, :, 0])  # line 0 skipped
    for i in range(nl):
        for j in range(3):
            L = i * 3 + j
            if L == 0:
                continue
            ld = getattr(S, olu)(data[L]) if olu else data[L]
            if olu:
                assert rel(dd[L], ld) < 1e-13
            ref = getattr(S, osol)(b[i, :, j], ld, 0)
            assert rel(got[i, :, j], ref) < TOL
            assert rel(got[i, :, j], np.linalg.solve(S.dense(mats[L], n), b[i, :, j])) < 1e-11


@pytest.mark.parametrize("kind", sorted(BANDS))
def test_band_patterns_constraint_row0(sfa, kind):
    """apply_constraints((0, val)): row 0 becomes the unit row and b[0] = val
    (la.py:575-586, 628-637, 674-684, 724-741)."""
    offs = BANDS[kind][0]
    n = 14
    rng = np.random.default_rng(7)
    d = _rand_bands(rng, offs, n)
    if kind == "tdma_o":
        d[1] = d[-1]  # get_solver takes TDMA_O for symmetric (-1, 0, 1) only
    M = sfa.SparseMatrix(d, (n, n))
    sol = sfa.la.Solver(M)
    assert type(sol).__name__.lower() == kind
    b = rng.standard_normal((n, 4))
    u = sol(dev(b), axis=0, constraints=((0, 0.25),))
    D = S.dense(d, n)
    D[0, :] = 0.0
    D[0, 0] = 1.0
    bb = b.copy()
    bb[0] = 0.25
    assert rel(u, np.linalg.solve(D, bb)) < 1e-11


@pytest.mark.parametrize("di", MATS)
@pytest.mark.parametrize("cplx", [False, True])
def test_xdma(sfa, di, cplx):
    print(f"[xdma {sorted(di)} cplx={cplx}]", flush=True)
    M = sfa.SparseMatrix(di, (N10, N10))
    sol = sfa.la.Solver(M)
    dense = S.dense(di, N10)
    rng = np.random.default_rng(0)
    b = rng.standard_normal(N10) + (1j * rng.standard_normal(N10) if cplx else 0)
    u = sol(dev(b))
    assert rel(u, np.linalg.solve(dense, b)) < TOL
    bh = rng.standard_normal((N10, N10, 3)) + (1j * rng.standard_normal((N10, N10, 3)) if cplx else 0)
    for axis in (0, 1):
        uh = sol(dev(bh), axis=axis)
        ref = np.moveaxis(np.linalg.solve(dense, np.moveaxis(bh, axis, 0).reshape(N10, -1))
                          .reshape(np.moveaxis(bh, axis, 0).shape), 0, axis)
        assert rel(uh, ref) < TOL


@pytest.mark.parametrize("n", [6, 7, 64, 255])
def test_tdma_pdma_vs_oracle(sfa, n):
    rng = np.random.default_rng(n)
    d3 = {-2: rng.uniform(0.1, 1, n - 2), 0: rng.uniform(3, 4, n), 2: rng.uniform(0.1, 1, n - 2)}
    d5 = {-4: rng.uniform(0.1, .3, n - 4), -2: rng.uniform(0.1, 1, n - 2), 0: rng.uniform(4, 5, n),
          2: rng.uniform(0.1, 1, n - 2), 4: rng.uniform(0.1, .3, n - 4)}
    b = rng.standard_normal((n, 9)) + 1j * rng.standard_normal((n, 9))
    u3 = sfa.la.TDMA(sfa.SparseMatrix(d3, (n, n)))(dev(b), axis=0)
    assert rel(u3, S.tdma_solve(b, S.tdma_lu(S.tdma_dia(d3, n)), 0)) < TOL
    u5 = sfa.la.PDMA(sfa.SparseMatrix(d5, (n, n)))(dev(b), axis=0)
    assert rel(u5, S.pdma_solve(b, S.pdma_lu(S.pdma_dia(d5, n)), 0)) < TOL


@pytest.mark.parametrize("bc", ["D", "N"])
@pytest.mark.parametrize("shape,axis", [((20,), 0), ((32, 7, 5), 0), ((6, 40, 5), 1), ((4, 5, 66), 2),
                                        ((256, 16, 9), 0), ((1024,), 0), ((4096, 3), 0), ((2, 2050), 1)])
def test_helmholtz(sfa, bc, shape, axis):
    """Chebyshev Helmholtz LU/solve/matvec vs the oracle; N >= 514 with few
    lines takes the LDS-staged long-chain solve."""
    print(f"[helmholtz {bc} {shape} axis={axis}]", flush=True)
    N = shape[axis]
    rng = np.random.default_rng(N)
    mshape = list(shape)
    mshape[axis] = 1
    alfa = -np.ones(mshape) * rng.uniform(0.5, 2)
    beta = rng.uniform(0.5, 50, mshape)
    if bc == "N":
        beta.flat[0] = 0.0  # the k = 0 Neumann gauge
    V = sfa.FunctionSpace(N, "chebyshev", bc=bc)
    V.axis = axis

    class _T:  # minimal tensor-product stand-in carrying the shape
        def __init__(self, s):
            self._s = s

        def shape(self, fo=True):
            return tuple(self._s)

    A = sfa.stiffness_matrix(V)
    B = sfa.mass_matrix(V)
    if len(shape) > 1:
        V.tensorproductspace = _T(shape)
    H = sfa.la.Helmholtz(A, B, alfa, beta)
    b = rng.standard_normal(shape) + 1j * rng.standard_normal(shape)
    u = H(dev(b), torch.zeros(shape, dtype=torch.complex128, device="cuda"))
    d0, d1, d2, L = S.helmholtz_lu(bc, N, np.moveaxis(alfa, axis, 0)[0], np.moveaxis(beta, axis, 0)[0])
    uref = S.helmholtz_solve(np.moveaxis(b, axis, 0), d0, d1, d2, L, neumann=(bc == "N"))
    uref = np.moveaxis(uref, 0, axis)
    assert rel(u, uref) < TOL
    # factors match the reference LU (la.pyx:717-763)
    for got, ref in zip((H.u0, H.u1, H.u2, H.L), (d0, d1, d2, L)):
        assert rel(torch.movedim(got, axis, 0), ref) < TOL
    # matvec of the solution reproduces b on the unknowns (except the gauge row)
    c = H.matvec(u, torch.zeros_like(u))
    cref = S.helmholtz_matvec(uref, np.moveaxis(alfa, axis, 0)[0] if len(shape) > 1 else alfa,
                              np.moveaxis(beta, axis, 0)[0] if len(shape) > 1 else beta, bc, axis)
    assert rel(c, cref) < TOL


@pytest.mark.parametrize("N", [8, 64, 255])
def test_poisson_add(sfa, N):
    V = sfa.FunctionSpace(N, "chebyshev", bc="D")
    A = sfa.stiffness_matrix(V)
    rng = np.random.default_rng(5)
    b = rng.standard_normal((N, 4)) + 1j * rng.standard_normal((N, 4))
    u = sfa.la.ADDSolver(A)(dev(b), torch.zeros(N, 4, dtype=torch.complex128, device="cuda"))
    ref = S.poisson_add_solve(b, N)
    assert rel(u[:N - 2], ref[:N - 2]) < TOL
    dense = S.dense(S.cheb_A_SD(N), N - 2)
    assert rel(u[:N - 2], np.linalg.solve(dense, b[:N - 2])) < 1e-11


@pytest.mark.parametrize("di", MATS)
@pytest.mark.parametrize("axis", [0, 1])
def test_sparse_matvec(sfa, di, axis):
    """SparseMatrix.matvec (matrixbase.py:87-144) along an axis vs numpy."""
    print(f"[matvec {sorted(di)} axis={axis}]", flush=True)
    M = sfa.SparseMatrix(di, (N10, N10), scale=1.5)
    rng = np.random.default_rng(3)
    shp = [12, 4, 3]
    shp[axis] = N10 + 2  # the vector is longer than the matrix: rows >= N come out zero
    v = rng.standard_normal(shp) + 1j * rng.standard_normal(shp)
    c = torch.full(shp, 7.0 + 1j, dtype=torch.complex128, device="cuda")
    M.matvec(dev(v), c, axis=axis)
    vm = np.moveaxis(v, axis, 0)
    ref = np.zeros_like(vm)
    ref[:N10] = np.tensordot(1.5 * S.dense(di, N10), vm[:N10], axes=(1, 0))
    assert rel(c, np.moveaxis(ref, 0, axis)) < 1e-14


def test_matvec_dense_tail_and_tpmatrix(sfa):
    """The Chebyshev stiffness matrix (N/2 diagonals) and TPMatrix.matvec with
    its per-mode scale (matrixbase.py:1437-1449)."""
    SD = sfa.FunctionSpace(40, "C", bc="D")
    K = sfa.FunctionSpace(8, "F", dtype="d")
    T = sfa.TensorProductSpace(None, (SD, K))
    A, B = sfa.helmholtz_tpmatrices(T, 2.0)
    rng = np.random.default_rng(4)
    shp = T.shape(True)
    u = rng.standard_normal(shp) + 1j * rng.standard_normal(shp)
    u[38:] = 0
    c = torch.zeros(shp, dtype=torch.complex128, device="cuda")
    B.matvec(dev(u), c)
    A.matvec(dev(u), c, accumulate=True)
    ref = np.zeros_like(u)
    for m in range(shp[1]):
        Mm = A.scale.reshape(-1)[0] * A.pmat.todense() * float(A.pmat.scale) + \
            B.scale.reshape(-1)[m] * B.pmat.todense() * float(B.pmat.scale)
        ref[:38, m] = Mm @ u[:38, m]
    assert rel(c, ref) < 1e-13


@pytest.mark.parametrize("case", ["legendre", "pg", "bc"])
def test_solver_generic_matvec_roundtrip(sfa, case):
    """SolverGeneric1ND.matvec then the solve returns u (the reference's
    G.matvec check, tests/test_matrices.py:594-621)."""
    print(f"[generic matvec {case}]", flush=True)
    bc = (0.5, -1.0) if case == "bc" else (0, 0)
    fam = "L" if case != "pg" else "C"
    SD = sfa.FunctionSpace(24, fam, bc=bc)
    T = sfa.TensorProductSpace(None, (SD, sfa.FunctionSpace(12, "F", dtype="d")))
    test = T.get_testspace("PG") if case == "pg" else None
    G = sfa.la.SolverGeneric1ND(sfa.helmholtz_tpmatrices(T, 1.5, test=test))
    rng = np.random.default_rng(5)
    shp = T.shape(True)
    u = rng.standard_normal(shp) + 1j * rng.standard_normal(shp)
    u[22:] = 0
    ud = dev(u)
    if case == "bc":
        SD.set_boundary_dofs(ud, final=True)
    c = torch.zeros(shp, dtype=torch.complex128, device="cuda")
    G.matvec(ud, c)
    uc = G(c, sfa.Function(T))
    assert rel(uc, ud.cpu().numpy()) < 1e-11
