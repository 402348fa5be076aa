"""GPU parity of the name-based backend (optimization/hip.py): every exported
Cython name, called with shenfun's arguments (NumPy arrays, as shenfun's own
classes pass them, and device tensors), matches the FP64 oracle; ANNSolver
along any axis; ThreeDMA's compatibility mode."""
import numpy as np
import pytest
import torch

from oracle import spectral as S

pytestmark = pytest.mark.gpu


def rel(a, b):
    a = a.cpu().numpy() if isinstance(a, torch.Tensor) else np.asarray(a)
    return np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300)


@pytest.fixture(scope="module")
def hip():
    from shenfun_amd.optimization import backend
    return backend("hip")


@pytest.fixture(scope="module")
def sfa():
    import shenfun_amd
    return shenfun_amd


def _rand_bands(rng, offs, n):
    d = {o: rng.uniform(0.2, 1.0, n - abs(o)) * (1 if o % 4 else -1) for o in offs}
    d[0] = rng.uniform(4.0, 5.0, n)
    return d


BANDS = [("TDMA", (-2, 0, 2), "tdma_lu", "tdma_solve"),
         ("PDMA", (-4, -2, 0, 2, 4), "pdma_lu", "pdma_solve"),
         ("TDMA_O", (-1, 0, 1), "tdma_o_lu", "tdma_o_solve"),
         ("FDMA", (-2, 0, 2, 4), "fdma_lu", "fdma_solve"),
         ("HeptaDMA", (-4, -2, 0, 2, 4, 6, 8), "heptadma_lu", "heptadma_solve"),
         ("TwoDMA", (0, 2), None, "twodma_solve"),
         ("ThreeDMA", (0, 2, 4), None, "threedma_solve"),
         ("DiagMA", (0,), None, "diagma_solve")]


@pytest.mark.parametrize("name,offs,lu,sol", BANDS)
@pytest.mark.parametrize("arrays", ["numpy", "torch"])
def test_banded_names(hip, name, offs, lu, sol, arrays):
    print(f"[dispatch {name} {arrays}]", flush=True)
    n = 17
    rng = np.random.default_rng(len(offs))
    d = _rand_bands(rng, offs, n)
    data = S.band_dia(d, n, offs) if name != "DiagMA" else d[0][None, :].copy()
    ref_data = getattr(S, lu)(data.copy()) if lu else data.copy()
    b = rng.standard_normal((n, 5)) + 1j * rng.standard_normal((n, 5))
    conv = (lambda a: torch.from_numpy(a).cuda()) if arrays == "torch" else (lambda a: a)
    dd = conv(data.copy())
    if lu:
        getattr(hip, f"{name}_LU")(dd)
        assert rel(dd, ref_data) < 1e-13
    kw = {"compat": True} if name == "ThreeDMA" else {}
    for axis in (0, 1):
        bb = b if axis == 0 else b.T.copy()
        u = conv(bb.copy())
        getattr(hip, f"{name}_Solve")(u, dd, axis)
        if name == "DiagMA":
            ref = S.diagma_solve(bb, ref_data[0], axis)
        else:
            ref = getattr(S, sol)(bb, ref_data, axis, **kw)
        assert rel(u, ref) < 1e-12, axis
    # one line
    line = conv(b[:, 0].copy())
    getattr(hip, f"{name}_inner_solve")(line, dd)
    ref = S.diagma_solve(b[:, 0], ref_data[0], 0) if name == "DiagMA" else \
        getattr(S, sol)(b[:, 0], ref_data, 0, **kw)
    assert rel(line, ref) < 1e-12


@pytest.mark.parametrize("naxes", [0, 1])
def test_solver_generic_1nd_solve_data(hip, naxes):
    """Per-mode TDMA solves of all modes but the first (la.pyx:371-426)."""
    print(f"[dispatch solve_data naxes={naxes}]", flush=True)
    n, m = 12, 7
    rng = np.random.default_rng(naxes)
    data = np.stack([S.tdma_lu(S.band_dia(_rand_bands(rng, (-2, 0, 2), n), n, (-2, 0, 2)))
                     for _ in range(m)])
    shape = (n, m) if naxes == 0 else (m, n)
    u0 = rng.standard_normal(shape) + 1j * rng.standard_normal(shape)
    u = u0.copy()
    hip.SolverGeneric1ND_solve_data(u, data, hip.TDMA_inner_solve, naxes, True)
    for i in range(m):
        line = u0[:, i] if naxes == 0 else u0[i]
        got = u[:, i] if naxes == 0 else u[i]
        ref = line if i == 0 else S.tdma_solve(line, data[i], 0)
        assert rel(got, ref) < 1e-12, i


@pytest.mark.parametrize("bc", ["D", "N"])
@pytest.mark.parametrize("axis", [0, 2])
def test_helmholtz_names(hip, sfa, bc, axis):
    """LU_Helmholtz / Solve_Helmholtz / Helmholtz(_Neumann)_matvec with
    shenfun's matrix objects and per-line scales, along axis 0 or 2 of a 3D
    array (la.pyx:708-973, Matvec.pyx:656-818)."""
    print(f"[dispatch helmholtz {bc} axis={axis}]", flush=True)
    N = 20
    V = sfa.FunctionSpace(N, "chebyshev", bc=(0, 0) if bc == "D" else
                          {"left": {"N": 0}, "right": {"N": 0}})
    A, B = sfa.stiffness_matrix(V), sfa.mass_matrix(V)
    shape = [5, 6, 4]
    shape[axis] = N
    ab_shape = list(shape)
    ab_shape[axis] = 1
    rng = np.random.default_rng(axis)
    alfa = -np.ones(ab_shape)
    beta = rng.uniform(1, 30, ab_shape)
    fac = [np.zeros(shape) for _ in range(4)]
    hip.LU_Helmholtz(A, B, alfa, beta, bc == "N", *fac, axis)
    ref = S.helmholtz_lu(bc, N, np.moveaxis(alfa, axis, 0)[0], np.moveaxis(beta, axis, 0)[0])
    for f, r in zip(fac, ref):
        assert np.allclose(np.moveaxis(f, axis, 0)[:N - 2], r[:N - 2], rtol=1e-13, atol=1e-13)
    b = rng.standard_normal(shape) + 1j * rng.standard_normal(shape)
    u = np.zeros_like(b)
    hip.Solve_Helmholtz(b, u, bc == "N", *fac, axis)
    uref = S.helmholtz_solve(b, *ref, neumann=(bc == "N"), axis=axis)
    sl = [slice(None)] * 3
    sl[axis] = slice(0, N - 2)
    assert rel(u[tuple(sl)], uref[tuple(sl)]) < 1e-12
    c = np.zeros_like(b)
    mv = hip.Helmholtz_Neumann_matvec if bc == "N" else hip.Helmholtz_matvec
    mv(u, c, alfa, beta, A, B, axis)
    cref = S.helmholtz_matvec(u, np.moveaxis(alfa, axis, 0)[0], np.moveaxis(beta, axis, 0)[0], bc, axis)
    assert rel(c[tuple(sl)], cref[tuple(sl)]) < 1e-12
    # matvec(solve(b)) == b on the unknowns (test_matrices.py:577-590 pattern)
    assert rel(c[tuple(sl)], b[tuple(sl)]) < 1e-11


def test_poisson_solve_add(hip, sfa):
    print("[dispatch Poisson_Solve_ADD]", flush=True)
    N = 24
    V = sfa.FunctionSpace(N, "chebyshev", bc=(0, 0))
    A = sfa.stiffness_matrix(V)
    b = np.random.default_rng(3).standard_normal((N, 3))
    u = b.copy()
    hip.Poisson_Solve_ADD(A, u, u, axis=0)
    ref = S.poisson_add_solve(b, N, float(A.scale), axis=0)
    assert rel(u[:N - 2], ref[:N - 2]) < 1e-12


@pytest.mark.parametrize("N", [8, 33, 128])
def test_recursive_transform_seam(hip, N):
    """scalar_product(input, output, x, w, axis, a) and
    evaluate_expansion_all(input, output, x, axis, a) with the Legendre
    recursion rows (N + 3 entries, 2 rows: anm, anp) as spectralbase.py:
    484-491, 552-558 pass them, along axis 1 of complex data."""
    print(f"[dispatch recursive N={N}]", flush=True)
    x, w = S.leg_points_weights(N)
    a = np.stack(S._leg_recursion_rows(N + 3))
    rng = np.random.default_rng(N)
    f = rng.standard_normal((3, N, 2)) + 1j * rng.standard_normal((3, N, 2))
    out = np.zeros_like(f)
    hip.scalar_product(f, out, x, w, 1, a)
    assert rel(out, S.leg_scalar_product(f, axis=1, x=x, w=w)) < 1e-12
    back = torch.zeros(f.shape, dtype=torch.complex128, device="cuda")
    hip.evaluate_expansion_all(torch.from_numpy(out).cuda(), back, x, 1, a)
    assert rel(back, S.leg_backward(out, axis=1, x=x)) < 1e-12


@pytest.mark.parametrize("axis", [0, 1, 2])
def test_ann_solver_any_axis(sfa, axis):
    """ANNSolver (chebyshev/la.py:74-128) on a 3D array along any axis vs the
    oracle's restatement."""
    print(f"[ANNSolver axis={axis}]", flush=True)
    N = 24
    V = sfa.FunctionSpace(N, "chebyshev", bc={"left": {"N": 0}, "right": {"N": 0}})
    A = sfa.stiffness_matrix(V)
    shape = [4, 5, 3]
    shape[axis] = N
    rng = np.random.default_rng(axis)
    b = rng.standard_normal(shape) + 1j * rng.standard_normal(shape)
    sol = sfa.la.ANNSolver(A)
    u = sol(torch.from_numpy(b.copy()).cuda(), axis=axis)
    ref = S.ann_solve(b, N, float(A.scale), axis=axis)
    assert rel(u, ref) < 1e-12


def test_threedma_compat_class(sfa):
    """la.ThreeDMA(compat=True) reproduces the reference's reading of the +4
    band (la.py:688); the default solves the matrix."""
    print("[ThreeDMA compat]", flush=True)
    n = 14
    rng = np.random.default_rng(5)
    d = {0: rng.uniform(4, 5, n), 2: rng.uniform(0.2, 1, n - 2), 4: rng.uniform(-1, -0.2, n - 4)}
    M = sfa.SparseMatrix(d, (n, n))
    b = rng.standard_normal((n, 3))
    data = S.band_dia(d, n, (0, 2, 4))
    u = sfa.la.ThreeDMA(M, compat=True)(torch.from_numpy(b.copy()).cuda())
    assert rel(u, S.threedma_solve(b, data, 0, compat=True)) < 1e-13
    u = sfa.la.ThreeDMA(M)(torch.from_numpy(b.copy()).cuda())
    assert rel(u, np.linalg.solve(S.dense(d, n), b)) < 1e-12
