"""GPU parity: end-to-end tensor-product pipelines vs the FP64 oracle."""
import numpy as np
import pytest
import torch

from oracle import spectral as S

pytestmark = pytest.mark.gpu


def rel(a, b):
    a = a.cpu().numpy() if isinstance(a, torch.Tensor) else a
    return np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300)


def dev(a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


@pytest.fixture(scope="module")
def sfa():
    import shenfun_amd
    return shenfun_amd


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_poisson2d_legendre_manufactured(sfa, seed):
    """2D Poisson -lap u = f, Legendre SD(64) x R2C(64): scalar product ->
    SolverGeneric1ND -> backward reproduces the exact discrete solution."""
    print(f"[poisson2d seed={seed}]", flush=True)
    x, y, f, u_exact, uh_exact = S.poisson2d_manufactured(64, 64, seed)
    SD = sfa.FunctionSpace(64, "legendre", bc=(0, 0))
    K = sfa.FunctionSpace(64, "F", dtype="d")
    T = sfa.TensorProductSpace(None, (SD, K))
    fj = dev(f)
    f_hat = T.scalar_product(fj)
    _, fh_ref, uh_ref = S.legendre_fourier_2d(f, 0.0)
    assert rel(f_hat, fh_ref) < 1e-12
    H = sfa.la.SolverGeneric1ND(sfa.helmholtz_tpmatrices(T, 0.0))
    u_hat = H(f_hat, sfa.Function(T))
    assert rel(u_hat, uh_exact) < 1e-12
    uq = T.backward(u_hat)
    assert rel(uq, u_exact) < 1e-12


@pytest.mark.parametrize("bc", ["D", "N"])
@pytest.mark.parametrize("shape", [(16, 12, 10), (32, 24, 16), (64, 64, 64), (128, 24, 20),
                                   (256, 12, 10), (256, 32, 64), (512, 16, 32), (512, 128, 64)])
def test_helmholtz3d_dropin_and_fused(sfa, bc, shape):
    """(alpha - lap) u = f on Chebyshev x C2C x R2C: the drop-in chain
    T.scalar_product -> Helmholtz -> T.backward and the fused pipeline agree
    with the oracle."""
    print(f"[helm3d {bc} {shape}]", flush=True)
    Nx, Ny, Nz = shape
    alpha = 1.0
    f = S.helmholtz3d_rhs(Nx, Ny, Nz, seed=3)
    u_ref, fh_ref, uh_ref = S.helmholtz3d_cheb_fourier(f, alpha, bc)
    SD = sfa.FunctionSpace(Nx, "chebyshev", bc=bc)
    K1 = sfa.FunctionSpace(Ny, "F", dtype="D")
    K2 = sfa.FunctionSpace(Nz, "F", dtype="d")
    T = sfa.TensorProductSpace(None, (SD, K1, K2))
    fj = dev(f)
    f_hat = T.scalar_product(fj)
    assert rel(f_hat, fh_ref) < 1e-12
    H = sfa.la.Helmholtz(*sfa.helmholtz_tpmatrices(T, alpha))
    u_hat = H(f_hat, sfa.Function(T))
    assert rel(u_hat, uh_ref) < 1e-12
    u = T.backward(u_hat)
    assert rel(u, u_ref) < 1e-12
    fused = sfa.HelmholtzSolver3D(T, alpha)
    u2 = fused(fj)
    assert rel(u2, u_ref) < 1e-12


def test_helmholtz3d_fused_repeat_identical(sfa):
    """Two calls of the fused pipeline on the same input give identical bits."""
    SD = sfa.FunctionSpace(32, "chebyshev", bc="D")
    T = sfa.TensorProductSpace(None, (SD, sfa.FunctionSpace(16, "F", dtype="D"),
                                      sfa.FunctionSpace(16, "F", dtype="d")))
    f = dev(S.helmholtz3d_rhs(32, 16, 16, seed=9))
    H = sfa.HelmholtzSolver3D(T, 2.5)
    a = H(f).clone()
    b = H(f)
    assert torch.equal(a, b)


@pytest.mark.parametrize("precision,shape,bc", [("fp64", (32, 24, 16), "D"), ("fp64", (64, 20, 16), "N"),
                                               ("fp64", (64, 16, 128), "D"), ("mixed", (64, 32, 64), "D"),
                                               ("mixed", (128, 16, 64), "N")])
def test_helmholtz3d_many_equals_single(sfa, precision, shape, bc):
    """sfa_helmholtz3d_many on a stack of 3 right-hand sides: each result has
    the bits of its single solve (generic and fused x-pass, padded kz pitch,
    line groups straddling two arrays), and matches the oracle."""
    print(f"[helm3d many {precision} {bc} {shape}]", flush=True)
    Nx, Ny, Nz = shape
    fs = [S.helmholtz3d_rhs(Nx, Ny, Nz, seed=20 + b) for b in range(3)]
    T = sfa.TensorProductSpace(None, (sfa.FunctionSpace(Nx, "chebyshev", bc=bc),
                                      sfa.FunctionSpace(Ny, "F", dtype="D"),
                                      sfa.FunctionSpace(Nz, "F", dtype="d")))
    H = sfa.HelmholtzSolver3D(T, 1.5, precision=precision)
    F = dev(np.stack(fs))
    U = H.solve_many(F)
    for b in range(3):
        assert torch.equal(U[b], H(F[b])), f"array {b}"
    u_ref, _, _ = S.helmholtz3d_cheb_fourier(fs[1], 1.5, bc)
    assert rel(U[1], u_ref) < (1e-12 if precision == "fp64" else 1e-10)


@pytest.mark.parametrize("bc", ["D", "N"])
@pytest.mark.parametrize("shape,planes", [((64, 64, 64), 0), ((128, 32, 64), 0), ((256, 64, 64), 0),
                                          ((64, 16, 32), 12), ((256, 256, 256), 0),
                                          ((512, 32, 64), 0)])
def test_helmholtz3d_mixed(sfa, bc, shape, planes):
    """precision='mixed' (FP64 low kz planes, FP32 transforms above, FP64
    solve) stays within 1e-10 relative max error of the FP64 oracle on the
    smooth workload right-hand side."""
    print(f"[helm3d mixed {bc} {shape} planes={planes}]", flush=True)
    Nx, Ny, Nz = shape
    f = S.helmholtz3d_rhs(Nx, Ny, Nz, seed=5)
    u_ref, _, _ = S.helmholtz3d_cheb_fourier(f, 1.0, bc)
    T = sfa.TensorProductSpace(None, (sfa.FunctionSpace(Nx, "chebyshev", bc=bc),
                                      sfa.FunctionSpace(Ny, "F", dtype="D"),
                                      sfa.FunctionSpace(Nz, "F", dtype="d")))
    H = sfa.HelmholtzSolver3D(T, 1.0, precision="mixed", fp64_planes=planes)
    u = H(dev(f))
    err = rel(u, u_ref)
    print(f"  max rel err {err:.3e}", flush=True)
    assert err < 1e-10
    # all-FP64 solve of the same input agrees with the oracle to round-off
    u64 = sfa.HelmholtzSolver3D(T, 1.0)(dev(f))
    assert rel(u64, u_ref) < 1e-12


ALLAXES3D = {0: (0, 1, 2), 1: (1, 0, 2), 2: (2, 0, 1)}


@pytest.mark.parametrize("axis", [0, 1, 2])
def test_helmholtz3d_any_axis(sfa, axis):
    """tests/test_matrices.py:559-593 of the reference: the Chebyshev axis at
    position 0, 1 or 2 (transform order ``axes``, the R2C axis transformed
    first); H(H.matvec(uh)) == uh, and backward -> forward round trips
    against the oracle chain."""
    print(f"[helmholtz3d axis={axis}]", flush=True)
    a = ALLAXES3D[axis]
    N = (8, 9, 10)
    bases = [None] * 3
    bases[a[0]] = sfa.FunctionSpace(N[a[0]], "C", bc=(0, 0))
    bases[a[1]] = sfa.FunctionSpace(N[a[1]], "F", dtype="D")
    bases[a[2]] = sfa.FunctionSpace(N[a[2]], "F", dtype="d")
    T = sfa.TensorProductSpace(None, bases, axes=a)
    H = sfa.la.Helmholtz(*sfa.helmholtz_tpmatrices(T, 0.0))
    rng = np.random.default_rng(axis)
    shp = T.shape(True)
    # coefficients of real data (Hermitian across the Fourier axes): the
    # forward transform of a random real field
    uh = T.forward(dev(rng.standard_normal(T.shape(False)))).cpu().numpy()
    f = H.matvec(dev(uh), torch.empty(shp, dtype=torch.complex128, device="cuda"))
    uc = H(f, torch.zeros(shp, dtype=torch.complex128, device="cuda"))
    assert rel(uc, uh) < 1e-12
    kinds = [None] * 3
    kinds[a[0]] = ("chebyshev", "D")
    kinds[a[1]] = ("fourier-c2c", None)
    kinds[a[2]] = ("fourier-r2c", None)
    u = T.backward(dev(uh))
    ref = uh
    for ax in a:  # backward order
        ref = S.basis_backward(kinds[ax][0], kinds[ax][1], ref, ax, N=N[ax])
    assert u.dtype == torch.float64 and rel(u, ref) < 1e-12
    assert rel(T.forward(u), uh) < 1e-11


@pytest.mark.parametrize("axis", [0, 1])
@pytest.mark.parametrize("fam", ["C", "L"])
def test_helmholtz2d_any_axis(sfa, axis, fam):
    """tests/test_matrices.py:594-621: 2D with the non-periodic axis first or
    second (R2C transformed first); f = (sA A + sB B) uh per mode from the
    dense matrices on the host, then the solve (la.Helmholtz for Chebyshev,
    SolverGeneric1ND for Legendre) returns uh."""
    print(f"[helmholtz2d {fam} axis={axis}]", flush=True)
    N = (8, 10)
    a = (0, 1) if axis == 0 else (1, 0)
    bases = [None] * 2
    bases[axis] = sfa.FunctionSpace(N[axis], fam, bc=(0, 0))
    bases[1 - axis] = sfa.FunctionSpace(N[1 - axis], "F", dtype="d")
    T = sfa.TensorProductSpace(None, bases, axes=a)
    mats = sfa.helmholtz_tpmatrices(T, 0.5)
    rng = np.random.default_rng(7)
    shp = T.shape(True)
    n = N[axis] - 2
    uh = np.zeros(shp, dtype=complex)
    um = np.moveaxis(uh, axis, 0)
    um[:n] = rng.random((n, shp[1 - axis])) + 1j * rng.random((n, shp[1 - axis]))
    fm = np.zeros_like(um)
    for m in range(shp[1 - axis]):
        M = sum(np.moveaxis(t.scale, axis, 0).reshape(-1)[m] * float(t.pmat.scale) * t.pmat.todense()
                for t in mats)
        fm[:n, m] = M @ um[:n, m]
    f = np.moveaxis(fm, 0, axis)
    sol = sfa.la.Helmholtz(*mats) if fam == "C" else sfa.la.SolverGeneric1ND(mats)
    uc = sol(dev(f), torch.zeros(shp, dtype=torch.complex128, device="cuda"))
    assert rel(uc, uh) < 1e-12
