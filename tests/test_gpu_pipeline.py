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
                                   (256, 12, 10), (256, 32, 64), (512, 16, 32)])
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
