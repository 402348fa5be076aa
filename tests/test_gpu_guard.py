"""GPU parity of the guarded mixed-precision 3D Helmholtz solve: on right-hand
sides whose spectrum decays fast, slowly or not at all, and on adversarial
ones (the solution dominated by a high-ky mode, all the data above the cut),
precision='mixed' stays within 1e-10 of the FP64 oracle, keeps the default
cut kz >= 17 in complex64 on the smooth workload and falls back to FP64 when
no cut qualifies."""
import numpy as np
import pytest
import torch

from oracle import spectral as S

pytestmark = pytest.mark.gpu

SHAPE = (128, 128, 128)


def rel(a, b):
    a = a.cpu().numpy() if isinstance(a, torch.Tensor) else a
    return np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300)


def dev(a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


@pytest.fixture(scope="module")
def sfa():
    import shenfun_amd
    return shenfun_amd


def _from_coeffs(c, Ny, Nz):
    """Real f on the Chebyshev-GC x uniform x uniform grid from coefficients
    c[kx, ky, kz] (C2C ky, R2C kz) of T_kx exp(i ky y) exp(i kz z)."""
    g = np.fft.irfftn(c, s=(Ny, Nz), axes=(1, 2)) * (Ny * Nz)
    return np.ascontiguousarray(S.cheb_backward(g, axis=0))


def rhs(kind, shape, seed=0):
    Nx, Ny, Nz = shape
    rng = np.random.default_rng(seed)
    shp = (Nx, Ny, Nz // 2 + 1)
    kx = np.arange(Nx, dtype=float)[:, None, None]
    ky = S.c2c_wavenumbers(Ny).astype(float)[None, :, None]
    kz = S.r2c_wavenumbers(Nz).astype(float)[None, None, :]
    k2 = kx ** 2 + ky ** 2 + kz ** 2
    if kind == "white":
        return rng.standard_normal(shape)
    c = rng.standard_normal(shp) + 1j * rng.standard_normal(shp)
    if kind.startswith("decay"):
        return _from_coeffs(c * (1 + k2) ** -float(kind[5:]), Ny, Nz)
    if kind == "ky_dominant":
        # u dominated by one high-ky mode (u ~ 1/(1 + 60^2) of f there), plus
        # a small slowly decaying part with content in every kz plane
        c = 1e-4 * c * (1 + k2) ** -1.0
        c[0, 60, 0] += 1.0
        return _from_coeffs(c, Ny, Nz)
    if kind == "high_kz":
        # all the data in kz = 40..60, above every admissible cut but one
        c = c * (1 + k2) ** -1.0
        c[:, :, :40] = 0.0
        c[:, :, 61:] = 0.0
        return _from_coeffs(c, Ny, Nz)
    raise ValueError(kind)


def _space(sfa, shape, bc):
    Nx, Ny, Nz = shape
    return sfa.TensorProductSpace(None, (sfa.FunctionSpace(Nx, "chebyshev", bc=bc),
                                         sfa.FunctionSpace(Ny, "F", dtype="D"),
                                         sfa.FunctionSpace(Nz, "F", dtype="d")))


@pytest.mark.parametrize("bc", ["D", "N"])
@pytest.mark.parametrize("kind", ["decay2", "decay1.5", "decay1", "decay0.5", "white",
                                  "ky_dominant", "high_kz"])
def test_guarded_mixed_any_spectrum(sfa, kind, bc):
    print(f"[guard {kind} {bc} {SHAPE}]", flush=True)
    Nx, Ny, Nz = SHAPE
    Nzh = Nz // 2 + 1
    f = rhs(kind, SHAPE, seed=3)
    u_ref, _, _ = S.helmholtz3d_cheb_fourier(f, 1.0, bc)
    H = sfa.HelmholtzSolver3D(_space(sfa, SHAPE, bc), 1.0, precision="mixed")
    u = H(dev(f))
    err = rel(u, u_ref)
    print(f"  kh={H.last_kh} predicted={H.last_predicted_error:.3e} err={err:.3e}", flush=True)
    assert err <= 1e-10
    # the prediction bounds the error (FP64 round-off floor 1e-12)
    assert err <= max(H.last_predicted_error, 1e-12)
    if kind == "decay2":
        assert H.last_kh == 17  # the smooth workload keeps the default cut
    if kind in ("white", "high_kz"):
        assert H.last_kh == Nzh  # no cut qualifies: FP64
    if H.last_kh < Nzh:
        assert 0.0 < H.last_predicted_error <= 1e-10


def test_guard_fallback_streak_and_retry(sfa):
    """After an FP64 fallback the solver runs FP64 directly for FP64_STREAK
    calls, then tries the guarded mixed form again; every result matches the
    oracle."""
    print("[guard streak]", flush=True)
    shape = (64, 64, 64)
    T = _space(sfa, shape, "D")
    H = sfa.HelmholtzSolver3D(T, 1.0, precision="mixed")
    fw = rhs("white", shape, seed=1)
    fs = rhs("decay2", shape, seed=1)
    uw_ref = S.helmholtz3d_cheb_fourier(fw, 1.0, "D")[0]
    us_ref = S.helmholtz3d_cheb_fourier(fs, 1.0, "D")[0]
    u = H(dev(fw))
    assert H.last_kh == 33 and rel(u, uw_ref) < 1e-12
    for _ in range(H.FP64_STREAK):
        u = H(dev(fs))
        assert H.last_kh == 33 and rel(u, us_ref) < 1e-12
    u = H(dev(fs))
    assert H.last_kh == 17 and rel(u, us_ref) < 1e-10


def test_guarded_solve_many(sfa):
    """One guard decision for a stack: a white-noise member forces FP64 for
    all; a smooth stack keeps the cut and equals the single solves bitwise."""
    print("[guard many]", flush=True)
    shape = (64, 32, 64)
    T = _space(sfa, shape, "N")
    fs = [rhs("decay2", shape, seed=s) for s in range(3)]
    H = sfa.HelmholtzSolver3D(T, 1.5, precision="mixed")
    U = H.solve_many(dev(np.stack(fs)))
    assert H.last_kh == 17
    for b in range(3):
        assert torch.equal(U[b], H(dev(fs[b])))
        assert rel(U[b], S.helmholtz3d_cheb_fourier(fs[b], 1.5, "N")[0]) < 1e-10
    fs[1] = rhs("white", shape, seed=5)
    H2 = sfa.HelmholtzSolver3D(T, 1.5, precision="mixed")
    U = H2.solve_many(dev(np.stack(fs)))
    assert H2.last_kh == 33
    for b in range(3):
        assert rel(U[b], S.helmholtz3d_cheb_fourier(fs[b], 1.5, "N")[0]) < 1e-12


@pytest.mark.parametrize("shape", [(64, 64, 64), (64, 512, 64)])
def test_guard_statistics_follow_each_call(sfa, shape):
    """The predicted error comes from the current call's statistics: calls
    alternating between two right-hand sides reproduce each one's prediction
    from a fresh solver (up to the order of the atomic sums), so no call reads
    the previous call's sums.  Ny = 512 takes the inverse y pass's two-launch
    form, whose statistics go through a kernel of their own."""
    print("[guard sequence]", shape, flush=True)
    T = _space(sfa, shape, "D")
    fa, fb = dev(rhs("decay2", shape, seed=3)), dev(rhs("decay2", shape, seed=4))
    ref = []
    for f in (fa, fb):
        H = sfa.HelmholtzSolver3D(T, 1.0, precision="mixed")
        H(f)
        assert H.last_kh == 17
        ref.append(H.last_predicted_error)
    assert abs(ref[0] - ref[1]) > 1e-6 * ref[0]
    H = sfa.HelmholtzSolver3D(T, 1.0, precision="mixed")
    for i in range(6):
        u = H(fa if i % 2 == 0 else fb)
        assert H.last_kh == 17
        assert abs(H.last_predicted_error - ref[i % 2]) <= 1e-9 * ref[i % 2]
    f = (fa, fb)[1]
    assert rel(u, S.helmholtz3d_cheb_fourier(f.cpu().numpy(), 1.0, "D")[0]) < 1e-10
