"""GPU parity: padded (3/2-rule) and dealias_direct transforms vs the oracle."""
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


# (oracle family, shenfun family, bc, dtype)
SPACES = [
    ("chebyshev", "C", None, "d"), ("chebyshev", "C", "D", "d"), ("chebyshev", "C", "N", "d"),
    ("legendre", "L", None, "d"), ("legendre", "L", "D", "d"),
    ("fourier-r2c", "F", None, "d"), ("fourier-c2c", "F", None, "D"),
]


def _coeffs(fam, N, shape_other, rng):
    if fam == "fourier-r2c":
        c = rng.standard_normal((N // 2 + 1,) + shape_other) + 1j * rng.standard_normal(
            (N // 2 + 1,) + shape_other)
        c[0] = c[0].real
        c[N // 2] = c[N // 2].real
        return c
    c = rng.standard_normal((N,) + shape_other) + 1j * rng.standard_normal((N,) + shape_other)
    return c


@pytest.mark.parametrize("ofam,fam,bc,dt", SPACES)
@pytest.mark.parametrize("N,axis", [(16, 0), (24, 1), (20, 2)])
def test_padded_1d_transforms(sfa, ofam, fam, bc, dt, N, axis):
    """backward / scalar_product / forward of a padded space along every axis
    of a 3D array vs the oracle (pad / truncate restatements + M-point
    transforms)."""
    print(f"[padded {ofam} {bc} N={N} axis={axis}]", flush=True)
    if ofam == "fourier-r2c" and axis != 2:
        pytest.skip("R2C is the last axis")
    V = sfa.FunctionSpace(N, fam, bc=bc, dtype=dt)
    P = V.get_dealiased(1.5)
    assert P.is_padded and P.shape(False) == 3 * N // 2 and P.shape(True) == V.shape(True)
    P.axis = axis
    P.dimensions = 3
    M = 3 * N // 2
    rng = np.random.default_rng(N + axis)
    other = (3, 5)
    c = np.moveaxis(_coeffs(ofam, N, other, rng), 0, axis)
    if ofam in ("chebyshev", "legendre"):
        c = c.real.copy()
        if bc is not None:
            sl = [slice(None)] * 3
            sl[axis] = slice(N - 2, N)
            c[tuple(sl)] = 0
    u = P.backward(dev(c))
    ref_u = S.padded_backward(ofam, bc, c, N, M, axis)
    assert u.shape[axis] == M and rel(u, ref_u) < TOL
    s = P.scalar_product(dev(ref_u))
    ref_s = S.padded_scalar_product(ofam, bc, ref_u, N, axis)
    assert rel(s, ref_s) < TOL
    # forward of the padded backward recovers the coefficients
    f = P.forward(dev(ref_u))
    assert rel(f, c) < 1e-11


@pytest.mark.parametrize("ofam,fam,bc,dt", SPACES)
@pytest.mark.parametrize("N", [12, 17])
def test_dealias_direct_backward(sfa, ofam, fam, bc, dt, N):
    print(f"[dealias_direct {ofam} {bc} N={N}]", flush=True)
    V = sfa.FunctionSpace(N, fam, bc=bc, dtype=dt).get_dealiased(1, dealias_direct=True)
    rng = np.random.default_rng(N)
    c = _coeffs(ofam, N, (4,), rng)
    if ofam in ("chebyshev", "legendre"):
        c = c.real.copy()
        if bc is not None:
            c[N - 2:] = 0
    cd = dev(c)
    u = V.backward(cd)
    ref = S.basis_backward(ofam, bc, S.dealias_direct_zero(c, ofam, N), 0, N=N)
    assert rel(u, ref) < TOL
    assert np.array_equal(cd.cpu().numpy(), c)  # the caller's coefficients are left as they were


def test_tensor_product_dealiased_product(sfa):
    """TensorProductSpace.get_dealiased: backward on the padded grid, a
    pointwise product, forward back: matches the oracle chain, and for the
    Fourier axes the 3/2 rule's exact truncated convolution property holds."""
    print("[tps dealiased]", flush=True)
    N = (12, 16, 20)
    SD = sfa.FunctionSpace(N[0], "C", bc="D")
    K1 = sfa.FunctionSpace(N[1], "F", dtype="D")
    K2 = sfa.FunctionSpace(N[2], "F", dtype="d")
    T = sfa.TensorProductSpace(None, (SD, K1, K2))
    Tp = T.get_dealiased(1.5)
    assert Tp is T.get_dealiased((1.5, 1.5, 1.5))
    assert Tp.shape(False) == (18, 24, 30) and Tp.shape(True) == T.shape(True)
    rng = np.random.default_rng(5)
    f = rng.standard_normal((N[0], N[1], N[2]))
    bases = [("chebyshev", "D"), ("fourier-c2c", None), ("fourier-r2c", None)]
    a_hat = S.tps_scalar_product(f, bases)
    a_hat[N[0] - 2:] = 0
    a_hat[:, N[1] // 2] = 0
    a_hat[:, :, N[2] // 2] = 0
    u = Tp.backward(dev(a_hat))
    ref = a_hat
    Ms = Tp.shape(False)
    for ax, (fam, bc) in enumerate(bases):
        ref = S.padded_backward(fam, bc, ref, N[ax], Ms[ax], ax)
    assert rel(u, ref) < TOL
    w = Tp.convolve(dev(a_hat), dev(a_hat))
    wr = ref * ref
    for ax in range(2, -1, -1):
        fam, bc = bases[ax]
        wr = S.padded_scalar_product(fam, bc, wr, N[ax], ax)
        if fam == "chebyshev":
            B = S.dense(S.cheb_B_SD(N[ax]), N[ax] - 2)
            wm = np.moveaxis(wr, ax, 0)
            sol = np.zeros_like(wm)
            sol[:N[ax] - 2] = np.linalg.solve(B, wm[:N[ax] - 2].reshape(N[ax] - 2, -1)).reshape(
                wm[:N[ax] - 2].shape)
            wr = np.moveaxis(sol, 0, ax)
    assert rel(w, wr) < 1e-11


@pytest.mark.parametrize("N", [9, 15, 27, 13])
@pytest.mark.parametrize("axis", [0, 1])
def test_r2c_odd_lengths(sfa, N, axis):
    """Odd R2C / C2R lengths (a full-length complex FFT) vs numpy's rfft/irfft."""
    print(f"[r2c odd N={N} axis={axis}]", flush=True)
    V = sfa.FunctionSpace(N, "F", dtype="d")
    V.axis = axis
    V.dimensions = 2
    rng = np.random.default_rng(N)
    f = rng.standard_normal((N, 6) if axis == 0 else (6, N))
    c = V.scalar_product(dev(f))
    assert rel(c, S.basis_scalar_product("fourier-r2c", None, f, axis)) < 1e-13
    u = V.backward(c)
    assert rel(u, f) < 1e-13
    assert rel(u, S.basis_backward("fourier-r2c", None, c.cpu().numpy(), axis, N=N)) < 1e-13


@pytest.mark.parametrize("fam,dt", [("fourier-r2c", "d"), ("fourier-c2c", "D")])
@pytest.mark.parametrize("N", [8, 9, 16, 33])
def test_convolve(sfa, fam, dt, N):
    """R2C/C2C.convolve: direct sums vs the oracle's restatement, and the
    padded (3/2 + 1/N, the reference test's factor) fast path vs the oracle
    and vs the truncated direct sums (test_transforms.py:84-116)."""
    print(f"[convolve {fam} N={N}]", flush=True)
    rng = np.random.default_rng(N)
    M0 = N // 2 + 1 if fam == "fourier-r2c" else N
    u0 = rng.random(M0) + 1j * rng.random(M0)
    u1 = rng.random(M0) + 1j * rng.random(M0)
    if fam == "fourier-r2c":
        u0[0] = u0[0].real
        u1[0] = u1[0].real
        if N % 2 == 0:
            u0[-1] = u0[-1].real
            u1[-1] = u1[-1].real
    V = sfa.FunctionSpace(N, "F", dtype=dt)
    uv1 = V.convolve(dev(u0), dev(u1), fast=False)
    ref1 = S.convolve_direct(u0, u1, N, fam == "fourier-r2c")
    assert rel(uv1, ref1) < 1e-13
    P = sfa.FunctionSpace(N, "F", dtype=dt, padding_factor=1.5 + 1.00001 / N)
    M = P.shape(False)
    uv2 = P.convolve(dev(u0), dev(u1), fast=True)
    ref2 = S.padded_scalar_product(fam, None, S.padded_backward(fam, None, u0, N, M)
                                   * S.padded_backward(fam, None, u1, N, M), N)
    assert rel(uv2, ref2) < 1e-12
    uv3 = np.zeros_like(ref2)
    if fam == "fourier-r2c":
        uv3[:] = ref1[:N // 2 + 1]
        if N % 2 == 0:
            uv3[-1] = uv3[-1].real * 2
    else:
        uv3[:N // 2 + 1] = ref1[:N // 2 + 1]
        uv3[-(N // 2):] += ref1[-(N // 2):]
    assert rel(uv2, uv3) < 1e-12
    with pytest.raises(ValueError):
        V.convolve(dev(u0), dev(u1), fast=True)  # unpadded: aliasing


@pytest.mark.parametrize("fam", ["C", "L"])
@pytest.mark.parametrize("bc", [(0, 0, 0, 0), {"left": {"D": 0}, "right": {"N": 0}},
                                {"left": {"N": 0.4}, "right": {"D": -0.3}}],
                         ids=["SB", "DN", "ND-nonhomogeneous"])
def test_padded_stencil_spaces(sfa, fam, bc):
    """Biharmonic and mixed bases padded by 3/2: backward evaluates the same
    series on the fine grid (eval at those points), forward on the fine grid
    recovers the coefficients; dealias_direct zeroes the top third."""
    print(f"[padded stencil {fam} {bc}]", flush=True)
    N = 18
    V = sfa.FunctionSpace(N, fam, bc=bc)
    Pd = V.get_dealiased(1.5)
    rng = np.random.default_rng(11)
    c = sfa.Function(V)
    c[:V.dim()] = torch.from_numpy(rng.standard_normal(V.dim())).cuda()
    xp = np.asarray(Pd.mesh())
    assert len(xp) == 27
    up = Pd.backward(c)
    ref = V.eval(xp, c).cpu().numpy()
    assert rel(up, ref) < 1e-12
    assert rel(Pd.forward(up), c.cpu().numpy()) < 1e-11
    D = V.get_dealiased(1, dealias_direct=True)
    cz = c.clone()
    lo = V._direct_range()[0]
    cz[lo:V.dim()] = 0
    assert rel(D.backward(c), V.backward(cz).cpu().numpy()) < 1e-12
