"""GPU parity: per-axis transforms (C ABI via shenfun_amd) vs the FP64 oracle."""
import numpy as np
import pytest
import torch

from oracle import spectral as S

pytestmark = pytest.mark.gpu

TOL = 1e-12  # max-norm relative error vs the FP64 oracle


def rel(a, b):
    a = a.cpu().numpy() if isinstance(a, torch.Tensor) else a
    return np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300)


def dev(a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


@pytest.fixture(scope="module")
def sfa():
    import shenfun_amd
    return shenfun_amd


@pytest.mark.parametrize("n", [8, 16, 30, 34, 64, 96, 100, 128, 254, 256, 512, 1000, 4096])
def test_fft_c2c_1d(sfa, n):
    print(f"[c2c n={n}]", flush=True)
    rng = np.random.default_rng(n)
    x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    B = sfa.FunctionSpace(n, "F", dtype="D")
    out = B.scalar_product(dev(x))
    assert rel(out, np.fft.fft(x) / n) < TOL
    back = B.backward(dev(np.fft.fft(x)))
    assert rel(back, np.fft.ifft(np.fft.fft(x)) * n) < TOL


@pytest.mark.parametrize("shape,axis", [((16, 12, 10), 0), ((16, 12, 10), 1), ((16, 12, 10), 2),
                                        ((6, 256, 5), 1), ((64, 64), 0), ((64, 64), 1)])
def test_fft_c2c_axes(sfa, shape, axis):
    print(f"[c2c {shape} axis={axis}]", flush=True)
    rng = np.random.default_rng(1)
    x = rng.standard_normal(shape) + 1j * rng.standard_normal(shape)
    B = sfa.FunctionSpace(shape[axis], "F", dtype="D")
    B.axis = axis
    out = B.scalar_product(dev(x))
    assert rel(out, np.fft.fft(x, axis=axis) / shape[axis]) < TOL
    inplace = dev(x)
    B.backward(inplace, inplace)
    assert rel(inplace, np.fft.ifft(x, axis=axis) * shape[axis]) < TOL


@pytest.mark.parametrize("shape,axis", [((64,), 0), ((10,), 0), ((34,), 0), ((16, 12, 10), 2),
                                        ((16, 12, 10), 0), ((8, 20, 6), 1), ((64, 256), 1),
                                        ((4, 4, 4096), 2)])
def test_fft_r2c_c2r(sfa, shape, axis):
    print(f"[r2c {shape} axis={axis}]", flush=True)
    rng = np.random.default_rng(2)
    x = rng.standard_normal(shape)
    n = shape[axis]
    B = sfa.FunctionSpace(n, "F", dtype="d")
    B.axis = axis
    X = B.scalar_product(dev(x))
    ref = np.fft.rfft(x, axis=axis) / n
    assert rel(X, ref) < TOL
    # c2r ignores the imaginary parts of the k = 0 and Nyquist entries
    Xc = np.fft.rfft(x, axis=axis)
    s0 = [slice(None)] * x.ndim
    s0[axis] = 0
    Xc[tuple(s0)] += 0.25j
    back = B.backward(dev(Xc))
    assert rel(back, np.fft.irfft(Xc, n=n, axis=axis) * n) < TOL


@pytest.mark.parametrize("bc", [None, "D", "N"])
@pytest.mark.parametrize("N", [8, 9, 33, 64, 256])
@pytest.mark.parametrize("cplx", [False, True])
def test_chebyshev_1d(sfa, bc, N, cplx):
    print(f"[cheb bc={bc} N={N} cplx={cplx}]", flush=True)
    rng = np.random.default_rng(N)
    f = rng.standard_normal(N) + (1j * rng.standard_normal(N) if cplx else 0)
    V = sfa.FunctionSpace(N, "chebyshev", bc=bc)
    for kind in ("fast", "vandermonde"):
        sp = V.scalar_product(dev(f), kind=kind)
        ref = S.basis_scalar_product("chebyshev", bc, f)
        assert rel(sp, ref) < TOL, kind
        c = rng.standard_normal(N) + (1j * rng.standard_normal(N) if cplx else 0)
        if bc is not None:
            c[-2:] = 0
        bw = V.backward(dev(c), kind=kind)
        assert rel(bw, S.basis_backward("chebyshev", bc, c)) < TOL, kind


@pytest.mark.parametrize("bc", [None, "D", "N"])
@pytest.mark.parametrize("shape,axis", [((64, 33), 0), ((12, 64), 1), ((256, 20, 6), 0),
                                        ((5, 128, 7), 1), ((6, 5, 256), 2)])
def test_chebyshev_axes(sfa, bc, shape, axis):
    print(f"[cheb axes bc={bc} {shape} axis={axis}]", flush=True)
    rng = np.random.default_rng(7)
    f = rng.standard_normal(shape) + 1j * rng.standard_normal(shape)
    V = sfa.FunctionSpace(shape[axis], "chebyshev", bc=bc)
    V.axis = axis
    sp = V.scalar_product(dev(f))
    assert rel(sp, S.basis_scalar_product("chebyshev", bc, f, axis)) < TOL
    bw = V.backward(dev(f))
    fc = f.copy()
    if bc is not None:
        s = [slice(None)] * f.ndim
        s[axis] = slice(shape[axis] - 2, None)
        fc[tuple(s)] = 0
    assert rel(bw, S.basis_backward("chebyshev", bc, fc, axis)) < TOL


@pytest.mark.parametrize("bc", [None, "D", "N"])
@pytest.mark.parametrize("N", [8, 33, 64, 200, 512])
def test_legendre(sfa, bc, N):
    print(f"[leg bc={bc} N={N}]", flush=True)
    rng = np.random.default_rng(N + 1)
    f = rng.standard_normal((N, 5)) + 1j * rng.standard_normal((N, 5))
    V = sfa.FunctionSpace(N, "legendre", bc=bc)
    sp = V.scalar_product(dev(f))
    assert rel(sp, S.basis_scalar_product("legendre", bc, f, 0)) < TOL
    c = f.copy()
    if bc is not None:
        c[-2:] = 0
    bw = V.backward(dev(c))
    assert rel(bw, S.basis_backward("legendre", bc, c, 0)) < TOL


@pytest.mark.parametrize("family", ["chebyshev", "legendre"])
@pytest.mark.parametrize("bc", ["D", "N"])
def test_forward_projection_roundtrip(sfa, family, bc):
    """forward(backward(u)) == u for u in the space (projection with the mass solve)."""
    print(f"[proj {family} {bc}]", flush=True)
    N = 40
    rng = np.random.default_rng(3)
    u = rng.standard_normal((N, 3)) + 1j * rng.standard_normal((N, 3))
    u[-2:] = 0
    V = sfa.FunctionSpace(N, family, bc=bc)
    phys = V.backward(dev(u))
    uh = V.forward(phys)
    assert rel(uh, u) < 1e-12


@pytest.mark.parametrize("cplx", [False, True])
@pytest.mark.parametrize("shape,axis", [((97, 1), 0), ((3, 130), 1), ((130, 4), 0), ((2, 150, 3), 1),
                                        ((5, 2, 61), 2), ((17, 70), 1), ((64, 40), 1), ((97, 40), 0),
                                        ((70, 5, 9), 0), ((128, 300), 0), ((33, 17), 0)])
def test_line_matmul_few_and_many_lines(sfa, shape, axis, cplx):
    """sfa_line_matmul (dense per-line contraction: the row-parallel form for
    up to 16 lines, the MFMA GEMM form for more lines along a strided axis
    whose other dims flatten, the tiled form otherwise) vs numpy, strided and
    contiguous lines, nout != nin."""
    from shenfun_amd import _lib
    print(f"[line_matmul {shape} axis={axis} cplx={cplx}]", flush=True)
    rng = np.random.default_rng(sum(shape) + axis)
    x = rng.standard_normal(shape) + (1j * rng.standard_normal(shape) if cplx else 0)
    nin = shape[axis]
    nout = nin - 3
    M = rng.standard_normal((nout, nin))
    ref = np.moveaxis(np.tensordot(M, x, axes=([1], [axis])), 0, axis)
    X = dev(x)
    oshape = list(shape)
    oshape[axis] = nout
    Y = torch.empty(oshape, dtype=X.dtype, device="cuda")
    Md = dev(M)
    _lib.call("sfa_line_matmul", _lib.ptr(X), _lib.ptr(Y), _lib.C128 if cplx else _lib.F64, X.dim(),
              _lib.arr(X.shape), _lib.arr(X.stride()), _lib.arr(Y.stride()), axis, nout, _lib.ptr(Md),
              _lib.stream_ptr())
    assert rel(Y, ref) < 1e-13


@pytest.mark.parametrize("bc", [None, "D", "N"])
@pytest.mark.parametrize("N", [2, 5, 9, 33, 65, 100, 257])
@pytest.mark.parametrize("cplx", [False, True])
def test_chebyshev_gl_1d(sfa, bc, N, cplx):
    """Gauss-Lobatto scalar product / backward (DCT-I through the FFT of the
    even extension, length 2(N-1); 100 -> 198 = 2 x 9 x 11) vs the oracle,
    and the dense (Vandermonde) kind on the GL nodes."""
    if bc is not None and N < 3:
        pytest.skip("composite basis needs N >= 3")
    print(f"[cheb GL bc={bc} N={N} cplx={cplx}]", flush=True)
    rng = np.random.default_rng(N + 7)
    f = rng.standard_normal((N, 3)) + (1j * rng.standard_normal((N, 3)) if cplx else 0)
    V = sfa.FunctionSpace(N, "chebyshev", bc=bc, quad="GL")
    c = rng.standard_normal((N, 3)) + (1j * rng.standard_normal((N, 3)) if cplx else 0)
    if bc is not None:
        c[-2:] = 0
    for kind in ("fast", "vandermonde"):
        sp = V.scalar_product(dev(f), kind=kind)
        assert rel(sp, S.basis_scalar_product("chebyshev", bc, f, 0, quad="GL")) < TOL, kind
        bw = V.backward(dev(c), kind=kind)
        assert rel(bw, S.basis_backward("chebyshev", bc, c, 0, quad="GL")) < TOL, kind


@pytest.mark.parametrize("bc", [None, "D", "N"])
def test_chebyshev_gl_forward_roundtrip(sfa, bc):
    """forward(backward(u)) == u with the GL mass matrix (last norm pi)."""
    print(f"[cheb GL roundtrip {bc}]", flush=True)
    N = 40
    rng = np.random.default_rng(5)
    u = rng.standard_normal((N, 4)) + 1j * rng.standard_normal((N, 4))
    if bc is not None:
        u[-2:] = 0
    V = sfa.FunctionSpace(N, "chebyshev", bc=bc, quad="GL")
    uh = V.forward(V.backward(dev(u)))
    assert rel(uh, u) < 1e-12


@pytest.mark.parametrize("bc", ["D", "N"])
def test_helmholtz2d_gl_dropin(sfa, bc):
    """(alpha - lap) u = f on Chebyshev-GL x Fourier: T.scalar_product ->
    la.Helmholtz (GL mass matrix) -> T.backward vs the oracle chain."""
    print(f"[helm2d GL {bc}]", flush=True)
    Nx, Ny, alpha = 36, 16, 2.0
    rng = np.random.default_rng(11)
    f = rng.standard_normal((Nx, Ny))
    T = sfa.TensorProductSpace(None, (sfa.FunctionSpace(Nx, "chebyshev", bc=bc, quad="GL"),
                                      sfa.FunctionSpace(Ny, "F", dtype="d")))
    fh = T.scalar_product(dev(f))
    H = sfa.la.Helmholtz(*sfa.helmholtz_tpmatrices(T, alpha))
    uh = H(fh, sfa.Function(T))
    u = T.backward(uh)
    fh_ref = S.basis_scalar_product("chebyshev", bc, S.basis_scalar_product("fourier-r2c", None, f, 1),
                                    0, quad="GL")
    ky = np.fft.rfftfreq(Ny, 1.0 / Ny)
    d0, d1, d2, L = S.helmholtz_lu(bc, Nx, np.full(ky.shape, -1.0), alpha + ky ** 2, quad="GL")
    uh_ref = S.helmholtz_solve(fh_ref, d0, d1, d2, L, neumann=(bc == "N"))
    u_ref = S.basis_backward("fourier-r2c", None,
                             S.basis_backward("chebyshev", bc, uh_ref, 0, quad="GL"), 1, N=Ny)
    assert rel(uh, uh_ref) < 1e-12
    assert rel(u, u_ref) < 1e-12


def test_planned_arrays_semantics(sfa):
    """spectralbase.py:235-313 / :602-648: transforms own planned device
    arrays; without arguments they transform input_array into output_array
    and return it; forward shares the scalar product's output array; an
    explicit output_array receives the result instead."""
    print("[planned arrays]", flush=True)
    N = 24
    V = sfa.FunctionSpace(N, "chebyshev", bc="D")
    rng = np.random.default_rng(1)
    f = rng.standard_normal(N)
    V.forward.input_array[:] = dev(f)
    out = V.forward()
    assert out is V.forward.output_array and out is V.scalar_product.output_array
    ref = S.basis_scalar_product("chebyshev", "D", f)
    assert rel(V.scalar_product(dev(f)), ref) < TOL
    assert V.scalar_product(dev(f)) is V.scalar_product.output_array  # same planned array again
    mine = torch.empty(N, dtype=torch.float64, device="cuda")
    assert V.scalar_product(dev(f), mine) is mine and rel(mine, ref) < TOL
    T = sfa.TensorProductSpace(None, (sfa.FunctionSpace(16, "chebyshev", bc="N"),
                                      sfa.FunctionSpace(12, "F", dtype="d")))
    g = rng.standard_normal((16, 12))
    T.forward.input_array[:] = dev(g)
    gh = T.forward()
    assert gh is T.forward.output_array and tuple(gh.shape) == (16, 7)
    gh0 = gh.cpu().numpy()
    u = T.backward(gh.clone())
    assert u is T.backward.output_array
    again = T.forward(u)  # overwrites the planned output array (gh) in place
    assert again is gh and rel(again, gh0) < 1e-12


def test_project(sfa):
    """project (forms/project.py:16-172): a physical array, a callable on the
    mesh and a Function of another space all project onto T; projecting a
    member of T returns its coefficients."""
    print("[project]", flush=True)
    Nx, Ny = 24, 16
    T = sfa.TensorProductSpace(None, (sfa.FunctionSpace(Nx, "chebyshev", bc="D"),
                                      sfa.FunctionSpace(Ny, "F", dtype="d")))
    rng = np.random.default_rng(9)
    uh = sfa.Function(T)
    uh[:Nx - 2] = dev(rng.standard_normal((Nx - 2, Ny // 2 + 1)) + 1j * rng.standard_normal((Nx - 2, Ny // 2 + 1)))
    uh[:, 0] = uh[:, 0].real + 0j
    uh[:, -1] = uh[:, -1].real + 0j
    u = T.backward(uh).clone()
    assert rel(sfa.project(u, T), uh.cpu().numpy()) < 1e-12
    # a callable in the space: (1 - x^2) cos(y) = phi_0-like member of SD x R2C
    f = lambda x, y: (1 - x ** 2) * np.cos(y)
    pf = sfa.project(f, T)
    assert rel(T.backward(pf), f(*T.mesh())) < 1e-12
    # a Function of the orthogonal space with the same mesh
    C = sfa.TensorProductSpace(None, (sfa.FunctionSpace(Nx, "chebyshev"),
                                      sfa.FunctionSpace(Ny, "F", dtype="d")))
    ch = C.forward(u).clone()
    out = sfa.Function(T)
    assert sfa.project(ch, T, output_array=out, uh_space=C) is out
    assert rel(out, uh.cpu().numpy()) < 1e-12
