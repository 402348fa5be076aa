"""Seeded synthetic inputs of the benchmark workloads (host numpy, FP64)."""
import numpy as np
import scipy.fft


def helmholtz3d_rhs(Nx, Ny, Nz, seed=0):
    """Smooth random f for (alpha - lap) u = f on Chebyshev x C2C x R2C.

    Spectral coefficients complex N(0,1) (1 + kx^2 + ky^2 + kz^2)^-2 on
    T_kx x exp(i ky y) x exp(i kz z) (kx = 0..Nx-1, ky = fftfreq, kz = rfftfreq);
    draw order real parts then imaginary parts, shape (Nx, Ny, Nz//2+1).
    Hermitian symmetry comes from synthesising with an FP64 inverse real FFT
    (c2c along y, c2r along z); the Chebyshev sum over kx is evaluated at the
    Gauss nodes with an unnormalised DCT-III.
    """
    rng = np.random.default_rng(seed)
    shp = (Nx, Ny, Nz // 2 + 1)
    c = rng.standard_normal(shp) + 1j * rng.standard_normal(shp)
    kx = np.arange(Nx, dtype=float)[:, None, None]
    ky = np.fft.fftfreq(Ny, 1.0 / Ny)[None, :, None]
    kz = np.fft.rfftfreq(Nz, 1.0 / Nz)[None, None, :]
    c *= (1 + kx ** 2 + ky ** 2 + kz ** 2) ** -2.0
    g = np.fft.irfftn(c, s=(Ny, Nz), axes=(1, 2)) * (Ny * Nz)
    f = 0.5 * scipy.fft.dct(g, type=3, axis=0) + g[0:1] / 2
    return np.ascontiguousarray(f)


def batch2d_problems(n=4096, seed=0):
    """Sizes and alphas of the batched uneven 2D Legendre x Fourier workload.

    Nx = 2 round(e^U(ln 32, ln 512) / 2), Ny = 2 round(e^U(ln 32, ln 256) / 2),
    alpha ~ U(0, 100); draw order: the n uniforms of Nx, then of Ny, then alpha.
    """
    rng = np.random.default_rng(seed)
    nx = 2 * np.round(np.exp(rng.uniform(np.log(32), np.log(512), n)) / 2).astype(np.int64)
    ny = 2 * np.round(np.exp(rng.uniform(np.log(32), np.log(256), n)) / 2).astype(np.int64)
    alpha = rng.uniform(0.0, 100.0, n)
    return nx, ny, alpha, rng


def batch2d_rhs(Nx, Ny, rng):
    """Smooth random f on Legendre-Gauss(Nx) x uniform(Ny) nodes: coefficients
    complex N(0,1) / (1 + kx^2 + ky^2) on L_kx(x) e^{i ky y} (kx < Nx,
    ky = 0..Ny/2; draw order real then imaginary parts, shape (Nx, Ny/2+1)),
    synthesised with the Legendre Vandermonde matrix and an FP64 inverse real
    FFT (which takes the Hermitian part)."""
    nyh = Ny // 2 + 1
    c = rng.standard_normal((Nx, nyh)) + 1j * rng.standard_normal((Nx, nyh))
    kx = np.arange(Nx, dtype=float)[:, None]
    ky = np.arange(nyh, dtype=float)[None, :]
    c /= 1 + kx ** 2 + ky ** 2
    g = _legvander_gauss(Nx) @ c
    return np.ascontiguousarray(np.fft.irfft(g, n=Ny, axis=1) * Ny)


_VANDER = {}


def _legvander_gauss(N):
    """Legendre Vandermonde matrix L_k(x_j) at the N Gauss nodes (cached)."""
    V = _VANDER.get(N)
    if V is None:
        from numpy.polynomial import legendre as npleg

        from .bases import leggauss
        V = _VANDER[N] = npleg.legvander(leggauss(N)[0], N - 1)
    return V
