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
