"""Timed CPU baseline of the 3D Helmholtz workload (test oracle / baseline only).

The reference's own CPU path cannot run here (mpi4py, mpi4py_fft/FFTW and
numba are absent), so the baseline is this restatement of it:
scipy.fft (pocketfft; same unnormalised DFT/DCT definitions as FFTW's
r2c/c2c/REDFT10/REDFT01) for the transforms, numpy for the composite stencils,
and the plain-C Solve_Helmholtz loop of oracle/banded.c (la.pyx:871-947
order) for the banded solve.  ``workers``/``nthreads`` = 1 matches shenfun's
defaults (FFTW threads=1, serial Cython loops).
"""
import ctypes
import os

import numpy as np
import scipy.fft as sf

from . import spectral as S

_LIB = os.path.join(os.path.dirname(os.path.abspath(__file__)), "_build", "liboracle.so")
_lib = None


def lib():
    global _lib
    if _lib is None:
        L = ctypes.CDLL(_LIB)
        L.orc_helmholtz_solve_axis0.argtypes = [ctypes.c_int, ctypes.c_long] + [ctypes.c_void_p] * 2 + \
            [ctypes.c_int] + [ctypes.c_void_p] * 4 + [ctypes.c_int]
        L.orc_helmholtz_solve_axis0.restype = None
        _lib = L
    return _lib


def prepare(shape, alpha, bc="D"):
    """LU factors per (ky, kz) mode (excluded from the timed solve)."""
    Nx, Ny, Nz = shape
    ky = S.c2c_wavenumbers(Ny).astype(float)
    kz = S.r2c_wavenumbers(Nz).astype(float)
    beta = alpha + ky[:, None] ** 2 + kz[None, :] ** 2
    d0, d1, d2, L = S.helmholtz_lu(bc, Nx, -np.ones_like(beta), beta)
    return [np.ascontiguousarray(a) for a in (d0, d1, d2, L)]


def helmholtz3d(f, factors, bc="D", workers=1, nthreads=1):
    """u for (alpha - lap) u = f, Chebyshev(bc) x C2C x R2C, FP64 on the CPU."""
    Nx, Ny, Nz = f.shape
    g = sf.rfft(f, axis=2, workers=workers) / Nz
    g = sf.fft(g, axis=1, workers=workers, overwrite_x=True) / Ny
    t = (sf.dct(g.real, type=2, axis=0, workers=workers)
         + 1j * sf.dct(g.imag, type=2, axis=0, workers=workers)) * (np.pi / (2 * Nx))
    K2 = S.stencil_k2("chebyshev", bc, Nx)
    t[:Nx - 2] += K2[:, None, None] * t[2:]
    u_hat = np.zeros_like(t)
    d0, d1, d2, L = factors
    lib().orc_helmholtz_solve_axis0(Nx, Ny * (Nz // 2 + 1), t.ctypes.data, u_hat.ctypes.data,
                                    int(bc == "N"), d0.ctypes.data, d1.ctypes.data, d2.ctypes.data,
                                    L.ctypes.data, nthreads)
    o = np.zeros_like(u_hat)
    o[:Nx - 2] += u_hat[:Nx - 2]
    o[2:] += K2[:, None, None] * u_hat[:Nx - 2]
    b = 0.5 * (sf.dct(o.real, type=3, axis=0, workers=workers)
               + 1j * sf.dct(o.imag, type=3, axis=0, workers=workers)) + o[0:1] / 2
    b = sf.ifft(b, axis=1, workers=workers, overwrite_x=True) * Ny
    return sf.irfft(b, n=Nz, axis=2, workers=workers) * Nz
