"""Write the golden fixtures of this folder (run in the build container).

* ``fastgl_leggauss.npz``: Legendre-Gauss nodes/weights from the reference's
  own fastgl.cpp (shenfun/legendre/fastgl), compiled where it lies into
  oracle/_ref/libfastgl_ref.so by oracle/build.py, in the ascending order of
  legendre/fastgl/fastgl_wrap.pyx:22-29.
* ``kat.json``: known answers printed in the reference's doctests
  (chebyshev/matrices.py:37-41, legendre/matrices.py:34-38,
  spectralbase.py:2304-2308, legendre/dlt.py:87-93).
* ``poisson2d_seed{0,1}.npz``: the 2D Legendre-Dirichlet x Fourier Poisson
  configuration: seeded right-hand side on the LG x uniform mesh and the exact
  discrete solution (closed form u = (2/3) phi_0(x) sum a_k cos(k y + phi_k)).
"""
import ctypes
import json
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
FASTGL = os.path.join(ROOT, "oracle", "_ref", "libfastgl_ref.so")
NS = [1, 2, 3, 4, 5, 17, 33, 64, 100, 101, 102, 173, 256, 333, 512]


def fastgl():
    lib = ctypes.CDLL(FASTGL)
    out = {}
    for N in NS:
        x = np.zeros(N)
        w = np.zeros(N)
        lib.ref_leggauss(N, x.ctypes.data_as(ctypes.c_void_p), w.ctypes.data_as(ctypes.c_void_p))
        out[f"x{N}"] = x
        out[f"w{N}"] = w
    np.savez(os.path.join(HERE, "fastgl_leggauss.npz"), Ns=np.array(NS), **out)


def kat():
    pi = np.pi
    d = {
        # chebyshev/matrices.py:37-41, ShenDirichlet(10) mass matrix
        "cheb_BSDSD_10": {"-2": [-pi / 2], "0": [1.5 * pi] + [pi] * 7, "2": [-pi / 2]},
        # spectralbase.py:2304-2308 / forms/inner.py:155-160, ShenDirichlet(6)
        "cheb_BSDSD_6": {"-2": [-pi / 2], "0": [1.5 * pi, pi, pi, pi], "2": [-pi / 2]},
        # legendre/matrices.py:34-38, ShenDirichlet(10), to 1e-7
        "leg_BSDSD_10": {
            "-2": [-0.4, -0.28571429, -0.22222222, -0.18181818, -0.15384615, -0.13333333],
            "0": [2.4, 0.95238095, 0.62222222, 0.46753247, 0.37606838, 0.31515152, 0.27149321,
                  0.23859649],
            "2": [-0.4, -0.28571429, -0.22222222, -0.18181818, -0.15384615, -0.13333333]},
        # legendre/dlt.py:87-93: backward of u = L_1 on LG(4) = the nodes, to 1e-8
        "leg_backward_L1_4": [-0.86113631, -0.33998104, 0.33998104, 0.86113631],
    }
    with open(os.path.join(HERE, "kat.json"), "w") as f:
        json.dump(d, f, indent=1)


def poisson2d(seed, Nx=64, Ny=64):
    """Same draw order as the configuration: a_k = N(0,1)/k^2 (8), phi_k = U(0, 2pi) (8)."""
    import sys
    sys.path.insert(0, ROOT)
    from oracle import spectral as S
    rng = np.random.default_rng(seed)
    k = np.arange(1, 9, dtype=float)
    a = rng.standard_normal(8) / k ** 2
    phi = rng.uniform(0.0, 2 * np.pi, 8)
    x = np.array(np.load(os.path.join(HERE, "fastgl_leggauss.npz"))[f"x{Nx}"])
    y = np.arange(Ny) * 2 * np.pi / Ny
    X, Y = x[:, None], y[None, :]
    u = sum((1 - X ** 2) * ak * np.cos(kk * Y + pk) for kk, ak, pk in zip(k, a, phi))
    f = sum((2 + kk ** 2 * (1 - X ** 2)) * ak * np.cos(kk * Y + pk) for kk, ak, pk in zip(k, a, phi))
    uh = np.zeros((Nx, Ny // 2 + 1), dtype=complex)
    for kk, ak, pk in zip(k.astype(int), a, phi):
        uh[0, kk] = (2.0 / 3.0) * ak * np.exp(1j * pk) / 2
    assert np.allclose(S.poisson2d_manufactured(Nx, Ny, seed)[2], f, rtol=0, atol=1e-12)
    np.savez(os.path.join(HERE, f"poisson2d_seed{seed}.npz"), x=x, y=y, f=f, u=u, u_hat=uh,
             a=a, phi=phi)


if __name__ == "__main__":
    fastgl()
    kat()
    poisson2d(0)
    poisson2d(1)
    print("golden fixtures written")
