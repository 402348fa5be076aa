"""Configuration dict mirroring the keys of shenfun/config.py:9-94 that the hot
path reads, with the HIP backend as the only executor.

* ``config['optimization']['mode']`` -- the backend that
  ``optimization.optimizer`` resolves shenfun's Cython names in (shenfun:
  'cython' | 'numba', optimization/__init__.py:47-80); here 'hip'.  The
  banded solvers' static LU/Solve methods are resolved when the module is
  imported, the Helmholtz and per-mode solve methods at every call; a mode
  that names no backend makes those raise (there is no host implementation).
  ``verbose`` reports names the backend lacks.
* ``config['transforms']['kind'][family]`` -- transform flavour per family
  (shenfun: 'fast' | 'recursive' | 'vandermonde', config.py:16-30).  'fast'
  runs the LDS FFT/DCT kernels, 'recursive' and 'vandermonde' run the dense
  per-line contraction on FP64 MFMA (the Legendre default, like shenfun's
  'recursive'); 'auto' (the Chebyshev default here) takes the dense
  contraction up to 128 points and the DCT kernels above (the faster of the
  two on the MI355X at each size).
* ``config['precision']`` -- 'fp64' or 'mixed': the default precision of
  ``HelmholtzSolver3D`` when none is given.
"""
config = {
    "optimization": {"mode": "hip", "verbose": False},
    "transforms": {
        "kind": {
            "chebyshev": "auto",
            "legendre": "recursive",
            "fourier": "fast",
        }
    },
    "precision": "fp64",
}
