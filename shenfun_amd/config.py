"""Configuration dict mirroring the keys of shenfun/config.py:9-94 that the hot
path reads, with the HIP backend as the only executor.

* ``config['optimization']['mode']`` -- backend name resolved by the solvers
  (shenfun: 'cython' | 'numba', optimization/__init__.py:47-80); here 'hip'.
* ``config['transforms']['kind'][family]`` -- transform flavour per family
  (shenfun: 'fast' | 'recursive' | 'vandermonde', config.py:16-30).  'fast'
  runs the LDS FFT/DCT kernels, 'recursive' and 'vandermonde' run the dense
  per-line contraction (the Legendre default, like shenfun's 'recursive').
* ``config['precision']`` -- 'fp64' or 'mixed' for the fused 3D pipeline.
"""
config = {
    "optimization": {"mode": "hip", "verbose": False},
    "transforms": {
        "kind": {
            "chebyshev": "fast",
            "legendre": "recursive",
            "fourier": "fast",
        }
    },
    "precision": "fp64",
}
