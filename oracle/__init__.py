"""FP64 CPU oracle for shenfun's spectral-Galerkin Helmholtz hot path.

TEST INFRASTRUCTURE, NOT PRODUCT CODE.  Only ``tests/``,
``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of ``bench.py`` may
import, call, link or execute anything in this package, and only as the
checker (or as the timed CPU baseline).  The product path
(``shenfun_amd``) never imports it and fails loudly when its HIP library is
missing.

Contents
--------
``spectral``  numpy/scipy restatement of the reference algorithms (transforms,
              stencils, Galerkin matrices, banded LU/solves, Helmholtz LU/solve/
              matvec, tensor-product pipelines).  Every function cites the
              reference file:line it restates.
``banded.c``  plain-C restatement of the serial banded loops (TDMA/PDMA/
              Helmholtz along an axis) used as the timed CPU baseline; built by
              ``oracle/build.py`` into ``oracle/_build/``.
``_ref``      (built, never committed) ``fastgl.cpp`` from the reference,
              compiled where it lies, used to pin the Legendre-Gauss nodes.

Pinning (see tests/test_oracle_*.py): each oracle function is checked against
the reference's doctest known answers, closed forms, library routines that
share the definition (scipy.fft DCT = FFTW REDFT10/01, numpy.fft, numpy
chebvander/legvander/leggauss), dense linear algebra on small sizes, and the
manufactured-solution closed form of the 2D Legendre x Fourier Poisson case.
No oracle function is "parity unpinned".
"""
from . import spectral  # noqa: F401
