"""shenfun_amd: MI355X-native spectral-Galerkin Helmholtz hot path.

A drop-in for shenfun's per-axis transforms (``forward``/``backward``/
``scalar_product``), its tensor-product pipeline and its banded solvers
(``la.TDMA``/``PDMA``/``SolverGeneric1ND``, ``chebyshev.la.Helmholtz``), with
all compute in hand-written HIP kernels for gfx950 behind the C ABI of
``include/shenfun_amd.h``.  torch-ROCm tensors are the buffer type.
"""
from . import io, la  # noqa: F401
from .io import Checkpoint  # noqa: F401
from ._lib import lib as _load_lib  # noqa: F401
from .bases import (C2C, R2C, ChebyshevOrthogonal, ChebyshevShenDirichlet,  # noqa: F401
                    ChebyshevShenNeumann, FunctionSpace, LegendreOrthogonal,
                    LegendreShenDirichlet, LegendreShenNeumann)
from .config import config  # noqa: F401
from .matrices import (SparseMatrix, TPMatrix, biharmonic_tpmatrices, helmholtz_tpmatrices, mass_matrix,  # noqa: F401
                       stiffness_matrix)
from .batched import BatchedHelmholtz2D  # noqa: F401
from .pipeline import HelmholtzSolver3D  # noqa: F401
from .tensorproductspace import Array, Function, TensorProductSpace, project  # noqa: F401

__version__ = "0.1.0"
