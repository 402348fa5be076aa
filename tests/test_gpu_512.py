"""GPU parity of the 3D Helmholtz solve at 512 x 512 x 512 (the mixed-precision
configuration's shape): the FP64 fused pipeline within 1e-12 and the guarded
mixed form within 1e-10 of the FP64 oracle, Dirichlet and Neumann.  The oracle
is the FP64 CPU restatement (scipy.fft transforms + the C Helmholtz solve)."""
import os

import numpy as np
import pytest
import torch

from oracle import cpu_pipeline as C
from oracle import spectral as S

pytestmark = pytest.mark.gpu

N = 512


def rel(a, b):
    return float(np.max(np.abs(a - b)) / np.max(np.abs(b)))


@pytest.fixture(scope="module")
def sfa():
    import shenfun_amd
    return shenfun_amd


@pytest.fixture(scope="module")
def f512():
    print(f"[512^3 rhs]", flush=True)
    return S.helmholtz3d_rhs(N, N, N, seed=0)


@pytest.mark.parametrize("bc", ["D", "N"])
def test_helmholtz3d_512(sfa, f512, bc):
    print(f"[helm3d 512^3 {bc}]", flush=True)
    workers = min(16, len(os.sched_getaffinity(0)))
    if not os.path.exists(C._LIB):
        from oracle import build as obuild
        obuild.build()
    u_ref = C.helmholtz3d(f512, C.prepare(f512.shape, 1.0, bc), bc, workers=workers,
                          nthreads=workers)
    print("  oracle done", flush=True)
    T = sfa.TensorProductSpace(None, (sfa.FunctionSpace(N, "chebyshev", bc=bc),
                                      sfa.FunctionSpace(N, "F", dtype="D"),
                                      sfa.FunctionSpace(N, "F", dtype="d")))
    f = torch.from_numpy(f512).cuda()
    H64 = sfa.HelmholtzSolver3D(T, 1.0)
    u = H64(f).cpu().numpy()
    e64 = rel(u, u_ref)
    del H64
    Hmx = sfa.HelmholtzSolver3D(T, 1.0, precision="mixed")
    u = Hmx(f).cpu().numpy()
    emx = rel(u, u_ref)
    print(f"  fp64 err {e64:.3e}  mixed err {emx:.3e} (kh={Hmx.last_kh}, "
          f"predicted {Hmx.last_predicted_error:.3e})", flush=True)
    assert e64 < 1e-12
    assert emx < 1e-10
    assert Hmx.last_kh == 17
