"""Fused forward -> Helmholtz solve -> backward on the GPU.

``HelmholtzSolver3D`` runs (alpha - lap) u = f on a Chebyshev(SD/SN) x C2C x
R2C TensorProductSpace in one C-ABI call (sfa_helmholtz3d): the sequence
T.scalar_product -> chebyshev.la.Helmholtz -> T.backward of the reference
(tensorproductspace.py:146-190, chebyshev/la.py:288-317,
tensorproductspace.py:46-96), with the Helmholtz LU factorised once per distinct beta = alpha + ky^2 + kz^2
at construction (like chebyshev.la.Helmholtz.__init__).

precision='mixed' keeps the lowest kz planes (the dominant part of smooth
data) in complex128 with FP64 transforms and stores/transforms the rest as
complex64; the solve and the first/last (physical) transforms stay FP64.  Its
error depends on the data (on the decay of the spectrum along z), so every
mixed solve is guarded (sfa_helmholtz3d_guarded): the kernels sum the spectral
energy above the cut, the host predicts the resulting relative max error
after the x-pass, and a prediction above ``tol`` redoes the solve with more
complex128 planes or in FP64.  ``last_kh`` / ``last_predicted_error`` report
the cut that produced the returned u.  guard=False (or precision='mixed' with
an explicit ``fp64_planes`` and guard=False) runs the unchecked form.

Nonhomogeneous boundary values on the Chebyshev axis (constant, varying along
y and z, or time-dependent) give u = u_0 + L: u_0 the fused solve with
homogeneous values, L the lift computed once by the drop-in chain on f = 0
(see ``lift``) and added by the solve's final z C2R as it stores u.
"""
import ctypes

import torch

from . import _lib
from .bases import C2C, ChebyshevBase, R2C


class HelmholtzSolver3D:
    # after a solve that needed FP64, this many calls run FP64 directly before
    # the mixed form is tried again
    FP64_STREAK = 8

    def __init__(self, T, alpha, precision=None, fp64_planes=0, tol=1e-10, guard=True,
                 own_workspace=True):
        if precision is None:
            from .config import config
            precision = config["precision"]
        b = T.bases
        if not (len(b) == 3 and isinstance(b[0], ChebyshevBase) and b[0].bc_kind in ("D", "N")
                and isinstance(b[1], C2C) and isinstance(b[2], R2C)):
            raise ValueError("expects Chebyshev(SD/SN) x C2C x R2C")
        if tuple(getattr(T, "axes", (0, 1, 2))) != (0, 1, 2):
            raise NotImplementedError("the fused solver takes the natural transform order axes=(0, 1, 2)")
        if b[0].quad != "GC":
            raise NotImplementedError("the fused solver is built for Chebyshev-Gauss (GC) points; "
                                      "use T.scalar_product -> la.Helmholtz -> T.backward for GL")
        if any(x.is_padded for x in b):
            raise NotImplementedError("the fused solver takes unpadded spaces; use "
                                      "T.scalar_product -> la.Helmholtz -> T.backward")
        if b[0].has_nonhomogeneous_bcs and b[0].bc_kind == "N" and float(alpha) == 0.0:
            raise NotImplementedError("nonhomogeneous Neumann values need alpha > 0 in the fused "
                                      "solver; use T.scalar_product -> la.Helmholtz -> T.backward")
        self.T = T
        self.alpha = float(alpha)
        self.neumann = b[0].bc_kind == "N"
        self.precision = {"fp64": 0, "mixed": 1}[precision]
        self.fp64_planes = int(fp64_planes)
        self.tol = float(tol)
        self.guard = bool(guard) and self.precision == 1
        self._kh_hint = self.fp64_planes
        self._fp64_left = 0
        self.last_kh = None
        self.last_predicted_error = None
        self._lift = None
        self._lift_key = None
        self.shape = T.shape(False)
        Nx, Ny, Nz = self.shape
        self._ws_bytes = int(_lib.lib().sfa_helmholtz3d_workspace(Nx, Ny, Nz, self.precision))
        # own_workspace=False: every call brings its own ``work`` (the serving
        # loop's per-slot workspaces)
        self.work = (torch.empty(self._ws_bytes, dtype=torch.uint8, device=T.device)
                     if own_workspace else None)
        # factorise once (chebyshev/la.py:223-276): LU factors per distinct beta
        pbytes = _lib.lib().sfa_helmholtz3d_plan_bytes(Nx, Ny, Nz)
        self.plan = torch.empty(int(pbytes), dtype=torch.uint8, device=T.device)
        _lib.call("sfa_helmholtz3d_plan", Nx, Ny, Nz, self.alpha, int(self.neumann),
                  _lib.ptr(self.plan), self.plan.numel(), _lib.stream_ptr())
        self._planned = True

    def __del__(self):
        # drop the library's registration of the plan buffer before its memory
        # can be handed to another tensor
        if getattr(self, "_planned", False):
            try:
                _lib.lib().sfa_helmholtz3d_plan_release(_lib.ptr(self.plan))
            except Exception:
                pass

    def __call__(self, f, u=None, work=None):
        """u for (alpha - lap) u = f.  work: optional uint8 CUDA tensor of at
        least ``workspace_bytes`` bytes used instead of the solver's own
        workspace (lets several solves of one solver run on different streams)."""
        if not (f.is_cuda and f.dtype == torch.float64 and f.is_contiguous()
                and tuple(f.shape) == tuple(self.shape)):
            raise ValueError(f"f must be a contiguous float64 CUDA tensor of shape {self.shape}")
        if u is None:
            u = torch.empty_like(f)
        if not (u.is_contiguous() and u.shape == f.shape and u.dtype == f.dtype):
            raise ValueError("u must match f")
        w = self.work if work is None else work
        if w is None or w.numel() < self._ws_bytes or not w.is_cuda:
            raise ValueError(f"work must hold at least {self._ws_bytes} bytes on the device")
        self._run(f, u, 1, w)
        return u

    def _run(self, f, u, nb, w):
        if self.T.bases[0].has_nonhomogeneous_bcs:
            self.lift()  # registered with the plan: the solve's last kernel adds it
        self._solve(f, u, nb, w)

    def lift(self):
        """The lift L of the x-basis' boundary values on the grid (FP64):
        (alpha - lap) L = 0 with L (Dirichlet) or L_x (Neumann) equal to the
        given values at x = -1, 1.  A solve returns u = u_0 + L, u_0 the fused
        solve of f with homogeneous values: the same discrete solution as the
        drop-in chain, where the bc matrices move the boundary basis' part to
        the right-hand side (chebyshev/la.py:303-307) and backward adds the
        boundary basis (tensorproductspace.py:2070-2089).  L is that chain on
        f = 0, computed when first needed and again when the values change
        (time-dependent values after ``bc.update``), and registered with the
        plan (sfa_helmholtz3d_set_lift) so that the solve's final z C2R adds
        it as it stores u.  The guard's error prediction covers u_0."""
        b0 = self.T.bases[0]
        key = repr(b0.bc_values)
        if self._lift is None or key != self._lift_key:
            from . import la
            from .matrices import helmholtz_tpmatrices
            from .tensorproductspace import Function
            H = la.Helmholtz(*helmholtz_tpmatrices(self.T, self.alpha))
            zero = torch.zeros(tuple(self.T.shape(True)), dtype=torch.complex128,
                               device=self.T.device)
            L = torch.empty(tuple(self.shape), dtype=torch.float64, device=self.T.device)
            L = self.T.backward(H(zero, Function(self.T)), L)
            _lib.call("sfa_helmholtz3d_set_lift", _lib.ptr(self.plan), _lib.ptr(L))
            self._lift = L
            self._lift_key = key
        return self._lift

    def _solve(self, f, u, nb, w):
        Nx, Ny, Nz = self.shape
        if self.guard and self._fp64_left == 0:
            used, nxt, err = ctypes.c_int(0), ctypes.c_int(0), ctypes.c_double(0.0)
            _lib.call("sfa_helmholtz3d_guarded", _lib.ptr(f), _lib.ptr(u), nb, Nx, Ny, Nz,
                      self.alpha, int(self.neumann), self._kh_hint, self.tol, _lib.ptr(self.plan),
                      _lib.ptr(w), w.numel(), ctypes.addressof(used), ctypes.addressof(nxt),
                      ctypes.addressof(err), _lib.stream_ptr())
            self.last_kh, self.last_predicted_error = used.value, err.value
            if nxt.value > 0:
                self._kh_hint = max(nxt.value, self.fp64_planes)
            else:
                self._fp64_left = self.FP64_STREAK
            return
        precision = self.precision
        if self.guard:  # FP64 streak after a fallback
            self._fp64_left -= 1
            precision = 0
        if nb == 1:
            _lib.call("sfa_helmholtz3d", _lib.ptr(f), _lib.ptr(u), Nx, Ny, Nz, self.alpha,
                      int(self.neumann), precision, self.fp64_planes, _lib.ptr(self.plan),
                      _lib.ptr(w), w.numel(), _lib.stream_ptr())
        else:
            _lib.call("sfa_helmholtz3d_many", _lib.ptr(f), _lib.ptr(u), nb, Nx, Ny, Nz, self.alpha,
                      int(self.neumann), precision, self.fp64_planes, _lib.ptr(self.plan),
                      _lib.ptr(w), w.numel(), _lib.stream_ptr())
        self.last_kh = (Nz // 2 + 1) if precision == 0 else self.fp64_plane_count()
        self.last_predicted_error = 0.0 if precision == 0 else None

    def solve_many(self, F, U=None, work=None):
        """Solves of a stack F (B, Nx, Ny, Nz) of right-hand sides in one pass
        of launches (sfa_helmholtz3d_many); each U[b] is bitwise equal to
        self(F[b]).  work: at least B * ``workspace_bytes`` bytes (allocated
        when None)."""
        if not (F.is_cuda and F.dtype == torch.float64 and F.is_contiguous() and F.dim() == 4
                and tuple(F.shape[1:]) == tuple(self.shape)):
            raise ValueError(f"F must be a contiguous float64 CUDA tensor of shape (B, {self.shape})")
        if U is None:
            U = torch.empty_like(F)
        if not (U.is_contiguous() and U.shape == F.shape and U.dtype == F.dtype):
            raise ValueError("U must match F")
        B = F.shape[0]
        need = B * self._ws_bytes
        if work is None:
            work = self.work if (B == 1 and self.work is not None) else \
                torch.empty(need, dtype=torch.uint8, device=F.device)
        if work.numel() < need or not work.is_cuda:
            raise ValueError(f"work must hold at least {need} bytes on the device")
        self._run(F, U, B, work)
        return U

    @property
    def workspace_bytes(self):
        return self._ws_bytes

    def stage_times(self, f, u, nrep):
        """Milliseconds per stage (z R2C, y C2C, x-pass, y C2C inverse, z C2R)
        averaged over nrep runs, from HIP events on the launch stream.  A
        guarded mixed solver times its kernels (with their statistics) at the
        cut of its last solve."""
        ms = (ctypes.c_float * 5)()
        Nx, Ny, Nz = self.shape
        precision, planes = self.precision, self.fp64_planes
        if self.guard:
            kh = self.last_kh if self.last_kh is not None else self.fp64_plane_count()
            precision, planes = (2, kh) if kh < Nz // 2 + 1 else (0, 0)
        _lib.call("sfa_helmholtz3d_stage_times", _lib.ptr(f), _lib.ptr(u), Nx, Ny, Nz, self.alpha,
                  int(self.neumann), precision, planes, _lib.ptr(self.plan),
                  _lib.ptr(self._work_for_timing()), self._ws_bytes, int(nrep), ms,
                  _lib.stream_ptr())
        names = ("z_r2c", "y_c2c", "x_pass", "y_c2c_inv", "z_c2r")
        return {n: ms[i] / nrep for i, n in enumerate(names)}

    def _work_for_timing(self):
        if self.work is None:
            self.work = torch.empty(self._ws_bytes, dtype=torch.uint8, device=self.T.device)
        return self.work

    def fp64_plane_count(self):
        """kz planes the next mixed solve keeps in complex128 (after clamping):
        the cut the guard found sufficient on the last input, else
        fp64_planes (0: 17)."""
        Nzh = self.shape[2] // 2 + 1
        k = self._kh_hint if self._kh_hint > 0 else 17
        return max(2, min(k, Nzh - 2))
