"""Continuous-batching request loop for Helmholtz/Poisson solves of mixed size.

A ``Server`` owns S slots (one HIP stream each, default priority).  Requests
arrive with an arrival time; the loop admits pending requests as slots and
device memory free up, so work of several requests overlaps on the device:

* 1D Chebyshev (SD/SN, N) requests of one class (N, bc) are stacked into one
  (B, N) array with their own alphas: one LU launch for the batch's factors
  (one line per request), then one scalar-product, one Helmholtz-solve and one
  backward launch.  A batch whose requests share one alpha that has been seen
  before uses the cached factors of (N, bc, alpha) instead, broadcast to every
  line; an alpha seen once fills no cache entry.  Lengths with a prime factor above 31 (whose mixed-radix
  DCT degenerates to an O(N p) DFT) use the cached dense transform matrices
  instead (the matrix kind of spectralbase.py:484-491, 550-558).  Reference
  path per request: chebyshev.bases scalar_product
  (chebyshev/bases.py:340-353) -> chebyshev.la.Helmholtz (chebyshev/la.py:
  223-317) -> backward (:315-338).
* 2D Legendre x Fourier requests of any sizes go together into one launch of
  the batched persistent kernel (sfa_batch2d_plan_ext + sfa_batch2d_solve)
  over a resident cache of Legendre/DFT matrices, or of its tile-parallel
  form (sfa_batch2d_wide_*) when one problem dominates the batch.
* 3D Chebyshev x Fourier^2 requests of one class (shape, bc, alpha) go
  through the fused solver with a per-slot workspace: alone as they are
  (sfa_helmholtz3d), or up to max_batch_3d of them stacked into one pass of
  launches (sfa_helmholtz3d_many) when several wait.

A request's right-hand side may be a host tensor (uploaded on the slot's
stream when its batch is launched, so the upload is inside its latency: pinned
inputs of 64 Ki elements or more by direct asynchronous copies, the others
through the slot's pinned staging buffer) or a device tensor.
Factorisations (1D Galerkin diagonals and stencils, Legendre/DFT matrices, 3D
plans) live in an LRU cache keyed by (basis, sizes, bc[, alpha]) with a byte
budget; each entry is built on the stream of the batch that missed and
carries an event that batches on other streams wait for.  Admission reserves
each batch's device bytes (inputs, outputs, staging, workspace) against a
memory budget; a batch that does not fit is shortened or waits.

Every request ends with a status: "ok", "invalid: ..." (malformed, refused at
arrival), "refused: ..." (needs more device memory than the budget),
"error: ..." (its launch failed; the other requests go on), "timeout" (its
deadline passed before admission) or "cancelled" (Server.cancel before
admission).  Latency of a request is the host-observed completion of its
batch minus its arrival time.
"""
import collections
import gc
import heapq
import itertools
import math
import time

import numpy as np
import torch

from . import _lib
from .bases import FunctionSpace
from .matrices import mass_matrix, stiffness_matrix

FAMILIES = ("cheb1d", "leg2d", "cheb3d")


class Request:
    __slots__ = ("rid", "family", "key", "f", "arrival", "t_admit", "t_done", "u", "status",
                 "deadline")

    def __init__(self, rid, family, key, f, arrival=0.0, deadline=None):
        self.rid = rid
        self.family = family  # "cheb1d" | "leg2d" | "cheb3d"
        self.key = key        # cheb1d: (N, bc, alpha); leg2d: (Nx, Ny, bc, alpha); cheb3d: ((Nx,Ny,Nz), bc, alpha)
        self.f = f
        self.arrival = arrival
        self.deadline = deadline  # seconds from the start of the run, or None
        self.t_admit = None
        self.t_done = None
        self.u = None
        self.status = None


def validate(r):
    """Raise ValueError with the reason if the request is malformed."""
    if r.family not in FAMILIES:
        raise ValueError(f"unknown family {r.family!r}")
    k = r.key
    if not isinstance(k, tuple):
        raise ValueError("key must be a tuple")
    if r.family == "cheb1d":
        if len(k) != 3:
            raise ValueError("cheb1d key is (N, bc, alpha)")
        shape = (int(k[0]),)
        bc, alpha = k[1], k[2]
        if shape[0] < 8:
            raise ValueError("cheb1d needs N >= 8")
    elif r.family == "leg2d":
        if len(k) != 4:
            raise ValueError("leg2d key is (Nx, Ny, bc, alpha)")
        shape = (int(k[0]), int(k[1]))
        bc, alpha = k[2], k[3]
        if shape[0] % 2 or shape[1] % 2 or not (8 <= shape[0] <= 1024 and 4 <= shape[1] <= 1024):
            raise ValueError("leg2d needs even Nx in [8, 1024] and even Ny in [4, 1024]")
    else:
        if len(k) != 3 or len(k[0]) != 3:
            raise ValueError("cheb3d key is ((Nx, Ny, Nz), bc, alpha)")
        shape = tuple(int(v) for v in k[0])
        bc, alpha = k[1], k[2]
        if shape[0] < 8 or shape[1] < 1 or shape[2] < 6 or shape[2] % 2:
            raise ValueError("cheb3d needs Nx >= 8, Ny >= 1 and even Nz >= 6")
    if bc not in ("D", "N"):
        raise ValueError(f"bc must be 'D' or 'N', not {bc!r}")
    if not (isinstance(alpha, (int, float)) and math.isfinite(alpha) and alpha >= 0):
        raise ValueError(f"alpha must be a finite number >= 0, not {alpha!r}")
    if r.family != "cheb1d" and bc == "N" and alpha == 0:
        raise ValueError("Neumann with alpha = 0 is singular")
    f = r.f
    if not isinstance(f, torch.Tensor):
        raise ValueError("f must be a torch tensor")
    if f.dtype != torch.float64 or tuple(f.shape) != shape:
        raise ValueError(f"f must be float64 of shape {shape}, not {f.dtype} {tuple(f.shape)}")


class FactorCache:
    """LRU map key -> (object, nbytes, ready event) with a byte budget;
    entries in use by an in-flight batch (refcount > 0) are not evicted."""

    def __init__(self, budget_bytes):
        self.budget = int(budget_bytes)
        self.items = collections.OrderedDict()
        self.nbytes = 0
        self.refs = collections.Counter()
        self.hits = 0
        self.misses = 0

    def lookup(self, key, build, stream=None):
        """(object, event): build() runs on `stream` on a miss (its kernels
        ordered on that stream) and an event recorded after it; a user on
        another stream waits for that event."""
        if key in self.items:
            self.items.move_to_end(key)
            self.hits += 1
            obj, _, ev = self.items[key]
            return obj, ev
        self.misses += 1
        ev = None
        if stream is not None:
            with torch.cuda.stream(stream):
                obj, nb = build()
                ev = torch.cuda.Event()
                ev.record(stream)
        else:
            obj, nb = build()
        self.items[key] = (obj, nb, ev)
        self.nbytes += nb
        self._evict()
        return obj, ev

    def get(self, key, build):
        return self.lookup(key, build)[0]

    def acquire(self, keys):
        for k in keys:
            self.refs[k] += 1

    def release(self, keys):
        for k in keys:
            self.refs[k] -= 1
            if self.refs[k] <= 0:
                del self.refs[k]
        self._evict()

    def _evict(self):
        for k in list(self.items):
            if self.nbytes <= self.budget:
                break
            if self.refs.get(k, 0) > 0:
                continue
            _, nb, _ = self.items.pop(k)
            self.nbytes -= nb


def _bcc(bc):
    return 0 if bc in ("D", 0) else 1


def _largest_prime_factor(n):
    p, f = 2, 1
    while p * p <= n:
        while n % p == 0:
            f, n = p, n // p
        p += 1
    return max(f, n)


class _RecentSet:
    """The last `cap` keys added (a bounded set)."""

    def __init__(self, cap):
        self.cap = int(cap)
        self.keys = collections.OrderedDict()

    def __contains__(self, k):
        return k in self.keys

    def add(self, k):
        self.keys[k] = None
        self.keys.move_to_end(k)
        while len(self.keys) > self.cap:
            self.keys.popitem(last=False)


def _use(stream, obj_ev):
    """The cached object, with `stream` ordered after its build."""
    obj, ev = obj_ev
    if ev is not None and stream is not None:
        stream.wait_event(ev)
    return obj


class Server:
    def __init__(self, slots=4, mem_budget=8 << 30, cache_budget=16 << 30, max_batch_1d=256,
                 max_batch_2d=1024, max_batch_3d=8, device="cuda", streams=None):
        self.device = torch.device(device)
        if streams is None:
            streams = [torch.cuda.Stream(device=self.device) for _ in range(int(slots))]
        self.streams = list(streams)
        self.free_slots = list(range(len(self.streams)))
        self.mem_budget = int(mem_budget)
        self.mem_used = 0
        self.cache = FactorCache(cache_budget)
        self.max_batch_1d = int(max_batch_1d)
        self.max_batch_2d = int(max_batch_2d)
        self.max_batch_3d = int(max_batch_3d)
        self.cus = (torch.cuda.get_device_properties(self.device).multi_processor_count
                    if self.device.type == "cuda" else 1)
        self.inflight = []  # (event, slot, requests, bytes, cache keys, keepalive)
        # pending requests: arrival order (finished entries dropped lazily from
        # the front) plus per-class FIFOs, so forming a batch costs O(batch)
        self.pending = collections.deque()
        self._q1d = {}                    # (N, bc) -> deque of 1D requests
        self._q2d = collections.deque()   # 2D requests
        self._q3d = {}                    # (shape, bc, alpha) -> deque of 3D requests
        self._npending = 0
        self._deadlines = []              # heap of (deadline, seq, request)
        self._seq = itertools.count()
        self._byrid = {}
        self.done = []
        self.batches = collections.Counter()
        self._work3d = {}   # slot -> workspace tensor (grown on demand)
        self._stage = {}    # slot -> pinned host staging tensor (grown on demand)
        self._seen1d = _RecentSet(4096)  # 1D (N, bc, alpha) factor keys seen once
        self._now = lambda: 0.0

    # ------------------------------------------------------------ engines
    # host inputs at least this many elements, already in pinned memory, are
    # copied to the device directly; smaller ones are gathered in the slot's
    # pinned staging buffer and copied together (one copy call per batch)
    DIRECT_UPLOAD = 1 << 16

    def _stage_upload(self, reqs, slot, stream):
        """One float64 device array holding the requests' f back to back
        (flattened), uploaded on `stream`: large pinned host inputs by direct
        host-to-device copies, other host inputs through the slot's pinned
        staging buffer, device inputs by a device copy.  Returns (device
        array, element offsets)."""
        sizes = [r.f.numel() for r in reqs]
        off = np.concatenate([[0], np.cumsum(sizes)]).astype(np.int64)
        total = int(off[-1])
        direct = [i for i, r in enumerate(reqs)
                  if not r.f.is_cuda and sizes[i] >= self.DIRECT_UPLOAD and r.f.is_pinned()]
        dset = set(direct)
        staged = [i for i, r in enumerate(reqs) if not r.f.is_cuda and i not in dset]
        with torch.cuda.stream(stream):
            D = torch.empty(total, dtype=torch.float64, device=self.device)
            for i in direct:
                D[off[i]:off[i + 1]].copy_(reqs[i].f.reshape(-1), non_blocking=True)
            if staged:
                st = self._stage.get(slot)
                if st is None or st.numel() < total:
                    st = torch.empty(max(total, 1 << 16), dtype=torch.float64).pin_memory()
                    self._stage[slot] = st
                for i in staged:
                    st[off[i]:off[i + 1]].copy_(reqs[i].f.reshape(-1))
                if len(staged) == len(reqs):
                    D.copy_(st[:total], non_blocking=True)
                else:
                    for i in staged:
                        D[off[i]:off[i + 1]].copy_(st[off[i]:off[i + 1]], non_blocking=True)
            for i, r in enumerate(reqs):
                if r.f.is_cuda:
                    D[off[i]:off[i + 1]].copy_(r.f.reshape(-1))
        return D, off

    def _cheb1d_class(self, N, bc, stream):
        """Galerkin diagonals, stencil and (long prime lengths) dense transform
        matrices of the class (N, bc); the factors are per batch."""
        def build():
            from .la import _helm_diags
            V = FunctionSpace(N, "chebyshev", bc=bc)
            diags = [torch.from_numpy(np.ascontiguousarray(d, dtype=float)).to(self.device)
                     for d in _helm_diags(stiffness_matrix(V), mass_matrix(V))]
            K2 = torch.from_numpy(np.ascontiguousarray(V.stencil_k2())).to(self.device)
            dense = None
            nb = 8 * (6 * N + N)
            if _largest_prime_factor(N) > 31:
                dense = (V._analysis_matrix(), V._synthesis_matrix())
                nb += 16 * N * N
            return (V, diags, K2, dense), nb
        key = ("cheb1d", N, bc)
        return _use(stream, self.cache.lookup(key, build, stream)), key

    def _leg_matrices(self, Nx, bc, stream):
        def build():
            L = _lib.lib()
            m = torch.empty(int(L.sfa_batch2d_leg_elems(Nx)), dtype=torch.float64, device=self.device)
            _lib.call("sfa_batch2d_leg_matrices", Nx, bc, _lib.ptr(m), _lib.stream_ptr())
            return m, m.numel() * 8
        return _use(stream, self.cache.lookup(("leg", Nx, bc), build, stream))

    def _dft_matrices(self, Ny, stream):
        def build():
            L = _lib.lib()
            m = torch.empty(int(L.sfa_batch2d_dft_elems(Ny)), dtype=torch.float64, device=self.device)
            _lib.call("sfa_batch2d_dft_matrices", Ny, _lib.ptr(m), _lib.stream_ptr())
            return m, m.numel() * 8
        return _use(stream, self.cache.lookup(("dft", Ny), build, stream))

    def _cheb3d_space(self, shape, bc):
        def build():
            from .tensorproductspace import TensorProductSpace
            T = TensorProductSpace(None, (FunctionSpace(shape[0], "chebyshev", bc=bc),
                                          FunctionSpace(shape[1], "F", dtype="D"),
                                          FunctionSpace(shape[2], "F", dtype="d")))
            return T, 0
        return self.cache.get(("tps3d", shape, bc), build)

    def _cheb3d_solver(self, key, stream):
        shape, bc, alpha = key

        def build():
            from .pipeline import HelmholtzSolver3D
            H = HelmholtzSolver3D(self._cheb3d_space(shape, bc), float(alpha), own_workspace=False)
            return H, H.plan.numel()
        return _use(stream, self.cache.lookup(("cheb3d",) + key, build, stream))

    def _launch_cheb1d(self, reqs, stream, slot):
        N, bc = reqs[0].key[0], reqs[0].key[1]
        (V, diags, K2, dense), ckey = self._cheb1d_class(N, bc, stream)
        neumann = int(bc == "N")
        B = len(reqs)
        shp = _lib.arr([B, N])
        st = _lib.arr([N, 1])
        F, _ = self._stage_upload(reqs, slot, stream)
        alphas = {float(r.key[2]) for r in reqs}
        keys = [ckey]
        abh = ab = None
        if len(alphas) == 1:
            # one alpha seen before: the cached factors of (N, bc, alpha),
            # broadcast to every line (an alpha seen once is factorised with
            # the batch, so requests with alphas of their own fill no cache)
            fkey = ("cheb1d_lu", N, bc, next(iter(alphas)))
            if fkey in self.cache.items or fkey in self._seen1d:
                fac, fkey = self._cheb1d_factors(N, bc, next(iter(alphas)), diags, stream)
                fst = _lib.arr([0, 1])
                keys.append(fkey)
            else:
                self._seen1d.add(fkey)
        with torch.cuda.stream(stream):
            F = F.view(B, N)
            fh = torch.empty_like(F)
            U = torch.empty_like(F)
            sp = _lib.stream_ptr(stream)
            if len(keys) == 1:
                # the batch's factors, one line per request (chebyshev/la.py:223-276):
                # alfa = -1 and beta = alpha per request, staged in pinned memory and
                # copied without a host wait (the host array is kept with the batch)
                abh = torch.tensor([-1.0] * B + [float(r.key[2]) for r in reqs],
                                   dtype=torch.float64).pin_memory()
                ab = abh.to(self.device, non_blocking=True)
                fac = torch.zeros((4, B, N), dtype=torch.float64, device=self.device)
                fst = st
                A0, A2, A4, Bm2, B0, B2 = diags
                _lib.call("sfa_helmholtz_lu", neumann, N, _lib.ptr(A0), _lib.ptr(A2), _lib.ptr(A4),
                          _lib.ptr(Bm2), _lib.ptr(B0), _lib.ptr(B2), _lib.ptr(ab), _lib.ptr(ab[B:]), 1,
                          _lib.ptr(fac[0]), _lib.ptr(fac[1]), _lib.ptr(fac[2]), _lib.ptr(fac[3]), 2,
                          shp, st, 1, sp)
            if dense is None:
                _lib.call("sfa_cheb_scalar_product", _lib.ptr(F), _lib.ptr(fh), _lib.F64, 2, shp, st, st,
                          1, _lib.ptr(K2), sp)
            else:
                _lib.call("sfa_line_matmul", _lib.ptr(F), _lib.ptr(fh), _lib.F64, 2, shp, st, st, 1, N,
                          _lib.ptr(dense[0]), sp)
            _lib.call("sfa_helmholtz_solve", _lib.ptr(fh), _lib.ptr(fh), _lib.F64, 2, shp, st, st, 1,
                      neumann, _lib.ptr(fac[0]), _lib.ptr(fac[1]), _lib.ptr(fac[2]), _lib.ptr(fac[3]), fst,
                      sp)
            if dense is None:
                _lib.call("sfa_cheb_backward", _lib.ptr(fh), _lib.ptr(U), _lib.F64, 2, shp, st, st, 1,
                          _lib.ptr(K2), sp)
            else:
                _lib.call("sfa_line_matmul", _lib.ptr(fh), _lib.ptr(U), _lib.F64, 2, shp, st, st, 1, N,
                          _lib.ptr(dense[1]), sp)
        for i, r in enumerate(reqs):
            r.u = U[i]
        return (F, fh, U, fac, ab, abh), keys

    def _cheb1d_factors(self, N, bc, alpha, diags, stream):
        """Helmholtz LU factors (4, 1, N) of the class (N, bc) at one alpha
        (chebyshev/la.py:223-276), cached."""
        def build():
            abh = torch.tensor([-1.0, float(alpha)], dtype=torch.float64).pin_memory()
            abd = abh.to(self.device, non_blocking=True)
            fac = torch.zeros((4, 1, N), dtype=torch.float64, device=self.device)
            A0, A2, A4, Bm2, B0, B2 = diags
            _lib.call("sfa_helmholtz_lu", int(bc == "N"), N, _lib.ptr(A0), _lib.ptr(A2),
                      _lib.ptr(A4), _lib.ptr(Bm2), _lib.ptr(B0), _lib.ptr(B2), _lib.ptr(abd),
                      _lib.ptr(abd[1:]), 1, _lib.ptr(fac[0]), _lib.ptr(fac[1]), _lib.ptr(fac[2]),
                      _lib.ptr(fac[3]), 2, _lib.arr([1, N]), _lib.arr([N, 1]), 1, _lib.stream_ptr())
            fac.pinned_source = abh  # the pinned copy source lives as long as the factors
            return fac, fac.numel() * 8
        key = ("cheb1d_lu", N, bc, float(alpha))
        return _use(stream, self.cache.lookup(key, build, stream)), key

    def _wide_2d(self, Nx, Ny):
        """Tile-parallel mode when one problem would dominate the persistent
        kernel's makespan: its contraction flops exceed twice the per-CU share."""
        m = Nx.astype(np.float64)
        y = Ny.astype(np.float64)
        fl = m * y * (4 * m + 2) + 2 * m * m * (y + 2)
        return bool(fl.max() * self.cus > 2.0 * fl.sum())

    def _launch_leg2d(self, reqs, stream, slot):
        L = _lib.lib()
        n = len(reqs)
        Nx = np.array([r.key[0] for r in reqs], dtype=np.int32)
        Ny = np.array([r.key[1] for r in reqs], dtype=np.int32)
        bc = np.array([_bcc(r.key[2]) for r in reqs], dtype=np.int32)
        al = np.array([r.key[3] for r in reqs], dtype=np.float64)
        # the batch's distinct matrices, looked up once each
        lk = list(dict.fromkeys(zip(Nx.tolist(), bc.tolist())))
        dk = list(dict.fromkeys(Ny.tolist()))
        legm = [self._leg_matrices(a, b, stream) for a, b in lk]
        dftm = [self._dft_matrices(y, stream) for y in dk]
        keys = [("leg", a, b) for a, b in lk] + [("dft", y) for y in dk]
        mats = legm + dftm
        # one arena view: offsets relative to the lowest matrix address
        base = min(m.data_ptr() for m in mats)
        top = max(m.data_ptr() + m.numel() * 8 for m in mats)
        lo = {k: (m.data_ptr() - base) // 8 for k, m in zip(lk, legm)}
        do = {k: (m.data_ptr() - base) // 8 for k, m in zip(dk, dftm)}
        phi = np.array([lo[k] for k in zip(Nx.tolist(), bc.tolist())], dtype=np.int64)
        dft = np.array([do[k] for k in Ny.tolist()], dtype=np.int64)
        sizes = Nx.astype(np.int64) * Ny.astype(np.int64)
        F, foff = self._stage_upload(reqs, slot, stream)
        foff = np.ascontiguousarray(foff[:-1])
        total = int(sizes.sum())
        nb = min(n, self.cus)
        ptr = lambda a: _lib.ctypes.c_void_p(a.ctypes.data)
        wide = self._wide_2d(Nx, Ny)
        with torch.cuda.stream(stream):
            U = torch.empty(total, dtype=torch.float64, device=self.device)
            sp = _lib.stream_ptr(stream)
            if wide:
                plan = torch.empty(int(L.sfa_batch2d_wide_plan_bytes(n, ptr(Nx), ptr(Ny))),
                                   dtype=torch.uint8, device=self.device)
                work = torch.empty(int(L.sfa_batch2d_wide_workspace(n, ptr(Nx), ptr(Ny))),
                                   dtype=torch.uint8, device=self.device)
                _lib.call("sfa_batch2d_wide_plan", n, ptr(Nx), ptr(Ny), ptr(al), ptr(bc), ptr(foff),
                          ptr(foff), total, total, _lib.ctypes.c_void_p(base), ptr(phi), ptr(dft),
                          (top - base) // 8, _lib.ptr(plan), plan.numel(), sp)
                _lib.call("sfa_batch2d_wide_solve", _lib.ptr(F), _lib.ptr(U), _lib.ptr(plan),
                          _lib.ptr(work), work.numel(), sp)
            else:
                plan = torch.empty(int(L.sfa_batch2d_plan_ext_bytes(n, nb)), dtype=torch.uint8,
                                   device=self.device)
                work = torch.empty(int(L.sfa_batch2d_workspace(n, ptr(Nx), ptr(Ny), nb)),
                                   dtype=torch.uint8, device=self.device)
                _lib.call("sfa_batch2d_plan_ext", n, ptr(Nx), ptr(Ny), ptr(al), ptr(bc), ptr(foff),
                          ptr(foff), total, total, _lib.ctypes.c_void_p(base), ptr(phi), ptr(dft),
                          (top - base) // 8, nb, _lib.ptr(plan), plan.numel(), sp)
                _lib.call("sfa_batch2d_solve", _lib.ptr(F), _lib.ptr(U), _lib.ptr(plan),
                          _lib.ptr(work), work.numel(), 0, sp)
        for r, part in zip(reqs, U.split(sizes.tolist())):
            r.u = part.view(r.key[0], r.key[1])
        return (F, U, plan, work, mats), keys

    def _launch_cheb3d(self, reqs, stream, slot):
        """One request: the fused solver on its own f.  Several requests of one
        class: their right-hand sides stacked and solved in one pass of
        launches (sfa_helmholtz3d_many); the results are bitwise those of the
        single solves."""
        H = self._cheb3d_solver(reqs[0].key, stream)
        B = len(reqs)
        need = B * int(_lib.lib().sfa_helmholtz3d_workspace(*H.shape, 0))
        w = self._work3d.get(slot)
        if w is None or w.numel() < need:
            with torch.cuda.stream(stream):
                w = torch.empty(need, dtype=torch.uint8, device=self.device)
            self._work3d[slot] = w
        F, _ = self._stage_upload(reqs, slot, stream) if (B > 1 or not reqs[0].f.is_cuda) \
            else (reqs[0].f, None)
        with torch.cuda.stream(stream):
            if B == 1:
                F = F.view(H.shape)
                U = torch.empty_like(F)
                H(F, U, work=w)
                reqs[0].u = U
            else:
                F = F.view((B,) + tuple(H.shape))
                U = H.solve_many(F, work=w)
                for i, r in enumerate(reqs):
                    r.u = U[i]
        return (F, U), [("cheb3d",) + reqs[0].key]

    # ------------------------------------------------------------ admission
    @staticmethod
    def _bytes(reqs):
        """Device bytes a batch reserves: input and output arrays (plus the
        intermediate and factors for 1D, the per-problem workspace for 2D,
        the workspace for 3D)."""
        fam = reqs[0].family
        n = sum(r.f.numel() for r in reqs) * 8
        if fam == "cheb1d":
            return 7 * n
        if fam == "leg2d":
            mx = max((r.key[0] * 2) * (r.key[1] + 2) for r in reqs) * 8
            return 2 * n + mx * min(len(reqs), 256)
        shp = reqs[0].key[0]
        ws = int(_lib.lib().sfa_helmholtz3d_workspace(*shp, 0))  # workspace per solve
        return 2 * n + len(reqs) * ws

    def _finish(self, r, status, t):
        r.status = status
        r.t_done = t
        self.done.append(r)

    def _push(self, r):
        """Queue an arriving request, or finish it at once if it is malformed
        or larger than the memory budget."""
        try:
            validate(r)
        except (ValueError, TypeError) as e:
            self._finish(r, f"invalid: {e}", self._now())
            return
        need = self._bytes([r])
        if need > self.mem_budget:
            self._finish(r, f"refused: needs {need} bytes of device memory, budget "
                            f"{self.mem_budget}", self._now())
            return
        self._byrid[r.rid] = r
        self.pending.append(r)
        if r.family == "cheb1d":
            self._q1d.setdefault(r.key[:2], collections.deque()).append(r)
        elif r.family == "leg2d":
            self._q2d.append(r)
        else:
            self._q3d.setdefault(r.key, collections.deque()).append(r)
        if r.deadline is not None:
            heapq.heappush(self._deadlines, (r.deadline, next(self._seq), r))
        self._npending += 1

    def cancel(self, rid):
        """Cancel a request that is still pending; False if it was admitted or
        finished already."""
        r = self._byrid.get(rid)
        if r is None or r.t_admit is not None or r.status is not None:
            return False
        self._drop(r, "cancelled")
        return True

    def _drop(self, r, status):
        self._npending -= 1
        self._byrid.pop(r.rid, None)
        self._finish(r, status, self._now())

    def _expire(self, t):
        while self._deadlines and self._deadlines[0][0] <= t:
            _, _, r = heapq.heappop(self._deadlines)
            if r.t_admit is None and r.status is None:
                self._drop(r, "timeout")

    def _queue(self, head):
        if head.family == "cheb1d":
            return self._q1d[head.key[:2]]
        return self._q2d if head.family == "leg2d" else self._q3d[head.key]

    def _cap(self, family):
        return {"cheb1d": self.max_batch_1d, "leg2d": self.max_batch_2d,
                "cheb3d": self.max_batch_3d}[family]

    @staticmethod
    def _live(r):
        return r.t_admit is None and r.status is None

    def _next_batch(self):
        """Oldest pending request and the requests that can join its launch
        (the front of its class FIFO, oldest first), removed from the class
        FIFO; finished (timed-out, cancelled) entries are dropped on the way."""
        while self.pending and not self._live(self.pending[0]):
            self.pending.popleft()
        if not self.pending:
            return None
        head = self.pending[0]
        q = self._queue(head)
        batch = []
        cap = self._cap(head.family)
        while q and len(batch) < cap:
            r = q.popleft()
            if self._live(r):
                batch.append(r)
        self._prune_queue(head)
        return batch

    def _prune_queue(self, head):
        q = self._queue(head)
        while q and not self._live(q[0]):
            q.popleft()
        if not q and head.family == "cheb1d":
            del self._q1d[head.key[:2]]
        elif not q and head.family == "cheb3d":
            del self._q3d[head.key]

    def _untake(self, batch):
        """Put requests of a formed batch back at the front of their FIFO."""
        head = batch[0]
        if head.family == "cheb1d":
            q = self._q1d.setdefault(head.key[:2], collections.deque())
        elif head.family == "leg2d":
            q = self._q2d
        else:
            q = self._q3d.setdefault(head.key, collections.deque())
        for r in reversed(batch):
            q.appendleft(r)

    def _take(self, batch, t):
        """Mark an admitted batch (from _next_batch)."""
        self._npending -= len(batch)
        for r in batch:
            r.t_admit = t
            self._byrid.pop(r.rid, None)

    def _poll(self, now):
        still = []
        for ev, slot, reqs, nbytes, keys, keep in self.inflight:
            if ev.query():
                t = now()
                for r in reqs:
                    self._finish(r, "ok", t)
                self.free_slots.append(slot)
                self.mem_used -= nbytes
                self.cache.release(keys)
                if reqs[0].family == "leg2d":  # drop the plan's host-side registry entry
                    _lib.lib().sfa_batch2d_release(_lib.ptr(keep[2]))
            else:
                still.append((ev, slot, reqs, nbytes, keys, keep))
        self.inflight = still

    def _launch(self, batch, stream, slot):
        fam = batch[0].family
        if fam == "cheb1d":
            return self._launch_cheb1d(batch, stream, slot)
        if fam == "leg2d":
            return self._launch_leg2d(batch, stream, slot)
        return self._launch_cheb3d(batch, stream, slot)

    def _record(self, stream):
        ev = torch.cuda.Event()
        ev.record(stream)
        return ev

    def _admit(self, now):
        while self.free_slots and self._npending:
            batch = self._next_batch()
            if not batch:
                continue
            nbytes = self._bytes(batch)
            # shorten a batch that does not fit; a single request always fits
            # an idle device (larger ones were refused at arrival)
            while len(batch) > 1 and self.mem_used + nbytes > self.mem_budget:
                self._untake(batch[-1:])
                batch.pop()
                nbytes = self._bytes(batch)
            if self.mem_used + nbytes > self.mem_budget and self.inflight:
                self._untake(batch)
                return  # wait for memory
            slot = self.free_slots.pop(0)
            stream = self.streams[slot]
            self._take(batch, now())
            fam = batch[0].family
            try:
                keep, keys = self._launch(batch, stream, slot)
            except Exception as e:  # a failed launch ends its own requests only
                t = now()
                for r in batch:
                    r.u = None
                    self._finish(r, f"error: {type(e).__name__}: {e}", t)
                self.free_slots.insert(0, slot)
                continue
            self.cache.acquire(keys)
            ev = self._record(stream)
            self.batches[fam] += 1
            self.mem_used += nbytes
            self.inflight.append((ev, slot, batch, nbytes, keys, keep))

    # ------------------------------------------------------------ loop
    def run(self, requests, realtime=True):
        """Serve ``requests`` (sorted by arrival, seconds from the start).
        realtime: release each request at its arrival time (open loop); else
        all are pending at once.  Returns the requests; each ends with a
        status."""
        reqs = sorted(requests, key=lambda r: r.arrival)
        for r in reqs:
            r.t_admit = r.t_done = r.u = r.status = None
        self.done = []
        self._sync()
        # the objects alive before the run are moved out of the collector's
        # reach for its duration, so a collection pauses the loop only for the
        # objects the run itself creates (a full collection over a large heap
        # stalls every request in flight)
        freeze = gc.isenabled()
        if freeze:
            gc.collect()
            gc.freeze()
        try:
            return self._loop(reqs, realtime)
        finally:
            if freeze:
                gc.unfreeze()

    def _loop(self, reqs, realtime):
        t0 = time.perf_counter()
        now = lambda: time.perf_counter() - t0
        self._now = now
        i = 0
        n = len(reqs)
        while len(self.done) < n:
            t = now() if realtime else float("inf")
            while i < n and reqs[i].arrival <= t:
                self._push(reqs[i])
                i += 1
            self._expire(now())
            self._poll(now)
            self._admit(now)
            if not self.inflight and not self._npending and i < n and realtime:
                dt = reqs[i].arrival - now()
                if dt > 2e-4:
                    time.sleep(dt - 1e-4)
        self._sync()
        return reqs

    def _sync(self):
        if self.device.type == "cuda":
            torch.cuda.synchronize(self.device)

    @staticmethod
    def stats(reqs):
        """Latency percentiles and throughput over the requests served ("ok"),
        and the count of every status."""
        ok = [r for r in reqs if r.status == "ok"]
        counts = collections.Counter(r.status.split(":")[0] for r in reqs if r.status)
        out = {"requests": len(reqs), "served": len(ok), "status": dict(counts)}
        if ok:
            lat = np.array([r.t_done - r.arrival for r in ok]) * 1e3
            span = max(r.t_done for r in ok) - min(r.arrival for r in ok)
            up = sum(r.f.numel() * 8 for r in ok if not r.f.is_cuda)
            out.update({"p50_ms": float(np.percentile(lat, 50)),
                        "p99_ms": float(np.percentile(lat, 99)), "mean_ms": float(lat.mean()),
                        "solves_per_s": len(ok) / span if span > 0 else float("inf"),
                        "uploaded_GB_per_s": up / span / 1e9 if span > 0 else 0.0})
        return out


# ---------------------------------------------------------------- workload
def cheb1d_rhs(N, rng):
    """Smooth f on the N Chebyshev-Gauss nodes: sum_k c_k T_k(x_j),
    c_k ~ N(0,1)/(1+k^2)."""
    k = np.arange(N)
    c = rng.standard_normal(N) / (1 + k * k)
    x = np.cos(np.pi * (2 * np.arange(N) + 1) / (2 * N))
    return np.polynomial.chebyshev.chebval(x, c)


def request_stream(n, rate, seed=0, classes=(32, 64, 16), device="cuda", upload=True,
                   alpha_per_request=True):
    """n requests with Poisson arrivals (rate per second) from the serving mix:
    40% 1D Chebyshev N in [64, 4096], 40% 2D Legendre x Fourier up to 512 x 256,
    20% 3D Chebyshev x Fourier^2 up to 256^3 (each axis in {32, 64, 128, 256}),
    Dirichlet/Neumann 50/50, alpha ~ U(0, 50).  Sizes and bc come from a fixed
    number of classes per family, each with one seeded smooth right-hand side.
    upload: the right-hand sides are pinned host tensors that each request
    uploads when admitted (else device-resident).  alpha_per_request: alpha is
    drawn per request (every 1D batch factorises, every 3D request plans
    anew), else per class (the factor cache can hit)."""
    from .workloads import batch2d_rhs, helmholtz3d_rhs
    rng = np.random.default_rng(seed)
    c1, c2, c3 = classes
    bcs = lambda m: [("D", "N")[b] for b in rng.integers(0, 2, m)]
    n1 = 2 * np.round(np.exp(rng.uniform(np.log(64), np.log(4096), c1)) / 2).astype(int)
    k1 = list(zip(n1.tolist(), bcs(c1), rng.uniform(0, 50, c1).tolist()))
    nx = 2 * np.round(np.exp(rng.uniform(np.log(32), np.log(512), c2)) / 2).astype(int)
    ny = 2 * np.round(np.exp(rng.uniform(np.log(32), np.log(256), c2)) / 2).astype(int)
    k2 = list(zip(nx.tolist(), ny.tolist(), bcs(c2), rng.uniform(0, 50, c2).tolist()))
    sz = rng.choice([32, 64, 128, 256], size=(c3, 3))
    k3 = [(tuple(int(v) for v in s), b, a) for s, b, a in zip(sz, bcs(c3), rng.uniform(0, 50, c3))]
    place = (lambda a: torch.from_numpy(a).pin_memory()) if upload else \
        (lambda a: torch.from_numpy(a).to(device))
    f1 = [place(cheb1d_rhs(k[0], rng)) for k in k1]
    f2 = [place(batch2d_rhs(k[0], k[1], rng)) for k in k2]
    f3 = [place(helmholtz3d_rhs(*k[0], seed=int(rng.integers(1 << 30)))) for k in k3]
    fam = rng.choice(3, size=n, p=[0.4, 0.4, 0.2])
    cls = rng.integers(0, 1 << 30, n)
    arrivals = np.cumsum(rng.exponential(1.0 / rate, n))
    alphas = rng.uniform(0, 50, n)
    out = []
    for i in range(n):
        a = float(alphas[i])
        if fam[i] == 0:
            j = cls[i] % c1
            key = (k1[j][0], k1[j][1], a if alpha_per_request else k1[j][2])
            out.append(Request(i, "cheb1d", key, f1[j], arrivals[i]))
        elif fam[i] == 1:
            j = cls[i] % c2
            key = k2[j][:3] + (a if alpha_per_request else k2[j][3],)
            out.append(Request(i, "leg2d", key, f2[j], arrivals[i]))
        else:
            j = cls[i] % c3
            key = (k3[j][0], k3[j][1], a if alpha_per_request else float(k3[j][2]))
            out.append(Request(i, "cheb3d", key, f3[j], arrivals[i]))
    return out
