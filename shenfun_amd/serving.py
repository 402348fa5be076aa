"""Continuous-batching request loop for Helmholtz/Poisson solves of mixed size.

A ``Server`` owns S slots (one HIP stream each, default priority).  Requests
arrive with an arrival time; the loop admits pending requests as slots and
device memory free up, so work of several requests overlaps on the device:

* 1D Chebyshev (SD/SN, N) requests of one class (N, bc, alpha) are stacked
  into one (B, N) array: one scalar-product, one Helmholtz-solve and one
  backward launch for the batch (the solve broadcasts the class's LU factors
  with a zero line stride).  Lengths with a prime factor above 31 (whose
  mixed-radix DCT degenerates to an O(N p) DFT) use the cached dense
  transform matrices instead (the matrix kind of spectralbase.py:484-491,
  550-558).  Reference path per request:
  chebyshev.bases scalar_product (chebyshev/bases.py:340-353) ->
  chebyshev.la.Helmholtz (chebyshev/la.py:288-317) -> backward (:315-338).
* 2D Legendre x Fourier requests of any sizes go together into one launch of
  the batched persistent kernel (sfa_batch2d_plan_ext + sfa_batch2d_solve)
  over a resident cache of Legendre/DFT matrices, or of its tile-parallel
  form (sfa_batch2d_wide_*) when one problem dominates the batch.
* 3D Chebyshev x Fourier^2 requests of one class (shape, bc, alpha) go
  through the fused solver with a per-slot workspace: alone as they are
  (sfa_helmholtz3d), or up to max_batch_3d of them stacked into one pass of
  launches (sfa_helmholtz3d_many) when several wait.

Factorisations (1D LU factors, Legendre/DFT matrices, 3D plans) live in an
LRU cache keyed by (basis, sizes, bc, alpha) with a byte budget.  Admission
reserves each batch's device bytes (outputs, staging, workspace) against a
memory budget; a batch that does not fit waits.  Latency of a request is the
host-observed completion of its batch minus its arrival time.
"""
import collections
import itertools
import time

import numpy as np
import torch

from . import _lib
from .bases import FunctionSpace
from .matrices import mass_matrix, stiffness_matrix


class Request:
    __slots__ = ("rid", "family", "key", "f", "arrival", "t_admit", "t_done", "u")

    def __init__(self, rid, family, key, f, arrival=0.0):
        self.rid = rid
        self.family = family  # "cheb1d" | "leg2d" | "cheb3d"
        self.key = key        # cheb1d: (N, bc, alpha); leg2d: (Nx, Ny, bc, alpha); cheb3d: ((Nx,Ny,Nz), bc, alpha)
        self.f = f
        self.arrival = arrival
        self.t_admit = None
        self.t_done = None
        self.u = None


class FactorCache:
    """LRU map key -> (object, nbytes) with a byte budget; entries in use by
    an in-flight batch (refcount > 0) are not evicted."""

    def __init__(self, budget_bytes):
        self.budget = int(budget_bytes)
        self.items = collections.OrderedDict()
        self.nbytes = 0
        self.refs = collections.Counter()
        self.hits = 0
        self.misses = 0

    def get(self, key, build):
        if key in self.items:
            self.items.move_to_end(key)
            self.hits += 1
            return self.items[key][0]
        self.misses += 1
        obj, nb = build()
        self.items[key] = (obj, nb)
        self.nbytes += nb
        self._evict()
        return obj

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
            _, nb = self.items.pop(k)
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


class Server:
    def __init__(self, slots=4, mem_budget=8 << 30, cache_budget=16 << 30, max_batch_1d=256,
                 max_batch_2d=1024, max_batch_3d=8, device="cuda"):
        self.device = torch.device(device)
        self.streams = [torch.cuda.Stream(device=self.device) for _ in range(int(slots))]
        self.free_slots = list(range(int(slots)))
        self.mem_budget = int(mem_budget)
        self.mem_used = 0
        self.cache = FactorCache(cache_budget)
        self.max_batch_1d = int(max_batch_1d)
        self.max_batch_2d = int(max_batch_2d)
        self.max_batch_3d = int(max_batch_3d)
        self.cus = torch.cuda.get_device_properties(self.device).multi_processor_count
        self.inflight = []  # (event, slot, requests, bytes, cache keys, keepalive)
        # pending requests: arrival order (admitted entries dropped lazily from
        # the front) plus per-class FIFOs, so forming a batch costs O(batch)
        self.pending = collections.deque()
        self._q1d = {}                    # (N, bc, alpha) -> deque of 1D requests
        self._q2d = collections.deque()   # 2D requests
        self._q3d = {}                    # (shape, bc, alpha) -> deque of 3D requests
        self._npending = 0
        self.done = []
        self.batches = collections.Counter()
        self._work3d = {}  # slot -> workspace tensor (grown on demand)

    # ------------------------------------------------------------ engines
    def _cheb1d_factors(self, key):
        N, bc, alpha = key

        def build():
            V = FunctionSpace(N, "chebyshev", bc=bc)
            from .la import Helmholtz
            H = Helmholtz(stiffness_matrix(V), mass_matrix(V), -1.0, float(alpha), device=self.device)
            K2 = torch.from_numpy(np.ascontiguousarray(V.stencil_k2())).to(self.device)
            dense = None
            nb = 8 * (4 * N + N)
            if _largest_prime_factor(N) > 31:
                dense = (V._analysis_matrix(), V._synthesis_matrix())
                nb += 16 * N * N
            return (V, H, K2, dense), nb
        return self.cache.get(("cheb1d",) + key, build)

    def _leg_matrices(self, Nx, bc):
        def build():
            L = _lib.lib()
            m = torch.empty(int(L.sfa_batch2d_leg_elems(Nx)), dtype=torch.float64, device=self.device)
            _lib.call("sfa_batch2d_leg_matrices", Nx, bc, _lib.ptr(m), _lib.stream_ptr())
            return m, m.numel() * 8
        return self.cache.get(("leg", Nx, bc), build)

    def _dft_matrices(self, Ny):
        def build():
            L = _lib.lib()
            m = torch.empty(int(L.sfa_batch2d_dft_elems(Ny)), dtype=torch.float64, device=self.device)
            _lib.call("sfa_batch2d_dft_matrices", Ny, _lib.ptr(m), _lib.stream_ptr())
            return m, m.numel() * 8
        return self.cache.get(("dft", Ny), build)

    def _cheb3d_solver(self, key):
        shape, bc, alpha = key

        def build():
            from .pipeline import HelmholtzSolver3D
            from .tensorproductspace import TensorProductSpace
            T = TensorProductSpace(None, (FunctionSpace(shape[0], "chebyshev", bc=bc),
                                          FunctionSpace(shape[1], "F", dtype="D"),
                                          FunctionSpace(shape[2], "F", dtype="d")))
            H = HelmholtzSolver3D(T, float(alpha))
            return H, H.plan.numel() + H.work.numel()
        return self.cache.get(("cheb3d",) + key, build)

    def _launch_cheb1d(self, reqs, stream):
        N, bc, alpha = reqs[0].key
        V, H, K2, dense = self._cheb1d_factors(reqs[0].key)
        B = len(reqs)
        shp = _lib.arr([B, N])
        st = _lib.arr([N, 1])
        fst = _lib.arr([0, 1])
        with torch.cuda.stream(stream):
            F = torch.stack([r.f for r in reqs])
            fh = torch.empty_like(F)
            U = torch.empty_like(F)
            sp = _lib.stream_ptr(stream)
            if dense is None:
                _lib.call("sfa_cheb_scalar_product", _lib.ptr(F), _lib.ptr(fh), _lib.F64, 2, shp, st, st,
                          1, _lib.ptr(K2), sp)
            else:
                _lib.call("sfa_line_matmul", _lib.ptr(F), _lib.ptr(fh), _lib.F64, 2, shp, st, st, 1, N,
                          _lib.ptr(dense[0]), sp)
            _lib.call("sfa_helmholtz_solve", _lib.ptr(fh), _lib.ptr(fh), _lib.F64, 2, shp, st, st, 1,
                      int(H.neumann), _lib.ptr(H.u0), _lib.ptr(H.u1), _lib.ptr(H.u2), _lib.ptr(H.L), fst,
                      sp)
            if dense is None:
                _lib.call("sfa_cheb_backward", _lib.ptr(fh), _lib.ptr(U), _lib.F64, 2, shp, st, st, 1,
                          _lib.ptr(K2), sp)
            else:
                _lib.call("sfa_line_matmul", _lib.ptr(fh), _lib.ptr(U), _lib.F64, 2, shp, st, st, 1, N,
                          _lib.ptr(dense[1]), sp)
        for i, r in enumerate(reqs):
            r.u = U[i]
        return (F, fh, U), [("cheb1d",) + reqs[0].key]

    def _wide_2d(self, Nx, Ny):
        """Tile-parallel mode when one problem would dominate the persistent
        kernel's makespan: its contraction flops exceed twice the per-CU share."""
        m = Nx.astype(np.float64)
        y = Ny.astype(np.float64)
        fl = m * y * (4 * m + 2) + 2 * m * m * (y + 2)
        return bool(fl.max() * self.cus > 2.0 * fl.sum())

    def _launch_leg2d(self, reqs, stream):
        L = _lib.lib()
        n = len(reqs)
        Nx = np.array([r.key[0] for r in reqs], dtype=np.int32)
        Ny = np.array([r.key[1] for r in reqs], dtype=np.int32)
        bc = np.array([_bcc(r.key[2]) for r in reqs], dtype=np.int32)
        al = np.array([r.key[3] for r in reqs], dtype=np.float64)
        # the batch's distinct matrices, looked up once each
        lk = list(dict.fromkeys(zip(Nx.tolist(), bc.tolist())))
        dk = list(dict.fromkeys(Ny.tolist()))
        legm = [self._leg_matrices(a, b) for a, b in lk]
        dftm = [self._dft_matrices(y) for y in dk]
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
        off = np.concatenate([[0], np.cumsum(sizes)[:-1]]).astype(np.int64)
        total = int(sizes.sum())
        # the requests' f read in place: element offsets from the lowest address
        fs = [r.f if r.f.is_contiguous() and r.f.dtype == torch.float64
              else r.f.to(torch.float64).contiguous() for r in reqs]
        fp = [f.data_ptr() for f in fs]
        fbase = min(fp)
        foff = np.array([(a - fbase) // 8 for a in fp], dtype=np.int64)
        fel = int(max(a + 8 * int(z) for a, z in zip(fp, sizes)) - fbase) // 8
        nb = min(n, self.cus)
        ptr = lambda a: _lib.ctypes.c_void_p(a.ctypes.data)
        wide = self._wide_2d(Nx, Ny)
        with torch.cuda.stream(stream):
            F = _lib.ctypes.c_void_p(fbase)
            U = torch.empty(total, dtype=torch.float64, device=self.device)
            sp = _lib.stream_ptr(stream)
            if wide:
                plan = torch.empty(int(L.sfa_batch2d_wide_plan_bytes(n, ptr(Nx), ptr(Ny))),
                                   dtype=torch.uint8, device=self.device)
                work = torch.empty(int(L.sfa_batch2d_wide_workspace(n, ptr(Nx), ptr(Ny))),
                                   dtype=torch.uint8, device=self.device)
                _lib.call("sfa_batch2d_wide_plan", n, ptr(Nx), ptr(Ny), ptr(al), ptr(bc), ptr(foff),
                          ptr(off), fel, total, _lib.ctypes.c_void_p(base), ptr(phi), ptr(dft),
                          (top - base) // 8, _lib.ptr(plan), plan.numel(), sp)
                _lib.call("sfa_batch2d_wide_solve", F, _lib.ptr(U), _lib.ptr(plan),
                          _lib.ptr(work), work.numel(), sp)
            else:
                plan = torch.empty(int(L.sfa_batch2d_plan_ext_bytes(n, nb)), dtype=torch.uint8,
                                   device=self.device)
                work = torch.empty(int(L.sfa_batch2d_workspace(n, ptr(Nx), ptr(Ny), nb)),
                                   dtype=torch.uint8, device=self.device)
                _lib.call("sfa_batch2d_plan_ext", n, ptr(Nx), ptr(Ny), ptr(al), ptr(bc), ptr(foff),
                          ptr(off), fel, total, _lib.ctypes.c_void_p(base), ptr(phi), ptr(dft),
                          (top - base) // 8, nb, _lib.ptr(plan), plan.numel(), sp)
                _lib.call("sfa_batch2d_solve", F, _lib.ptr(U), _lib.ptr(plan),
                          _lib.ptr(work), work.numel(), 0, sp)
        for r, part in zip(reqs, U.split(sizes.tolist())):
            r.u = part.view(r.key[0], r.key[1])
        return (fs, U, plan, work, mats), keys

    def _launch_cheb3d(self, reqs, stream, slot):
        """One request: the fused solver on its own f.  Several requests of one
        class: their right-hand sides stacked and solved in one pass of
        launches (sfa_helmholtz3d_many); the results are bitwise those of the
        single solves."""
        H = self._cheb3d_solver(reqs[0].key)
        B = len(reqs)
        w = self._work3d.get(slot)
        if w is None or w.numel() < B * H.workspace_bytes:
            with torch.cuda.stream(stream):
                w = torch.empty(B * H.workspace_bytes, dtype=torch.uint8, device=self.device)
            self._work3d[slot] = w
        with torch.cuda.stream(stream):
            if B == 1:
                F = reqs[0].f
                U = torch.empty_like(F)
                H(F, U, work=w)
                reqs[0].u = U
            else:
                F = torch.stack([r.f for r in reqs])
                U = H.solve_many(F, work=w)
                for i, r in enumerate(reqs):
                    r.u = U[i]
        return (F, U), [("cheb3d",) + reqs[0].key]

    # ------------------------------------------------------------ admission
    @staticmethod
    def _bytes(reqs):
        fam = reqs[0].family
        n = sum(r.f.numel() for r in reqs) * 8
        if fam == "cheb1d":
            return 3 * n
        if fam == "leg2d":
            mx = max((r.key[0] * 2) * (r.key[1] + 2) for r in reqs) * 8
            return 2 * n + mx * min(len(reqs), 256)
        shp = reqs[0].key[0]
        ws = shp[0] * shp[1] * (shp[2] // 2 + 8) * 16  # workspace per solve (padded kz pitch)
        return (n if len(reqs) == 1 else 2 * n) + len(reqs) * ws

    def _push(self, r):
        self.pending.append(r)
        if r.family == "cheb1d":
            self._q1d.setdefault(r.key, collections.deque()).append(r)
        elif r.family == "leg2d":
            self._q2d.append(r)
        else:
            self._q3d.setdefault(r.key, collections.deque()).append(r)
        self._npending += 1

    def _queue(self, head):
        if head.family == "cheb1d":
            return self._q1d[head.key]
        return self._q2d if head.family == "leg2d" else self._q3d[head.key]

    def _cap(self, family):
        return {"cheb1d": self.max_batch_1d, "leg2d": self.max_batch_2d,
                "cheb3d": self.max_batch_3d}[family]

    def _next_batch(self):
        """Oldest pending request and the requests that can join its launch
        (the front of its class FIFO, oldest first)."""
        while self.pending and self.pending[0].t_admit is not None:
            self.pending.popleft()
        if not self.pending:
            return None
        head = self.pending[0]
        return list(itertools.islice(self._queue(head), self._cap(head.family)))

    def _take(self, batch, t):
        """Remove an admitted batch (from _next_batch) from the queues."""
        q = self._queue(batch[0])
        for _ in batch:
            q.popleft()
        if not q and batch[0].family == "cheb1d":
            del self._q1d[batch[0].key]
        elif not q and batch[0].family == "cheb3d":
            del self._q3d[batch[0].key]
        self._npending -= len(batch)
        for r in batch:
            r.t_admit = t

    def _poll(self, now):
        still = []
        for ev, slot, reqs, nbytes, keys, keep in self.inflight:
            if ev.query():
                t = now()
                for r in reqs:
                    r.t_done = t
                self.done.extend(reqs)
                self.free_slots.append(slot)
                self.mem_used -= nbytes
                self.cache.release(keys)
                if reqs[0].family == "leg2d":  # drop the plan's host-side registry entry
                    _lib.lib().sfa_batch2d_release(_lib.ptr(keep[2]))
            else:
                still.append((ev, slot, reqs, nbytes, keys, keep))
        self.inflight = still

    def _admit(self, now):
        while self.free_slots and self._npending:
            batch = self._next_batch()
            nbytes = self._bytes(batch)
            if self.mem_used + nbytes > self.mem_budget and self.inflight:
                return  # wait for memory
            slot = self.free_slots.pop(0)
            stream = self.streams[slot]
            self._take(batch, now())
            fam = batch[0].family
            if fam == "cheb1d":
                keep, keys = self._launch_cheb1d(batch, stream)
            elif fam == "leg2d":
                keep, keys = self._launch_leg2d(batch, stream)
            else:
                keep, keys = self._launch_cheb3d(batch, stream, slot)
            self.cache.acquire(keys)
            ev = torch.cuda.Event()
            ev.record(stream)
            self.batches[fam] += 1
            self.mem_used += nbytes
            self.inflight.append((ev, slot, batch, nbytes, keys, keep))

    # ------------------------------------------------------------ loop
    def run(self, requests, realtime=True):
        """Serve ``requests`` (sorted by arrival, seconds from the start).
        realtime: release each request at its arrival time (open loop); else
        all are pending at once.  Returns the served requests."""
        reqs = sorted(requests, key=lambda r: r.arrival)
        for r in reqs:
            r.t_admit = r.t_done = r.u = None
        self.done = []
        torch.cuda.synchronize(self.device)
        t0 = time.perf_counter()
        now = lambda: time.perf_counter() - t0
        i = 0
        n = len(reqs)
        while len(self.done) < n:
            t = now() if realtime else float("inf")
            while i < n and reqs[i].arrival <= t:
                self._push(reqs[i])
                i += 1
            self._poll(now)
            self._admit(now)
            if not self.inflight and not self._npending and i < n and realtime:
                dt = reqs[i].arrival - now()
                if dt > 2e-4:
                    time.sleep(dt - 1e-4)
        torch.cuda.synchronize(self.device)
        return reqs

    @staticmethod
    def stats(reqs):
        lat = np.array([r.t_done - r.arrival for r in reqs]) * 1e3
        span = max(r.t_done for r in reqs) - min(r.arrival for r in reqs)
        return {"requests": len(reqs), "p50_ms": float(np.percentile(lat, 50)),
                "p99_ms": float(np.percentile(lat, 99)), "mean_ms": float(lat.mean()),
                "solves_per_s": len(reqs) / span if span > 0 else float("inf")}


# ---------------------------------------------------------------- workload
def cheb1d_rhs(N, rng):
    """Smooth f on the N Chebyshev-Gauss nodes: sum_k c_k T_k(x_j),
    c_k ~ N(0,1)/(1+k^2)."""
    k = np.arange(N)
    c = rng.standard_normal(N) / (1 + k * k)
    x = np.cos(np.pi * (2 * np.arange(N) + 1) / (2 * N))
    return np.polynomial.chebyshev.chebval(x, c)


def request_stream(n, rate, seed=0, classes=(32, 64, 16), device="cuda"):
    """n requests with Poisson arrivals (rate per second) from the serving mix:
    40% 1D Chebyshev N in [64, 4096], 40% 2D Legendre x Fourier up to 512 x 256,
    20% 3D Chebyshev x Fourier^2 up to 256^3 (each axis in {32, 64, 128, 256}),
    Dirichlet/Neumann 50/50, alpha ~ U(0, 50).  Requests are drawn from a fixed
    number of classes per family (classes = sizes, bc, alpha), each with one
    seeded smooth right-hand side resident on the device."""
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
    f1 = [torch.from_numpy(cheb1d_rhs(k[0], rng)).to(device) for k in k1]
    f2 = [torch.from_numpy(batch2d_rhs(k[0], k[1], rng)).to(device) for k in k2]
    f3 = [torch.from_numpy(helmholtz3d_rhs(*k[0], seed=int(rng.integers(1 << 30)))).to(device) for k in k3]
    fam = rng.choice(3, size=n, p=[0.4, 0.4, 0.2])
    cls = rng.integers(0, 1 << 30, n)
    arrivals = np.cumsum(rng.exponential(1.0 / rate, n))
    out = []
    for i in range(n):
        if fam[i] == 0:
            j = cls[i] % c1
            out.append(Request(i, "cheb1d", k1[j], f1[j], arrivals[i]))
        elif fam[i] == 1:
            j = cls[i] % c2
            out.append(Request(i, "leg2d", k2[j], f2[j], arrivals[i]))
        else:
            j = cls[i] % c3
            out.append(Request(i, "cheb3d", k3[j], f3[j], arrivals[i]))
    return out
