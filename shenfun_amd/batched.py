"""Batched 2D Legendre x Fourier Helmholtz solves in one persistent launch.

``BatchedHelmholtz2D`` solves nprob independent problems
(alpha_i - lap) u_i = f_i on LegendreSD/SN(Nx_i) x R2C(Ny_i): per problem the
reference sequence TensorProductSpace.scalar_product -> SolverGeneric1ND ->
TensorProductSpace.backward (tensorproductspace.py:146-190, la.py:1390,
tensorproductspace.py:46-96), all problems in one launch of
sfa_batch2d_solve.  Inputs and outputs are packed row-major in one float64
tensor each (``pack``/``unpack``).  ``autotune`` times the kernel's block-tile
variants and workgroup counts on the actual batch and keeps the fastest.

``mode="wide"`` runs the tile-parallel form (sfa_batch2d_wide_solve: each
contraction one launch over all (problem, tile) tasks, consecutive tasks
sharing a problem's operands in the L2); its outputs are bitwise equal to the
persistent form's, and ``autotune`` times it beside the persistent variants.
"""
import json
import os
import time

import numpy as np
import torch

from . import _lib


def _bc_code(b):
    if b in ("D", 0, "Dirichlet"):
        return 0
    if b in ("N", 1, "Neumann"):
        return 1
    raise ValueError(f"bc must be 'D' or 'N', got {b!r}")


def tile_variants():
    """[(BM, BN, BK)] block tiles and K steps of the batched kernel, indexed by variant id."""
    L = _lib.lib()
    n = L.sfa_batch2d_variants(None, None, 0)
    bm = (_lib.ctypes.c_int * n)()
    bn = (_lib.ctypes.c_int * n)()
    L.sfa_batch2d_variants(bm, bn, n)
    return [(bm[i], bn[i] // 1000, bn[i] % 1000) for i in range(n)]


class BatchedHelmholtz2D:
    def __init__(self, Nx, Ny, alpha, bc="D", nblocks=None, variant=0, device="cuda",
                 mode="persistent"):
        if mode not in ("persistent", "wide"):
            raise ValueError(f"mode must be 'persistent' or 'wide', got {mode!r}")
        self.mode = mode
        self.Nx = np.ascontiguousarray(np.asarray(Nx, dtype=np.int32).reshape(-1))
        self.Ny = np.ascontiguousarray(np.asarray(Ny, dtype=np.int32).reshape(-1))
        n = self.Nx.size
        if self.Ny.size != n:
            raise ValueError("Nx and Ny must have the same length")
        self.alpha = np.ascontiguousarray(np.broadcast_to(np.asarray(alpha, dtype=np.float64), (n,)))
        if isinstance(bc, (str, int)):
            bc = [bc] * n
        self.bc = np.ascontiguousarray(np.array([_bc_code(b) for b in bc], dtype=np.int32))
        if self.bc.size != n:
            raise ValueError("bc must be one value or one per problem")
        sizes = self.Nx.astype(np.int64) * self.Ny.astype(np.int64)
        self.offsets = np.ascontiguousarray(np.concatenate([[0], np.cumsum(sizes)[:-1]]).astype(np.int64))
        self.sizes = sizes
        self.total = int(sizes.sum())
        self.nprob = n
        self.device = torch.device(device)
        self.cus = torch.cuda.get_device_properties(self.device).multi_processor_count
        self.variant = int(variant)
        self.plan = None
        if mode == "wide":
            self._build_wide()
        else:
            self.build(nblocks if nblocks is not None else self.cus)

    # ------------------------------------------------------------ plan
    def build(self, nblocks):
        """(Re)build the device plan for a workgroup count (schedule + matrices)."""
        if self.mode != "persistent":
            raise ValueError("build(nblocks) applies to the persistent mode")
        L = _lib.lib()
        ptr = lambda a: _lib.ctypes.c_void_p(a.ctypes.data)
        nb = int(nblocks)
        pbytes = L.sfa_batch2d_plan_bytes(self.nprob, ptr(self.Nx), ptr(self.Ny), ptr(self.bc), nb)
        if pbytes < 0:
            raise ValueError("invalid problem sizes (even Nx in [8, 1024], even Ny in [4, 1024])")
        if self.plan is not None:
            L.sfa_batch2d_release(_lib.ptr(self.plan))
        self.plan = torch.empty(int(pbytes), dtype=torch.uint8, device=self.device)
        _lib.call("sfa_batch2d_plan", self.nprob, ptr(self.Nx), ptr(self.Ny), ptr(self.alpha),
                  ptr(self.bc), ptr(self.offsets), ptr(self.offsets), self.total, self.total, nb,
                  _lib.ptr(self.plan), self.plan.numel(), _lib.stream_ptr())
        wbytes = L.sfa_batch2d_workspace(self.nprob, ptr(self.Nx), ptr(self.Ny), nb)
        self.work = torch.empty(int(wbytes), dtype=torch.uint8, device=self.device)
        self.nblocks = nb

    def _build_wide(self):
        """Matrix arena (one Phi/P block per distinct (Nx, bc), one Df/Di block
        per distinct Ny) and the tile-parallel plan."""
        L = _lib.lib()
        if self.plan is not None:
            L.sfa_batch2d_release(_lib.ptr(self.plan))
            self.plan = None
        ptr = lambda a: _lib.ctypes.c_void_p(a.ctypes.data)
        if L.sfa_batch2d_plan_bytes(self.nprob, ptr(self.Nx), ptr(self.Ny), ptr(self.bc), 1) < 0:
            raise ValueError("invalid problem sizes (even Nx in [8, 1024], even Ny in [4, 1024])")
        leg, dft, o = {}, {}, 0
        for nx, b in sorted(set(zip(self.Nx.tolist(), self.bc.tolist()))):
            leg[(nx, b)] = o
            o += -(-int(L.sfa_batch2d_leg_elems(nx)) // 32) * 32
        for ny in sorted(set(self.Ny.tolist())):
            dft[ny] = o
            o += -(-int(L.sfa_batch2d_dft_elems(ny)) // 32) * 32
        self.mats = torch.empty(o, dtype=torch.float64, device=self.device)
        sp = _lib.stream_ptr()
        for (nx, b), off in leg.items():
            _lib.call("sfa_batch2d_leg_matrices", nx, b, _lib.ptr(self.mats[off:]), sp)
        for ny, off in dft.items():
            _lib.call("sfa_batch2d_dft_matrices", ny, _lib.ptr(self.mats[off:]), sp)
        phi = np.array([leg[(int(a), int(b))] for a, b in zip(self.Nx, self.bc)], dtype=np.int64)
        df = np.array([dft[int(b)] for b in self.Ny], dtype=np.int64)
        pbytes = L.sfa_batch2d_wide_plan_bytes(self.nprob, ptr(self.Nx), ptr(self.Ny))
        self.plan = torch.empty(int(pbytes), dtype=torch.uint8, device=self.device)
        _lib.call("sfa_batch2d_wide_plan", self.nprob, ptr(self.Nx), ptr(self.Ny), ptr(self.alpha),
                  ptr(self.bc), ptr(self.offsets), ptr(self.offsets), self.total, self.total,
                  _lib.ptr(self.mats), ptr(phi), ptr(df), o, _lib.ptr(self.plan), self.plan.numel(), sp)
        wbytes = L.sfa_batch2d_wide_workspace(self.nprob, ptr(self.Nx), ptr(self.Ny))
        self.work = torch.empty(int(wbytes), dtype=torch.uint8, device=self.device)
        self.nblocks = None

    def __del__(self):
        try:
            if self.plan is not None:
                _lib.lib().sfa_batch2d_release(_lib.ptr(self.plan))
        except Exception:
            pass

    # ------------------------------------------------------------ data
    def pack(self, arrays):
        """One float64 tensor holding the problems' (Nx_i, Ny_i) arrays back to back."""
        if len(arrays) != self.nprob:
            raise ValueError(f"expected {self.nprob} arrays")
        parts = []
        for i, a in enumerate(arrays):
            t = torch.as_tensor(a, dtype=torch.float64)
            if tuple(t.shape) != (int(self.Nx[i]), int(self.Ny[i])):
                raise ValueError(f"problem {i}: shape {tuple(t.shape)} != ({self.Nx[i]}, {self.Ny[i]})")
            parts.append(t.reshape(-1))
        return torch.cat(parts).to(self.device)

    def unpack(self, U):
        """Views of the problems' (Nx_i, Ny_i) arrays inside a packed tensor."""
        return [U[int(o):int(o) + int(s)].view(int(nx), int(ny))
                for o, s, nx, ny in zip(self.offsets, self.sizes, self.Nx, self.Ny)]

    # ------------------------------------------------------------ solve
    def __call__(self, F, U=None):
        if not (F.is_cuda and F.dtype == torch.float64 and F.is_contiguous() and F.numel() == self.total):
            raise ValueError(f"F must be a contiguous float64 CUDA tensor of {self.total} elements")
        if U is None:
            U = torch.empty_like(F)
        if not (U.is_contiguous() and U.dtype == torch.float64 and U.numel() == self.total):
            raise ValueError("U must match F")
        if U.data_ptr() == F.data_ptr():
            raise ValueError("U must not alias F")
        if self.mode == "wide":
            _lib.call("sfa_batch2d_wide_solve", _lib.ptr(F), _lib.ptr(U), _lib.ptr(self.plan),
                      _lib.ptr(self.work), self.work.numel(), _lib.stream_ptr())
        else:
            _lib.call("sfa_batch2d_solve", _lib.ptr(F), _lib.ptr(U), _lib.ptr(self.plan),
                      _lib.ptr(self.work), self.work.numel(), self.variant, _lib.stream_ptr())
        return U

    # ------------------------------------------------------------ tuning
    def _use(self, cfg):
        """Switch to a tuned configuration: {"mode", "nblocks", "variant"}."""
        if cfg.get("mode", "persistent") == "wide":
            if self.mode != "wide" or self.plan is None:
                self.mode = "wide"
                self._build_wide()
        else:
            if self.mode != "persistent" or cfg["nblocks"] != self.nblocks:
                self.mode = "persistent"
                self.build(cfg["nblocks"])
            self.variant = cfg["variant"]

    def autotune(self, F, blocks_per_cu=(1, 2), variants=None, reps=3, cache=None, wide=True):
        """Time every (tile variant, workgroups per CU) pair of the persistent
        form, and the tile-parallel form (``wide``), on this batch and keep the
        fastest; both forms give the same bits.  Returns {"best": {...},
        "table": [...]}.  With ``cache`` (a JSON path) a previous result for the
        same batch signature is reused and new results are stored."""
        key = self.signature()
        if cache and os.path.exists(cache):
            try:
                with open(cache) as fh:
                    hit = json.load(fh).get(key)
            except (OSError, ValueError):
                hit = None
            if hit:
                self._use(hit)
                return {"best": hit, "table": [], "cached": True}
        U = torch.empty_like(F)

        def timed():
            self(F, U)
            torch.cuda.synchronize()
            best = float("inf")
            for _ in range(reps):
                t0 = time.perf_counter()
                self(F, U)
                torch.cuda.synchronize()
                best = min(best, time.perf_counter() - t0)
            return best * 1e3

        tiles = tile_variants()
        variants = range(len(tiles)) if variants is None else variants
        table = []
        self.mode = "persistent"
        for bpc in blocks_per_cu:
            self.build(bpc * self.cus)
            for v in variants:
                self.variant = v
                table.append({"mode": "persistent", "variant": v, "tile": list(tiles[v]),
                              "nblocks": self.nblocks, "ms": timed(),
                              "label": "x".join(map(str, tiles[v])) + f"@{self.nblocks}"})
        if wide:
            self.mode = "wide"
            self._build_wide()
            table.append({"mode": "wide", "variant": None, "tile": None, "nblocks": None,
                          "ms": timed(), "label": "wide"})
        win = min(table, key=lambda r: r["ms"])
        self._use(win)
        if cache:
            data = {}
            if os.path.exists(cache):
                try:
                    with open(cache) as fh:
                        data = json.load(fh)
                except (OSError, ValueError):
                    data = {}
            data[key] = win
            with open(cache, "w") as fh:
                json.dump(data, fh, indent=1)
        return {"best": win, "table": table, "cached": False}

    def signature(self):
        """Batch signature for the tuning cache: device, problem count, size histogram."""
        h = np.histogram(np.log2(self.sizes.astype(float)), bins=16, range=(4, 20))[0]
        name = torch.cuda.get_device_properties(self.device).name
        return f"{name}|{self.nprob}|{int(self.total)}|" + ",".join(map(str, h))

    def algorithmic_flops(self):
        """Algorithmic FP64 flops per batch: the Legendre analysis and
        synthesis contractions over the Ny + 2 real columns of the R2C
        spectrum, 2 (Nx - 2) Nx (Ny + 2) each (the recursive transforms of
        transforms.pyx:219-323 do this work), plus a real FFT pair along y,
        2.5 Ny log2(Ny) per line and direction.  The kernel's dense DFT
        contractions along y do more (``flops``)."""
        nx = self.Nx.astype(float)
        ny = self.Ny.astype(float)
        leg = 4.0 * (nx - 2) * nx * (ny + 2)
        fft = 2 * 2.5 * nx * ny * np.log2(ny)
        return float(np.sum(leg + fft))

    def flops(self):
        """FP64 flops the kernel executes per batch: the four contractions
        (two Legendre, two dense real DFT), 2 per multiply-add."""
        nx = self.Nx.astype(float)
        ny = self.Ny.astype(float)
        m = nx - 2
        L = ny + 2
        return float(np.sum(2 * m * nx * ny + 2 * m * ny * L + 2 * nx * m * L + 2 * nx * L * ny))

    def bytes_moved(self):
        """Compulsory HBM bytes per batch: f read and u written once."""
        return float(16 * self.total)
