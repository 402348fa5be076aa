"""Banded solvers of the hot path on the GPU (shenfun/la.py,
shenfun/chebyshev/la.py interfaces; compute via the C ABI).

* ``DiagMA``, ``TwoDMA``, ``TDMA``, ``TDMA_O``, ``ThreeDMA``, ``FDMA``, ``PDMA``,
  ``HeptaDMA``: ``__call__(b, u=None, axis=0, constraints=())``
  (la.py:216-249, 291-820); LU on the device once, solve along any axis in
  place on ``u``.
* ``Solver(mats)``: band pattern -> solver class (la.py:17-40).
* ``SolverGeneric1ND(mats)``: one non-periodic axis, per-mode factors
  (la.py:1183-1434), mode 0 solved by its own solver (with constraints).
* ``Helmholtz``: Chebyshev alfa*A + beta*B solver (chebyshev/la.py:131-351),
  ``__call__(b, u=None, constraints=())`` and ``matvec(v, c)``.
* ``ADDSolver`` / ``ANNSolver``: Chebyshev Poisson (chebyshev/la.py:11-128).
"""
import numpy as np
import torch

from . import _lib
from .matrices import SparseMatrix, TPMatrix, _split_bc, apply_bc_matrices
from .optimization import optimizer, runtimeoptimizer


def _no_host(*args, **kwargs):
    raise NotImplementedError("no host implementation: config['optimization']['mode'] must name a "
                              "backend ('hip')")


def _geom(t):
    return _lib.arr(t.shape), _lib.arr(t.stride())


def _dtype_code(t):
    if t.dtype == torch.complex128:
        return _lib.C128
    if t.dtype == torch.float64:
        return _lib.F64
    raise TypeError(f"unsupported dtype {t.dtype} (float64 / complex128 only)")


def _check(t, name):
    if not isinstance(t, torch.Tensor) or not t.is_cuda:
        raise TypeError(f"{name} must be a CUDA (HIP) torch tensor")
    if t.dim() < 1 or t.dim() > 3:
        raise ValueError(f"{name} must have 1..3 dims")


def _prepare(b, u):
    _check(b, "b")
    if u is None:
        return b
    _check(u, "u")
    if u.shape != b.shape or u.dtype != b.dtype:
        raise ValueError("u and b must have the same shape and dtype")
    if u.data_ptr() != b.data_ptr():
        u.copy_(b)
    return u


def _sum_mats(mats):
    if isinstance(mats, (list, tuple)):
        m = mats[0]
        for x in mats[1:]:
            m = m + x
        return m
    return mats


class _Banded:
    """Common part of the DIA-data banded solvers (BandedMatrixSolver, la.py:251-288)."""

    offsets = ()
    lu_fn = None
    solve_fn = None

    def __init__(self, mats):
        mat = _sum_mats(mats)
        assert isinstance(mat, SparseMatrix)
        assert mat.shape[0] == mat.shape[1]
        self.mat = mat
        n = mat.shape[0]
        data = np.zeros((len(self.offsets), n))
        for r, off in enumerate(self.offsets):
            if off in mat:
                v = mat.diagonal_array(off) * float(mat.scale)
                if off >= 0:
                    data[r, off:] = v
                else:
                    data[r, :n + off] = v
        self._data = data  # host DIA rows (offsets sorted), scipy dia layout
        self._lu = None
        self._device = None
        self._pinned = False

    def _dev(self, device):
        if self._lu is None or self._device != device:
            d = torch.from_numpy(self._data.copy()).to(device)
            self._call_lu(d, d.shape[1], 1)
            self._lu = d
            self._device = device
        return self._lu

    def _call_lu(self, d, n, nbatch):
        if self.lu_fn is not None:
            if self._dispatch and nbatch == 1:
                type(self).LU(d)
            else:
                _lib.call(self.lu_fn, _lib.ptr(d), n, nbatch, _lib.stream_ptr())

    def _call_solve(self, u, axis, data, n, dls, skip):
        shp, st = _geom(u)
        _lib.call(self.solve_fn, _lib.ptr(u), _dtype_code(u), u.dim(), shp, st, axis, _lib.ptr(data),
                  n, dls, skip, _lib.stream_ptr())

    def perform_lu(self, device="cuda"):
        return self._dev(torch.device(device))

    def apply_constraints(self, b, constraints, axis=0):
        if len(constraints) > 0:
            assert len(constraints) == 1
            assert constraints[0][0] == 0, "Can only fix first row"
            if not self._pinned:  # once: the factors then keep the unit row
                self._pin_row0()
                self._pinned = True
            s = [slice(None)] * b.dim()
            s[axis] = 0
            b[tuple(s)] = constraints[0][1]
        return b

    def _pin_row0(self):
        raise NotImplementedError

    # solves go through the static Solve resolved by the backend dispatch
    _dispatch = True

    def __call__(self, b, u=None, axis=0, constraints=()):
        u = _prepare(b, u)
        self.apply_constraints(u, constraints, axis)
        data = self._dev(u.device)
        if self._dispatch:
            type(self).Solve(u, data, axis)
        else:
            self._call_solve(u, axis, data, data.shape[1], 0, 0)
        return u

    solve = __call__


class DiagMA(_Banded):
    offsets = (0,)
    lu_fn = None
    solve_fn = "sfa_diagma_solve"

    def _pin_row0(self):
        self._data[0, 0] = 1.0
        self._lu = None


class TDMA(_Banded):
    offsets = (-2, 0, 2)
    lu_fn = "sfa_tdma_lu"
    solve_fn = "sfa_tdma_solve"

    def _pin_row0(self):
        # la.py:355-364
        self._data[1, 0] = 1.0
        self._data[2, 2] = 0.0
        self._lu = None


class PDMA(_Banded):
    offsets = (-4, -2, 0, 2, 4)
    lu_fn = "sfa_pdma_lu"
    solve_fn = "sfa_pdma_solve"

    def _pin_row0(self):
        # la.py:459-474
        self._data[2, 0] = 1.0
        self._data[3, 2] = 0.0
        self._data[4, 4] = 0.0
        self._lu = None


class TDMA_O(_Banded):
    """Tridiagonal solver, offsets -1, 0, 1 (la.py:396-447): one chain per line."""
    offsets = (-1, 0, 1)
    lu_fn = "sfa_tdma_o_lu"
    solve_fn = "sfa_tdma_o_solve"

    def _pin_row0(self):
        # row 0 replaced by the unit row (SparseMatrixSolver.apply_constraints, la.py:77-120)
        self._data[1, 0] = 1.0
        self._data[2, 1] = 0.0
        self._lu = None


class TwoDMA(_Banded):
    """Upper 2-diagonal solver, offsets 0, 2 (la.py:615-659); no LU."""
    offsets = (0, 2)
    lu_fn = None
    solve_fn = "sfa_twodma_solve"

    def _pin_row0(self):
        # la.py:628-637
        self._data[0, 0] = 1.0
        self._data[1, 2] = 0.0
        self._lu = None


class ThreeDMA(_Banded):
    """Upper 3-diagonal solver, offsets 0, 2, 4 (la.py:661-709); no LU.  The +4
    band enters the back substitution from its own DIA row.  The reference
    reads the +2 row there (la.py:688, la.pyx:574), which solves the matrix
    only when the two bands coincide; compat=True reproduces that reading, so
    that results equal shenfun's for any matrix."""
    offsets = (0, 2, 4)
    lu_fn = None
    solve_fn = "sfa_threedma_solve"
    _dispatch = False

    def __init__(self, mats, compat=False):
        super().__init__(mats)
        self.compat = bool(compat)

    def _dev(self, device):
        d = super()._dev(device)
        if self.compat:
            d = d.clone()
            d[2, 4:] = d[1, 4:]
        return d

    def _pin_row0(self):
        # la.py:674-684
        self._data[0, 0] = 1.0
        self._data[1, 2] = 0.0
        self._data[2, 4] = 0.0
        self._lu = None


class FDMA(_Banded):
    """4-diagonal solver, offsets -2, 0, 2, 4 (la.py:545-613)."""
    offsets = (-2, 0, 2, 4)
    lu_fn = "sfa_fdma_lu"
    solve_fn = "sfa_fdma_solve"

    def _pin_row0(self):
        # la.py:575-586
        self._data[1, 0] = 1.0
        self._data[2, 2] = 0.0
        self._data[3, 4] = 0.0
        self._lu = None


class HeptaDMA(_Banded):
    """Heptadiagonal solver, offsets -4, -2, 0, 2, 4, 6, 8 (la.py:711-820)."""
    offsets = (-4, -2, 0, 2, 4, 6, 8)
    lu_fn = "sfa_heptadma_lu"
    solve_fn = "sfa_heptadma_solve"

    def _pin_row0(self):
        # la.py:724-741
        for r, off in ((2, 0), (3, 2), (4, 4), (5, 6), (6, 8)):
            self._data[r, off] = 1.0 if off == 0 else 0.0
        self._lu = None


# shenfun's static LU/Solve methods of the banded solvers (la.py:291-820),
# resolved by name through the configured backend (optimization/hip.py:
# TDMA_LU, TDMA_Solve, ...) like the reference's @optimizer methods
for _cls, _lu_name in ((DiagMA, None), (TDMA, "LU"), (PDMA, "LU"), (TDMA_O, "LU"), (TwoDMA, None),
                       (ThreeDMA, None), (FDMA, "LU"), (HeptaDMA, "LU")):
    def _mk(cls, name):
        f = lambda *a, **k: _no_host(*a, **k)  # noqa: E731
        f.__name__ = name
        f.__qualname__ = f"{cls.__name__}.{name}"
        return staticmethod(optimizer(f))
    _cls.Solve = _mk(_cls, "Solve")
    _cls.inner_solve = _mk(_cls, "inner_solve")
    if _lu_name:
        _cls.LU = _mk(_cls, "LU")
del _cls, _lu_name, _mk


# band patterns of SparseMatrix.get_solver (matrixbase.py:439-474), narrowest
# first: SolverGeneric1ND takes the first whose offsets cover every mode's keys
_PATTERNS = (DiagMA, TwoDMA, TDMA_O, TDMA, ThreeDMA, FDMA, PDMA, HeptaDMA)


class Solve(_Banded):
    """The reference's general fallback (SuperLU, la.py:823-844) for the band
    patterns without a dedicated solver, e.g. the -2..2 mass matrices of the
    mixed Dirichlet/Neumann bases: banded LU without pivoting on the device
    (sfa_band_lu / sfa_band_solve), offsets -nl..nu within -4..4.  The
    Galerkin mass and Helmholtz matrices are SPD, so no pivoting is needed; a
    zero or non-finite pivot raises."""

    _dispatch = False

    def __init__(self, mats):
        mat = _sum_mats(mats)
        if float(mat.scale) == 0:
            raise ValueError("Solve: the matrix is zero")
        self.nl, self.nu = _band_extent(mat.keys())
        self.offsets = tuple(range(-self.nl, self.nu + 1))
        super().__init__(mat)

    def _call_lu(self, d, n, nbatch):
        _band_lu(d, n, self.nl, self.nu, nbatch)

    def _pin_row0(self):
        # row 0 := the unit row (SparseMatrixSolver.apply_constraints, la.py:216-249)
        n = self._data.shape[1]
        self._data[self.nl, 0] = 1.0
        for q in range(1, self.nu + 1):
            if q < n:
                self._data[self.nl + q, q] = 0.0
        self._lu = None

    def _call_solve(self, u, axis, data, n, dls, skip):
        shp, st = _geom(u)
        _lib.call("sfa_band_solve", _lib.ptr(u), _dtype_code(u), u.dim(), shp, st, axis,
                  _lib.ptr(data), n, dls, skip, self.nl, self.nu, _lib.stream_ptr())


def _band_extent(keys):
    keys = [int(k) for k in keys]
    nl, nu = max(0, -min(keys)), max(1, max(keys))
    if nl > 4 or nu > 4:
        raise NotImplementedError(f"band offsets {sorted(keys)} outside -4..4")
    return nl, nu


def _band_lu(d, n, nl, nu, nbatch):
    """In-place banded LU of nbatch DIA blocks; checks the pivots."""
    _lib.call("sfa_band_lu", _lib.ptr(d), n, nl, nu, nbatch, _lib.stream_ptr())
    piv = d.view(nbatch, nl + 1 + nu, n)[:, nl, :]
    if not bool(torch.isfinite(d).all()) or bool((piv == 0).any()):
        raise ZeroDivisionError("banded LU without pivoting met a zero pivot")


def Solver(mats):
    """Return the solver matching the band pattern of ``mats`` (la.py:17-40)."""
    mat = _sum_mats(mats)
    return mat.get_solver()(mat)


class SolverGeneric1ND:
    """Per-mode banded solves along the single non-periodic axis (la.py:1183-1434)."""

    def __init__(self, mats):
        assert isinstance(mats, (list, tuple)) and all(isinstance(m, TPMatrix) for m in mats)
        mats, self.bc_mats = _split_bc(list(mats))
        self.mats = list(mats)
        self.naxes = mats[0].naxes[0]
        self.space = mats[0].space
        shape = list(self.space.shape(True))
        ax = self.naxes
        n = mats[0].pmat.shape[0]
        self.n = n
        mshape = shape[:ax] + shape[ax + 1:]
        self.nmodes = int(np.prod(mshape)) if mshape else 1
        keys = set()
        for m in mats:
            keys |= set(m.pmat.keys())
        keys = tuple(sorted(keys))
        self._band = None
        for cls in _PATTERNS:
            if set(keys) <= set(cls.offsets):
                break
        else:
            # no dedicated pattern: the general banded fallback (Solve)
            cls = Solve
            self._band = _band_extent(keys)
        self.kind = cls.__name__.lower().replace("diagma", "diag")
        self.offsets = cls.offsets if self._band is None else \
            tuple(range(-self._band[0], self._band[1] + 1))
        self._lu_fn, self._solve_fn = cls.lu_fn, cls.solve_fn
        # per-mode DIA data (nmodes, rows, n), C order over the non-axis dims
        data = np.zeros((self.nmodes, len(self.offsets), n))
        for m in mats:
            sc = np.broadcast_to(m.scale, [s if i != ax else 1 for i, s in enumerate(shape)])
            sc = np.ascontiguousarray(np.moveaxis(sc, ax, -1)).reshape(-1) * float(m.pmat.scale)
            for r, off in enumerate(self.offsets):
                if off not in m.pmat:
                    continue
                v = m.pmat.diagonal_array(off)
                if off >= 0:
                    data[:, r, off:] += sc[:, None] * v[None, :]
                else:
                    data[:, r, :n + off] += sc[:, None] * v[None, :]
        self._data = data
        self._dev = None
        # mode 0 (Fourier wavenumber 0 / (0, 0)) has its own solver
        mat0 = SparseMatrix({off: data[0, r, max(off, 0):n + min(off, 0)]
                             for r, off in enumerate(self.offsets)}, (n, n))
        self.sol0 = cls(mat0)

    def matvec(self, u, c):
        """c = sum of the matrices times u (la.py:1212-1229), plus the boundary
        matrices times u's boundary slots when there are any."""
        for j, m in enumerate(self.mats):
            m.matvec(u, c, accumulate=j > 0)
        if self.bc_mats:
            apply_bc_matrices(self.bc_mats, c, u, sign=+1.0)
        return c

    @staticmethod
    @runtimeoptimizer
    def solve_data(u, data, sol, naxes, is_rank_zero):
        """Per-mode solves of all modes but (0[, 0]) (la.py:1321-1344),
        resolved to the backend's SolverGeneric1ND_solve_data."""
        _no_host()

    def perform_lu(self, device):
        if self._dev is None or self._dev.device != device:
            d = torch.from_numpy(self._data).to(device)
            if self._band is not None:
                _band_lu(d, self.n, self._band[0], self._band[1], self.nmodes)
            elif self._lu_fn is not None:
                _lib.call(self._lu_fn, _lib.ptr(d), self.n, self.nmodes, _lib.stream_ptr())
            self._dev = d
        return self._dev

    def __call__(self, b, u=None, constraints=(), fast=True):
        if self.bc_mats:
            # la.py:1212-1229: u's boundary slots set, b -= bc matrices x slots
            apply_bc_matrices(self.bc_mats, b, b if u is None else u)
        ax = self.naxes
        if u is not None and u.shape != b.shape:
            # a Petrov-Galerkin right-hand side lives on the test space (longer
            # along the axis): its first n entries are the equations
            # (BandedMatrixSolver.solve, la.py:256-260: u[sl] = b[sl])
            _check(b, "b")
            _check(u, "u")
            if b.dim() != u.dim() or any(b.shape[i] != u.shape[i] for i in range(b.dim()) if i != ax) \
                    or b.shape[ax] < self.n or b.dtype != u.dtype:
                raise ValueError("b and u must agree off the solve axis")
            u.narrow(ax, 0, self.n).copy_(b.narrow(ax, 0, self.n))
            b = u
        u = _prepare(b, u)
        data = self.perform_lu(u.device)
        # mode 0 line: index 0 along every non-axis dim
        s = [0] * u.dim()
        s[ax] = slice(None)
        line0 = u[tuple(s)]
        self.sol0(line0.unsqueeze(0) if line0.dim() == 0 else line0, axis=0,
                  constraints=constraints)
        shp, st = _geom(u)
        if self._band is not None:
            _lib.call("sfa_band_solve", _lib.ptr(u), _dtype_code(u), u.dim(), shp, st, ax,
                      _lib.ptr(data), self.n, data.shape[1] * data.shape[2], 1, self._band[0],
                      self._band[1], _lib.stream_ptr())
        elif self.kind == "threedma":
            _lib.call(self._solve_fn, _lib.ptr(u), _dtype_code(u), u.dim(), shp, st, ax,
                      _lib.ptr(data), self.n, data.shape[1] * data.shape[2], 1, _lib.stream_ptr())
        else:
            # all modes but the first, one launch (la.py:1346-1364 -> la.pyx:371)
            name = {"diag": "DiagMA"}.get(self.kind, self.sol0.__class__.__name__) + "_inner_solve"
            self.solve_data(u, data, name, ax, True)
        if self.bc_mats:
            self.bc_mats[0].pmat.space.set_boundary_dofs(u, final=True)
        return u


def _helm_diags(A, B):
    """Unscaled diagonals (A0, A2, A4, Bm2, B0, B2) of the Chebyshev SD/SN
    Galerkin matrices, as LU_Helmholtz_1D reads them (la.pyx:717-733)."""
    M = A.shape[0]

    def full(mat, k, n):
        if n <= 0:
            return np.zeros(1)
        if k not in mat:
            return np.zeros(n)
        return mat.diagonal_array(k)[:n]

    return [full(A, 0, M), full(A, 2, M - 2), full(A, 4, M - 4), full(B, -2, M - 2),
            full(B, 0, M), full(B, 2, M - 2)]


class Helmholtz:
    """Chebyshev Helmholtz solver alfa*A + beta*B (chebyshev/la.py:131-351).

    ``Helmholtz(A, B)`` with the scales carried by the (TP)matrices, or
    ``Helmholtz(A, B, alfa, beta)``; A is the stiffness (ASDSDmat/ASNSNmat),
    B the mass matrix (BSDSDmat/BSNSNmat).
    """

    def __init__(self, *args, device="cuda"):
        if len(args) == 1:
            args = args[0]
        args, self.bc_mats = _split_bc(list(args))
        assert len(args) in (2, 4)
        Am, Bm = args[:2]
        byk = {m.get_key(): m for m in (Am, Bm)}
        self.A = byk.get("ASDSDmat", byk.get("ASNSNmat"))
        self.B = byk.get("BSDSDmat", byk.get("BSNSNmat"))
        assert self.A is not None and self.B is not None
        if len(args) == 2:
            alfa, beta = self.A.scale, self.B.scale
        else:
            alfa, beta = args[2], args[3]
        Apm = self.A.pmat if isinstance(self.A, TPMatrix) else self.A
        Bpm = self.B.pmat if isinstance(self.B, TPMatrix) else self.B
        if isinstance(self.A, TPMatrix):
            alfa = np.asarray(alfa) * float(Apm.scale)
            beta = np.asarray(beta) * float(Bpm.scale)
        space = Apm.testfunction[0]
        self.neumann = space.boundary_condition() == "Neumann"
        self.axis = self.A.axis if isinstance(self.A, TPMatrix) else space.axis
        T = self.A.tensorproductspace if isinstance(self.A, TPMatrix) else space.tensorproductspace
        self.T = T
        N = Apm.shape[0] + 2
        self.N = N
        if T is not None:
            shape = list(T.shape(True))
        else:
            shape = [N]
        ab_shape = list(shape)
        ab_shape[self.axis] = 1
        alfa = np.broadcast_to(np.atleast_1d(np.asarray(alfa, float)), ab_shape)
        beta = np.broadcast_to(np.atleast_1d(np.asarray(beta, float)), ab_shape)
        self.alfa = np.ascontiguousarray(alfa)
        self.beta = np.ascontiguousarray(beta)
        dev = torch.device(device)
        ax = self.axis
        self._ab = [torch.from_numpy(np.array(np.moveaxis(x, ax, -1), dtype=float).reshape(-1)).to(dev)
                    for x in (self.alfa, self.beta)]
        self._diags = [torch.from_numpy(np.ascontiguousarray(d, dtype=float)).to(dev)
                       for d in _helm_diags(Apm, Bpm)]
        # matvec tail: A[k][k+2j] = At[k] * w_{k+2j}, w = 1 (SD) or (k+2j)^2 (SN)
        M = N - 2
        A2 = _helm_diags(Apm, Bpm)[1]
        if self.neumann:
            k = np.arange(M - 2, dtype=float)
            At = A2 / (k + 2) ** 2
        else:
            At = A2
        self._At = torch.from_numpy(np.ascontiguousarray(At)).to(dev)
        self.shape = shape
        self.u0 = torch.zeros(shape, dtype=torch.float64, device=dev)
        self.u1 = torch.zeros(shape, dtype=torch.float64, device=dev)
        self.u2 = torch.zeros(shape, dtype=torch.float64, device=dev)
        self.L = torch.zeros(shape, dtype=torch.float64, device=dev)
        self._Apm, self._Bpm = Apm, Bpm
        self.LU_Helmholtz(Apm, Bpm, self.alfa, self.beta, self.neumann, self.u0, self.u1, self.u2,
                          self.L, ax)

    # shenfun's static methods (chebyshev/la.py:278-286, 343-351), resolved at
    # every call through the configured backend (optimization/hip.py)
    @staticmethod
    @runtimeoptimizer
    def LU_Helmholtz(A, B, As, Bs, neumann, u0, u1, u2, L, axis=0):
        _no_host()

    @staticmethod
    @runtimeoptimizer
    def Solve_Helmholtz(b, u, neumann, u0, u1, u2, L, axis=0):
        _no_host()

    @staticmethod
    @runtimeoptimizer
    def Helmholtz_matvec(v, b, alfa, beta, A, B, axis=0):
        _no_host()

    @staticmethod
    @runtimeoptimizer
    def Helmholtz_Neumann_matvec(v, b, alfa, beta, A, B, axis=0):
        _no_host()

    def __call__(self, b, u=None, constraints=()):
        _check(b, "b")
        if u is None:
            u = b
        _check(u, "u")
        if tuple(b.shape) != tuple(self.shape) or u.shape != b.shape:
            raise ValueError(f"arrays must have shape {tuple(self.shape)}")
        # nonhomogeneous boundary values: b -= bc matrices x u's boundary slots
        apply_bc_matrices(self.bc_mats, b, u)
        self.Solve_Helmholtz(b, u, self.neumann, self.u0, self.u1, self.u2, self.L, self.axis)
        if constraints != ():
            assert len(constraints) == 1
            u[(0,) * u.dim()] = constraints[0][1]
        return u

    def matvec(self, v, c):
        _check(v, "v")
        _check(c, "c")
        shp, vst = _geom(v)
        _, cst = _geom(c)
        A0, _, _, Bm2, B0, B2 = self._diags
        _lib.call("sfa_helmholtz_matvec", _lib.ptr(v), _lib.ptr(c), _dtype_code(v), v.dim(), shp, vst,
                  cst, self.axis, int(self.neumann), _lib.ptr(A0), _lib.ptr(self._At), _lib.ptr(Bm2),
                  _lib.ptr(B0), _lib.ptr(B2), _lib.ptr(self._ab[0]), _lib.ptr(self._ab[1]), 1,
                  _lib.stream_ptr())
        return c


class ADDSolver:
    """Chebyshev Dirichlet Poisson A u = b (chebyshev/la.py:11-63)."""

    def __init__(self, mats, device="cuda"):
        A = _sum_mats(mats)
        assert A.get_key() == "ASDSDmat"
        self.mat = A
        M = A.shape[0]
        dev = torch.device(device)
        self._d = torch.from_numpy(A.diagonal_array(0)).to(dev)
        self._d1 = torch.from_numpy(A.diagonal_array(2) if M > 2 else np.zeros(1)).to(dev)
        self.scale = float(A.scale)

    def __call__(self, b, u=None, axis=0, constraints=()):
        _check(b, "b")
        if u is None:
            u = b
        shp, bst = _geom(b)
        _, ust = _geom(u)
        _lib.call("sfa_poisson_add_solve", _lib.ptr(b), _lib.ptr(u), _dtype_code(u), u.dim(), shp,
                  bst, ust, axis, _lib.ptr(self._d), _lib.ptr(self._d1), self.scale, _lib.stream_ptr())
        return u


class ANNSolver:
    """Chebyshev Neumann Poisson (chebyshev/la.py:66-128) along any axis of a
    1-3 dimensional array: the Neumann Helmholtz system with beta = 0, whose
    k = 0 gauge fixes u_0 (la.pyx:735-738), factorised once for the 1D
    operator and broadcast over the lines (zero factor stride); each
    constraint (row, value), default (0, 0), then sets that row along the
    axis, as the reference's u[row] = value after moving the axis first."""

    def __init__(self, mats, device="cuda"):
        A = _sum_mats(mats)
        assert A.get_key() == "ASNSNmat"
        from .matrices import mass_matrix
        space = A.testfunction[0]
        self.mat = A
        self.N = A.shape[0] + 2
        self._H = Helmholtz(A, mass_matrix(space), float(A.scale), 0.0, device=device)

    def __call__(self, b, u=None, axis=0, constraints=((0, 0),)):
        _check(b, "b")
        if u is None:
            u = b
        _check(u, "u")
        ax = axis % b.dim()
        if b.shape[ax] != self.N or u.shape != b.shape or u.dtype != b.dtype:
            raise ValueError(f"b and u need {self.N} entries along axis {ax} and equal shapes")
        H = self._H
        shp, bst = _geom(b)
        _, ust = _geom(u)
        fst = [0] * b.dim()
        fst[ax] = 1
        _lib.call("sfa_helmholtz_solve", _lib.ptr(b), _lib.ptr(u), _dtype_code(u), u.dim(), shp, bst,
                  ust, ax, 1, _lib.ptr(H.u0), _lib.ptr(H.u1), _lib.ptr(H.u2), _lib.ptr(H.L),
                  _lib.arr(fst), _lib.stream_ptr())
        for row, val in constraints:
            u.select(ax, int(row)).fill_(val)
        u.narrow(ax, self.N - 2, 2).zero_()  # homogeneous boundary slots (set_boundary_dofs)
        return u


class Biharmonic:
    """Chebyshev Galerkin biharmonic solver a0 S + alfa A + beta B along one
    axis (chebyshev/la.py:354-558): ``Biharmonic(S, A, B)`` with the scales
    carried by the (TP)matrices (SSBSBmat, ASBSBmat, BSBSBmat in any order,
    ``biharmonic_tpmatrices`` gives them), or ``Biharmonic(S, A, B, a0, alfa,
    beta)``.  LU once per mode on the device (two parity chains, the dense
    upper tail of S kept in rank-two form), ``__call__(b, u=None)`` solves in
    one launch, ``matvec(v, c)`` multiplies.  The factor arrays u0, u1, u2,
    l0, l1, ak, bk have the reference's layout."""

    def __init__(self, *args, device="cuda"):
        if len(args) == 1:
            args = args[0]
        args, bc = _split_bc(list(args))
        if bc:
            raise NotImplementedError("Biharmonic with nonhomogeneous boundary values")
        assert len(args) in (3, 6)
        byk = {m.get_key(): m for m in args[:3]}
        S, A, B = byk["SSBSBmat"], byk["ASBSBmat"], byk["BSBSBmat"]
        if len(args) == 3:
            a0, alfa, beta = S.scale, A.scale, B.scale
            if isinstance(S, TPMatrix):
                a0 = np.asarray(a0) * float(S.pmat.scale)
                alfa = np.asarray(alfa) * float(A.pmat.scale)
                beta = np.asarray(beta) * float(B.pmat.scale)
        else:
            a0, alfa, beta = args[3:6]
        a0 = np.unique(np.asarray(a0, dtype=float))
        if a0.size != 1:
            raise ValueError("the biharmonic scale a0 must be one number")
        self.a0 = float(a0[0])
        Spm = S.pmat if isinstance(S, TPMatrix) else S
        Apm = A.pmat if isinstance(A, TPMatrix) else A
        Bpm = B.pmat if isinstance(B, TPMatrix) else B
        space = Spm.testfunction[0]
        self.axis = S.axis if isinstance(S, TPMatrix) else space.axis
        T = S.tensorproductspace if isinstance(S, TPMatrix) else space.tensorproductspace
        N = Spm.shape[0] + 4
        shape = list(T.shape(True)) if T is not None else [N]
        ab_shape = list(shape)
        ab_shape[self.axis] = 1
        self.alfa = np.ascontiguousarray(np.broadcast_to(np.asarray(alfa, dtype=float), ab_shape))
        self.beta = np.ascontiguousarray(np.broadcast_to(np.asarray(beta, dtype=float), ab_shape))
        self.shape = shape
        M = (N - 4 + 1) // 2
        fshape = [2] + shape
        fshape[1 + self.axis] = M
        dev = torch.device(device)
        self.u0, self.u1, self.u2, self.l0, self.l1, self.ak, self.bk = (
            torch.zeros(fshape, dtype=torch.float64, device=dev) for _ in range(7))
        diag = lambda m, k: np.ascontiguousarray(  # noqa: E731
            np.broadcast_to(np.asarray(m[k], dtype=float), (m.shape[0] - abs(k),)))
        self._diags = [torch.from_numpy(diag(m, k)).to(dev) for m, k in
                       ((Spm, 0), (Spm, 2), (Spm, 4), (Apm, -2), (Apm, 0), (Apm, 2),
                        (Bpm, -4), (Bpm, -2), (Bpm, 0), (Bpm, 2), (Bpm, 4))]
        self.LU_Biharmonic(self.a0, self.alfa, self.beta, *self._diags, self.u0, self.u1, self.u2,
                           self.l0, self.l1, self.axis)
        self.Biharmonic_factor_pr(self.ak, self.bk, self.l0, self.l1, self.axis)

    # shenfun's static methods (chebyshev/la.py:497-518), resolved at every call
    @staticmethod
    @runtimeoptimizer
    def LU_Biharmonic(a0, alfa, beta, sii, siu, siuu, ail, aii, aiu, bill, bil, bii, biu, biuu,
                      u0, u1, u2, l0, l1, axis):
        _no_host()

    @staticmethod
    @runtimeoptimizer
    def Biharmonic_factor_pr(ak, bk, l0, l1, axis):
        _no_host()

    @staticmethod
    @runtimeoptimizer
    def Biharmonic_Solve(b, u, u0, u1, u2, l0, l1, ak, bk, a0, axis=0):
        _no_host()

    @staticmethod
    @runtimeoptimizer
    def Biharmonic_matvec(v, b, a0, alfa, beta, sii, siu, siuu, ail, aii, aiu, bill, bil, bii,
                          biu, biuu, axis=0):
        _no_host()

    def __call__(self, b, u=None, **kw):
        _check(b, "b")
        if u is None:
            u = b
        _check(u, "u")
        if tuple(b.shape) != tuple(self.shape) or u.shape != b.shape:
            raise ValueError(f"arrays must have shape {tuple(self.shape)}")
        self.Biharmonic_Solve(b, u, self.u0, self.u1, self.u2, self.l0, self.l1, self.ak, self.bk,
                              self.a0, self.axis)
        return u

    def matvec(self, v, c):
        _check(v, "v")
        _check(c, "c")
        self.Biharmonic_matvec(v, c, self.a0, self.alfa, self.beta, *self._diags, self.axis)
        return c
