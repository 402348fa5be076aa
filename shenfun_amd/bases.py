"""1D function spaces of the hot path with device (HIP) transforms.

Mirrors the per-axis SpectralBase interface (spectralbase.py:200-313):
``scalar_product``, ``forward`` and ``backward`` with
``(input_array=None, output_array=None, kind=None)``, acting along
``self.axis`` of torch tensors on the GPU.  Spaces:

* Chebyshev ``Orthogonal`` (GC), ``ShenDirichlet``, ``ShenNeumann``
  (chebyshev/bases.py:118-465, 476-554, 863-945);
* Legendre ``Orthogonal`` (LG), ``ShenDirichlet``, ``ShenNeumann``
  (legendre/bases.py:76-313, 319-384, 449-510);
* Fourier ``R2C`` and ``C2C`` (fourier/bases.py:54-641).

Compute runs only through the C ABI of libshenfun_amd.so.
"""
import numpy as np
import torch

from . import _lib
from .config import config

PI = np.pi


def _dev(device=None):
    return torch.device(device) if device is not None else torch.device("cuda")


def _geom(t):
    return _lib.arr(t.shape), _lib.arr(t.stride())


def _check(t, name, dtype=None):
    if not isinstance(t, torch.Tensor) or not t.is_cuda:
        raise TypeError(f"{name} must be a CUDA (HIP) torch tensor")
    if dtype is not None and t.dtype != dtype:
        raise TypeError(f"{name} must have dtype {dtype}, got {t.dtype}")
    if t.dim() < 1 or t.dim() > 3:
        raise ValueError(f"{name} must have 1..3 dims")


class Transform:
    """Planned transform of a space (spectralbase.Transform,
    spectralbase.py:2371-2416; arrays planned by SpectralBase.plan :602-648):
    called as ``(input_array=None, output_array=None, kind=None)``.
    ``input_array`` and ``output_array`` are device arrays owned by the space,
    allocated on first use; a call without ``input_array`` transforms
    ``input_array``, and a call without ``output_array`` writes into and
    returns the planned output array (one per input shape and dtype), so a
    later call overwrites it, as in the reference.  ``forward`` shares the
    output array of ``scalar_product`` (forward.output_array aliases the
    scalar product's, spectralbase.py:643)."""

    def __init__(self, space, fn, name):
        self._space = space
        self._fn = fn
        self._name = name

    def __call__(self, input_array=None, output_array=None, kind=None, **kw):
        sp = self._space
        if input_array is None:
            input_array = self.input_array
        sp._active.append(self._name)
        try:
            return self._fn(sp, input_array, output_array, kind, **kw)
        finally:
            sp._active.pop()

    @property
    def input_array(self):
        shape, dtype = self._space._planned_shape(self._name, "in")
        return self._space._planned_array(self._name, "in", shape, dtype)

    @property
    def output_array(self):
        shape, dtype = self._space._planned_shape(self._name, "out")
        return self._space._planned_array(self._name, "out", shape, dtype)

    tmp_array = output_array


def _planned_key(name):
    return "scalar_product" if name == "forward" else name


class _Planned:
    """Planner plumbing shared by the 1D spaces and TensorProductSpace."""

    def _init_planner(self):
        self._active = []
        self._plans = {}
        for name in ("scalar_product", "forward", "backward"):
            setattr(self, name, Transform(self, getattr(type(self), name), name))

    def _planned_array(self, name, which, shape, dtype):
        k = (_planned_key(name), which, tuple(int(v) for v in shape), dtype)
        a = self._plans.get(k)
        if a is None:
            a = torch.zeros(k[2], dtype=dtype, device=self.device)
            self._plans[k] = a
        return a

    def _planned_out(self, shape, dtype):
        """Output array of the transform being called (no output_array given)."""
        name = self._active[-1] if self._active else "scalar_product"
        return self._planned_array(name, "out", shape, dtype)


class SpectralBase(_Planned):
    """Common part of the 1D spaces."""

    def __init__(self, N, dtype=np.float64, device=None, padding_factor=1, dealias_direct=False):
        self.N = int(N)
        self.dtype = np.dtype(dtype)
        self.axis = 0
        self.dimensions = 1
        self.tensorproductspace = None
        self.device = _dev(device)
        self._cache = {}
        # spectralbase.py:51-54: a padded space keeps N coefficients and
        # evaluates on floor(padding_factor N) points
        self.padding_factor = 1
        if padding_factor != 1:
            self.padding_factor = np.floor(self.N * padding_factor) / self.N if self.N > 0 else 1
        self.dealias_direct = bool(dealias_direct)
        if self.dealias_direct and self.padding_factor != 1:
            raise ValueError("dealias_direct cannot be combined with padding_factor != 1")
        self._twin_space = None
        self._init_planner()

    def _planned_shape(self, name, which):
        """Default planned arrays: the 1D line, physical <-> spectral."""
        phys = torch.complex128 if self.dtype.kind == "c" else torch.float64
        spec = torch.complex128 if isinstance(self, R2C) else phys
        physical = (which == "in") == (name != "backward")
        if physical:
            return (self.shape(False),), phys
        return (self.shape(True),), spec

    # ---------------------------------------------------------------- info
    def family(self):
        raise NotImplementedError

    def boundary_condition(self):
        return None

    def dim(self):
        return self.N

    def slice(self):
        return slice(0, self.dim())

    def shape(self, forward_output=True):
        """Spectral (N) or physical (padded: floor(padding_factor N)) length
        along the axis (spectralbase.py:860-871)."""
        return self.N if forward_output else self._phys_len()

    def mesh(self):
        return self.points_and_weights()[0]

    # ------------------------------------------------------------ dealiasing
    @property
    def is_padded(self):
        return self.padding_factor != 1

    def _phys_len(self):
        if self.padding_factor != 1:
            return int(np.floor(self.padding_factor * self.N))
        return self.N

    def get_dealiased(self, padding_factor=1.5, dealias_direct=False, **kw):
        """Space otherwise as self for dealiasing (spectralbase.py:1111-1140):
        padded by ``padding_factor`` or, with ``dealias_direct``, the top third
        of the wavenumbers zeroed before backward."""
        if padding_factor == 1 and not dealias_direct:
            return self
        return self._clone(self.N, padding_factor, dealias_direct)

    def _clone(self, N, padding_factor=1, dealias_direct=False):
        raise NotImplementedError

    def _twin(self):
        """Unpadded space of the padded length: runs this space's transforms
        on the floor(padding_factor N) points."""
        if self._twin_space is None:
            self._twin_space = self._clone(self._phys_len())
        t = self._twin_space
        t.axis = self.axis
        t.dimensions = self.dimensions
        return t

    def _tmp(self, key, shape, dtype):
        k = ("tmp", key, tuple(int(v) for v in shape), dtype)
        if k not in self._cache:
            self._cache[k] = torch.empty(k[2], dtype=dtype, device=self.device)
        return self._cache[k]

    _pad_kind = 0  # sfa_pad_axis family: 0 polynomial, 1 R2C, 2 C2C

    def _pad(self, inp, out, mode, nbc=0, z0=0, z1=0):
        """sfa_pad_axis: mode 0 pad, 1 truncate, 2 copy with [z0, z1) zeroed."""
        code = _lib.C128 if inp.dtype == torch.complex128 else _lib.F64
        ishp, ist = _geom(inp)
        oshp, ost = _geom(out)
        _lib.call("sfa_pad_axis", _lib.ptr(inp), _lib.ptr(out), code, inp.dim(), ishp, ist, oshp,
                  ost, self.axis, mode, self._pad_kind, self.N, nbc, z0, z1, _lib.stream_ptr())
        return out

    def _direct_range(self):
        """Wavenumbers zeroed by dealias_direct (spectralbase.py:1374-1376)."""
        return 2 * self.N // 3, self.N

    def _dealias_input(self, inp, nbc=0):
        """dealias_direct: backward of a copy with the top third zeroed (the
        reference zeroes its planned input array in place)."""
        z0, z1 = self._direct_range()
        return self._pad(inp, self._tmp("dd", inp.shape, inp.dtype), 2, nbc, z0, max(z0, z1))

    # ------------------------------------------------------------ evaluation
    def map_reference_domain(self, x):
        """The hot-path spaces live on their reference domains ([-1, 1], [0, 2 pi))."""
        return np.asarray(x, dtype=float)

    def evaluate_basis_all(self, x):
        """(len(x), N) trial basis functions at x (spectralbase.py:372,
        argument=1; composite spaces: _composite, spectralbase.py:1589-1603)."""
        raise NotImplementedError

    def _eval_weights(self):
        return None

    def _eval_matrix(self, x):
        """evaluate_basis_all as complex128, with the R2C weights folded in."""
        E = np.asarray(self.evaluate_basis_all(x), dtype=np.complex128)
        w = self._eval_weights()
        return E if w is None else E * w[None, :]

    def eval(self, x, u, output_array=None):
        """u(x) at the points x from the expansion coefficients u
        (spectralbase.py:493-514, kind='vandermonde'): the basis functions at
        x are built on the host, the sum runs on the device
        (sfa_tensor_eval).  1D spaces only: a multi-dimensional array is
        evaluated by TensorProductSpace.eval."""
        x = np.atleast_1d(self.map_reference_domain(x)).ravel()
        u = self._tensor(u)
        if u.dim() != 1 or u.shape[0] != (self.shape(True) if isinstance(self, FourierBase)
                                          else self.N):
            raise ValueError("eval takes the 1D coefficient array of this space")
        # real values for R2C, and for real coefficients of a polynomial space
        real = isinstance(self, R2C) or (not isinstance(self, FourierBase) and not u.is_complex())
        return tensor_eval(u, [self._eval_matrix(x)], x.size, real, output_array)

    @property
    def is_composite(self):
        return False

    def _kind(self, kind):
        if kind is None:
            kind = config["transforms"]["kind"][self.family()]
        if kind not in ("fast", "recursive", "vandermonde", "hip", "auto"):
            raise ValueError(f"unknown transform kind {kind!r}")
        return kind

    def _out(self, inp, output_array, shape_axis, dtype):
        shape = list(inp.shape)
        shape[self.axis] = shape_axis
        if output_array is None:
            out = self._planned_out(shape, dtype)
            if out.data_ptr() == inp.data_ptr():  # transforming the planned output array itself
                out = torch.empty(shape, dtype=dtype, device=inp.device)
            return out
        _check(output_array, "output_array", dtype)
        if list(output_array.shape) != shape:
            raise ValueError(f"output_array shape {tuple(output_array.shape)} != {tuple(shape)}")
        return output_array

    def _tensor(self, a):
        if isinstance(a, np.ndarray):
            return torch.from_numpy(np.ascontiguousarray(a)).to(self.device)
        return a

    def _dev_array(self, key, fn):
        if key not in self._cache:
            self._cache[key] = torch.from_numpy(np.ascontiguousarray(fn())).to(self.device)
        return self._cache[key]


class _PolyBase(SpectralBase):
    """Chebyshev/Legendre orthogonal or composite (Shen) space."""

    bc_kind = None  # None, 'D' or 'N'

    @property
    def is_composite(self):
        return self.bc_kind is not None

    def dim(self):
        return self.N - 2 if self.is_composite else self.N

    def boundary_condition(self):
        return {None: None, "D": "Dirichlet", "N": "Neumann"}[self.bc_kind]

    def short_name(self):
        return {None: "T" if self.family() == "chebyshev" else "L", "D": "SD", "N": "SN"}[self.bc_kind]

    def stencil_k2(self):
        """K2(k), k < N-2: phi_k = P_k + K2(k) P_{k+2}."""
        k = np.arange(self.N - 2, dtype=float)
        if self.bc_kind == "D":
            return -np.ones(self.N - 2)
        if self.family() == "chebyshev":
            return -(k / (k + 2)) ** 2  # chebyshev/bases.py:937
        return -k * (k + 1) / (k * k + 5 * k + 6)  # legendre/bases.py:502

    def _K2(self):
        if not self.is_composite:
            return None
        return self._dev_array("K2", self.stencil_k2)

    def _dtype_code(self, t):
        if t.dtype == torch.complex128:
            return _lib.C128
        if t.dtype == torch.float64:
            return _lib.F64
        raise TypeError(f"unsupported dtype {t.dtype} (float64 / complex128 only)")

    # dense per-line contraction: 'recursive' / 'vandermonde' kinds
    def _matmul(self, inp, out, M):
        shp, sin = _geom(inp)
        _, sout = _geom(out)
        _lib.call("sfa_line_matmul", _lib.ptr(inp), _lib.ptr(out), self._dtype_code(inp), inp.dim(),
                  shp, sin, sout, self.axis, int(M.shape[0]), _lib.ptr(M), _lib.stream_ptr())
        return out

    def _analysis_matrix(self):
        raise NotImplementedError

    def _synthesis_matrix(self):
        raise NotImplementedError

    # ------------------------------------------------------------ transforms
    def _clone(self, N, padding_factor=1, dealias_direct=False):
        return type(self)(N, quad=self.quad, bc=self._bc_orig, dtype=self.dtype,
                          device=self.device, padding_factor=padding_factor,
                          dealias_direct=dealias_direct)

    def stencil(self):
        """{offset: coefficients} of the basis in the orthogonal polynomials."""
        if not self.is_composite:
            return {0: np.ones(self.N)}
        return {0: np.ones(self.N - 2), 2: self.stencil_k2()}

    def stencil_matrix(self):
        """(N, N): row k < dim() holds phi_k's orthogonal coefficients."""
        S = np.zeros((self.N, self.N))
        k = np.arange(self.dim())
        for off, v in self.stencil().items():
            S[k, k + off] = v
        return S

    def ortho_norms(self, n):
        """Exact (P_m, P_m)_w, m < n: Chebyshev pi (m = 0), pi/2; Legendre 2/(2m+1)."""
        m = np.arange(n, dtype=float)
        if self.family() == "chebyshev":
            h = np.full(n, PI / 2)
            h[0] = PI
            return h
        return 2.0 / (2 * m + 1)

    def get_testspace(self, kind="Galerkin", **kw):
        """self for Galerkin; for 'PG' the Phi2 space on N + 2 points when self
        has two boundary conditions (spectralbase.py:1148-1204)."""
        if kind in ("Galerkin", "G"):
            return self
        if kind not in ("Petrov-Galerkin", "PG"):
            raise ValueError(f"unknown test space kind {kind!r}")
        if self.N - self.dim() != 2:
            return self._clone(self.N, self.padding_factor, self.dealias_direct)
        cls = ChebyshevPhi2 if self.family() == "chebyshev" else LegendrePhi2
        return cls(self.N + 2, quad=self.quad, dtype=self.dtype, device=self.device)

    # ------------------------------------------------------ boundary values
    @property
    def has_nonhomogeneous_bcs(self):
        # decided by the values as given: a time-dependent value that happens
        # to vanish at the current time still counts
        return self.is_composite and any(not (isinstance(v, float) and v == 0)
                                         for v in getattr(self, "_bc_orig", self.bc_values))

    def bc_basis(self):
        """Orthogonal coefficients (rows: left, right function) of the boundary
        basis, the lowest-degree polynomials with unit value (Dirichlet) or
        unit derivative (Neumann) at one end and zero at the other
        (get_bc_basis, utilities/findbasis.py:68-140): D: (P0 -+ P1)/2;
        N: P1/2 -+ P2/8 (Chebyshev), P1/2 -+ P2/6 (Legendre)."""
        if self.bc_kind == "D":
            return np.array([[0.5, -0.5, 0.0], [0.5, 0.5, 0.0]])
        if self.bc_kind == "DN":  # P0 (u(-1)), P0 + P1 (u'(1)): chebyshev/bases.py:1977-1978
            return np.array([[1.0, 0.0, 0.0], [1.0, 1.0, 0.0]])
        if self.bc_kind == "ND":  # -P0 + P1 (u'(-1)), P0 (u(1)): legendre/bases.py:1216-1217
            return np.array([[-1.0, 1.0, 0.0], [1.0, 0.0, 0.0]])
        c = 1.0 / 8.0 if self.family() == "chebyshev" else 1.0 / 6.0
        return np.array([[0.0, 0.5, -c], [0.0, 0.5, c]])

    def _vander_at(self, x, n, deriv=0):
        """P_k^(deriv)(x), k < n (host, numpy's polynomial classes)."""
        P = np.polynomial.Chebyshev if self.family() == "chebyshev" else np.polynomial.Legendre
        x = np.asarray(x, dtype=float)
        return np.stack([P.basis(k).deriv(deriv)(x) for k in range(n)], axis=1)

    def evaluate_basis_all(self, x):
        x = np.asarray(x, dtype=float).ravel()
        vander = (np.polynomial.chebyshev.chebvander if self.family() == "chebyshev"
                  else np.polynomial.legendre.legvander)
        V = vander(x, self.N - 1)
        d = self.dim()
        P = np.zeros((x.size, self.N))
        for off, v in self.stencil().items():
            P[:, :d] += V[:, off:off + d] * v[None, :]
        if self.bc_kind in ("D", "N", "DN", "ND"):
            # boundary slots: the boundary basis functions carrying the values
            P[:, d:] = self._bc_functions(x)
        return P

    def _bc_functions(self, x, deriv=0):
        """phi_bc_j^(deriv)(x), shape (len(x), 2)."""
        return self._vander_at(x, 3, deriv) @ self.bc_basis().T

    def _test_functions(self, x):
        """phi_k(x), k < dim(), shape (len(x), dim()) (the stencil on the
        orthogonal polynomials)."""
        return self.evaluate_basis_all(x)[:, :self.dim()]

    def bc_matrices(self):
        """Boundary mass and stiffness matrices (N-2, 2): B[k, j] =
        (phi_bc_j, phi_k)_w and, in the units of this space's stiffness matrix,
        A[k, j] = (phi_bc_j'', phi_k)_w (Chebyshev) or -(phi_bc_j'', phi_k)
        (Legendre, whose A is (phi', phi')); quadrature on the space's own N
        points is exact for these degrees (host, one-time, like the
        reference's matrix assembly)."""
        key = "bcmats"
        if key not in self._cache:
            x, w = self.points_and_weights(self.N)
            Phi = self._test_functions(x) * w[:, None]
            B = Phi.T @ self._bc_functions(x)
            A = Phi.T @ self._bc_functions(x, 2)
            if self.family() == "legendre":
                A = -A
            self._cache[key] = (A, B)
        return self._cache[key]

    def _bc_update(self, src, dst, W, nrows, scales=None):
        """sfa_boundary_update with host matrices W (nmat, nrows, 2)."""
        Wd = torch.from_numpy(np.ascontiguousarray(W, dtype=float)).to(dst.device)
        code = _lib.C128 if dst.dtype == torch.complex128 else _lib.F64
        sshp, sst = _geom(src)
        dshp, dst_st = _geom(dst)
        sc = None
        stride = 0
        if scales is not None:
            sc = torch.from_numpy(np.ascontiguousarray(scales, dtype=float)).to(dst.device)
            stride = sc.shape[-1]
        _lib.call("sfa_boundary_update", _lib.ptr(src), _lib.ptr(dst), code, dst.dim(), sshp, sst,
                  dshp, dst_st, self.axis, 2, nrows, W.shape[0], _lib.ptr(Wd),
                  _lib.ptr(sc) if sc is not None else None, stride, _lib.stream_ptr())
        return dst

    def set_boundary_dofs(self, u, final=False):
        """Boundary slots N-2, N-1 of ``u`` along the axis := the boundary
        values (BoundaryValues.set_boundary_dofs, tensorproductspace.py:2108-2128).
        In a tensor-product space a constant boundary value is the Fourier
        mode 0 of the axes already transformed (all later axes when
        ``final`` is False, every other axis when True), the same value along
        the others."""
        if not self.is_composite:
            return u
        for j, v in enumerate(self.bc_values):
            if not isinstance(v, float):
                u.narrow(self.axis, self.N - 2 + j, 1).copy_(self._bc_slot_array(j, u, final))
                continue
            sl = [slice(None)] * u.dim()
            sl[self.axis] = self.N - 2 + j
            u[tuple(sl)] = 0
            if v == 0:
                continue
            for i in range(u.dim()):
                if i != self.axis and (final or self._transformed_before(i)):
                    sl[i] = 0
            u[tuple(sl)] = v
        return u

    def _transformed_before(self, i):
        """Axis i is already in spectral space when forward reaches this axis."""
        T = self.tensorproductspace
        return T.transformed_before(i, self.axis) if T is not None else i > self.axis

    def _bc_slot_array(self, j, u, final):
        """Boundary slot j for a boundary value that varies along the other
        axes: the function on their mesh, forward-transformed along the axes
        already transformed (BoundaryValues.set_tensor_bcs,
        tensorproductspace.py:1909-2068, projects it onto the other bases the
        same way); shape of u with 1 along this axis."""
        T = self.tensorproductspace
        if T is None:
            raise ValueError("a boundary value that varies needs the space in a TensorProductSpace")
        key = ("bcslot", j, bool(final), tuple(u.shape), u.dtype)
        if key in self._cache:
            return self._cache[key]
        ax = self.axis
        nd = T.dimensions
        X = [np.asarray(b.mesh(), dtype=float) for b in T.bases]
        coords = [X[i].reshape([X[i].size if k == i else 1 for k in range(nd)]) for i in range(nd)]
        others = [i for i in range(nd) if i != ax]
        v = self.bc_values[j]
        if hasattr(v, "free_symbols"):
            import sympy as sp
            names = "xyz"[:nd]  # the coordinates of axes 0, 1, 2
            byname = {str(q): q for q in v.free_symbols}
            if any(n not in names or names.index(n) == ax for n in byname):
                raise ValueError(f"boundary value {v} must be a function of the other coordinates")
            fn = sp.lambdify([byname.get(names[i], sp.Symbol(names[i])) for i in others], v, "numpy")
        else:
            fn = v
        shape = [1 if i == ax else X[i].size for i in range(nd)]
        arr = np.array(np.broadcast_to(np.asarray(fn(*[coords[i] for i in others]), dtype=float),
                                       shape))
        t = torch.from_numpy(arr).to(u.device)
        for i in T.axes[::-1]:  # forward order: an R2C axis first
            if i == ax or not (final or self._transformed_before(i)):
                continue
            b = T.bases[i]
            if not isinstance(b, FourierBase):
                raise NotImplementedError("boundary values varying along a non-periodic axis")
            t = t if isinstance(b, R2C) else t.to(torch.complex128)
            shp = list(t.shape)
            shp[i] = b.shape(True)
            t = b.forward(t, torch.empty(shp, dtype=torch.complex128, device=u.device))
        t = t.to(u.dtype) if u.dtype.is_complex or not t.dtype.is_complex else t.real.contiguous()
        self._cache[key] = t
        return t

    def scalar_product(self, input_array=None, output_array=None, kind=None):
        kind = self._kind(kind)
        inp = self._tensor(input_array)
        _check(inp, "input_array")
        if inp.shape[self.axis] != self._phys_len():
            raise ValueError("input_array has the wrong length along axis")
        out = self._out(inp, output_array, self.N, inp.dtype)
        if self.is_padded:
            # padded-length scalar product, truncated to the N coefficients
            # (spectralbase.py:225-229, _truncation_forward :1362-1366)
            tmp = self._tmp("sp", inp.shape, inp.dtype)
            self._twin()._scalar_product(inp, tmp, kind)
            self._pad(tmp, out, 1, self.N - self.dim())
            return out
        self._scalar_product(inp, out, kind)
        return out

    def backward(self, input_array=None, output_array=None, kind=None):
        kind = self._kind(kind)
        inp = self._tensor(input_array)
        _check(inp, "input_array")
        if inp.shape[self.axis] != self.N:
            raise ValueError("input_array has the wrong length along axis")
        out = self._out(inp, output_array, self._phys_len(), inp.dtype)
        if out.data_ptr() == inp.data_ptr():
            raise ValueError("backward cannot run in place")
        nbc = self.N - self.dim()
        src = inp
        if self.is_padded:
            # zero-padded coefficients, boundary slots moved to the end
            # (_padding_backward, spectralbase.py:1368-1389), then the
            # padded-length synthesis
            shp = list(inp.shape)
            shp[self.axis] = self._phys_len()
            tmp = self._pad(inp, self._tmp("bw", shp, inp.dtype), 0, nbc)
            self._twin()._backward(tmp, out, kind)
        else:
            if self.dealias_direct:
                inp = self._dealias_input(inp, nbc)
            self._backward(inp, out, kind)
        if self.has_nonhomogeneous_bcs:
            # the boundary functions weighted by the input's boundary slots,
            # on the physical points (to_ortho + _add_to_orthogonal,
            # spectralbase.py:1616-1634, tensorproductspace.py:2070-2089)
            lift = self._cache.get("bclift")
            if lift is None:
                lift = self._bc_functions(self.mesh())[None]
                self._cache["bclift"] = lift
            self._bc_update(src, out, lift, self._phys_len())
        return out

    def forward(self, input_array=None, output_array=None, kind=None):
        """Projection: scalar product, then the inverse mass matrix along axis
        (spectralbase.py:235-263, apply_inverse_mass :560-577).  BC slots
        N-2, N-1 of a composite space are set to 0 (homogeneous BCs,
        tensorproductspace.py:2091-2106)."""
        out = self.scalar_product(input_array, output_array, kind)
        if self.is_composite:
            self.set_boundary_dofs(out)
            if self.has_nonhomogeneous_bcs:
                # _add_mass_rhs (tensorproductspace.py:2091-2106): minus the
                # boundary mass matrix times the boundary values
                self._bc_update(out, out, -self.bc_matrices()[1][None], self.N - 2)
        self._mass_solver()(out, axis=self.axis)
        return out

    def _mass_solver(self):
        if "mass" not in self._cache:
            from . import la
            from .matrices import mass_matrix, SparseMatrix
            if self.is_composite:
                B = mass_matrix(self)
            else:
                B = SparseMatrix({0: self.norm_sq()}, (self.N, self.N))
            self._cache["mass"] = B.get_solver()(B)
        return self._cache["mass"]


class ChebyshevBase(_PolyBase):
    """Chebyshev space on Gauss (GC, default) or Gauss-Lobatto (GL) points; fast
    transforms are the LDS DCT-II/III (GC) or DCT-I (GL) kernels."""

    def family(self):
        return "chebyshev"

    def points_and_weights(self, N=None):
        """chebyshev/bases.py:202-216.  GC: x_j = cos(pi(2j+1)/(2N)), w = pi/N;
        GL: x_j = cos(pi j/(N-1)), w = pi/(N-1) with both ends halved (both
        descending)."""
        N = self._phys_len() if N is None else N
        j = np.arange(N, dtype=float)
        if self.quad == "GL":
            w = np.full(N, PI / (N - 1))
            w[0] /= 2
            w[-1] /= 2
            return np.cos(PI * j / (N - 1)), w
        return np.cos(PI * (2 * j + 1) / (2 * N)), np.full(N, PI / N)

    def norm_sq(self):
        """Quadrature norms (T_k, T_k)_w, chebyshev/bases.py:271-277: pi/2,
        pi for k = 0 and, with GL, for k = N-1."""
        h = np.full(self.N, PI / 2)
        h[0] = PI
        if self.quad == "GL":
            h[-1] = PI
        return h

    def _vander(self):
        x, _ = self.points_and_weights(self.N)
        j = np.arange(self.N)
        return np.cos(np.outer(np.arccos(x), j))  # T_k(x_j) = cos(k theta_j)

    def _analysis_matrix(self):
        def build():
            V = self._vander()
            _, w = self.points_and_weights(self.N)
            M = (V * w[:, None]).T.copy()  # M[k, j] = w_j T_k(x_j)
            if self.is_composite:
                K2 = self.stencil_k2()
                M[:-2] = M[:-2] + K2[:, None] * M[2:]
            return M
        return self._dev_array("Ma", build)

    def _synthesis_matrix(self):
        def build():
            V = self._vander()
            if self.is_composite:
                K2 = self.stencil_k2()
                S = np.zeros_like(V)
                S[:, :-2] = V[:, :-2] + K2[None, :] * V[:, 2:]
                return S
            return V
        return self._dev_array("Ms", build)

    # 'auto': up to this many points the transforms run as dense contractions
    # on FP64 MFMA (sfa_line_matmul); above it as DCTs on the complex FFT
    # kernels.  Measured on the MI355X along axis 0 of (N, N, N/2 + 1) arrays:
    # N = 128 scalar product 21.7 us dense vs 26.0 us DCT, backward 21.7 vs
    # 32.5; N = 256 scalar product 197 vs 164 (the DCT wins).
    DENSE_MAX = 128

    def _auto(self, kind, inp, out):
        if kind != "auto":
            return kind
        dense = self._phys_len() <= self.DENSE_MAX and out.data_ptr() != inp.data_ptr()
        return "vandermonde" if dense else "fast"

    def _scalar_product(self, inp, out, kind):
        kind = self._auto(kind, inp, out)
        if kind in ("fast", "hip"):
            shp, sin = _geom(inp)
            _, sout = _geom(out)
            fn = "sfa_cheb_scalar_product_gl" if self.quad == "GL" else "sfa_cheb_scalar_product"
            _lib.call(fn, _lib.ptr(inp), _lib.ptr(out), self._dtype_code(inp),
                      inp.dim(), shp, sin, sout, self.axis, _lib.ptr(self._K2()), _lib.stream_ptr())
        else:
            if out.data_ptr() == inp.data_ptr():
                raise ValueError("dense transform cannot run in place")
            self._matmul(inp, out, self._analysis_matrix())

    def _backward(self, inp, out, kind):
        kind = self._auto(kind, inp, out)
        if kind in ("fast", "hip"):
            shp, sin = _geom(inp)
            _, sout = _geom(out)
            fn = "sfa_cheb_backward_gl" if self.quad == "GL" else "sfa_cheb_backward"
            _lib.call(fn, _lib.ptr(inp), _lib.ptr(out), self._dtype_code(inp),
                      inp.dim(), shp, sin, sout, self.axis, _lib.ptr(self._K2()), _lib.stream_ptr())
        else:
            self._matmul(inp, out, self._synthesis_matrix())


class LegendreBase(_PolyBase):
    """Legendre-Gauss space; transforms are dense per-line contractions with
    matrices built on the device from the three-term recurrence (the default
    'recursive' kind, spectralbase.py:484-491, 550-558)."""

    def family(self):
        return "legendre"

    def points_and_weights(self, N=None):
        """Legendre-Gauss nodes (ascending) and weights (legendre/bases.py:121-126)."""
        N = self._phys_len() if N is None else N
        key = ("pw", N)
        if key not in self._cache:
            self._cache[key] = leggauss(N)
        return self._cache[key]

    def norm_sq(self):
        k = np.arange(self.N, dtype=float)
        return 2.0 / (2 * k + 1)

    def _xw(self):
        x, w = self.points_and_weights(self.N)
        return (self._dev_array("x", lambda: x), self._dev_array("w", lambda: w))

    def _analysis_matrix(self):
        if "Ma" not in self._cache:
            x, w = self._xw()
            M = torch.empty((self.N, self.N), dtype=torch.float64, device=self.device)
            _lib.call("sfa_legendre_matrix", _lib.ptr(x), _lib.ptr(w), self.N, self.N,
                      _lib.ptr(self._K2()), 1, _lib.ptr(M), _lib.stream_ptr())
            self._cache["Ma"] = M
        return self._cache["Ma"]

    def _synthesis_matrix(self):
        if "Ms" not in self._cache:
            x, w = self._xw()
            M = torch.empty((self.N, self.N), dtype=torch.float64, device=self.device)
            _lib.call("sfa_legendre_matrix", _lib.ptr(x), _lib.ptr(w), self.N, self.N,
                      _lib.ptr(self._K2()), 0, _lib.ptr(M), _lib.stream_ptr())
            self._cache["Ms"] = M
        return self._cache["Ms"]

    def _scalar_product(self, inp, out, kind):
        if out.data_ptr() == inp.data_ptr():
            raise ValueError("Legendre transforms cannot run in place")
        self._matmul(inp, out, self._analysis_matrix())

    def _backward(self, inp, out, kind):
        self._matmul(inp, out, self._synthesis_matrix())


def _legendre_pd(n, x):
    p0 = np.ones_like(x)
    p1 = x.copy()
    for k in range(2, n + 1):
        p0, p1 = p1, ((2 * k - 1) * x * p1 - (k - 1) * p0) / k
    dp = n * (x * p1 - p0) / (x * x - 1)
    return p1, dp


def leggauss(N):
    """Legendre-Gauss nodes (ascending) and weights in float64.

    Newton iteration on the three-term recurrence in extended precision
    (np.longdouble) from a cosine initial guess, then rounded; symmetric.
    """
    if N == 1:
        return np.array([0.0]), np.array([2.0])
    ld = np.longdouble
    k = np.arange(1, N + 1, dtype=ld)
    x = np.cos(ld(np.pi) * (k - ld(0.25)) / (N + ld(0.5)))
    for _ in range(100):
        p, dp = _legendre_pd(N, x)
        dx = p / dp
        x = x - dx
        if np.max(np.abs(dx)) < ld(1e-19):
            break
    _, dp = _legendre_pd(N, x)
    w = 2 / ((1 - x * x) * dp * dp)
    o = np.argsort(x)
    x = x[o].astype(np.float64)
    w = w[o].astype(np.float64)
    return 0.5 * (x - x[::-1]), 0.5 * (w + w[::-1])


class FourierBase(SpectralBase):
    """Fourier exponentials on [0, 2 pi): scalar product = FFT/N, backward =
    unnormalised inverse (fourier/bases.py:131-162, 380-407, 526-533)."""

    def family(self):
        return "fourier"

    def boundary_condition(self):
        return "Periodic"

    def points_and_weights(self, N=None):
        N = self._phys_len() if N is None else N
        return np.arange(N, dtype=float) * 2 * PI / N, np.array([1.0 / N])

    def _clone(self, N, padding_factor=1, dealias_direct=False):
        return type(self)(N, device=self.device, padding_factor=padding_factor,
                          dealias_direct=dealias_direct)

    def get_testspace(self, kind="Galerkin", **kw):
        """self for Galerkin; a copy for 'PG' (Fourier spaces are their own test space)."""
        if kind in ("Galerkin", "G"):
            return self
        return self._clone(self.N, self.padding_factor, self.dealias_direct)

    def forward(self, input_array=None, output_array=None, kind=None):
        return self.scalar_product(input_array, output_array, kind)

    def evaluate_basis_all(self, x):
        """exp(i k x) for the stored wavenumbers k (fourier/bases.py:186-197)."""
        x = np.asarray(x, dtype=float).ravel()
        return np.exp(1j * x[:, None] * self.wavenumbers()[None, :].astype(float))

    def convolve(self, u, v, uv=None, fast=True):
        """Convolution of the coefficient vectors u and v (R2C fourier/bases.py:437-475,
        C2C :600-640).  fast: backward both on the padded grid, multiply,
        forward (the space must be padded; with padding_factor >= 3/2 + 1/N
        the result is alias-free); otherwise the direct O(N^2) sums, N + 1
        (R2C) or 2N (C2C) entries."""
        u = self._tensor(u)
        v = self._tensor(v)
        _check(u, "u", torch.complex128)
        _check(v, "v", torch.complex128)
        if fast:
            if not self.padding_factor > 1.0:
                raise ValueError("padding factor must be > 3/2+1/N to perform convolution "
                                 "without aliasing")
            shp = list(u.shape)
            shp[self.axis] = self._phys_len()
            dt = torch.float64 if isinstance(self, R2C) else torch.complex128
            a = self.backward(u, self._tmp("cv_a", shp, dt))
            b = self.backward(v, self._tmp("cv_b", shp, dt))
            return self.forward(pointwise_mul(a, b, self._tmp("cv_ab", shp, dt)), uv)
        if u.dim() != 1 or v.shape != u.shape:
            raise ValueError("the direct convolution is 1D")
        real = isinstance(self, R2C)
        n_out = self.N + 1 if real else 2 * self.N
        if uv is None:
            uv = torch.zeros(n_out, dtype=torch.complex128, device=u.device)
        _check(uv, "uv", torch.complex128)
        if tuple(uv.shape) != (n_out,) or not (u.is_contiguous() and v.is_contiguous()
                                                and uv.is_contiguous()):
            raise ValueError(f"uv must be a contiguous ({n_out},) complex128 array")
        if u.shape[0] != self.shape(True):
            raise ValueError("u has the wrong length")
        _lib.call("sfa_convolve_direct", _lib.ptr(u), _lib.ptr(v), _lib.ptr(uv), self.N, int(real),
                  _lib.stream_ptr())
        return uv


def pointwise_mul(a, b, out):
    """out = a * b on the device (contiguous arrays of one dtype)."""
    for t in (a, b, out):
        if not t.is_contiguous() or t.shape != a.shape or t.dtype != a.dtype:
            raise ValueError("pointwise_mul takes contiguous arrays of one shape and dtype")
    code = _lib.C128 if a.dtype == torch.complex128 else _lib.F64
    _lib.call("sfa_pointwise_mul", _lib.ptr(a), _lib.ptr(b), _lib.ptr(out), code, a.numel(),
              _lib.stream_ptr())
    return out


def tensor_eval(c, mats, npts, real, output_array=None):
    """sum over k of c[k0, k1, k2] E0[p, k0] E1[p, k1] E2[p, k2] for each
    point p (one sfa_tensor_eval launch); ``mats`` are host (npts, n_d)
    basis-function matrices for the leading axes of the device array ``c``.
    Returns the real part when ``real``."""
    if c.dim() != len(mats) or not 1 <= c.dim() <= 3:
        raise ValueError("one evaluation matrix per axis of the coefficient array")
    for d, E in enumerate(mats):
        if E.shape != (npts, c.shape[d]):
            raise ValueError(f"axis {d}: evaluation matrix {E.shape} does not match "
                             f"({npts}, {c.shape[d]})")
    dev = c.device
    cc = c.to(torch.complex128).contiguous()
    n = list(c.shape) + [1] * (3 - c.dim())
    Es = [torch.from_numpy(np.ascontiguousarray(E, dtype=np.complex128)).to(dev) for E in mats]
    Es += [torch.ones((npts, 1), dtype=torch.complex128, device=dev)] * (3 - len(Es))
    out = torch.empty(npts, dtype=torch.complex128, device=dev)
    _lib.call("sfa_tensor_eval", _lib.ptr(cc), n[0], n[1], n[2], _lib.ptr(Es[0]), _lib.ptr(Es[1]),
              _lib.ptr(Es[2]), _lib.ptr(out), npts, _lib.stream_ptr())
    res = out.real if real else out
    if output_array is not None:
        output_array.copy_(res.reshape(output_array.shape))
        return output_array
    return res.contiguous()


class R2C(FourierBase):
    _pad_kind = 1

    def __init__(self, N, device=None, padding_factor=1, dealias_direct=False):
        super().__init__(N, np.float64, device, padding_factor, dealias_direct)

    def short_name(self):
        return "R2C"

    def dim(self):
        return self.N // 2 + 1

    def shape(self, forward_output=True):
        return self.N // 2 + 1 if forward_output else self._phys_len()

    def wavenumbers(self):
        return np.fft.rfftfreq(self.N, 1.0 / self.N).astype(int)  # fourier/bases.py:334-342

    def _direct_range(self):
        return self.N // 3, self.N // 2 + 1  # fourier/bases.py:430-433

    def _eval_weights(self):
        """u(x) = Re sum_k w_k c_k e^{ikx} over the stored half: w = 2 but for
        k = 0 and the Nyquist mode of even N (fourier/bases.py:380-394)."""
        w = np.full(self.N // 2 + 1, 2.0)
        w[0] = 1.0
        if self.N % 2 == 0:
            w[-1] = 1.0
        return w

    def scalar_product(self, input_array=None, output_array=None, kind=None):
        inp = self._tensor(input_array)
        _check(inp, "input_array", torch.float64)
        if inp.shape[self.axis] != self._phys_len():
            raise ValueError("input_array has the wrong length along axis")
        out = self._out(inp, output_array, self.N // 2 + 1, torch.complex128)
        if self.is_padded:
            # FFT on the padded points (normalised by 1/M), truncated with the
            # Nyquist rule (fourier/bases.py:409-418)
            shp = list(inp.shape)
            shp[self.axis] = self._phys_len() // 2 + 1
            tmp = self._twin().scalar_product(inp, self._tmp("sp", shp, torch.complex128))
            return self._pad(tmp, out, 1)
        shp, sin = _geom(inp)
        _, sout = _geom(out)
        _lib.call("sfa_fft_r2c", _lib.ptr(inp), _lib.ptr(out), inp.dim(), shp, sin, sout, self.axis,
                  1.0 / self.N, _lib.stream_ptr())
        return out

    def backward(self, input_array=None, output_array=None, kind=None):
        inp = self._tensor(input_array)
        _check(inp, "input_array", torch.complex128)
        if inp.shape[self.axis] != self.N // 2 + 1:
            raise ValueError("input_array has the wrong length along axis")
        out = self._out(inp, output_array, self._phys_len(), torch.float64)
        if self.is_padded:
            shp = list(inp.shape)
            shp[self.axis] = self._phys_len() // 2 + 1
            tmp = self._pad(inp, self._tmp("bw", shp, torch.complex128), 0)
            return self._twin().backward(tmp, out)
        if self.dealias_direct:
            inp = self._dealias_input(inp)
        shp = list(inp.shape)
        shp[self.axis] = self.N
        _, sin = _geom(inp)
        _, sout = _geom(out)
        _lib.call("sfa_fft_c2r", _lib.ptr(inp), _lib.ptr(out), inp.dim(), _lib.arr(shp), sin, sout,
                  self.axis, 1.0, _lib.stream_ptr())
        return out


class C2C(FourierBase):
    _pad_kind = 2

    def __init__(self, N, device=None, padding_factor=1, dealias_direct=False):
        super().__init__(N, np.complex128, device, padding_factor, dealias_direct)

    def short_name(self):
        return "C2C"

    def wavenumbers(self):
        return np.fft.fftfreq(self.N, 1.0 / self.N).astype(int)  # fourier/bases.py:535-543

    def _c2c(self, input_array, output_array, sign, scale):
        inp = self._tensor(input_array)
        _check(inp, "input_array", torch.complex128)
        out = self._out(inp, output_array, self.N, torch.complex128)
        shp, sin = _geom(inp)
        _, sout = _geom(out)
        _lib.call("sfa_fft_c2c", _lib.ptr(inp), _lib.ptr(out), inp.dim(), shp, sin, sout, self.axis,
                  sign, scale, _lib.stream_ptr())
        return out

    def _direct_range(self):
        z = slice(self.N // 3, -(self.N // 3) + 1).indices(self.N)  # fourier/bases.py:595-598
        return z[0], z[1]

    def scalar_product(self, input_array=None, output_array=None, kind=None):
        if not self.is_padded:
            return self._c2c(input_array, output_array, -1, 1.0 / self.N)
        # padded-length FFT (1/M), wavenumbers folded back (fourier/bases.py:570-578)
        inp = self._tensor(input_array)
        _check(inp, "input_array", torch.complex128)
        if inp.shape[self.axis] != self._phys_len():
            raise ValueError("input_array has the wrong length along axis")
        out = self._out(inp, output_array, self.N, torch.complex128)
        tmp = self._twin().scalar_product(inp, self._tmp("sp", inp.shape, torch.complex128))
        return self._pad(tmp, out, 1)

    def backward(self, input_array=None, output_array=None, kind=None):
        if not (self.is_padded or self.dealias_direct):
            return self._c2c(input_array, output_array, +1, 1.0)
        inp = self._tensor(input_array)
        _check(inp, "input_array", torch.complex128)
        if inp.shape[self.axis] != self.N:
            raise ValueError("input_array has the wrong length along axis")
        if self.dealias_direct:
            return self._c2c(self._dealias_input(inp), output_array, +1, 1.0)
        out = self._out(inp, output_array, self._phys_len(), torch.complex128)
        shp = list(inp.shape)
        shp[self.axis] = self._phys_len()
        tmp = self._pad(inp, self._tmp("bw", shp, torch.complex128), 0)
        return self._twin().backward(tmp, out)


# ---------------------------------------------------------------- factories
def _poly(cls_base, bc_kind, name):
    return type(name, (cls_base,), {"bc_kind": bc_kind})


class BoundaryValues:
    """``space.bc`` (BoundaryValues.update, tensorproductspace.py:1859-1907):
    ``update(time)`` substitutes ``time`` for the sympy symbol t in the
    boundary values as given and makes the results the current values; the
    boundary slots of later forward transforms, Functions and solves use them
    (cached slot arrays of varying values are dropped).  Without ``time`` the
    last time is kept (initially 0)."""

    def __init__(self, base):
        self.base = base
        self.bc_time = 0.0

    def update(self, time=None, update_tensor=False):
        vals = []
        changed = bool(update_tensor)
        if time is not None:
            self.bc_time = float(time)
        for v in self.base._bc_orig:
            if hasattr(v, "free_symbols"):
                ts = [q for q in v.free_symbols if q.name == "t"]
                if ts:
                    v = v.subs(ts[0], self.bc_time)
                    changed = True
            vals.append(_bcval(v))
        self.base.bc_values = tuple(vals)
        if changed:
            cache = self.base._cache
            for k in [k for k in cache if isinstance(k, tuple) and k[:1] == ("bcslot",)]:
                del cache[k]
        return self.base.bc_values


class _Init:
    def __init__(self, N, quad=None, bc=None, dtype=np.float64, device=None, padding_factor=1,
                 dealias_direct=False, **kw):
        self.quad = quad
        # boundary values (left, right) of a composite space; a string or
        # dict bc is parsed by FunctionSpace
        self._bc_orig = tuple(_bcval(v) for v in bc) if isinstance(bc, (tuple, list)) \
            else (0.0, 0.0)
        self.bc_values = self._bc_orig
        SpectralBase.__init__(self, N, dtype, device, padding_factor, dealias_direct)
        # time-dependent values (sympy expressions in t) start at t = 0
        self.bc = BoundaryValues(self)
        self.bc.update()
        if self.is_padded and quad == "GL":
            raise NotImplementedError("padding is supported on Gauss (GC / LG) points")


class ChebyshevOrthogonal(_Init, ChebyshevBase):
    bc_kind = None


class ChebyshevShenDirichlet(_Init, ChebyshevBase):
    bc_kind = "D"


class ChebyshevShenNeumann(_Init, ChebyshevBase):
    bc_kind = "N"


class LegendreOrthogonal(_Init, LegendreBase):
    bc_kind = None


class LegendreShenDirichlet(_Init, LegendreBase):
    bc_kind = "D"


class LegendreShenNeumann(_Init, LegendreBase):
    bc_kind = "N"


def _stencil_dict(family, kind, k):
    """{offset: coefficient(k)} of phi_k in the orthogonal polynomials.
    ShenBiharmonic (u = u' = 0 at both ends): chebyshev/bases.py:1175-1260
    (``_stencil``), legendre/bases.py ShenBiharmonic; Phi2 (test space of
    two-condition bases): chebyshev/bases.py:1441-1445, legendre/bases.py:631;
    Phi4 (test space of four-condition bases): chebyshev/bases.py Phi4,
    legendre/bases.py Phi4."""
    n = np.asarray(k, dtype=float)
    if kind in ("DN", "ND"):
        # phi_k = P_k + b_k P_{k+1} + c_k P_{k+2}: DN u(-1) = u'(1) = 0, ND
        # u'(-1) = u(1) = 0 (chebyshev/bases.py DirichletNeumann /
        # NeumannDirichlet, legendre/bases.py:1084-1272)
        sgn = 1.0 if kind == "DN" else -1.0
        if family == "chebyshev":
            q = 2 * n * n + 6 * n + 5
            return {0: np.ones_like(n), 1: sgn * 4 * (n + 1) / q, 2: -(2 * n * n + 2 * n + 1) / q}
        return {0: np.ones_like(n), 1: sgn * (2 * n + 3) / (n + 2) ** 2,
                2: -((n + 1) / (n + 2)) ** 2}
    if kind == "SB":
        if family == "chebyshev":
            return {0: np.ones_like(n), 2: -(2 * n + 4) / (n + 3), 4: (n + 1) / (n + 3)}
        return {0: np.ones_like(n), 2: -(4 * n + 10) / (2 * n + 7), 4: (2 * n + 3) / (2 * n + 7)}
    if kind == "P2":
        if family == "chebyshev":
            return {0: 1 / (2 * PI * (n + 1) * (n + 2)), 2: -1 / (PI * (n + 1) * (n + 3)),
                    4: 1 / (2 * PI * (n + 2) * (n + 3))}
        return {0: 1 / (2 * (2 * n + 3)), 2: -(2 * n + 5) / (2 * n + 7) / (2 * n + 3),
                4: 1 / (2 * (2 * n + 7))}
    if family == "chebyshev":  # P4
        return {0: 1 / (8 * PI * (n + 1) * (n + 2) * (n + 3) * (n + 4)),
                2: -1 / (2 * PI * (n + 1) * (n + 3) * (n + 4) * (n + 5)),
                4: 3 / (4 * PI * (n + 2) * (n + 3) * (n + 5) * (n + 6)),
                6: -1 / (2 * PI * (n + 3) * (n + 4) * (n + 5) * (n + 7)),
                8: 1 / (8 * PI * (n + 4) * (n + 5) * (n + 6) * (n + 7))}
    return {0: 1 / (2 * (8 * n ** 3 + 60 * n ** 2 + 142 * n + 105)),
            2: -2 / (8 * n ** 3 + 84 * n ** 2 + 262 * n + 231),
            4: 3 * (2 * n + 9) / ((2 * n + 5) * (2 * n + 7) * (2 * n + 11) * (2 * n + 13)),
            6: -2 / (8 * n ** 3 + 132 * n ** 2 + 694 * n + 1155),
            8: 1 / (2 * (8 * n ** 3 + 156 * n ** 2 + 1006 * n + 2145))}


class _StencilSpace:
    """Composite spaces given by a stencil of several bands: the
    ShenBiharmonic trial space (4 boundary slots) and the Petrov-Galerkin
    test spaces Phi2 (4) and Phi4 (8) that get_testspace('PG') builds on N + 2
    / N + 4 points (spectralbase.py:1148-1204).  Transforms are dense
    per-line contractions with the stencil folded into the quadrature
    matrices (built once on the host); forward of ShenBiharmonic solves its
    pentadiagonal mass matrix."""

    kind = None
    _nbc = {"SB": 4, "P2": 4, "P4": 8, "DN": 2, "ND": 2}
    _names = {"SB": "Biharmonic", "P2": "Biharmonic", "P4": "Biharmonic",
              "DN": "DirichletNeumann", "ND": "NeumannDirichlet"}

    @property
    def bc_kind(self):
        return self.kind

    def dim(self):
        return self.N - self._nbc[self.kind]

    def boundary_condition(self):
        return self._names[self.kind]

    def short_name(self):
        return self.kind

    def stencil(self):
        return _stencil_dict(self.family(), self.kind, np.arange(self.dim()))

    def _ortho_vander(self):
        x, w = self.points_and_weights(self.N)
        return self._vander_at(x, self.N), w

    def _analysis_matrix(self):
        def build():
            V, w = self._ortho_vander()
            return self.stencil_matrix() @ (V * w[:, None]).T  # M[k, j] = w_j phi_k(x_j)
        return self._dev_array("STa", build)

    def _synthesis_matrix(self):
        def build():
            V, _ = self._ortho_vander()
            return V @ self.stencil_matrix().T  # M[j, k] = phi_k(x_j)
        return self._dev_array("STs", build)

    def _scalar_product(self, inp, out, kind):
        if out.data_ptr() == inp.data_ptr():
            raise ValueError("dense transform cannot run in place")
        self._matmul(inp, out, self._analysis_matrix())

    def _backward(self, inp, out, kind):
        self._matmul(inp, out, self._synthesis_matrix())

    def forward(self, input_array=None, output_array=None, kind=None):
        if self.kind in ("DN", "ND"):
            # boundary slots and their mass right-hand side as the D/N bases
            return super().forward(input_array, output_array, kind)
        if self.kind in ("P2", "P4"):
            raise NotImplementedError(f"{self.kind} is a test space: scalar_product and backward only")
        out = self.scalar_product(input_array, output_array, kind)
        sl = [slice(None)] * out.dim()
        sl[self.axis] = slice(self.dim(), self.N)
        out[tuple(sl)] = 0
        self._mass_solver()(out, axis=self.axis)
        return out

    def _mass_solver(self):
        if "mass" not in self._cache:
            from .matrices import GalerkinMatrix
            B = GalerkinMatrix(self, self, 0, f"B{self.kind}{self.kind}mat")
            self._cache["mass"] = B.get_solver()(B)
        return self._cache["mass"]

    def get_testspace(self, kind="Galerkin", **kw):
        if kind in ("Galerkin", "G"):
            return self
        if kind not in ("Petrov-Galerkin", "PG"):
            raise ValueError(f"unknown test space kind {kind!r}")
        if self.kind in ("DN", "ND"):
            raise NotImplementedError("Petrov-Galerkin test spaces of the mixed bases")
        if self.kind != "SB":
            return self
        cls = ChebyshevPhi4 if self.family() == "chebyshev" else LegendrePhi4
        return cls(self.N + 4, quad=self.quad, dtype=self.dtype, device=self.device)

    def _clone(self, N, padding_factor=1, dealias_direct=False):
        if (padding_factor != 1 or dealias_direct) and self.kind in ("P2", "P4"):
            raise NotImplementedError("dealiasing of the Petrov-Galerkin test spaces")
        # the stencils depend on k only, so a padded space runs on an
        # unpadded twin of the padded length like the D/N bases
        return type(self)(N, quad=self.quad, bc=self._bc_orig, dtype=self.dtype, device=self.device,
                          padding_factor=padding_factor, dealias_direct=dealias_direct)


class ChebyshevDirichletNeumann(_StencilSpace, _Init, ChebyshevBase):
    kind = "DN"


class LegendreDirichletNeumann(_StencilSpace, _Init, LegendreBase):
    kind = "DN"


class ChebyshevNeumannDirichlet(_StencilSpace, _Init, ChebyshevBase):
    kind = "ND"


class LegendreNeumannDirichlet(_StencilSpace, _Init, LegendreBase):
    kind = "ND"


class ChebyshevPhi2(_StencilSpace, _Init, ChebyshevBase):
    kind = "P2"


class LegendrePhi2(_StencilSpace, _Init, LegendreBase):
    kind = "P2"


class ChebyshevPhi4(_StencilSpace, _Init, ChebyshevBase):
    kind = "P4"


class LegendrePhi4(_StencilSpace, _Init, LegendreBase):
    kind = "P4"


class ChebyshevShenBiharmonic(_StencilSpace, _Init, ChebyshevBase):
    kind = "SB"


class LegendreShenBiharmonic(_StencilSpace, _Init, LegendreBase):
    kind = "SB"


def _is_biharmonic(bc):
    if isinstance(bc, str):
        return bc.lower() in ("biharmonic", "b")
    if isinstance(bc, (tuple, list)) and len(bc) == 4:
        if any(float(v) != 0 for v in bc):
            raise NotImplementedError("biharmonic spaces take homogeneous boundary values")
        return True
    if isinstance(bc, dict) and set(bc) == {"left", "right"} and \
            all(set(bc[sd]) == {"D", "N"} for sd in bc):
        if any(float(v) != 0 for sd in bc for v in bc[sd].values()):
            raise NotImplementedError("biharmonic spaces take homogeneous boundary values")
        return True
    return False


def _bcval(v):
    """A boundary value: a number (float), or a function of the other
    coordinates (a Python callable taking them in axis order, or a sympy
    expression in x, y, z) kept as given."""
    if callable(v) and not hasattr(v, "free_symbols"):
        return v
    if hasattr(v, "free_symbols"):
        if not v.free_symbols:
            return float(v)
        return v
    return float(v)


def _parse_bc(bc):
    """Boundary conditions -> (kind, (left value, right value)).

    ``'Dirichlet'`` / ``'Neumann'`` (homogeneous), ``(a, b)`` (Dirichlet
    u(-1) = a, u(1) = b) or the dict form ``{'left': {'D': a}, 'right': {'D': b}}``
    with 'D' or 'N' at both ends (spectralbase.py BoundaryConditions); values
    are numbers.
    """
    if bc is None:
        return None, (0.0, 0.0)
    if isinstance(bc, str):
        b = bc.lower()
        if b in ("dirichlet", "d"):
            return "D", (0.0, 0.0)
        if b in ("neumann", "n"):
            return "N", (0.0, 0.0)
        raise ValueError(f"unsupported bc {bc!r}")
    if isinstance(bc, (tuple, list)):
        if len(bc) == 2:
            return "D", (_bcval(bc[0]), _bcval(bc[1]))
        raise ValueError(f"unsupported bc {bc!r}")
    if isinstance(bc, dict):
        if set(bc) != {"left", "right"} or any(len(bc[s]) != 1 for s in bc):
            raise NotImplementedError("one condition at each end, 'left' and 'right'")
        (kl, vl), = bc["left"].items()
        (kr, vr), = bc["right"].items()
        if kl != kr or kl not in ("D", "N"):
            raise NotImplementedError("Dirichlet or Neumann at both ends")
        return kl, (_bcval(vl), _bcval(vr))
    raise ValueError(f"unsupported bc {bc!r}")


def _mixed_bc(bc):
    """('DN', (a, b)) for {'left': {'D': a}, 'right': {'N': b}}, ('ND', (a, b))
    for the mirror (DirichletNeumann / NeumannDirichlet), None otherwise."""
    if not isinstance(bc, dict) or set(bc) != {"left", "right"} or \
            any(not isinstance(bc[s], dict) or len(bc[s]) != 1 for s in bc):
        return None
    (kl, vl), = bc["left"].items()
    (kr, vr), = bc["right"].items()
    if {kl, kr} != {"D", "N"}:
        return None
    return kl + kr, (_bcval(vl), _bcval(vr))


def _bc_kind(bc):
    return _parse_bc(bc)[0]


def FunctionSpace(N, family="Fourier", bc=None, dtype="d", quad=None, domain=None, device=None,
                  **kw):
    """shenfun.FunctionSpace (forms/arguments.py:30) for the hot-path spaces."""
    fam = family.lower()
    if domain is not None:
        ref = (0.0, 2 * PI) if fam in ("fourier", "f") else (-1.0, 1.0)
        if not np.allclose([float(d) for d in domain], ref, rtol=0, atol=1e-14):
            # other domains scale the wavenumbers / derivatives (domain_factor);
            # refused rather than silently solved on the reference domain
            raise NotImplementedError(f"only the reference domain {ref} is on the hot path")
    pad = dict(padding_factor=kw.get("padding_factor", 1),
               dealias_direct=kw.get("dealias_direct", False))
    if fam in ("fourier", "f"):
        if np.dtype(dtype).kind == "c":
            return C2C(N, device=device, **pad)
        return R2C(N, device=device, **pad)
    if bc is not None and _is_biharmonic(bc):
        if fam in ("chebyshev", "c"):
            return ChebyshevShenBiharmonic(N, quad=quad or "GC", device=device)
        if fam in ("legendre", "l"):
            return LegendreShenBiharmonic(N, quad="LG", device=device)
        raise NotImplementedError(f"biharmonic {family!r} space")
    mixed = _mixed_bc(bc)
    if mixed is not None:
        cls = {("chebyshev", "DN"): ChebyshevDirichletNeumann, ("c", "DN"): ChebyshevDirichletNeumann,
               ("chebyshev", "ND"): ChebyshevNeumannDirichlet, ("c", "ND"): ChebyshevNeumannDirichlet,
               ("legendre", "DN"): LegendreDirichletNeumann, ("l", "DN"): LegendreDirichletNeumann,
               ("legendre", "ND"): LegendreNeumannDirichlet, ("l", "ND"): LegendreNeumannDirichlet}
        if (fam, mixed[0]) not in cls:
            raise NotImplementedError(f"mixed boundary conditions on {family!r}")
        return cls[(fam, mixed[0])](N, quad=quad or ("GC" if fam in ("chebyshev", "c") else "LG"),
                                    bc=mixed[1], device=device)
    kind, vals = _parse_bc(bc)
    if fam in ("chebyshev", "c"):
        if quad not in (None, "GC", "GL"):
            raise NotImplementedError("Chebyshev-Gauss (GC) and Gauss-Lobatto (GL) quadratures only")
        cls = {None: ChebyshevOrthogonal, "D": ChebyshevShenDirichlet, "N": ChebyshevShenNeumann}[kind]
        return cls(N, quad=quad or "GC", bc=vals, device=device, **pad)
    if fam in ("legendre", "l"):
        if quad not in (None, "LG"):
            raise NotImplementedError("Legendre-Gauss (LG) quadrature only")
        cls = {None: LegendreOrthogonal, "D": LegendreShenDirichlet, "N": LegendreShenNeumann}[kind]
        return cls(N, quad="LG", bc=vals, device=device, **pad)
    raise NotImplementedError(f"family {family!r} is not on the hot path")
