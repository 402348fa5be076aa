"""Tensor-product spaces on one GPU (tensorproductspace.py:262-469 mirror).

One GPU holds the whole array (no pencils, no transfers): the forward and
scalar-product transforms run axis by axis from the last axis to the first,
the backward transform from the first to the last
(ForwardTransform/ScalarTransform/BackwardTransform,
tensorproductspace.py:45-259), each axis being one call of the 1D space's
HIP transform on the strided device array.
"""
import numpy as np
import torch

from .bases import FourierBase, R2C, _Planned, pointwise_mul, tensor_eval


class TensorProductSpace(_Planned):
    def __init__(self, comm, bases, axes=None, dtype=None, slab=False, device=None, **kw):
        """``comm`` is accepted for signature compatibility and ignored: one
        GPU holds the full array."""
        self.bases = list(bases)
        self.dimensions = len(self.bases)
        if not 1 <= self.dimensions <= 3:
            raise ValueError("1..3 dimensional spaces are supported")
        # transform order (tensorproductspace.py:314-469): backward runs
        # axes[0], axes[1], ..., forward and scalar_product the reverse, so an
        # R2C basis (real physical data) must be axes[-1], transformed first
        self.axes = tuple(int(a) for a in axes) if axes is not None else tuple(range(self.dimensions))
        if sorted(self.axes) != list(range(self.dimensions)):
            raise ValueError(f"axes {axes} is not a permutation of the dimensions")
        for i, b in enumerate(self.bases):
            if isinstance(b, R2C) and i != self.axes[-1]:
                raise ValueError("an R2C space must be the first axis transformed forward, axes[-1]")
            b.axis = i
            b.dimensions = self.dimensions
            b.tensorproductspace = self
        self.device = torch.device(device) if device is not None else torch.device("cuda")
        self._work = {}
        self._padded = {}
        self._init_planner()

    def _planned_shape(self, name, which):
        """Planned arrays (tensorproductspace.py:45-259): physical shape and
        dtype on the physical side, spectral on the other."""
        physical = (which == "in") == (name != "backward")
        if physical:
            return self.shape(False), self.physical_dtype()
        return self.shape(True), self.spectral_dtype()

    # ---------------------------------------------------------------- info
    def shape(self, forward_output=True):
        return tuple(b.shape(forward_output) for b in self.bases)

    def global_shape(self, forward_output=True):
        return self.shape(forward_output)

    def local_slice(self, forward_output=True):
        return tuple(slice(0, s) for s in self.shape(forward_output))

    def physical_dtype(self):
        return torch.float64 if any(isinstance(b, R2C) for b in self.bases) or not any(
            isinstance(b, FourierBase) for b in self.bases) else torch.complex128

    def spectral_dtype(self):
        return torch.complex128 if any(isinstance(b, FourierBase) for b in self.bases) \
            else torch.float64

    def nonperiodic_axis(self):
        ax = [i for i, b in enumerate(self.bases) if not isinstance(b, FourierBase)]
        if len(ax) != 1:
            raise ValueError("exactly one non-periodic axis expected")
        return ax[0]

    def wavenumbers(self):
        return [b.wavenumbers() if isinstance(b, FourierBase) else np.arange(b.N)
                for b in self.bases]

    local_wavenumbers = wavenumbers

    def get_testspace(self, kind="Galerkin", **kw):
        """self for Galerkin; for 'PG' the space with each basis replaced by
        its Petrov-Galerkin test space (Phi2 on N + 2 points for the
        two-condition axis, tensorproductspace.py get_testspace)."""
        if kind in ("Galerkin", "G"):
            return self
        key = ("testspace", kind)
        if key not in self._padded:
            bases = [b.get_testspace(kind) if hasattr(b, "get_testspace") else b._clone(b.N)
                     for b in self.bases]
            self._padded[key] = TensorProductSpace(None, bases, axes=self.axes, device=self.device)
        return self._padded[key]

    def get_dealiased(self, padding_factor=1.5, dealias_direct=False):
        """Space otherwise as self for dealiasing (tensorproductspace.py:543-578):
        every basis padded by ``padding_factor`` (a number or one per axis), or
        the top third of each axis' wavenumbers zeroed with ``dealias_direct``."""
        if np.ndim(padding_factor) == 0:
            if padding_factor == 1 and not dealias_direct:
                return self
            padding_factor = (padding_factor,) * self.dimensions
        padding_factor = tuple(padding_factor)
        if len(padding_factor) != self.dimensions:
            raise ValueError("one padding factor per axis")
        key = (padding_factor, bool(dealias_direct))
        if key not in self._padded:
            bases = [b.get_dealiased(padding_factor=pf, dealias_direct=dealias_direct)
                     for b, pf in zip(self.bases, padding_factor)]
            bases = [b._clone(b.N) if b is o else b for b, o in zip(bases, self.bases)]
            self._padded[key] = TensorProductSpace(None, bases, axes=self.axes, device=self.device)
        return self._padded[key]

    # ------------------------------------------------------------ helpers
    def _buf(self, key, shape, dtype):
        k = (key, tuple(shape), dtype)
        if k not in self._work:
            self._work[k] = torch.empty(shape, dtype=dtype, device=self.device)
        return self._work[k]

    def transformed_before(self, i, ax):
        """True if forward transforms axis i before axis ax."""
        return self.axes.index(i) > self.axes.index(ax)

    def _chain(self, inp, output_array, method, order, kind):
        cur = inp
        n = len(order)
        for j, ax in enumerate(order):
            b = self.bases[ax]
            last = j == n - 1
            k = kind[b.family()] if isinstance(kind, dict) else kind
            if method == "backward":
                shp = list(cur.shape)
                shp[ax] = self.shape(False)[ax]
                dt = torch.float64 if (isinstance(b, R2C)) else cur.dtype
            else:
                shp = list(cur.shape)
                shp[ax] = self.shape(True)[ax]
                dt = torch.complex128 if isinstance(b, R2C) else cur.dtype
            if last and output_array is not None:
                out = output_array
            elif last:
                out = self._planned_out(shp, dt)
                if out.data_ptr() == cur.data_ptr():
                    out = torch.empty(shp, dtype=dt, device=self.device)
            else:
                out = self._buf((method, j % 2), shp, dt)
                if out.data_ptr() == cur.data_ptr():
                    out = self._buf((method, 2), shp, dt)
            getattr(b, method)(cur, out, kind=k)
            cur = out
        return cur

    def mesh(self):
        """Broadcastable quadrature meshes, one per axis (numpy, host)."""
        out = []
        for i, b in enumerate(self.bases):
            x = np.asarray(b.mesh(), dtype=float)
            s = [1] * self.dimensions
            s[i] = x.size
            out.append(x.reshape(s))
        return out

    # ---------------------------------------------------------- transforms
    def scalar_product(self, input_array=None, output_array=None, kind=None, **kw):
        """(tensorproductspace.py:146-190): last axis first."""
        return self._chain(input_array, output_array, "scalar_product", list(self.axes[::-1]), kind)

    def forward(self, input_array=None, output_array=None, kind=None, **kw):
        """(tensorproductspace.py:205-259): last axis first."""
        return self._chain(input_array, output_array, "forward", list(self.axes[::-1]), kind)

    def backward(self, input_array=None, output_array=None, kind=None, **kw):
        """(tensorproductspace.py:46-96): first axis first."""
        return self._chain(input_array, output_array, "backward", list(self.axes), kind)

    def convolve(self, a_hat, b_hat, ab_hat=None):
        """Convolution of two spectral arrays (tensorproductspace.py:637-666):
        backward both (on this space's, possibly padded, grid), pointwise
        product, forward into ``ab_hat``; alias-free when the Fourier bases
        are padded by at least 3/2 (``get_dealiased``)."""
        shp, dt = self.shape(False), self.physical_dtype()
        a = self.backward(a_hat, self._buf(("conv", 0), shp, dt))
        b = self.backward(b_hat, self._buf(("conv", 1), shp, dt))
        return self.forward(pointwise_mul(a, b, self._buf(("conv", 2), shp, dt)), ab_hat)


    def eval(self, points, coefficients, output_array=None, method=1):
        """Function values at scattered points (tensorproductspace.py:668-760):
        ``points`` of shape (D, npts).  Each axis' basis functions at the
        points' coordinates are built on the host (R2C weights folded in),
        the sum over all coefficients runs on the device in one launch
        (sfa_tensor_eval); ``method`` is accepted and ignored."""
        pts = np.atleast_2d(np.asarray(points, dtype=float))
        if pts.shape[0] != self.dimensions:
            raise ValueError(f"points must have shape ({self.dimensions}, npts)")
        npts = pts.shape[1]
        c = coefficients
        if tuple(c.shape) != tuple(self.shape(True)):
            raise ValueError(f"coefficients of shape {tuple(c.shape)}, expected {self.shape(True)}")
        if not isinstance(c, torch.Tensor):
            c = torch.from_numpy(np.ascontiguousarray(c)).to(self.device)
        mats = [b._eval_matrix(b.map_reference_domain(pts[i])) for i, b in enumerate(self.bases)]
        real = any(isinstance(b, R2C) for b in self.bases) or \
            (self.physical_dtype() == torch.float64 and not c.is_complex())
        return tensor_eval(c, mats, npts, real, output_array)


def Array(space, buffer=None, val=None):
    """Physical-space array of ``space`` (torch tensor on the GPU)."""
    if isinstance(space, TensorProductSpace):
        shape, dtype = space.shape(False), space.physical_dtype()
        dev = space.device
    else:
        shape = (space.shape(False),)
        dtype = torch.float64 if space.dtype.kind == "f" else torch.complex128
        dev = space.device
    if buffer is not None:
        return _from_buffer(buffer, shape, dtype, dev)
    t = torch.zeros(shape, dtype=dtype, device=dev)
    if val is not None:
        t.fill_(val)
    return t


def Function(space, buffer=None, val=None):
    """Spectral-space array of ``space`` (torch tensor on the GPU); the
    boundary slots of a space with nonhomogeneous boundary values hold those
    values (BoundaryValues.set_boundary_dofs(final=True), as the reference's
    Function does on creation)."""
    t = _function(space, buffer, val)
    bases = space.bases if isinstance(space, TensorProductSpace) else [space]
    if buffer is None:
        for b in bases:
            if getattr(b, "has_nonhomogeneous_bcs", False):
                b.set_boundary_dofs(t, final=True)
    return t


def _function(space, buffer=None, val=None):
    if isinstance(space, TensorProductSpace):
        shape, dtype = space.shape(True), space.spectral_dtype()
        dev = space.device
    else:
        shape = (space.shape(True) if isinstance(space, FourierBase) else space.N,)
        dtype = torch.complex128 if isinstance(space, FourierBase) or space.dtype.kind == "c" \
            else torch.float64
        dev = space.device
    if buffer is not None:
        return _from_buffer(buffer, shape, dtype, dev)
    t = torch.zeros(shape, dtype=dtype, device=dev)
    if val is not None:
        t.fill_(val)
    return t


def _from_buffer(buffer, shape, dtype, dev):
    if isinstance(buffer, torch.Tensor):
        t = buffer.to(device=dev, dtype=dtype)
    else:
        a = np.asarray(buffer)
        if not a.flags.writeable:
            a = a.copy()  # read-only views (broadcast meshes) are copied
        t = torch.as_tensor(a, device=dev).to(dtype)
    return t.reshape(shape).contiguous()


def project(uh, T, output_array=None, fill=True, uh_space=None):
    """Galerkin projection onto ``T`` (forms/project.py:16-172 for the
    hot-path cases): find u in T with (u - uh, v)_w = 0 for all v in T.

    * ``uh`` a physical array on T's quadrature mesh: u = T.forward(uh)
      (scalar product, then the inverse mass matrix);
    * ``uh`` coefficients of another space ``uh_space`` (a Function of it):
      u = T.forward(uh_space.backward(uh)) -- the meshes must agree;
    * ``uh`` a Python callable f(x0, x1, ...) of the mesh coordinates:
      evaluated on T's mesh, then projected.

    Returns ``output_array`` (a new Function of T when None); with ``fill``
    it is zeroed first, as the reference does."""
    if output_array is None:
        output_array = Function(T)
    elif fill:
        output_array.zero_()
    if callable(uh) and not isinstance(uh, torch.Tensor):
        xs = T.mesh() if isinstance(T, TensorProductSpace) else [np.asarray(T.mesh(), dtype=float)]
        shape = T.shape(False) if isinstance(T, TensorProductSpace) else (T.N,)
        vals = np.broadcast_to(np.asarray(uh(*xs)), shape)
        uh = Array(T, buffer=np.ascontiguousarray(vals))
    elif uh_space is not None:
        uh = uh_space.backward(uh).clone()
    return T.forward(uh, output_array)