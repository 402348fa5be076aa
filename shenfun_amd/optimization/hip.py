"""HIP backend of the name-based dispatch: shenfun's Cython kernels by name and
signature, computing on the device through the C ABI.

Every function takes the arguments of its namesake in shenfun's
optimization/cython modules (la.pyx, transforms.pyx, Matvec.pyx) and works in
place like it.  Arrays may be torch tensors on the device (used in place, no
copies) or NumPy arrays (copied to the device, the result copied back into
the same array) -- the latter is what shenfun's own classes pass.

Reference entry points replaced (file:line in shenfun/optimization/cython):
  la.pyx:29 ThreeDMA_Solve, :45 TwoDMA_Solve, :61 PDMA_Solve, :77 TDMA_Solve,
  :93 TDMA_O_Solve, :109 DiagMA_Solve, :165 FDMA_Solve, :181 HeptaDMA_Solve,
  :199 FDMA_LU, :213 PDMA_LU, :243 TDMA_LU, :254 TDMA_O_LU, :265 HeptaDMA_LU,
  :371 SolverGeneric1ND_solve_data, :462-592 the *_inner_solve line solves,
  :614 Poisson_Solve_ADD, :708 LU_Helmholtz, :845 Solve_Helmholtz;
  transforms.pyx:175 scalar_product, :199 evaluate_expansion_all;
  Matvec.pyx:697 Helmholtz_matvec, :722-818 Helmholtz_Neumann_matvec;
  la.pyx:975 LU_Biharmonic, :1195 Biharmonic_factor_pr, :1235 Biharmonic_Solve,
  Matvec.pyx:908 Biharmonic_matvec.
"""
import collections
import hashlib

import numpy as np
import torch

from .. import _lib

__all__ = [
    "TDMA_LU", "TDMA_Solve", "TDMA_inner_solve", "PDMA_LU", "PDMA_Solve", "PDMA_inner_solve",
    "TDMA_O_LU", "TDMA_O_Solve", "TDMA_O_inner_solve", "FDMA_LU", "FDMA_Solve",
    "FDMA_inner_solve", "HeptaDMA_LU", "HeptaDMA_Solve", "HeptaDMA_inner_solve",
    "TwoDMA_Solve", "TwoDMA_inner_solve", "ThreeDMA_Solve", "ThreeDMA_inner_solve",
    "DiagMA_Solve", "DiagMA_inner_solve", "SolverGeneric1ND_solve_data", "LU_Helmholtz",
    "Solve_Helmholtz", "Helmholtz_matvec", "Helmholtz_Neumann_matvec", "Poisson_Solve_ADD",
    "scalar_product", "evaluate_expansion_all",
]


# ---------------------------------------------------------------- arrays
class _Dev:
    """Device view of an argument: a CUDA tensor as it is, anything else
    (NumPy array, host tensor) as a contiguous device copy whose result
    ``back()`` writes into the original."""

    def __init__(self, a, device="cuda"):
        self.orig = a
        if isinstance(a, torch.Tensor):
            self.t = a if a.is_cuda else a.to(device)
        else:
            self.t = torch.from_numpy(np.ascontiguousarray(a)).to(device)

    def back(self):
        a = self.orig
        if isinstance(a, torch.Tensor):
            if not a.is_cuda:
                a.copy_(self.t.cpu())
        else:
            a[...] = self.t.cpu().numpy()


def _dtype_code(t):
    if t.dtype == torch.complex128:
        return _lib.C128
    if t.dtype == torch.float64:
        return _lib.F64
    raise TypeError(f"unsupported dtype {t.dtype} (float64 / complex128 only)")


def _geom(t):
    return _lib.arr(t.shape), _lib.arr(t.stride())


def _rows(data):
    d = _Dev(data)
    if not d.t.is_contiguous() or d.t.dtype != torch.float64 or d.t.dim() < 2:
        raise TypeError("data must be C-contiguous float64 DIA rows")
    return d


# ---------------------------------------------------------------- banded LU
def _lu(fn, data):
    d = _rows(data)
    n = d.t.shape[-1]
    nb = int(np.prod(d.t.shape[:-2])) if d.t.dim() > 2 else 1
    _lib.call(fn, _lib.ptr(d.t), n, nb, _lib.stream_ptr())
    d.back()


def TDMA_LU(data):
    """In-place LU of the (3, n) DIA rows, offsets -2, 0, 2 (la.pyx:243-252)."""
    _lu("sfa_tdma_lu", data)


def PDMA_LU(data):
    """In-place LU of the (5, n) DIA rows, offsets -4..4 (la.pyx:213-241)."""
    _lu("sfa_pdma_lu", data)


def TDMA_O_LU(data):
    """In-place LU of the (3, n) DIA rows, offsets -1, 0, 1 (la.pyx:254-263)."""
    _lu("sfa_tdma_o_lu", data)


def FDMA_LU(data):
    """In-place LU of the (4, n) DIA rows, offsets -2, 0, 2, 4 (la.pyx:199-211)."""
    _lu("sfa_fdma_lu", data)


def HeptaDMA_LU(data):
    """In-place LU of the (7, n) DIA rows, offsets -4..8 (la.pyx:265-303)."""
    _lu("sfa_heptadma_lu", data)


# ---------------------------------------------------------------- banded solves
def _solve(fn, u, data, axis, line_stride=0, skip=0):
    du = _Dev(u)
    dd = _rows(data)
    t = du.t
    shp, st = _geom(t)
    ax = axis % t.dim()
    _lib.call(fn, _lib.ptr(t), _dtype_code(t), t.dim(), shp, st, ax, _lib.ptr(dd.t),
              dd.t.shape[-1], line_stride, skip, _lib.stream_ptr())
    du.back()
    return u


def TDMA_Solve(u, data, axis):
    """u := A^-1 u along axis with TDMA_LU'd rows (la.pyx:77-91)."""
    return _solve("sfa_tdma_solve", u, data, axis)


def PDMA_Solve(u, data, axis):
    """la.pyx:61-75."""
    return _solve("sfa_pdma_solve", u, data, axis)


def TDMA_O_Solve(u, data, axis):
    """la.pyx:93-107."""
    return _solve("sfa_tdma_o_solve", u, data, axis)


def FDMA_Solve(u, data, axis):
    """la.pyx:165-179."""
    return _solve("sfa_fdma_solve", u, data, axis)


def HeptaDMA_Solve(u, data, axis):
    """la.pyx:181-197."""
    return _solve("sfa_heptadma_solve", u, data, axis)


def TwoDMA_Solve(u, data, axis):
    """Upper bands 0, 2 (la.pyx:45-59)."""
    return _solve("sfa_twodma_solve", u, data, axis)


def _threedma_rows(data):
    """The rows the reference's ThreeDMA back substitution reads: the +4
    coefficients from row 1 (la.pyx:574, la.py:688 read u2 = data[1, 4:]),
    placed where the kernel reads them (row 2)."""
    d = _rows(data)
    rows = d.t.clone()
    n = rows.shape[-1]
    rows[..., 2, 4:] = rows[..., 1, 4:n]
    return rows


def ThreeDMA_Solve(u, data, axis):
    """Upper bands 0, 2, 4 (la.pyx:29-43).  Reproduces the reference's reading
    of the DIA rows: the +4 coefficients come from row 1 (la.pyx:574), so the
    result equals shenfun's for any data; the class la.ThreeDMA solves the
    matrix its rows describe."""
    return _solve("sfa_threedma_solve", u, _threedma_rows(data), axis)


def DiagMA_Solve(u, data, axis):
    """la.pyx:109-163."""
    return _solve("sfa_diagma_solve", u, data, axis)


def TDMA_inner_solve(u, data):
    """One line (la.pyx:516-532)."""
    return TDMA_Solve(u, data, 0)


def PDMA_inner_solve(u, data):
    return PDMA_Solve(u, data, 0)


def TDMA_O_inner_solve(u, data):
    return TDMA_O_Solve(u, data, 0)


def FDMA_inner_solve(u, data):
    return FDMA_Solve(u, data, 0)


def HeptaDMA_inner_solve(u, data):
    return HeptaDMA_Solve(u, data, 0)


def TwoDMA_inner_solve(u, data):
    return TwoDMA_Solve(u, data, 0)


def ThreeDMA_inner_solve(u, data):
    return ThreeDMA_Solve(u, data, 0)


def DiagMA_inner_solve(u, data):
    return DiagMA_Solve(u, data, 0)


_INNER = {"TDMA_inner_solve": "sfa_tdma_solve", "PDMA_inner_solve": "sfa_pdma_solve",
          "TDMA_O_inner_solve": "sfa_tdma_o_solve", "FDMA_inner_solve": "sfa_fdma_solve",
          "ThreeDMA_inner_solve": "sfa_threedma_solve", "TwoDMA_inner_solve": "sfa_twodma_solve",
          "DiagMA_inner_solve": "sfa_diagma_solve", "HeptaDMA_inner_solve": "sfa_heptadma_solve"}


def SolverGeneric1ND_solve_data(u, data, sol, naxes, is_zero_index):
    """Per-mode banded solves along axis ``naxes`` of a 2D/3D complex u
    (la.pyx:371-426): ``data`` holds the LU'd DIA rows of every mode,
    (M0[, M1], rows, n) in C order over the non-axis dims; ``sol`` is the
    line solver (its __name__, e.g. TDMA_inner_solve, picks the band
    pattern, la.pyx:307-325); mode 0 (or (0, 0)) is skipped when
    is_zero_index.  One launch for all modes."""
    name = sol if isinstance(sol, str) else getattr(sol, "__name__", "")
    fn = _INNER.get(name)
    if fn is None:
        raise ValueError(f"no device solver for {name!r}")
    d = _rows(data)
    rows = d.t
    if name == "ThreeDMA_inner_solve":
        rows = _threedma_rows(data)
    nrow, n = rows.shape[-2], rows.shape[-1]
    return _solve(fn, u, rows, naxes, line_stride=nrow * n, skip=1 if is_zero_index else 0)


# ---------------------------------------------------------------- Chebyshev Helmholtz
def _diag(M, k, n):
    """Diagonal k of a dict-like banded matrix as a length-n float64 array
    (scalars broadcast; missing diagonals are zero; n <= 0 gives one zero)."""
    if n <= 0:
        return np.zeros(1)
    if k not in M:
        return np.zeros(n)
    return np.ascontiguousarray(np.broadcast_to(np.asarray(M[k], dtype=float), (n,))[:n])


def _per_line(s, shape, axis):
    """Scale s (scalar, or array broadcastable to shape with 1 along axis) as
    one value per line, C order over the non-axis dims."""
    ab_shape = list(shape)
    ab_shape[axis] = 1
    a = np.asarray(s, dtype=float)
    if a.ndim == 0:
        a = a.reshape([1] * len(shape))
    a = np.broadcast_to(a, ab_shape)
    return torch.from_numpy(np.ascontiguousarray(np.moveaxis(a, axis, -1)).reshape(-1)).cuda()


def _helm_diags(A, B):
    M = A.shape[0]
    return [torch.from_numpy(np.array(d, dtype=float)).cuda() for d in
            (_diag(A, 0, M), _diag(A, 2, M - 2), _diag(A, 4, M - 4), _diag(B, -2, M - 2),
             _diag(B, 0, M), _diag(B, 2, M - 2))], M


def LU_Helmholtz(A, B, A_s, B_s, neumann, d0, d1, d2, L, axis=0):
    """Factors d0, d1, d2, L (arrays of u's shape, N along axis) of
    A_s A + B_s B per line (la.pyx:708-843); A, B the Chebyshev SD/SN
    stiffness and mass matrices (dicts of diagonals), A_s/B_s scalars or
    arrays with 1 along axis."""
    outs = [_Dev(x) for x in (d0, d1, d2, L)]
    t = outs[0].t
    ax = axis % t.dim()
    diags, M = _helm_diags(A, B)
    if t.shape[ax] != M + 2:
        raise ValueError(f"factor arrays need {M + 2} entries along axis {ax}")
    alfa = _per_line(A_s, t.shape, ax)
    beta = _per_line(B_s, t.shape, ax)
    shp, st = _geom(t)
    _lib.call("sfa_helmholtz_lu", int(bool(neumann)), M + 2, *[_lib.ptr(x) for x in diags],
              _lib.ptr(alfa), _lib.ptr(beta), 1, *[_lib.ptr(o.t) for o in outs], t.dim(), shp, st,
              ax, _lib.stream_ptr())
    for o in outs:
        o.back()


def Solve_Helmholtz(b, u, neumann, d0, d1, d2, L, axis=0):
    """u = (A_s A + B_s B)^-1 b along axis with LU_Helmholtz's factors
    (la.pyx:845-973); u may be b."""
    db = _Dev(b)
    du = db if u is b else _Dev(u)
    fs = [_Dev(x) for x in (d0, d1, d2, L)]
    t = db.t
    ax = axis % t.dim()
    shp, bst = _geom(t)
    _, ust = _geom(du.t)
    _, fst = _geom(fs[0].t)
    _lib.call("sfa_helmholtz_solve", _lib.ptr(t), _lib.ptr(du.t), _dtype_code(t), t.dim(), shp, bst,
              ust, ax, int(bool(neumann)), *[_lib.ptr(f.t) for f in fs], fst, _lib.stream_ptr())
    du.back()
    return u


def _helmholtz_matvec(v, b, alfa, beta, A, B, axis, neumann):
    dv = _Dev(v)
    db = _Dev(b)
    t = dv.t
    ax = axis % t.dim()
    diags, M = _helm_diags(A, B)
    A0, A2 = diags[0], diags[1]
    if neumann:
        k = torch.arange(M - 2, dtype=torch.float64, device=A2.device)
        At = A2 / (k + 2) ** 2
    else:
        At = A2
    al = _per_line(alfa, t.shape, ax)
    be = _per_line(beta, t.shape, ax)
    shp, vst = _geom(t)
    _, cst = _geom(db.t)
    _lib.call("sfa_helmholtz_matvec", _lib.ptr(t), _lib.ptr(db.t), _dtype_code(t), t.dim(), shp, vst,
              cst, ax, int(neumann), _lib.ptr(A0), _lib.ptr(At), _lib.ptr(diags[3]),
              _lib.ptr(diags[4]), _lib.ptr(diags[5]), _lib.ptr(al), _lib.ptr(be), 1,
              _lib.stream_ptr())
    db.back()
    return b


def Helmholtz_matvec(v, b, alfa, beta, A, B, axis=0):
    """b = (alfa A + beta B) v along axis, Chebyshev SD (Matvec.pyx:656-720)."""
    return _helmholtz_matvec(v, b, alfa, beta, A, B, axis, False)


def Helmholtz_Neumann_matvec(v, b, alfa, beta, A, B, axis=0):
    """b = (alfa A + beta B) v along axis, Chebyshev SN (Matvec.pyx:722-818)."""
    return _helmholtz_matvec(v, b, alfa, beta, A, B, axis, True)


def Poisson_Solve_ADD(A, b, u, axis=0):
    """Chebyshev Dirichlet Poisson A u = b (la.pyx:614-706); A's diagonal 0
    and its row-constant tail A[2]; u /= A.scale unless it is 1."""
    db = _Dev(b)
    du = db if u is b else _Dev(u)
    t = db.t
    M = A.shape[0]
    A0 = torch.from_numpy(_diag(A, 0, M)).cuda()
    A2 = torch.from_numpy(_diag(A, 2, M - 2)).cuda()
    ax = axis % t.dim()
    shp, bst = _geom(t)
    _, ust = _geom(du.t)
    scale = float(np.asarray(getattr(A, "scale", 1.0)).reshape(-1)[0])
    _lib.call("sfa_poisson_add_solve", _lib.ptr(t), _lib.ptr(du.t), _dtype_code(t), t.dim(), shp, bst,
              ust, ax, _lib.ptr(A0), _lib.ptr(A2), scale, _lib.stream_ptr())
    du.back()
    return u


# ---------------------------------------------------------------- recursive transforms
_MATRICES = collections.OrderedDict()
_MAX_MATRICES = 64


def _recursion_matrix(x, w, a, n, analysis):
    """Device transform matrix of the recursion rows ``a`` at nodes x (and
    weights w), n modes; cached by content (the reference recomputes the
    recurrence per line and call)."""
    x = np.ascontiguousarray(x, dtype=float)
    a = np.ascontiguousarray(a, dtype=float)
    w = None if w is None else np.ascontiguousarray(w, dtype=float)
    h = hashlib.blake2b(digest_size=16)
    for arr in (x, a) + (() if w is None else (w,)):
        h.update(arr.tobytes())
    key = (h.hexdigest(), a.shape, n, analysis, torch.cuda.current_device())
    M = _MATRICES.get(key)
    if M is not None:
        _MATRICES.move_to_end(key)
        return M
    if a.ndim != 2 or a.shape[0] not in (2, 3):
        raise ValueError("a must be the (2 or 3, Mx) DIA rows of the recursion matrix")
    Nx = x.shape[0]
    xd = torch.from_numpy(x).cuda()
    ad = torch.from_numpy(a).cuda()
    wd = torch.from_numpy(w).cuda() if w is not None else None
    M = torch.empty((n, Nx) if analysis else (Nx, n), dtype=torch.float64, device="cuda")
    _lib.call("sfa_recursion_matrix", _lib.ptr(xd), _lib.ptr(wd), Nx, _lib.ptr(ad), a.shape[0],
              a.shape[1], n, int(analysis), _lib.ptr(M), _lib.stream_ptr())
    _MATRICES[key] = M
    while len(_MATRICES) > _MAX_MATRICES:
        _MATRICES.popitem(last=False)
    return M


def _contract(input_array, output_array, axis, M):
    di = _Dev(input_array)
    do = _Dev(output_array)
    t = di.t
    ax = axis % t.dim()
    shp, ist = _geom(t)
    _, ost = _geom(do.t)
    _lib.call("sfa_line_matmul", _lib.ptr(t), _lib.ptr(do.t), _dtype_code(t), t.dim(), shp, ist, ost, ax,
              do.t.shape[ax], _lib.ptr(M), _lib.stream_ptr())
    do.back()
    return output_array


def scalar_product(input_array, output_array, x, w, axis, a):
    """output_i = sum_j P_i(x_j) w_j input_j along axis, P from the recursion
    rows a (transforms.pyx:175-190, 270-323); the matrix is built once on the
    device (sfa_recursion_matrix) and applied as one batched contraction."""
    n = output_array.shape[axis]
    if input_array.shape[axis] != len(x):
        raise ValueError("input length along axis must equal the number of nodes")
    return _contract(input_array, output_array, axis, _recursion_matrix(x, w, a, n, True))


def evaluate_expansion_all(input_array, output_array, x, axis, a):
    """output_j = sum_i input_i P_i(x_j) along axis (transforms.pyx:199-267)."""
    if output_array.shape[axis] != len(x):
        raise ValueError("output length along axis must equal the number of nodes")
    n = input_array.shape[axis]
    return _contract(input_array, output_array, axis, _recursion_matrix(x, None, a, n, False))


# ---------------------------------------------------------------- Chebyshev biharmonic
def _diag_ptrs(diags):
    """Device copies of the 11 diagonals (sii, siu, siuu, ail, aii, aiu, bill,
    bil, bii, biu, biuu) and the C array of their addresses."""
    import ctypes
    ds = [d if isinstance(d, torch.Tensor) and d.is_cuda and d.dtype == torch.float64
          else torch.from_numpy(np.ascontiguousarray(np.atleast_1d(np.asarray(d, dtype=float)))).cuda()
          for d in diags]
    arr = (ctypes.c_void_p * 11)(*[d.data_ptr() for d in ds])
    return ds, ctypes.cast(arr, ctypes.c_void_p), arr


def LU_Biharmonic(a0, alfa, beta, sii, siu, siuu, ail, aii, aiu, bill, bil, bii, biu, biuu,
                  u0, u1, u2, l0, l1, axis=0):
    """Factors u0, u1, u2, l0, l1 ((2,) + u's shape with the axis length the
    even chain's) of a0 S + alfa A + beta B for the Chebyshev ShenBiharmonic
    basis (la.pyx:975-1138); alfa, beta scalars or arrays with 1 along axis."""
    outs = [_Dev(x) for x in (u0, u1, u2, l0, l1)]
    t = outs[0].t
    ndim = t.dim() - 1
    ax = axis % ndim
    N = (sii.numel() if isinstance(sii, torch.Tensor) else int(np.size(sii))) + 4
    shape = list(t.shape[1:])
    shape[ax] = N
    al = _per_line(alfa, shape, ax)
    be = _per_line(beta, shape, ax)
    ds, dptr, _keep = _diag_ptrs((sii, siu, siuu, ail, aii, aiu, bill, bil, bii, biu, biuu))
    _lib.call("sfa_biharmonic_lu", ndim, _lib.arr(shape), ax, float(a0), _lib.ptr(al), _lib.ptr(be), 1,
              dptr, *[_lib.ptr(o.t) for o in outs], None, None, _lib.stream_ptr())
    for o in outs:
        o.back()


def Biharmonic_factor_pr(ak, bk, l0, l1, axis=0):
    """ak, bk of U's tail rows from l0, l1 (la.pyx:1141-1233)."""
    outs = [_Dev(x) for x in (ak, bk)]
    d0, d1 = _Dev(l0), _Dev(l1)
    t = outs[0].t
    ndim = t.dim() - 1
    ax = axis % ndim
    shape = list(t.shape[1:])
    shape[ax] = 2 * t.shape[1 + ax] + 4
    _lib.call("sfa_biharmonic_factor", ndim, _lib.arr(shape), ax, _lib.ptr(d0.t), _lib.ptr(d1.t),
              _lib.ptr(outs[0].t), _lib.ptr(outs[1].t), _lib.stream_ptr())
    for o in outs:
        o.back()


def Biharmonic_Solve(b, u, u0, u1, u2, l0, l1, ak, bk, a0, axis=0):
    """u = (a0 S + alfa A + beta B)^-1 b along axis with the factors above
    (la.pyx:1235-1310); u may be b; entries N-4.. of u are not written."""
    db = _Dev(b)
    du = db if u is b else _Dev(u)
    fs = [_Dev(x) for x in (u0, u1, u2, l0, l1, ak, bk)]
    t = db.t
    ax = axis % t.dim()
    shp, bst = _geom(t)
    _, ust = _geom(du.t)
    _lib.call("sfa_biharmonic_solve", _lib.ptr(t), _lib.ptr(du.t), _dtype_code(t), t.dim(), shp, bst, ust,
              ax, float(a0), *[_lib.ptr(f.t) for f in fs], _lib.stream_ptr())
    du.back()
    return u


def Biharmonic_matvec(v, b, a0, alfa, beta, sii, siu, siuu, ail, aii, aiu, bill, bil, bii, biu, biuu,
                      axis=0):
    """b = (a0 S + alfa A + beta B) v along axis (Matvec.pyx:820-923)."""
    dv = _Dev(v)
    db = _Dev(b)
    t = dv.t
    ax = axis % t.dim()
    al = _per_line(alfa, t.shape, ax)
    be = _per_line(beta, t.shape, ax)
    ds, dptr, _keep = _diag_ptrs((sii, siu, siuu, ail, aii, aiu, bill, bil, bii, biu, biuu))
    shp, vst = _geom(t)
    _, cst = _geom(db.t)
    _lib.call("sfa_biharmonic_matvec", _lib.ptr(t), _lib.ptr(db.t), _dtype_code(t), t.dim(), shp, vst, cst,
              ax, float(a0), _lib.ptr(al), _lib.ptr(be), 1, dptr, _lib.stream_ptr())
    db.back()
    return b


__all__ += ["LU_Biharmonic", "Biharmonic_factor_pr", "Biharmonic_Solve", "Biharmonic_matvec"]
