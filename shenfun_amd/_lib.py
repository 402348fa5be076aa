"""ctypes binding of libshenfun_amd.so (the C ABI declared in include/shenfun_amd.h).

The library is loaded lazily on first use, after torch (so both share one HIP
runtime).  There is no fallback: a missing or stale library raises.
"""
import ctypes
import os

import torch  # noqa: F401  (load torch's HIP runtime first)

_HERE = os.path.dirname(os.path.abspath(__file__))
LIBPATH = os.path.join(_HERE, "_lib", "libshenfun_amd.so")

F64 = 0
C128 = 1

_c_void_p = ctypes.c_void_p
_i64 = ctypes.c_int64
_int = ctypes.c_int
_dbl = ctypes.c_double
_p64 = ctypes.POINTER(ctypes.c_int64)

# name -> argtypes ('v' void*, 'i' int, 'l' int64, 'd' double, 'a' int64 array)
_SIGS = {
    "sfa_fft_c2c": "vviaaaiidv",
    "sfa_fft_r2c": "vviaaaidv",
    "sfa_fft_c2r": "vviaaaidv",
    "sfa_cheb_scalar_product": "vviiaaaivv",
    "sfa_cheb_backward": "vviiaaaivv",
    "sfa_cheb_scalar_product_gl": "vviiaaaivv",
    "sfa_cheb_backward_gl": "vviiaaaivv",
    "sfa_legendre_matrix": "vvllvivv",
    "sfa_line_matmul": "vviiaaailvv",
    "sfa_recursion_matrix": "vvlvlllivv",
    "sfa_tdma_lu": "vllv",
    "sfa_pdma_lu": "vllv",
    "sfa_tdma_solve": "viiaaivlllv",
    "sfa_pdma_solve": "viiaaivlllv",
    "sfa_diagma_solve": "viiaaivlllv",
    "sfa_pad_axis": "vviiaaaaiiillllv",
    "sfa_boundary_update": "vviiaaaailllvvlv",
    "sfa_pointwise_mul": "vvvilv",
    "sfa_axpy": "vvdllv",
    "sfa_helmholtz3d_set_lift": "vv",
    "sfa_convolve_direct": "vvvliv",
    "sfa_band_matvec": "vviiaaaaillivvviv",
    "sfa_tensor_eval": "vlllvvvvlv",
    "sfa_band_lu": "vliilv",
    "sfa_band_solve": "viiaaivllliiv",
    "sfa_tdma_o_lu": "vllv",
    "sfa_fdma_lu": "vllv",
    "sfa_heptadma_lu": "vllv",
    "sfa_tdma_o_solve": "viiaaivlllv",
    "sfa_twodma_solve": "viiaaivlllv",
    "sfa_threedma_solve": "viiaaivlllv",
    "sfa_fdma_solve": "viiaaivlllv",
    "sfa_heptadma_solve": "viiaaivlllv",
    "sfa_helmholtz_lu": "ilvvvvvvvvlvvvviaaiv",
    "sfa_helmholtz_solve": "vviiaaaiivvvvav",
    "sfa_helmholtz_matvec": "vviiaaaiivvvvvvvlv",
    "sfa_poisson_add_solve": "vviiaaaivvdv",
    "sfa_helmholtz3d_plan": "llldivlv",
    "sfa_helmholtz3d": "vvllldiiivvlv",
    "sfa_helmholtz3d_many": "vvlllldiiivvlv",
    "sfa_helmholtz3d_stage_times": "vvllldiiivvlivv",
    "sfa_helmholtz3d_guarded": "vvlllldiidvvlvvvv",
    "sfa_helmholtz3d_plan_release": "v",
    "sfa_helmholtz3d_noalloc": "vvllldiiivv",
    "sfa_biharmonic_lu": "iaidvvlvvvvvvvvv",
    "sfa_biharmonic_factor": "iaivvvvv",
    "sfa_biharmonic_solve": "vviiaaaidvvvvvvvv",
    "sfa_biharmonic_matvec": "vviiaaaidvvlvv",
    "sfa_batch2d_plan": "ivvvvvvllivlv",
    "sfa_batch2d_solve": "vvvvliv",
    "sfa_batch2d_release": "v",
    "sfa_batch2d_leg_matrices": "iivv",
    "sfa_batch2d_dft_matrices": "ivv",
    "sfa_batch2d_plan_ext": "ivvvvvvllvvvlivlv",
    "sfa_batch2d_wide_plan": "ivvvvvvllvvvlvlv",
    "sfa_batch2d_wide_solve": "vvvvlv",
    "sfa_mfma_f64_layout_probe": "vv",
}

_TYPES = {"v": _c_void_p, "i": _int, "l": _i64, "d": _dbl, "a": _p64}

_lib = None


def lib():
    """Return the loaded library (loads it on first call)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIBPATH):
            raise RuntimeError(
                f"shenfun_amd HIP library not found at {LIBPATH}; build it with "
                "`python -c 'import __graft_entry__ as g; g.build()'` (no CPU fallback exists)")
        L = ctypes.CDLL(LIBPATH)
        for name, sig in _SIGS.items():
            fn = getattr(L, name)
            fn.argtypes = [_TYPES[c] for c in sig]
            fn.restype = _int
        L.sfa_last_error.restype = ctypes.c_char_p
        L.sfa_last_error.argtypes = []
        L.sfa_version.restype = _int
        L.sfa_helmholtz3d_workspace.restype = _i64
        L.sfa_helmholtz3d_workspace.argtypes = [_i64, _i64, _i64, _int]
        L.sfa_helmholtz3d_noalloc_max_bytes.restype = _i64
        L.sfa_helmholtz3d_noalloc_max_bytes.argtypes = []
        L.sfa_helmholtz3d_plan_bytes.restype = _i64
        L.sfa_helmholtz3d_plan_bytes.argtypes = [_i64, _i64, _i64]
        L.sfa_batch2d_plan_bytes.restype = _i64
        L.sfa_batch2d_plan_bytes.argtypes = [_int, _c_void_p, _c_void_p, _c_void_p, _int]
        L.sfa_batch2d_workspace.restype = _i64
        L.sfa_batch2d_workspace.argtypes = [_int, _c_void_p, _c_void_p, _int]
        L.sfa_batch2d_variants.restype = _int
        L.sfa_batch2d_variants.argtypes = [_c_void_p, _c_void_p, _int]
        for name in ("sfa_batch2d_leg_elems", "sfa_batch2d_dft_elems"):
            getattr(L, name).restype = _i64
            getattr(L, name).argtypes = [_int]
        L.sfa_batch2d_plan_ext_bytes.restype = _i64
        L.sfa_batch2d_plan_ext_bytes.argtypes = [_int, _int]
        for name in ("sfa_batch2d_wide_plan_bytes", "sfa_batch2d_wide_workspace"):
            getattr(L, name).restype = _i64
            getattr(L, name).argtypes = [_int, _c_void_p, _c_void_p]
        _lib = L
    return _lib


def exported_symbols():
    return sorted(list(_SIGS) + ["sfa_version", "sfa_last_error", "sfa_helmholtz3d_workspace",
                                 "sfa_helmholtz3d_plan_bytes", "sfa_helmholtz3d_noalloc_max_bytes",
                                 "sfa_batch2d_plan_bytes",
                                 "sfa_batch2d_workspace", "sfa_batch2d_variants",
                                 "sfa_batch2d_leg_elems", "sfa_batch2d_dft_elems",
                                 "sfa_batch2d_plan_ext_bytes", "sfa_batch2d_wide_plan_bytes",
                                 "sfa_batch2d_wide_workspace"])


def arr(vals):
    """int64 C array from a sequence."""
    vals = [int(v) for v in vals]
    return (ctypes.c_int64 * len(vals))(*vals)


def ptr(t):
    """Device pointer of a tensor (None -> NULL)."""
    if t is None:
        return None
    return ctypes.c_void_p(t.data_ptr())


def stream_ptr(stream=None):
    if stream is None:
        stream = torch.cuda.current_stream()
    return ctypes.c_void_p(stream.cuda_stream)


def call(name, *args):
    rc = getattr(lib(), name)(*args)
    if rc != 0:
        msg = lib().sfa_last_error().decode()
        raise RuntimeError(f"{name} failed ({rc}): {msg}")
    return rc
