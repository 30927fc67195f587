"""ctypes binding of libcarcassonne_hip.so (the C ABI declared in include/carcassonne_hip.h).

The product path has no CPU fallback: if the library is missing or fails to
load, every entry point raises LibraryMissingError.
"""
import ctypes
import os

from .errors import HipError, LibraryMissingError

LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "libcarcassonne_hip.so")

CC_OK, CC_ERR_RANK, CC_ERR_DIM, CC_ERR_NOT_CONVERGED = 0, -1, -2, -3
CC_ERR_HIP, CC_ERR_OOM, CC_ERR_ARG, CC_ERR_UNSUPPORTED = -4, -5, -6, -7
CC_F64, CC_F32, CC_F16, CC_BF16 = 0, 1, 2, 3

_vp, _i64, _i32, _f64 = ctypes.c_void_p, ctypes.c_int64, ctypes.c_int, ctypes.c_double
_pi64, _pvp = ctypes.POINTER(ctypes.c_int64), ctypes.POINTER(ctypes.c_void_p)
_pi32, _pf64 = ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_double)

# name -> argtypes (all return int status)
SIGNATURES = {
    "cc_version": [],
    "cc_permute": [_vp, _i32, _vp, _vp, _i32, _pi64, _pi64],
    "cc_gemm_f64": [_vp, _i64, _i64, _i64, _i64, _f64, _vp, _i64, _i64, _i64, _vp, _i64, _i64, _i64, _f64,
                    _vp, _i64, _i64, _i64],
    "cc_axpby_f64": [_vp, _i64, _f64, _vp, _f64, _vp],
    "cc_fill_f64": [_vp, _i64, _f64, _vp],
    "cc_set_identity_f64": [_vp, _i64, _vp],
    "cc_dot_f64": [_vp, _i64, _vp, _vp, _vp],
    "cc_count_nonfinite_f64": [_vp, _i64, _vp, _vp],
    "cc_cast": [_vp, _i64, _i32, _vp, _i32, _vp],
    "cc_binary_f64": [_vp, _i32, _i64, _vp, _vp, _vp],
    "cc_unary_f64": [_vp, _i32, _i64, _vp, _f64, _vp],
    "cc_scale_rows_cols_f64": [_vp, _i32, _i64, _i64, _vp, _vp, _vp],
    "cc_svd_work_size": [_i64, _i64, _i64, _pi64],
    "cc_svd_f64": [_vp, _i64, _i64, _i64, _vp, _vp, _vp, _vp, _vp],
    "cc_fused_truncation_work_size": [_i64, _i64, _i64, _pi64],
    "cc_fused_truncation_f64": [_vp, _i64, _i64, _i64, _i64, _vp, _vp, _vp, _vp],
    "cc_sign_normalize_columns_f64": [_vp, _i64, _i64, _vp],
    "cc_stage3_matvec_generic_work_size": [_i64, _i64, _i64, _i64, _pi64],
    "cc_stage3_matvec_generic_f64": [_vp, _i32, _pvp, _pvp, _pvp, _i64, _i64, _i64, _i64, _vp, _vp, _vp],
    "cc_stage3_fused_supported": [_i64, _i64],
    "cc_stage3_packed_size": [_i64, _i64, _pi64],
    "cc_stage3_fused_work_size": [_i64, _i64, _i64, _pi64],
    "cc_stage3_pack_f64": [_vp, _i64, _i64, _vp, _vp],
    "cc_stage3_pack_strided_f64": [_vp, _i64, _i64, _i64, _i64, _vp, _vp, _vp],
    "cc_stage3_plan_create": [ctypes.POINTER(ctypes.c_void_p), _i32, _pvp, _i32, _pvp, _i32, _pi32, _pi32, _pi32,
                              _i32, _pf64, _i64, _i64, _i64, _vp],
    "cc_stage3_plan_apply": [_vp, _vp, _vp, _vp],
    "cc_stage3_plan_destroy": [_vp],
    "cc_profile_begin": [],
    "cc_profile_begin_sampled": [_i32],
    "cc_profile_end": [_i32, _pf64, _pi64],
    "cc_linalg_work_size": [_i64, _pi64],
    "cc_cholesky_f64": [_vp, _i64, _vp, _vp, _vp],
    "cc_trtri_lower_f64": [_vp, _i64, _vp, _vp, _vp],
    "cc_cholesky_inverse_f64": [_vp, _i64, _vp, _vp, _vp, _vp],
    "cc_transpose_f64": [_vp, _i64, _vp, _vp],
    "cc_skinny_f64": [_vp, _i64, _i32, _vp, _i32, _vp, _vp],
    "cc_lanczos_work_size": [_i64, _i32, _pi64],
    "cc_lanczos_f64": [_vp, _vp, _i64, _i64, _vp, _vp, _vp, _vp, _vp, _f64, _f64, _i32, _i32, _vp, _vp, _pf64],
    "cc_stage3_lp_slab_size": [_i64, _i64, _pi64],
    "cc_stage3_lp_convert": [_vp, _i32, _i64, _i64, _vp, _vp, _vp],
    "cc_stage3_lp_convert_strided": [_vp, _i32, _i64, _i64, _i64, _i64, _vp, _i64, _vp, _vp, _vp],
    "cc_stage3_lp_plan_chunks": [_vp, _pi32],
    "cc_stage3_lp_plan_diagonal_form": [_vp, _pi32],
    "cc_stage3_lp_plan_rescale": [_vp, _vp, _vp],
    "cc_stage3_packed_size_f32": [_i64, _i64, _i32, _pi64],
    "cc_calibration_lds_read": [_vp, _i64, _vp, _vp, _i32],
    "cc_calibration_mfma": [_vp, _i32, _i32, _i32, _vp],
    "cc_stage3_pack_strided_f32": [_vp, _i32, _i64, _i64, _i64, _i64, _vp, _vp, _vp],
    "cc_stage3_plan_create_f32": [ctypes.POINTER(ctypes.c_void_p), _i32, _pvp, _i32, _pvp, _i32, _pi32, _pi32, _pi32,
                                  _i32, _pf64, _i64, _i64, _i64, _vp],
    "cc_stage3_lp_work_size": [_i64, _i64, _i64, _i32, _pi64],
    "cc_stage3_lp_plan_create": [ctypes.POINTER(ctypes.c_void_p), _i32, _i32, _pvp, _i32, _pvp, _i32, _pi32, _pi32,
                                 ctypes.POINTER(ctypes.c_float), _i64, _i64, _i64, _i32, _i32, _vp, _vp, _vp],
    "cc_stage3_lp_plan_apply": [_vp, _vp, _vp, _vp],
    "cc_stage3_lp_plan_destroy": [_vp],
    "cc_lanczos_lp": [_vp, _vp, _i64, _i64, _vp, _vp, _vp, _vp, _vp, _f64, _f64, _i32, _i32, _vp, _vp, _pf64],
    "cc_basis_dots_f64": [_vp, _i32, _i64, _vp, _vp, _vp, _i32],
    "cc_basis_combine_f64": [_vp, _i32, _i64, _vp, _vp, _vp],
    "cc_small_geneig_f64": [_vp, _i32, _vp, _vp, _vp, _vp],
    "cc_jd_work_size": [_i64, _pi64],
    "cc_jd_correction_lp": [_vp, _vp, _i64, _i64, _vp, _vp, _vp, _vp, _vp, _vp, _i32, _vp, _vp],
    "cc_normalize_f64": [_vp, _i64, _vp, _vp],
    "cc_axpby_dev_f64": [_vp, _i64, _vp, _f64, _vp, _f64, _vp],
    "cc_mlp_param_count": [_i64, _i64, _i64, _pi64],
    "cc_mlp_work_size": [_i64, _i64, _i64, _pi64],
    "cc_mlp_features": [_vp, _i32, _i32, _vp, _f64, _i64, _vp],
    "cc_mlp_forward": [_vp, _i64, _i64, _i64, _i64, _vp, _vp, _vp, _vp, _vp, _vp],
    "cc_mlp_backward": [_vp, _i64, _i64, _i64, _i64, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp],
    "cc_mlp_adam": [_vp, _i64, _vp, _vp, _vp, _vp, _vp, _f64, _f64, _f64, _f64, _i64],
    "cc_stage3_matvec_fused_f64": [_vp, _i32, _pvp, _i32, _pvp, _i32, _pi32, _pi32, _pi32, _i32, _pf64, _i64, _i64,
                                   _i64, _vp, _vp, _vp],
    "cc_stage3_matvec_fused_f64_noalloc": [_vp, _i32, _pvp, _i32, _pvp, _i32, _pi32, _pi32, _pi32, _i32, _pf64, _i64,
                                           _i64, _i64, _vp, _vp],
}

_lib = None
_load_error = None


def load():
    global _lib, _load_error
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        _load_error = "native library not built: " + LIB_PATH + " (run __graft_entry__.build())"
        raise LibraryMissingError(_load_error)
    try:
        lib = ctypes.CDLL(LIB_PATH)
    except OSError as e:
        _load_error = str(e)
        raise LibraryMissingError("cannot load {}: {}".format(LIB_PATH, e))
    lib.cc_last_error.restype = ctypes.c_char_p
    lib.cc_last_error.argtypes = []
    for name, args in SIGNATURES.items():
        f = getattr(lib, name)
        f.restype = ctypes.c_int
        f.argtypes = args
    _lib = lib
    return lib


def exported_symbols():
    return list(SIGNATURES) + ["cc_last_error"]


def check(rc, what=""):
    if rc == CC_OK:
        return
    msg = load().cc_last_error().decode(errors="replace")
    from .errors import RelaxFailed
    if rc == CC_ERR_ARG or rc == CC_ERR_DIM or rc == CC_ERR_RANK:
        raise ValueError("{}: {}".format(what, msg))
    if rc == CC_ERR_NOT_CONVERGED:
        raise RelaxFailed(float("nan"), float("nan"))
    if rc == CC_ERR_UNSUPPORTED:
        raise NotImplementedError("{}: {}".format(what, msg))
    if rc == CC_ERR_OOM:
        raise MemoryError("{}: {}".format(what, msg))
    raise HipError("{} failed ({}): {}".format(what, rc, msg))


def call(name, *args):
    lib = load()
    check(getattr(lib, name)(*args), name)


def i64_array(values):
    values = [int(v) for v in values]
    return (ctypes.c_int64 * max(1, len(values)))(*values)


def i32_array(values):
    values = [int(v) for v in values]
    return (ctypes.c_int * max(1, len(values)))(*values)


def f64_array(values):
    values = [float(v) for v in values]
    return (ctypes.c_double * max(1, len(values)))(*values)


def f32_array(values):
    values = [float(v) for v in values]
    return (ctypes.c_float * max(1, len(values)))(*values)


def query(name, *args):
    """Call a size / capability query whose last argument is an int64 out-parameter."""
    out = i64_array([0])
    call(name, *args, out)
    return int(out[0])


def ptr_array(ptrs):
    return (ctypes.c_void_p * max(1, len(ptrs)))(*[ctypes.c_void_p(p) if p else None for p in ptrs])
