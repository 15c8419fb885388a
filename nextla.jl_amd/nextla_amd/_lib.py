"""ctypes binding of libnlahip (declared in include/nla.h).

The library is the only compute path: if it is missing, every entry point
raises instead of falling back to anything else.
"""
import ctypes
import os
import threading

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.normpath(os.path.join(_HERE, "..", "lib", "libnlahip.so"))

NLA_F32 = 0
NLA_F64 = 1

_c_char = ctypes.c_char
_i32 = ctypes.c_int32
_i64 = ctypes.c_int64
_dbl = ctypes.c_double
_vp = ctypes.c_void_p

class BatchItem(ctypes.Structure):
    """nla_batch_item of include/nla.h."""
    _fields_ = [("n", ctypes.c_int64), ("m", ctypes.c_int64), ("A", ctypes.c_void_p), ("lda", ctypes.c_int64),
                ("B", ctypes.c_void_p), ("ldb", ctypes.c_int64), ("alpha", ctypes.c_double)]


NLA_MIXED_MAX_ITER = 30


class BatchTaskDesc(ctypes.Structure):
    """nla_batch_task_desc of include/nla.h."""
    _fields_ = [("launch", ctypes.c_int32), ("kind", ctypes.c_int32), ("system", ctypes.c_int64),
                ("r0", ctypes.c_int64), ("rows", ctypes.c_int64), ("c0", ctypes.c_int64), ("cols", ctypes.c_int64),
                ("k0", ctypes.c_int64), ("kn", ctypes.c_int64), ("alpha", ctypes.c_double), ("beta", ctypes.c_double)]


class ServeRequest(ctypes.Structure):
    """nla_serve_request of include/nla.h."""
    _fields_ = [("side", ctypes.c_char), ("uplo", ctypes.c_char), ("trans", ctypes.c_char), ("func", ctypes.c_char),
                ("dtype", ctypes.c_int32), ("alpha", ctypes.c_double), ("n", ctypes.c_int64), ("m", ctypes.c_int64),
                ("A", ctypes.c_void_p), ("lda", ctypes.c_int64), ("B", ctypes.c_void_p), ("ldb", ctypes.c_int64),
                ("a_key", ctypes.c_int64)]


class ServeDone(ctypes.Structure):
    """nla_serve_done of include/nla.h."""
    _fields_ = [("ticket", ctypes.c_int64), ("status", ctypes.c_int32), ("iterations", ctypes.c_int32),
                ("backward_error", ctypes.c_double)]


class ServeStats(ctypes.Structure):
    """nla_serve_stats_t of include/nla.h."""
    _fields_ = [("budget_bytes", ctypes.c_int64), ("used_bytes", ctypes.c_int64), ("queued", ctypes.c_int64),
                ("running", ctypes.c_int64), ("completed_pending", ctypes.c_int64), ("resident", ctypes.c_int64)]


NLA_SERVE_PLAIN = 0
NLA_SERVE_REFINED = 1


class MixedInfo(ctypes.Structure):
    """nla_mixed_info of include/nla.h."""
    _fields_ = [("iterations", ctypes.c_int32), ("precision", ctypes.c_int32), ("backward_error", ctypes.c_double),
                ("history", ctypes.c_double * (NLA_MIXED_MAX_ITER + 1)),
                ("backward_error_frobenius", ctypes.c_double), ("residual", ctypes.c_int32)]


# name -> (restype, argtypes); must list every function of include/nla.h
SIGNATURES = {
    "nla_abi_version": (_i32, []),
    "nla_last_error": (ctypes.c_char_p, []),
    "nla_rectrxm": (_i32, [_c_char, _c_char, _c_char, _c_char, _i32, _dbl, _i64, _i64, _vp, _i64, _vp, _i64, _vp]),
    "nla_rectrxm_noalloc": (_i32, [_c_char, _c_char, _c_char, _c_char, _i32, _dbl, _i64, _i64, _vp, _i64, _vp, _i64,
                                   _vp]),
    "nla_trsm": (_i32, [_c_char, _c_char, _c_char, _c_char, _i32, _dbl, _i64, _i64, _vp, _i64, _vp, _i64, _vp]),
    "nla_trmm": (_i32, [_c_char, _c_char, _c_char, _c_char, _i32, _dbl, _i64, _i64, _vp, _i64, _vp, _i64, _vp]),
    "nla_gemm_add": (_i32, [_i32, _i64, _i64, _i64, _vp, _i64, _vp, _i64, _vp, _i64, _vp]),
    "nla_gemm_sub": (_i32, [_i32, _i64, _i64, _i64, _vp, _i64, _vp, _i64, _vp, _i64, _vp]),
    "nla_gemm": (_i32, [_i32, _c_char, _c_char, _i64, _i64, _i64, _dbl, _vp, _i64, _vp, _i64, _dbl, _vp, _i64, _vp]),
    "nla_leaf": (_i32, [_c_char, _c_char, _c_char, _i32, _i64, _i64, _vp, _i64, _vp, _i64, _vp]),
    "nla_batch_workspace_bytes": (_i64, [_c_char, _i32, ctypes.POINTER(BatchItem), _i64]),
    "nla_batch_plan_create": (_i32, [_c_char, _c_char, _c_char, _c_char, _i32, ctypes.POINTER(BatchItem), _i64, _vp,
                                     _i64, _vp, ctypes.POINTER(_vp)]),
    "nla_batch_plan_execute": (_i32, [_vp, _vp, _i32]),
    "nla_batch_plan_destroy": (_i32, [_vp]),
    "nla_batch_mixed_workspace_bytes": (_i64, [_c_char, ctypes.POINTER(BatchItem), _i64]),
    "nla_batch_mixed_plan_create": (_i32, [_c_char, _c_char, _c_char, ctypes.POINTER(BatchItem), _i64, _vp, _i64,
                                           _vp, ctypes.POINTER(_vp)]),
    "nla_batch_mixed_plan_execute": (_i32, [_vp, _dbl, _i32, _vp, _i32, ctypes.POINTER(_dbl),
                                            ctypes.POINTER(_i32)]),
    "nla_batch_mixed_plan_destroy": (_i32, [_vp]),
    "nla_batch_plan_describe": (_i64, [_c_char, _c_char, _c_char, _c_char, _i32, _vp, _i64, _vp, _i64]),
    "nla_serve_create": (_i32, [_i64, _i32, _i32, _i32, _i32, _i32, _vp]),
    "nla_serve_set_refinement": (_i32, [_vp, _dbl, _i32, _i32]),
    "nla_serve_submit": (_i32, [_vp, _vp, _vp, _vp]),
    "nla_serve_step": (_i32, [_vp, _vp, _i64, _vp]),
    "nla_serve_drain": (_i32, [_vp]),
    "nla_serve_stats": (_i32, [_vp, _vp]),
    "nla_serve_evicted": (_i32, [_vp, _vp, _i64, _vp]),
    "nla_serve_destroy": (_i32, [_vp]),
    "nla_getrf2": (_i32, [_i32, _i64, _i64, _vp, _i64, _vp, _vp, _vp]),
    "nla_lauum_workspace_bytes": (_i64, [_i32, _i64]),
    "nla_larfb": (_i32, [_c_char, _c_char, _c_char, _c_char, _i32, _i64, _i64, _i64, _vp, _i64, _vp, _i64, _vp, _i64,
                         _vp, _i64, _vp]),
    "nla_pamm": (_i32, [_c_char, _c_char, _c_char, _c_char, _i32, _i64, _i64, _i64, _i64, _vp, _i64, _vp, _i64, _vp,
                        _i64, _vp, _i64, _vp]),
    "nla_parfb": (_i32, [_c_char, _c_char, _c_char, _c_char, _i32, _i64, _i64, _i64, _i64, _i64, _i64, _vp, _i64, _vp,
                         _i64, _vp, _i64, _vp, _i64, _vp, _i64, _vp]),
    "nla_lauum": (_i32, [_c_char, _i32, _i64, _vp, _i64, _vp, _i64, _vp]),
    "nla_mixed_workspace_bytes": (_i64, [_c_char, _i64, _i64]),
    "nla_rectrxm_mixed": (_i32, [_c_char, _c_char, _c_char, _i32, _dbl, _i64, _i64, _vp, _i64, _vp, _i64, _dbl, _i32,
                                 _vp, _i64, _vp, ctypes.POINTER(MixedInfo)]),
    "nla_residual_workspace_bytes": (_i64, [_c_char, _i32, _i64, _i64]),
    "nla_residual": (_i32, [_c_char, _c_char, _c_char, _i32, _dbl, _i64, _i64, _vp, _i64, _vp, _i64, _vp, _i64, _vp,
                            _i64, _vp, _i64, _vp]),
    "nla_set_config": (_i32, [_i32, _i32]),
    "nla_prof_enable": (_i32, [_i32]),
    "nla_prof_summary": (_i32, [_i32, ctypes.POINTER(_i64), ctypes.POINTER(_dbl), ctypes.POINTER(_dbl),
                                ctypes.POINTER(_dbl)]),
}

NLA_KIND_GEMM = 0
NLA_KIND_LEAF = 1
NLA_KIND_LP_GEMM = 2
NLA_KIND_RESIDUAL = 3
NLA_KIND_RESIDUAL_GEMM = 4
NLA_KIND_LP_PAIR_GEMM = 5
NLA_LP_BF16 = 0
NLA_LP_FP16 = 1
NLA_LP_FP32 = 2
NLA_LP_AUTO = 3
NLA_LP_FP64 = 4
NLA_LP_FP16X2 = 5
NLA_ERR_NOT_CONVERGED = 5
NLA_ERR_WORKSPACE = 4
NLA_NOALLOC_SCRATCH_BYTES = 16 * 1024 * 1024 + 64 * 1024
NLA_CFG_LEAF_MODE = 0
NLA_CFG_PANEL = 1
NLA_CFG_PANEL_ROWS_F32 = 2
NLA_CFG_PANEL_ROWS_F64 = 3
NLA_CFG_PANEL_STRIP_F32 = 4
NLA_CFG_PANEL_STRIP_F64 = 5
NLA_CFG_BATCH_CLASSES = 6
NLA_CFG_PANEL_WAVES = 7
NLA_CFG_LU_SPLIT_ROWS = 8
NLA_CFG_RESIDUAL = 9
NLA_CFG_LU_SPIN_LIMIT = 10
NLA_CFG_OZ_VARIANT = 11
NLA_CFG_BATCH_ORDER = 12
NLA_RES_FP64 = 0
NLA_RES_INT8 = 1


def set_config(key, value):
    """Process-wide method switch below the recursion; returns the old value."""
    return load().nla_set_config(key, value)


def prof_enable(on=True):
    check(load().nla_prof_enable(1 if on else 0), "nla_prof_enable")


def prof_summary(kind):
    """(launches, ms, flops, bytes) of the kernels recorded since prof_enable."""
    n, ms, fl, by = _i64(), _dbl(), _dbl(), _dbl()
    check(load().nla_prof_summary(kind, ctypes.byref(n), ctypes.byref(ms), ctypes.byref(fl), ctypes.byref(by)),
          "nla_prof_summary")
    return n.value, ms.value, fl.value, by.value

_lock = threading.Lock()
_lib = None


class NLAError(RuntimeError):
    """A HIP/runtime failure reported by libnlahip (positive status)."""


def load(path=None):
    """Load libnlahip once and attach the C signatures."""
    global _lib
    with _lock:
        if _lib is not None and path is None:
            return _lib
        p = path or LIB_PATH
        if not os.path.exists(p):
            raise RuntimeError(
                f"libnlahip not built: {p} is missing (run `make -C nextla.jl_amd` or __graft_entry__.build())")
        lib = ctypes.CDLL(p)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        if path is None:
            _lib = lib
        return lib


def check(status, what):
    """Turn a C status into the reference's error behaviour (ArgumentError /
    ErrorException in Julia; ValueError / NLAError here)."""
    if status == 0:
        return
    msg = load().nla_last_error().decode(errors="replace")
    if status < 0:
        raise ValueError(f"{what}: {msg}")
    raise NLAError(f"{what}: status {status}: {msg}")


def ch(c):
    """One-character argument as the C ABI's `char`."""
    if not isinstance(c, str) or len(c) != 1:
        raise ValueError(f"expected a single character, got {c!r}")
    return c.encode("ascii")
