"""FP64-accurate triangular solve by low-precision correction solves and
FP64 iterative refinement (nla_rectrxm_mixed of include/nla.h).

The FP64 solution it converges to is the one unified_rectrxm!(side, uplo,
transpose, alpha, 'S', A, B) computes (src/rectrxm.jl:43-76); the correction
solves use the same recursion on an fp32 copy of A with bf16 / fp16 / fp32
GEMM updates, and each step's FP64 residual alpha*B - op(A)*X decides when to
stop: the normwise backward error of every system (column of B for side 'L',
row for 'R'), max_j ||r_j|| / (||A||_lb ||x_j|| + ||alpha*b_j||) with ||A||_lb
the largest column 2-norm of A (<= ||A||_2); the whole-matrix Frobenius form
is reported beside it as backward_error_frobenius.

precision "fp16x2" updates with fp16 pairs (each scaled fp32 operand as
hi + lo, three fp16 MFMA products: about 22 significant bits at the fp16
kernel's cost).  precision "fp64" takes no low-precision step: X0 is the
plain FP64 solve,
certified by the same residual and test (corrections, if ever needed, in
fp32).
"""
import ctypes

import torch

from . import _lib, _canonical, _require_device, _stream

PRECISIONS = {"bf16": _lib.NLA_LP_BF16, "fp16": _lib.NLA_LP_FP16, "fp32": _lib.NLA_LP_FP32,
              "auto": _lib.NLA_LP_AUTO, "fp64": _lib.NLA_LP_FP64, "fp16x2": _lib.NLA_LP_FP16X2}
_NAMES = {v: k for k, v in PRECISIONS.items()}
RESIDUALS = {"fp64": _lib.NLA_RES_FP64, "int8": _lib.NLA_RES_INT8}
RESIDUALS_NAMES = {v: k for k, v in RESIDUALS.items()}


def workspace_bytes(side, n, m):
    """Device workspace nla_rectrxm_mixed needs for an n x n A and B with
    other dimension m."""
    return int(_lib.load().nla_mixed_workspace_bytes(_lib.ch(side.upper()), int(n), int(m)))


def rectrxm_mixed_(side, uplo, transpose, alpha, A, B, precision="fp16", tol=1e-14, max_iter=20,
                   workspace=None, raise_on_stall=True):
    """Solve op(A)*X = alpha*B (side 'L') or X*op(A) = alpha*B (side 'R') in
    place to FP64 accuracy.  A, B: float64 tensors on one ROCm device (any
    layout _canonical accepts).  Returns (B, info) with info a dict of
    iterations, final precision, backward error and its history.

    workspace: optional uint8 device tensor of at least workspace_bytes()
    bytes (256-byte aligned start); allocated here when omitted."""
    for k, (v, ok) in enumerate(((side, "LRlr"), (uplo, "LUlu"), (transpose, "NTCntc")), 1):
        if not isinstance(v, str) or len(v) != 1 or v not in ok:
            raise ValueError(f"argument {k} has invalid value {v!r}")
    if precision not in PRECISIONS:
        raise ValueError(f"precision must be one of {sorted(PRECISIONS)}, got {precision!r}")
    if A.dtype != torch.float64 or B.dtype != torch.float64:
        raise ValueError("the refined solve takes float64 A and B")
    _require_device(A, B)
    side, uplo, tr, n, m, lda, ldb = _canonical(side.upper(), uplo.upper(), transpose.upper(), A, B)
    need = workspace_bytes(side, n, m)
    if workspace is None:
        workspace = torch.empty(need + 256, dtype=torch.uint8, device=B.device)
    base = workspace.data_ptr()
    off = (-base) % 256
    if workspace.numel() - off < need:
        raise ValueError(f"workspace holds {workspace.numel() - off} aligned bytes, {need} needed")
    info = _lib.MixedInfo()
    st = _lib.load().nla_rectrxm_mixed(_lib.ch(side), _lib.ch(uplo), _lib.ch("T" if tr else "N"),
                                       PRECISIONS[precision], float(alpha), n, m, A.data_ptr(), lda, B.data_ptr(),
                                       ldb, float(tol), int(max_iter), base + off, need, _stream(B),
                                       ctypes.byref(info))
    out = {"iterations": info.iterations, "precision": _NAMES.get(info.precision, info.precision),
           "residual": RESIDUALS_NAMES.get(info.residual, info.residual),
           "backward_error": info.backward_error, "backward_error_frobenius": info.backward_error_frobenius,
           "history": [info.history[i] for i in range(min(info.iterations, _lib.NLA_MIXED_MAX_ITER) + 1)]}
    if st == _lib.NLA_ERR_NOT_CONVERGED and not raise_on_stall:
        out["converged"] = False
        return B, out
    _lib.check(st, "rectrxm_mixed!")
    out["converged"] = True
    return B, out


def residual_workspace_bytes(side, method, n, m):
    """Device workspace nla_residual needs (method "fp64" or "int8")."""
    return int(_lib.load().nla_residual_workspace_bytes(_lib.ch(side.upper()), RESIDUALS[method], int(n), int(m)))


def residual(side, uplo, transpose, alpha, A, X, B, R, method="int8", workspace=None):
    """R = alpha*B - op(A)*X (side 'L') or alpha*B - X*op(A) (side 'R') to
    FP64 accuracy (nla_residual of include/nla.h).  method "int8": int8
    modular products, which first replace X by its value rounded to >= 55
    bits of each system's largest entry; "fp64": the FP64 MFMA product.  A,
    X, B, R: float64 tensors on one ROCm device, X, B and R column-major of one shape (A in any layout _canonical accepts).
    Returns R."""
    if method not in RESIDUALS:
        raise ValueError(f"method must be one of {sorted(RESIDUALS)}, got {method!r}")
    for t in (A, X, B, R):
        if t.dtype != torch.float64:
            raise ValueError("the residual takes float64 tensors")
    _require_device(A, X, B, R)
    side_u = side.upper()
    side_c, uplo_c, tr, n, m, lda, ldb = _canonical(side_u, uplo.upper(), transpose.upper(), A, B)
    for name, t in (("B", B), ("X", X), ("R", R)):
        if t.shape != B.shape or t.stride(0) != 1:
            raise ValueError(f"{name} must be column-major with B's shape")
    need = residual_workspace_bytes(side_c, method, n, m)
    if need < 0:
        raise ValueError(f"the {method} residual does not take n = {n}")
    if workspace is None:
        workspace = torch.empty(need + 256, dtype=torch.uint8, device=B.device)
    base = workspace.data_ptr()
    off = (-base) % 256
    if workspace.numel() - off < need:
        raise ValueError(f"workspace holds {workspace.numel() - off} aligned bytes, {need} needed")
    st = _lib.load().nla_residual(_lib.ch(side_c), _lib.ch(uplo_c), _lib.ch("T" if tr else "N"), RESIDUALS[method],
                                  float(alpha), n, m, A.data_ptr(), lda, X.data_ptr(), max(1, X.stride(1)),
                                  B.data_ptr(), ldb, R.data_ptr(), max(1, R.stride(1)), base + off, need,
                                  _stream(B))
    _lib.check(st, "residual")
    return R
