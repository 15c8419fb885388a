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
"""
import ctypes

import torch

from . import _lib, _canonical, _require_device, _stream

PRECISIONS = {"bf16": _lib.NLA_LP_BF16, "fp16": _lib.NLA_LP_FP16, "fp32": _lib.NLA_LP_FP32,
              "auto": _lib.NLA_LP_AUTO}
_NAMES = {v: k for k, v in PRECISIONS.items()}


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
           "backward_error": info.backward_error, "backward_error_frobenius": info.backward_error_frobenius,
           "history": [info.history[i] for i in range(min(info.iterations, _lib.NLA_MIXED_MAX_ITER) + 1)]}
    if st == _lib.NLA_ERR_NOT_CONVERGED and not raise_on_stall:
        out["converged"] = False
        return B, out
    _lib.check(st, "rectrxm_mixed!")
    out["converged"] = True
    return B, out
