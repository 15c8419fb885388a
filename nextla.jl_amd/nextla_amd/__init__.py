"""Host-side mirror of NextLA.jl's triangular solve / multiply interface on
MI355X (gfx950), backed by libnlahip.

Names follow the reference with Julia's `!` spelled `_`:

    unified_rectrxm_(side, uplo, transpose, alpha, func, A, B)  src/rectrxm.jl:43
    GEMM_ADD_(A, B, C)   C += A*B                               src/matmul.jl:69
    GEMM_SUB_(A, B, C)   A -= B*C                               src/matmul.jl:76
    LeftLowerTRSM_(A, B) ... RightUpperTRSM_(A, B)              src/trsm.jl:128-150
    LeftLowerTRMM_(A, B) ... RightUpperTRMM_(A, B)              src/trmm.jl:332-393
    trsm / trmm(side, uplo, transa, diag, A, B, alpha)          src/trsm.jl:186, src/trmm.jl:430
    getrf2_(A, ipiv=None) -> (A, ipiv, info)                    src/lu.jl:185-300
    lauum_(uplo, n, A, ib=None)                                 src/lauum.jl:52
    larfb_(side, trans, direct, storev, V, T, C, work=None)     src/larfb.jl:37,166

A and B are 2-D torch tensors on a ROCm device, float32 or float64.  Julia
matrices are column-major; a tensor with stride (1, ld) is used as is, and a
row-major tensor (stride (ld, 1)) is handled through the transpose identities
(X*op(A) = B  <=>  op(A)^T X^T = B^T), so no copy is made either way.  B is
updated in place and returned.  Work is queued on torch's current stream of
B's device; call torch.cuda.synchronize() (the reference's
KernelAbstractions.synchronize) before reading results on the host.

Deliberate differences from the reference: arguments are validated before any
device work (invalid characters or shapes raise ValueError and leave B
untouched), lower-case characters are accepted, and alpha == 0 sets B to zero
without reading it (BLAS semantics).
"""
import torch

from . import _lib
from ._lib import NLA_F32, NLA_F64, NLAError, load

__all__ = [
    "unified_rectrxm_", "GEMM_ADD_", "GEMM_SUB_",
    "LeftLowerTRSM_", "LeftUpperTRSM_", "RightLowerTRSM_", "RightUpperTRSM_",
    "LeftLowerTRMM_", "LeftUpperTRMM_", "RightLowerTRMM_", "RightUpperTRMM_",
    "trsm", "trmm", "getrf2_", "lauum_", "larfb_", "pamm_", "parfb_", "NLAError", "load", "NLA_F32", "NLA_F64", "colmajor",
]

_DT = {torch.float32: NLA_F32, torch.float64: NLA_F64}


def colmajor(t):
    """A column-major (Julia-layout) copy of a 2-D tensor."""
    return t.t().contiguous().t()


def _layout(t, name):
    """(transposed, ld): transposed=True when memory holds t^T column-major."""
    if t.dim() != 2:
        raise ValueError(f"{name} must be a matrix, got {t.dim()} dims")
    r, c = t.shape
    if r == 0 or c == 0:  # empty: strides are meaningless, nothing is read
        return False, max(r, 1)
    s0, s1 = t.stride()
    if s0 == 1 and (c <= 1 or s1 >= max(1, r)):
        return False, max(s1 if c > 1 else r, 1)
    if s1 == 1 and (r <= 1 or s0 >= max(1, c)):
        return True, max(s0 if r > 1 else c, 1)
    raise ValueError(f"{name} must have unit stride in one dimension, got strides {t.stride()}")


def _stream(t):
    """hipStream_t handle of torch's current stream on t's device."""
    return torch.cuda.current_stream(t.device).cuda_stream


def _require_device(*ts):
    """All operands on one ROCm device: host tensors are rejected, never
    computed on the CPU."""
    d = ts[0].device
    if not all(t.is_cuda and t.device == d for t in ts):
        raise ValueError("all operands must live on the same ROCm device")


def _check_pair(A, B):
    if A.dtype not in _DT:
        raise ValueError(f"unsupported element type {A.dtype} (Float32/Float64 only)")
    if B.dtype != A.dtype:
        raise ValueError(f"A is {A.dtype} but B is {B.dtype}")
    _require_device(A, B)


def _flip(c, a, b):
    return b if c == a else a


def unified_rectrxm_(side, uplo, transpose, alpha, func, A, B):
    """unified_rectrxm!(side, uplo, transpose, alpha, func, A, B) of
    src/rectrxm.jl:43-76.

    func 'S': solve op(A)*X = alpha*B ('L') or X*op(A) = alpha*B ('R').
    func 'M': B <- alpha*op(A)*B ('L') or alpha*B*op(A) ('R').
    Returns B (overwritten)."""
    for k, (v, ok) in enumerate(((side, "LRlr"), (uplo, "LUlu"), (transpose, "NTCntc"), (func, "SMsm")), 1):
        if not isinstance(v, str) or len(v) != 1 or v not in ok:
            raise ValueError(f"argument {k} has invalid value {v!r}")
    side, uplo, transpose, func = side.upper(), uplo.upper(), transpose.upper(), func.upper()
    _check_pair(A, B)
    side, uplo, tr, n, m, lda, ldb = _canonical(side, uplo, transpose, A, B)
    st = _lib.load().nla_rectrxm(_lib.ch(side), _lib.ch(uplo), _lib.ch("T" if tr else "N"), _lib.ch(func),
                                 _DT[A.dtype], float(alpha), n, m, A.data_ptr(), lda, B.data_ptr(), ldb,
                                 _stream(B))
    _lib.check(st, "unified_rectrxm!")
    return B


def _canonical(side, uplo, transpose, A, B):
    """The column-major call equivalent to (side, uplo, transpose) on the
    semantic matrices A, B whatever their memory layout:
    (side, uplo, transposed, n, m, lda, ldb) of the stored matrices."""
    n = A.shape[0]
    if A.dim() != 2 or A.shape[1] != n:
        raise ValueError(f"A must be square, got {tuple(A.shape)}")
    if B.dim() != 2:
        raise ValueError("B must be a matrix")
    if (side == "L" and B.shape[0] != n) or (side == "R" and B.shape[1] != n):
        raise ValueError(f"B {tuple(B.shape)} does not match A {tuple(A.shape)} for side {side!r}")
    a_t, lda = _layout(A, "A")
    b_t, ldb = _layout(B, "B")
    tr = transpose != "N"
    if a_t:  # memory holds A^T: same op with the transpose toggled, uplo flipped
        uplo = _flip(uplo, "L", "U")
        tr = not tr
    if b_t:  # memory holds B^T: solve/multiply on the transposed equation
        side = _flip(side, "L", "R")
        tr = not tr
    # the stored column-major matrix is n x m for side 'L', m x n for side 'R'
    rc, cc = (B.shape[1], B.shape[0]) if b_t else (B.shape[0], B.shape[1])
    m = cc if side == "L" else rc
    return side, uplo, tr, n, m, lda, ldb


def _gemm(alpha, X, Y, beta, C, what):
    """C = beta*C + alpha*X*Y on semantic matrices of any supported layout."""
    for t, nm in ((X, "X"), (Y, "Y"), (C, "C")):
        if t.dtype not in _DT or t.dtype != C.dtype:
            raise ValueError(f"{what}: {nm} has element type {t.dtype}, expected {C.dtype} (Float32/Float64)")
    _require_device(X, Y, C)
    N, R = X.shape
    R2, M = Y.shape
    if R2 != R or tuple(C.shape) != (N, M):
        raise ValueError(f"{what}: shapes {tuple(C.shape)} != {tuple(X.shape)} * {tuple(Y.shape)}")
    xt, ldx = _layout(X, "X")
    yt, ldy = _layout(Y, "Y")
    ct, ldc = _layout(C, "C")
    if not ct:  # C = b*C + a*op(X)*op(Y)
        args = ("T" if xt else "N", "T" if yt else "N", N, M, R, X, ldx, Y, ldy)
    else:  # C^T = b*C^T + a*Y^T*X^T
        args = ("N" if yt else "T", "N" if xt else "T", M, N, R, Y, ldy, X, ldx)
    tx, ty, mm, nn, kk, P, ldp, Q, ldq = args
    st = _lib.load().nla_gemm(_DT[C.dtype], _lib.ch(tx), _lib.ch(ty), mm, nn, kk, float(alpha), P.data_ptr(), ldp,
                              Q.data_ptr(), ldq, float(beta), C.data_ptr(), ldc, _stream(C))
    _lib.check(st, what)
    return C


def GEMM_ADD_(A, B, C):
    """GEMM_ADD!(A, B, C) of src/matmul.jl:69-73: C += A*B; returns C."""
    return _gemm(1.0, A, B, 1.0, C, "GEMM_ADD!")


def GEMM_SUB_(A, B, C):
    """GEMM_SUB!(A, B, C) of src/matmul.jl:76-80: A -= B*C; returns A."""
    return _gemm(-1.0, B, C, 1.0, A, "GEMM_SUB!")


def _leaf(func, side, uplo):
    def f(A, B):
        return unified_rectrxm_(side, uplo, "N", 1.0, func, A, B)
    return f


LeftLowerTRSM_ = _leaf("S", "L", "L")
LeftUpperTRSM_ = _leaf("S", "L", "U")
RightLowerTRSM_ = _leaf("S", "R", "L")
RightUpperTRSM_ = _leaf("S", "R", "U")
LeftLowerTRMM_ = _leaf("M", "L", "L")
LeftUpperTRMM_ = _leaf("M", "L", "U")
RightLowerTRMM_ = _leaf("M", "R", "L")
RightUpperTRMM_ = _leaf("M", "R", "U")

for _f, _doc in ((LeftLowerTRSM_, "A\\B, A lower (src/trsm.jl:128)"), (LeftUpperTRSM_, "A\\B, A upper (src/trsm.jl:134)"),
                 (RightLowerTRSM_, "B/A, A lower (src/trsm.jl:140)"), (RightUpperTRSM_, "B/A, A upper (src/trsm.jl:146)"),
                 (LeftLowerTRMM_, "A*B, A lower (src/trmm.jl:332)"), (LeftUpperTRMM_, "A*B, A upper (src/trmm.jl:349)"),
                 (RightLowerTRMM_, "B*A, A lower (src/trmm.jl:366)"), (RightUpperTRMM_, "B*A, A upper (src/trmm.jl:383)")):
    _f.__doc__ = f"In-place {_doc}; any n (the reference leaf is limited to one block)."


def _trxm_diag(func, side, uplo, transa, diag, A, B, alpha):
    for k, (v, ok) in enumerate(((side, "LRlr"), (uplo, "LUlu"), (transa, "NTCntc"), (diag, "NUnu")), 1):
        if not isinstance(v, str) or len(v) != 1 or v not in ok:
            raise ValueError(f"argument {k} has invalid value {v!r}")
    _check_pair(A, B)
    side, uplo, tr, n, m, lda, ldb = _canonical(side.upper(), uplo.upper(), transa.upper(), A, B)
    fn = _lib.load().nla_trsm if func == "S" else _lib.load().nla_trmm
    st = fn(_lib.ch(side), _lib.ch(uplo), _lib.ch("T" if tr else "N"), _lib.ch(diag.upper()), _DT[A.dtype],
            float(alpha), n, m, A.data_ptr(), lda, B.data_ptr(), ldb, _stream(B))
    _lib.check(st, "trsm" if func == "S" else "trmm")
    return B


def trsm(side, uplo, transa, diag, A, B, alpha=1.0):
    """trsm(side, uplo, transa, diag, A, B, alpha) of src/trsm.jl:186-204:
    solve op(A)*X = alpha*B ('L') or X*op(A) = alpha*B ('R') in place.  Unlike
    the reference wrapper (which ignores both), transa and diag are honoured:
    diag 'U' takes A's diagonal as 1 without reading it."""
    return _trxm_diag("S", side, uplo, transa, diag, A, B, alpha)


def trmm(side, uplo, transa, diag, A, B, alpha=1.0):
    """trmm(side, uplo, transa, diag, A, B, alpha) of src/trmm.jl:430-447:
    B <- alpha*op(A)*B ('L') or alpha*B*op(A) ('R'), transa and diag
    honoured (diag 'U': unit diagonal, not read)."""
    return _trxm_diag("M", side, uplo, transa, diag, A, B, alpha)


def getrf2_(A, ipiv=None):
    """getrf2!(A, ipiv, info) of src/lu.jl:185-300: LU factorization with
    partial pivoting, A = P*L*U, in place (unit-diagonal L below the
    diagonal, U on and above it).  A: column-major float32/float64 tensor on
    a ROCm device (Julia layout; a row-major tensor would be factored as its
    transpose, so it is rejected).  Returns (A, ipiv, info): ipiv an int64
    device tensor of min(m, n) 1-based interchanges, info 0 or the 1-based
    index of the first exactly zero pivot (read after synchronizing the
    stream; the factorization completes either way, as LAPACK's)."""
    if A.dtype not in _DT:
        raise ValueError(f"unsupported element type {A.dtype} (Float32/Float64 only)")
    _require_device(A)
    tr, lda = _layout(A, "A")
    m, n = A.shape
    if tr and m > 1 and n > 1:
        raise ValueError("A must be column-major (stride (1, ld)); use colmajor(A)")
    k = min(m, n)
    if ipiv is None:
        ipiv = torch.empty(max(k, 1), dtype=torch.int64, device=A.device)
    elif ipiv.dtype != torch.int64 or ipiv.device != A.device or ipiv.numel() < k or not ipiv.is_contiguous():
        raise ValueError("ipiv must be a contiguous int64 tensor of min(m, n) entries on A's device")
    info = torch.zeros(1, dtype=torch.int64, device=A.device)
    st = load().nla_getrf2(_DT[A.dtype], m, n, A.data_ptr(), lda, ipiv.data_ptr(), info.data_ptr(), _stream(A))
    _lib.check(st, "getrf2!")
    inf = int(info.item())
    if inf < 0:
        raise _lib.NLAError("getrf2!: a multi-work-group leaf timed out waiting at its grid barrier")
    return A, ipiv[:k], inf


def lauum_(uplo, n, A, ib=None, workspace=None):
    """lauum!(uplo, n, A, ib) of src/lauum.jl:52: A := U*U^T (uplo 'U') or
    L^T*L (uplo 'L') on the leading n x n block, in place, only that
    triangle read and written.  A: column-major float32/float64 tensor on a
    ROCm device.  ib (the reference's block size) is accepted and ignored:
    the blocking is internal.  workspace: optional uint8 device tensor of at
    least nla_lauum_workspace_bytes bytes (allocated here when omitted).
    Returns A."""
    if not isinstance(uplo, str) or uplo.upper() not in ("U", "L"):
        raise ValueError(f"uplo must be 'U' or 'L', got {uplo!r}")
    if not isinstance(n, int) or n < 0:
        raise ValueError(f"n must be non-negative, got {n!r}")
    if A.dtype not in _DT:
        raise ValueError(f"unsupported element type {A.dtype} (Float32/Float64 only)")
    _require_device(A)
    if A.dim() != 2 or A.shape[0] < n or A.shape[1] < n:
        raise ValueError(f"A {tuple(A.shape)} is smaller than n = {n}")
    tr, lda = _layout(A, "A")
    if tr and n > 1:
        raise ValueError("A must be column-major (stride (1, ld)); use colmajor(A)")
    lib = load()
    need = int(lib.nla_lauum_workspace_bytes(_DT[A.dtype], n))
    if workspace is None:
        workspace = torch.empty(max(need, 1), dtype=torch.uint8, device=A.device)
    elif workspace.numel() < need:
        raise ValueError(f"workspace holds {workspace.numel()} bytes, {need} needed")
    st = lib.nla_lauum(_lib.ch(uplo.upper()), _DT[A.dtype], n, A.data_ptr(), lda, workspace.data_ptr(), need,
                       _stream(A))
    _lib.check(st, "lauum!")
    return A


def larfb_(side, trans, direct, storev, V, T, C, work=None):
    """larfb!(side, trans, direct, storev, V, T, C) of src/larfb.jl:166 (the
    sized form, src/larfb.jl:37): C := H*C, H^T*C, C*H or C*H^T with
    H = I - V*T*V^T, forward columnwise reflectors (the reference's only
    form).  V (m x k for side 'L', n x k for 'R'; unit lower, diagonal and
    upper part not read), T (k x k upper) and C (m x n): column-major
    float32/float64 tensors on one ROCm device.  Returns C."""
    for k_, (v, ok) in enumerate(((side, "LRlr"), (trans, "NTCntc"), (direct, "FBfb"), (storev, "CRcr")), 1):
        if not isinstance(v, str) or len(v) != 1 or v not in ok:
            raise ValueError(f"argument {k_} has invalid value {v!r}")
    if direct.upper() != "F" or storev.upper() != "C":
        raise ValueError("larfb! supports only forward, columnwise reflectors (as the reference)")
    if C.dtype not in _DT or V.dtype != C.dtype or T.dtype != C.dtype:
        raise ValueError("V, T and C must share one element type, float32 or float64")
    _require_device(V, T, C)
    m, n = C.shape
    k = T.shape[0]
    lds = []
    for t, nm in ((V, "V"), (T, "T"), (C, "C")):
        tr, ld = _layout(t, nm)
        if tr and t.shape[0] > 1 and t.shape[1] > 1:
            raise ValueError(f"{nm} must be column-major (stride (1, ld)); use colmajor({nm})")
        lds.append(ld)
    wr = n if side.upper() == "L" else m
    if work is None:
        work = torch.empty((k, max(wr, 1)), dtype=C.dtype, device=C.device).t()
    tr, ldw = _layout(work, "work")
    st = load().nla_larfb(_lib.ch(side.upper()), _lib.ch(trans.upper()), b"F", b"C", _DT[C.dtype], m, n, k,
                          V.data_ptr(), lds[0], T.data_ptr(), lds[1], C.data_ptr(), lds[2], work.data_ptr(),
                          max(ldw, wr), _stream(C))
    _lib.check(st, "larfb!")
    return C


def _cm_ld(t, name):
    """Leading dimension of a column-major (stride (1, ld)) matrix."""
    tr, ld = _layout(t, name)
    if tr and t.shape[0] > 1 and t.shape[1] > 1:
        raise ValueError(f"{name} must be column-major (stride (1, ld)); use colmajor({name})")
    return ld


def _chars(*pairs):
    for k_, (v, ok) in enumerate(pairs, 1):
        if not isinstance(v, str) or len(v) != 1 or v not in ok:
            raise ValueError(f"argument {k_} has invalid value {v!r}")


def pamm_(op, side, storev, direct, m, n, k, l, A1, A2, V, W):
    """pamm!(op, side, storev, direct, m, n, k, l, A1, A2, V, W) of
    src/pamm.jl:50: op 'W': W = A1 + op(V)*A2 (side 'L') or A1 + A2*op(V)
    ('R'); op 'A': A2 -= op(V)*W or A2 -= W*op(V), with V's pentagonal
    structure (the l x l triangle through the device TRMM).  Column-major
    float32/float64 tensors on one ROCm device; A1 is not read for op 'A'.
    Returns W ('W') or A2 ('A')."""
    _chars((op, "WAwa"), (side, "LRlr"), (storev, "CRcr"), (direct, "FBfb"))
    for v, nm in ((m, "m"), (n, "n"), (k, "k"), (l, "l")):
        if not isinstance(v, int) or v < 0:
            raise ValueError(f"{nm} must be non-negative, got {v!r}")
    op, side, storev, direct = op.upper(), side.upper(), storev.upper(), direct.upper()
    ts = [A2, V, W] + ([A1] if op == "W" else [])
    if A2.dtype not in _DT or any(t.dtype != A2.dtype for t in ts):
        raise ValueError("A1, A2, V and W must share one element type, float32 or float64")
    _require_device(*ts)
    lda1 = _cm_ld(A1, "A1") if op == "W" else 1
    st = load().nla_pamm(_lib.ch(op), _lib.ch(side), _lib.ch(storev), _lib.ch(direct), _DT[A2.dtype], m, n, k, l,
                         A1.data_ptr() if op == "W" else None, lda1, A2.data_ptr(), _cm_ld(A2, "A2"), V.data_ptr(),
                         _cm_ld(V, "V"), W.data_ptr(), _cm_ld(W, "W"), _stream(A2))
    _lib.check(st, "pamm!")
    return W if op == "W" else A2


def parfb_(side, trans, direct, storev, A1, A2, V, T, work=None, k=None, l=None):
    """parfb!(side, trans, direct, storev, m1, n1, m2, n2, k, l, A1, A2, V, T,
    work) of src/parfb.jl:46 (LAPACK's tprfb): A1 and A2 := H*C, H^T*C, C*H
    or C*H^T for C = [A1; A2] / [A2; A1] (side 'L', forward / backward) or
    [A1 A2] / [A2 A1] (side 'R'), H = I - W*T*W^T with W the identity and the
    pentagonal V.  k defaults to T's order; l, the order of V's triangle,
    defaults to the reference's short form (src/parfb.jl:229: the number of
    columns of V), which is only meaningful there when it does not exceed
    k.  Column-major float32/float64 tensors on one ROCm device; work (k x n2
    for side 'L', m2 x k for 'R') is allocated when omitted.  Returns (A1, A2)."""
    _chars((side, "LRlr"), (trans, "NTCntc"), (direct, "FBfb"), (storev, "CRcr"))
    side, trans, direct, storev = side.upper(), trans.upper(), direct.upper(), storev.upper()
    ts = [A1, A2, V, T]
    if A1.dtype not in _DT or any(t.dtype != A1.dtype for t in ts):
        raise ValueError("A1, A2, V and T must share one element type, float32 or float64")
    _require_device(*ts)
    m1, n1 = A1.shape
    m2, n2 = A2.shape
    k = T.shape[0] if k is None else k
    l = V.shape[1] if l is None else l
    wr, wc = (k, n2) if side == "L" else (m2, k)
    if work is None:
        work = torch.empty((max(wc, 1), max(wr, 1)), dtype=A1.dtype, device=A1.device).t()
    _require_device(A1, work)
    st = load().nla_parfb(_lib.ch(side), _lib.ch(trans), _lib.ch(direct), _lib.ch(storev), _DT[A1.dtype], m1, n1,
                          m2, n2, k, l, A1.data_ptr(), _cm_ld(A1, "A1"), A2.data_ptr(), _cm_ld(A2, "A2"),
                          V.data_ptr(), _cm_ld(V, "V"), T.data_ptr(), _cm_ld(T, "T"), work.data_ptr(),
                          max(_cm_ld(work, "work"), wr), _stream(A1))
    _lib.check(st, "parfb!")
    return A1, A2
