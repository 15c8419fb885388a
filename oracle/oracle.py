"""CPU oracle for NextLA.jl's recursive triangular solve / multiply.

TEST INFRASTRUCTURE ONLY: imported by tests/, the smoke check and the bench's
CPU-baseline leg, and never by the product path (nextla.jl_amd/).

Two oracles, as in the reference's own tests (test/unified_rectrxm.jl:36-40):
  * `rectrxm` etc. — the C restatement of unified_rectrxm! (liboracle.so,
    rectrxm_oracle.c), following src/rectrxm.jl:43-198, src/matmul.jl:5-80,
    src/trsm.jl:5-150 and src/trmm.jl:43-393 operation by operation;
  * `blas` — LinearAlgebra.BLAS.trsm!/trmm! of the reference tests, here the
    same OpenBLAS 0.3.29 routines (dtrsm/strsm/dtrmm/strmm) through scipy.

Error measures: `backward_error` (TRSM, needs no reference solution) and
`product_error` (TRMM), both normwise in Frobenius norm and computed in FP64.

`lauum` restates lauum!(uplo, n, A, ib) (src/lauum.jl:52-186, lauu2!
src/lauu2.jl:48-107): U*U^T or L^T*L on the uplo triangle, the other
triangle untouched; pinned against LAPACK dlauum / slauum.

`getrf2` restates the reference's recursive LU with partial pivoting
(src/lu.jl:185-300, laswp src/lu.jl:470-520) in numpy; it is pinned against
LAPACK's dgetrf / sgetrf through scipy (same pivots and factors) and by
P·L·U = A (tests/test_oracle.py).
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None


def build():
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def load():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        lib = ctypes.CDLL(LIB_PATH)
        c, i64, dbl, vp, i32 = ctypes.c_char, ctypes.c_long, ctypes.c_double, ctypes.c_void_p, ctypes.c_int
        for sfx in ("d", "s"):
            f = getattr(lib, f"oracle_rectrxm_{sfx}")
            f.argtypes = [c, c, c, c, dbl, i64, i64, vp, i64, vp, i64, i32]
            f.restype = i32
            f = getattr(lib, f"oracle_rectrxm_thr_{sfx}")
            f.argtypes = [c, c, c, c, dbl, i64, i64, vp, i64, vp, i64, i64, i32]
            f.restype = i32
            f = getattr(lib, f"oracle_gemm_add_{sfx}")
            f.argtypes = [i64, i64, i64, vp, i64, vp, i64, vp, i64, i32]
            f.restype = None
            f = getattr(lib, f"oracle_gemm_sub_{sfx}")
            f.argtypes = [i64, i64, i64, vp, i64, vp, i64, vp, i64, i32]
            f.restype = None
            f = getattr(lib, f"oracle_leaf_{sfx}")
            f.argtypes = [c, c, c, i64, i64, vp, i64, vp, i64, i32]
            f.restype = None
        _lib = lib
    return _lib


def _sfx(a):
    if a.dtype == np.float64:
        return "d"
    if a.dtype == np.float32:
        return "s"
    raise TypeError(f"oracle supports float32/float64, got {a.dtype}")


def _f(a):
    """Fortran-ordered copy (column-major, like Julia)."""
    return np.array(a, order="F", copy=True)


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def rectrxm(side, uplo, trans, alpha, func, A, B, nthreads=1, threshold=None):
    """unified_rectrxm!(side, uplo, trans, alpha, func, A, B) restated on the
    CPU (src/rectrxm.jl:43-76); returns the new B (inputs are not modified).
    `threshold` overrides the reference's leaf thresholds (256 for 'S', 16 for
    'M', src/rectrxm.jl:52,63) to exercise the recursion on small inputs."""
    A = _f(A)
    X = _f(B)
    assert A.dtype == X.dtype
    n = A.shape[0]
    m = X.shape[1] if side == "L" else X.shape[0]
    lib = load()
    if threshold is None:
        fn = getattr(lib, f"oracle_rectrxm_{_sfx(A)}")
        fn(side.encode(), uplo.encode(), trans.encode(), func.encode(), float(alpha), n, m, _ptr(A), max(1, n),
           _ptr(X), max(1, X.shape[0]), int(nthreads))
    else:
        fn = getattr(lib, f"oracle_rectrxm_thr_{_sfx(A)}")
        fn(side.encode(), uplo.encode(), trans.encode(), func.encode(), float(alpha), n, m, _ptr(A), max(1, n),
           _ptr(X), max(1, X.shape[0]), int(threshold), int(nthreads))
    return X


def gemm_add(X, Y, C, nthreads=1):
    """GEMM_ADD!(X, Y, C): returns C + X*Y (src/matmul.jl:69-73)."""
    X, Y, C = _f(X), _f(Y), _f(C)
    N, R = X.shape
    M = Y.shape[1]
    getattr(load(), f"oracle_gemm_add_{_sfx(C)}")(N, R, M, _ptr(X), max(1, N), _ptr(Y), max(1, R), _ptr(C),
                                                  max(1, N), int(nthreads))
    return C


def gemm_sub(C, X, Y, nthreads=1):
    """GEMM_SUB!(C, X, Y): returns C - X*Y (src/matmul.jl:76-80)."""
    X, Y, C = _f(X), _f(Y), _f(C)
    N, R = X.shape
    M = Y.shape[1]
    getattr(load(), f"oracle_gemm_sub_{_sfx(C)}")(N, R, M, _ptr(C), max(1, N), _ptr(X), max(1, N), _ptr(Y),
                                                  max(1, R), int(nthreads))
    return C


def leaf(func, side, uplo, A, B, nthreads=1):
    """One reference leaf kernel (LeftLowerTRSM! ... RightUpperTRMM!) applied
    to the whole of A (src/trsm.jl:128-150, src/trmm.jl:332-393)."""
    A, X = _f(A), _f(B)
    n = A.shape[0]
    m = X.shape[1] if side == "L" else X.shape[0]
    getattr(load(), f"oracle_leaf_{_sfx(A)}")(func.encode(), side.encode(), uplo.encode(), n, m, _ptr(A),
                                              max(1, n), _ptr(X), max(1, X.shape[0]), int(nthreads))
    return X


def blas(side, uplo, trans, alpha, func, A, B, diag="N"):
    """LinearAlgebra.BLAS.trsm!/trmm!(side, uplo, trans, diag, alpha, A, B) as
    called by test/unified_rectrxm.jl:36-40 (diag 'N' there), through scipy's
    OpenBLAS; diag 'U' leaves A's diagonal unreferenced."""
    from scipy.linalg import blas as sb
    A = _f(A)
    B = _f(B)
    pre = "d" if A.dtype == np.float64 else "s"
    fn = getattr(sb, pre + ("trsm" if func == "S" else "trmm"))
    t = {"N": 0, "T": 1, "C": 2}[trans]
    return fn(alpha, A, B, side=0 if side == "L" else 1, lower=1 if uplo == "L" else 0, trans_a=t,
              diag=1 if diag == "U" else 0)


def op_matrix(uplo, trans, A):
    """op(A) restricted to its `uplo` triangle, in FP64."""
    A = np.asarray(A, dtype=np.float64)
    T = np.tril(A) if uplo == "L" else np.triu(A)
    return T.T if trans in ("T", "C") else T


def backward_error(side, uplo, trans, alpha, A, B, X):
    """Normwise backward error of a TRSM solution in FP64:
    ||alpha*B - op(A)X|| / (||op(A)|| ||X|| + |alpha| ||B||)  (Frobenius)."""
    T = op_matrix(uplo, trans, A)
    B = np.asarray(B, dtype=np.float64)
    X = np.asarray(X, dtype=np.float64)
    R = alpha * B - (T @ X if side == "L" else X @ T)
    den = np.linalg.norm(T) * np.linalg.norm(X) + abs(alpha) * np.linalg.norm(B)
    return float(np.linalg.norm(R) / den) if den > 0 else float(np.linalg.norm(R))


def product_error(side, uplo, trans, alpha, A, B, Y):
    """Normwise error of a TRMM result against the FP64 product:
    ||Y - alpha*op(A)B|| / (|alpha| ||op(A)|| ||B||)  (Frobenius)."""
    T = op_matrix(uplo, trans, A)
    B = np.asarray(B, dtype=np.float64)
    Y = np.asarray(Y, dtype=np.float64)
    Z = alpha * (T @ B if side == "L" else B @ T)
    den = abs(alpha) * np.linalg.norm(T) * np.linalg.norm(B)
    return float(np.linalg.norm(Y - Z) / den) if den > 0 else float(np.linalg.norm(Y - Z))


def rel_error(X, Y):
    """norm(X - Y)/norm(Y) of test/unified_rectrxm.jl:42."""
    X = np.asarray(X, dtype=np.float64)
    Y = np.asarray(Y, dtype=np.float64)
    d = np.linalg.norm(Y)
    return float(np.linalg.norm(X - Y) / d) if d > 0 else float(np.linalg.norm(X - Y))


def _laswp(A, k1, k2, ipiv):
    """Rows k and ipiv[k] swapped in order for k = k1..k2 (1-based),
    src/lu.jl:470-520 (incx = 1)."""
    for k in range(k1, k2 + 1):
        p = int(ipiv[k - 1])
        if p != k:
            A[[k - 1, p - 1], :] = A[[p - 1, k - 1], :]


def _getrf2_rec(A, ipiv):
    """One call of getrf2! on the view A (src/lu.jl:185-300); returns info
    (0 or the 1-based first zero pivot, LAPACK dgetrf2's meaning)."""
    m, n = A.shape
    if m == 0 or n == 0:
        return 0
    if m == 1:  # src/lu.jl:216-221
        ipiv[0] = 1
        return 1 if A[0, 0] == 0 else 0
    if n == 1:  # src/lu.jl:223-253
        sfmin = np.finfo(A.dtype).tiny
        col = A[:, 0]
        i = int(np.argmax(np.abs(col)))  # first index of the maximum, as idamax
        ipiv[0] = i + 1
        if col[i] == 0:
            return 1
        if i != 0:
            col[0], col[i] = col[i], col[0]
        if abs(col[0]) >= sfmin:
            col[1:] *= A.dtype.type(1) / col[0]
        else:
            col[1:] /= col[0]
        return 0
    n1 = min(m, n) // 2  # src/lu.jl:255-296
    n2 = n - n1
    info = _getrf2_rec(A[:, :n1], ipiv)
    _laswp(A[:, n1:], 1, n1, ipiv)
    L11 = np.tril(A[:n1, :n1], -1) + np.eye(n1, dtype=A.dtype)
    A[:n1, n1:] = _unit_lower_solve(L11, A[:n1, n1:])
    A[n1:, n1:] -= A[n1:, :n1] @ A[:n1, n1:]
    iinfo = _getrf2_rec(A[n1:, n1:], ipiv[n1:])
    if info == 0 and iinfo > 0:
        info = iinfo + n1
    ipiv[n1:min(m, n)] += n1
    _laswp(A[:, :n1], n1 + 1, min(m, n), ipiv)
    return info


def _unit_lower_solve(L, B):
    import scipy.linalg
    return scipy.linalg.solve_triangular(L, B, lower=True, unit_diagonal=True, check_finite=False)


def getrf2(A):
    """Recursive LU with partial pivoting of a copy of A: returns (LU, ipiv
    (1-based int64), info) as getrf2!(A, ipiv, info) of the reference."""
    A = np.array(A, order="F", copy=True)
    ipiv = np.zeros(min(A.shape), dtype=np.int64)
    info = _getrf2_rec(A, ipiv)
    return A, ipiv, info


def lu_reconstruct(LU, ipiv):
    """P·L·U from packed factors and 1-based interchanges (FP64)."""
    m, n = LU.shape
    k = min(m, n)
    L = np.tril(LU[:, :k].astype(np.float64), -1) + np.eye(m, k)
    U = np.triu(LU[:k, :].astype(np.float64))
    R = L @ U
    for i in range(k - 1, -1, -1):  # undo the interchanges, last first
        p = int(ipiv[i]) - 1
        if p != i:
            R[[i, p], :] = R[[p, i], :]
    return R


def lauum(uplo, n, A):
    """lauum!(uplo, n, A, ib) of src/lauum.jl:52-186 on a copy of A: the
    leading n x n triangle replaced by U*U^T (uplo 'U') or L^T*L (uplo 'L'),
    computed in FP64; everything else unchanged."""
    A = np.array(A, order="F", copy=True)
    T = A[:n, :n].astype(np.float64)
    if uplo == "U":
        U = np.triu(T)
        R = np.triu(U @ U.T)
        mask = np.triu(np.ones((n, n), dtype=bool))
    else:
        L = np.tril(T)
        R = np.tril(L.T @ L)
        mask = np.tril(np.ones((n, n), dtype=bool))
    blk = A[:n, :n]
    blk[mask] = R[mask].astype(A.dtype)
    return A


def larft(V, tau):
    """T of H(1)...H(k) = I - V T V^T (forward, columnwise; LAPACK dlarft),
    as the reference builds it with larft! for its tests."""
    k = V.shape[1]
    T = np.zeros((k, k), dtype=np.float64)
    Vd = V.astype(np.float64)
    for i in range(k):
        T[i, i] = tau[i]
        if i:
            T[:i, i] = -tau[i] * (T[:i, :i] @ (Vd[:, :i].T @ Vd[:, i]))
    return T.astype(V.dtype)


def larfb(side, trans, V, T, C):
    """larfb! (src/larfb.jl:37-142) from the closed form H = I - V T V^T (V
    unit lower), in FP64: H C, H^T C, C H or C H^T."""
    Vd = np.tril(V.astype(np.float64), -1)
    k = V.shape[1]
    Vd[np.arange(k), np.arange(k)] = 1.0
    Td = np.triu(T.astype(np.float64))
    H = np.eye(Vd.shape[0]) - Vd @ Td @ Vd.T
    Hop = H if trans.upper() == "N" else H.T
    Cd = C.astype(np.float64)
    return (Hop @ Cd if side.upper() == "L" else Cd @ Hop).astype(C.dtype)
