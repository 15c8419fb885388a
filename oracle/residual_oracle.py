"""Exact-arithmetic oracle of the FP64-accurate triangular residual
R = alpha*B - op(A)*X (side 'L') or alpha*B - X*op(A) (side 'R').

TEST INFRASTRUCTURE ONLY: imported by tests/, never by the product path.

The residual has no counterpart in the reference (its solve, which the
refinement converges to, is unified_rectrxm! of src/rectrxm.jl:43-76); what
fixes it is arithmetic, so this oracle computes it exactly with Python
integers:
  * `fixed_point_model` restates the int8 method's rounding rule: op(A) (for
    side 'R' its transpose, so that the product is Ahat * Xhat with Xhat's
    columns the systems) rounded to bA bits of each row's largest entry, X to
    bX bits of each system's largest entry, S = fl(alpha*B) scaled by the row
    and system exponents; R_int = S_int - Ahat_int Xhat_int exactly; and
    returns (R_int * 2^-(e_i + f_j) rounded once to FP64, X rounded);
  * `exact_residual` is alpha*B - op(A)*X of the unrounded operands, exactly
    (every double is an integer multiple of 2^-1074), rounded once to FP64;
  * `plan_bits(n)` restates the choice of the moduli count and bit split (the
    fewest moduli <= 256 whose product covers bA + bX + ceil(log2 n) + 3 bits
    with bA, bX >= 55).
The oracle's own checks (tests/test_oracle.py): the model with huge bit
counts equals exact_residual; exact_residual equals numpy's FP64 residual
within its rounding bound; a planted rounding mistake is caught.
"""
import math

import numpy as np

MODULI = [256, 255, 253, 251, 247, 241, 239, 233, 229, 227, 223, 217, 211, 199, 197, 193, 191, 181, 179, 173]
BITS_MIN = 55


def plan_bits(n):
    """(T, bA, bX) of an order-n residual."""
    lg = max(0, math.ceil(math.log2(n))) if n > 1 else 0
    logm = 0.0
    T = 0
    for T, p in enumerate(MODULI, 1):
        logm += math.log2(p)
        if math.floor(logm) - 3 - lg >= 2 * BITS_MIN:
            break
    budget = min(math.floor(logm) - 3 - lg, 124)
    bA = (budget + 1) // 2
    return T, bA, budget - bA


def _ahat_xhat(side, uplo, trans, A, X):
    """Ahat (triangle of op(A), or of op(A)^T for side 'R') and Xhat (systems
    as columns) as float64 arrays."""
    n = A.shape[0]
    T = np.tril(A) if uplo == "L" else np.triu(A)
    op = T.T if trans != "N" else T
    if side == "L":
        return op.astype(np.float64), X.astype(np.float64)
    return op.T.astype(np.float64), X.T.astype(np.float64)


def _row_exponents(M, bits):
    """e_i = bits - 1 - floor(log2 max_k |M_ik|), 0 for a zero row."""
    e = np.zeros(M.shape[0], dtype=np.int64)
    mx = np.abs(M).max(axis=1) if M.shape[1] else np.zeros(M.shape[0])
    for i, v in enumerate(mx):
        if v > 0:
            e[i] = bits - 1 - (math.frexp(float(v))[1] - 1)
    return e


def _to_int(v, e):
    """round(v * 2^e) (half to even, as rint) as a Python int, exactly."""
    num, den = float(v).as_integer_ratio()
    if e >= 0:
        num <<= e
    else:
        den <<= -e
    q, r = divmod(num, den)
    if 2 * r > den or (2 * r == den and q % 2 == 1):
        q += 1
    return q


def _ldexp_round(x, e):
    """The Python int (or rational) x * 2^e rounded once to a double."""
    from fractions import Fraction
    f = Fraction(x) * (Fraction(2) ** e)
    return float(f)


def fixed_point_model(side, uplo, trans, alpha, A, X, B, bA, bX):
    """(R, X_rounded) of the int8 method: exact residual of the rounded
    operands, rounded once to FP64 (R has B's shape)."""
    Ah, Xh = _ahat_xhat(side, uplo, trans, A, X)
    S = (np.float64(alpha) * B.astype(np.float64))
    Sh = S if side == "L" else S.T
    n, m = Xh.shape
    e = _row_exponents(Ah, bA)
    f = _row_exponents(Xh.T, bX)
    Ai = [[_to_int(Ah[i, k], int(e[i])) for k in range(n)] for i in range(n)]
    Xi = [[_to_int(Xh[k, j], int(f[j])) for k in range(n)] for j in range(m)]  # by system
    Xr = np.empty_like(Xh)
    for j in range(m):
        for k in range(n):
            Xr[k, j] = math.ldexp(float(Xi[j][k]), -int(f[j])) if Xi[j][k] != 0 else 0.0 * Xh[k, j]
    R = np.empty((n, m))
    for j in range(m):
        xj = Xi[j]
        for i in range(n):
            sc = int(e[i] + f[j])
            s_int = _to_int(Sh[i, j], sc)
            acc = sum(a * x for a, x in zip(Ai[i], xj) if a)
            R[i, j] = _ldexp_round(s_int - acc, -sc)
    if side == "L":
        return R, Xr
    return R.T.copy(), Xr.T.copy()


def exact_residual(side, uplo, trans, alpha, A, X, B):
    """alpha*B - op(A)*X (side 'L') or alpha*B - X*op(A) (side 'R') of the
    unrounded operands with S = fl(alpha*B), exactly, rounded once to FP64."""
    Ah, Xh = _ahat_xhat(side, uplo, trans, A, X)
    S = (np.float64(alpha) * B.astype(np.float64))
    Sh = S if side == "L" else S.T
    n, m = Xh.shape
    K = 1100  # every finite double is an integer multiple of 2^-1074
    Ai = [[_to_int(Ah[i, k], K) for k in range(n)] for i in range(n)]
    Xi = [[_to_int(Xh[k, j], K) for k in range(n)] for j in range(m)]
    R = np.empty((n, m))
    for j in range(m):
        for i in range(n):
            acc = sum(a * x for a, x in zip(Ai[i], Xi[j]) if a)
            R[i, j] = _ldexp_round(_to_int(Sh[i, j], 2 * K) - acc, -2 * K)
    return R if side == "L" else R.T.copy()
