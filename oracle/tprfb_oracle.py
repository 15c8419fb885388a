"""CPU oracle for the triangular-pentagonal block reflectors of tile QR:
pamm! and parfb! (src/pamm.jl, src/parfb.jl).

TEST INFRASTRUCTURE ONLY: imported by tests/, never by the product path
(nextla.jl_amd/).

  * `pamm` restates pamm! (src/pamm.jl:86-466) combination by combination in
    numpy (FP64), on 0-based slices: the reference's mp = min(m-l+1, m),
    kp = min(l+1, k) (forward) or mp = min(l+1, m), kp = min(k-l+1, k)
    (backward) become mp - 1 and kp - 1 here; generic_trimatmul! /
    generic_mattrimul! become products with np.triu / np.tril of the block.
  * `parfb` restates parfb! (src/parfb.jl:46-202): pamm 'W', the product by
    op(T)'s triangle (upper forward, lower backward), A1 -= W, pamm 'A'.
  * `tprfb_dense` is the closed form LAPACK's xTPRFB documents and the
    reference's test compares parfb! against (test/parfb.jl:43-46): the
    pentagonal V (V1 rectangular, V2 trapezoidal) as a dense matrix, W = the
    identity stacked with it, H = I - W T W^T (columnwise) or
    I - W^T T W (rowwise), applied to [A; B] / [B; A] (side 'L', forward /
    backward) or [A B] / [B A] (side 'R').

Pinning (tests/test_oracle.py): `tprfb_dense` against LAPACK's dtpmqrt
(which applies xTPRFB) through scipy for forward columnwise reflectors of a
dtpqrt factorization, both sides and both transposes; the backward forms
against the forward ones on reversed rows / reflectors, the rowwise forms
against the columnwise ones on V^T; H^T H = I for dtpqrt's V and T; and
`parfb` against `tprfb_dense` for every side/trans/direct/storev on random
V and T.
"""
import numpy as np


def _f64(a):
    return np.array(a, dtype=np.float64, copy=True)


def pamm(op, side, storev, direct, m, n, k, l, A1, A2, V, W):
    """pamm!(op, side, storev, direct, m, n, k, l, A1, A2, V, W) in FP64.
    Returns (W, A2) after the call (op 'W' writes W; op 'A' writes A2 and
    the triangle rows / columns of W)."""
    op, side, storev, direct = op.upper(), side.upper(), storev.upper(), direct.upper()
    W = _f64(W)
    A2 = _f64(A2)
    if m == 0 or n == 0 or k == 0:
        return W, A2
    A1 = _f64(A1) if A1 is not None else None
    V = _f64(V)
    left, col, fwd = side == "L", storev == "C", direct == "F"
    kp = min(l, k - 1) if fwd else min(k - l, k - 1)
    up, lo = np.triu, np.tril
    if op == "W":
        if left:
            mp = min(m - l, m - 1) if fwd else min(l, m - 1)
            if fwd:
                if l > 0:
                    W[0:l, 0:n] = A2[m - l:m, 0:n]
                    if col:
                        W[0:l, 0:n] = up(V[mp:mp + l, 0:l]).T @ W[0:l, 0:n]
                        if m > l:
                            W[0:l, 0:n] += V[0:m - l, 0:l].T @ A2[0:m - l, 0:n]
                    else:
                        W[0:l, 0:n] = lo(V[0:l, mp:mp + l]) @ W[0:l, 0:n]
                        if m > l:
                            W[0:l, 0:n] += V[0:l, 0:m - l] @ A2[0:m - l, 0:n]
                if k > l:
                    if col:
                        W[kp:kp + k - l, 0:n] = V[0:m, kp:kp + k - l].T @ A2[0:m, 0:n]
                    else:
                        W[kp:kp + k - l, 0:n] = V[kp:kp + k - l, 0:m] @ A2[0:m, 0:n]
            else:
                if l > 0:
                    W[k - l:k, 0:n] = A2[0:l, 0:n]
                    if col:
                        W[kp:kp + l, 0:n] = lo(V[0:l, kp:kp + l]).T @ W[kp:kp + l, 0:n]
                        if m > l:
                            W[kp:kp + l, 0:n] += V[mp:mp + m - l, kp:kp + l].T @ A2[mp:mp + m - l, 0:n]
                    else:
                        W[kp:kp + l, 0:n] = up(V[kp:kp + l, 0:l]) @ W[kp:kp + l, 0:n]
                        if m > l:
                            W[kp:kp + l, 0:n] += V[kp:kp + l, mp:mp + m - l] @ A2[mp:mp + m - l, 0:n]
                if k > l:
                    if col:
                        W[0:k - l, 0:n] = V[0:m, 0:k - l].T @ A2[0:m, 0:n]
                    else:
                        W[0:k - l, 0:n] = V[0:k - l, 0:m] @ A2[0:m, 0:n]
            W[0:k, 0:n] += A1[0:k, 0:n]
        else:
            np_ = min(n - l, n - 1) if fwd else min(l, n - 1)
            if fwd:
                if l > 0:
                    W[0:m, 0:l] = A2[0:m, n - l:n]
                    if col:
                        W[0:m, 0:l] = W[0:m, 0:l] @ up(V[np_:np_ + l, 0:l])
                        if n > l:
                            W[0:m, 0:l] += A2[0:m, 0:n - l] @ V[0:n - l, 0:l]
                    else:
                        W[0:m, 0:l] = W[0:m, 0:l] @ lo(V[0:l, np_:np_ + l]).T
                        if n > l:
                            W[0:m, 0:l] += A2[0:m, 0:n - l] @ V[0:l, 0:n - l].T
                if k > l:
                    if col:
                        W[0:m, kp:kp + k - l] = A2[0:m, 0:n] @ V[0:n, kp:kp + k - l]
                    else:
                        W[0:m, kp:kp + k - l] = A2[0:m, 0:n] @ V[kp:kp + k - l, 0:n].T
            else:
                if l > 0:
                    W[0:m, k - l:k] = A2[0:m, 0:l]
                    if col:
                        W[0:m, kp:kp + l] = W[0:m, kp:kp + l] @ lo(V[0:l, kp:kp + l])
                        if n > l:
                            W[0:m, kp:kp + l] += A2[0:m, np_:np_ + n - l] @ V[np_:np_ + n - l, kp:kp + l]
                    else:
                        W[0:m, kp:kp + l] = W[0:m, kp:kp + l] @ up(V[kp:kp + l, 0:l]).T
                        if n > l:
                            W[0:m, kp:kp + l] += A2[0:m, np_:np_ + n - l] @ V[kp:kp + l, np_:np_ + n - l].T
                if k > l:
                    if col:
                        W[0:m, 0:k - l] = A2[0:m, 0:n] @ V[0:n, 0:k - l]
                    else:
                        W[0:m, 0:k - l] = A2[0:m, 0:n] @ V[0:k - l, 0:n].T
            W[0:m, 0:k] += A1[0:m, 0:k]
        return W, A2
    # op 'A'
    if left:
        mp = min(m - l, m - 1) if fwd else min(l, m - 1)
        if fwd:
            if m > l:
                A2[0:m - l, 0:n] -= (V[0:m - l, 0:k] if col else V[0:k, 0:m - l].T) @ W[0:k, 0:n]
            if l > 0:
                if k > l:
                    Vb = V[mp:mp + l, kp:kp + k - l] if col else V[kp:kp + k - l, mp:mp + l].T
                    A2[mp:mp + l, 0:n] -= Vb @ W[kp:kp + k - l, 0:n]
                Tl = up(V[mp:mp + l, 0:l]) if col else lo(V[0:l, mp:mp + l]).T
                W[0:l, 0:n] = Tl @ W[0:l, 0:n]
                A2[m - l:m, 0:n] -= W[0:l, 0:n]
        else:
            if m > l:
                Vb = V[mp:mp + m - l, 0:k] if col else V[0:k, mp:mp + m - l].T
                A2[mp:mp + m - l, 0:n] -= Vb @ W[0:k, 0:n]
            if l > 0:
                if k > l:
                    Vb = V[0:l, 0:k - l] if col else V[0:k - l, 0:l].T
                    A2[0:l, 0:n] -= Vb @ W[0:k - l, 0:n]
                Tl = lo(V[0:l, kp:kp + l]) if col else up(V[kp:kp + l, 0:l]).T
                W[kp:kp + l, 0:n] = Tl @ W[kp:kp + l, 0:n]
                A2[0:l, 0:n] -= W[k - l:k, 0:n]
        return W, A2
    np_ = min(n - l, n - 1) if fwd else min(l, n - 1)
    if fwd:
        if n > l:
            A2[0:m, 0:n - l] -= W[0:m, 0:k] @ (V[0:n - l, 0:k].T if col else V[0:k, 0:n - l])
        if l > 0:
            if k > l:
                Vb = V[np_:np_ + l, kp:kp + k - l].T if col else V[kp:kp + k - l, np_:np_ + l]
                A2[0:m, np_:np_ + l] -= W[0:m, kp:kp + k - l] @ Vb
            Tr = up(V[np_:np_ + l, 0:l]).T if col else lo(V[0:l, np_:np_ + l])
            W[0:m, 0:l] = W[0:m, 0:l] @ Tr
            A2[0:m, n - l:n] -= W[0:m, 0:l]
    else:
        if n > l:
            Vb = V[np_:np_ + n - l, 0:k].T if col else V[0:k, np_:np_ + n - l]
            A2[0:m, np_:np_ + n - l] -= W[0:m, 0:k] @ Vb
        if l > 0:
            if k > l:
                Vb = V[0:l, 0:k - l].T if col else V[0:k - l, 0:l]
                A2[0:m, 0:l] -= W[0:m, 0:k - l] @ Vb
            Tr = lo(V[0:l, kp:kp + l]).T if col else up(V[kp:kp + l, 0:l])
            W[0:m, kp:kp + l] = W[0:m, kp:kp + l] @ Tr
            A2[0:m, 0:l] -= W[0:m, k - l:k]
    return W, A2


def parfb(side, trans, direct, storev, m1, n1, m2, n2, k, l, A1, A2, V, T, work):
    """parfb!(...) of src/parfb.jl:46-202 in FP64. Returns (A1, A2)."""
    side, trans, direct = side.upper(), trans.upper(), direct.upper()
    A1 = _f64(A1)
    A2 = _f64(A2)
    if m1 == 0 or n1 == 0 or m2 == 0 or n2 == 0 or k == 0:
        return A1, A2
    W, _ = pamm("W", side, storev, direct, m2, n2, k, l, A1, A2, V, work)
    Tt = np.triu(_f64(T)[0:k, 0:k]) if direct == "F" else np.tril(_f64(T)[0:k, 0:k])
    Top = Tt if trans == "N" else Tt.T
    if side == "L":
        W[0:k, 0:n2] = Top @ W[0:k, 0:n2]
        A1[0:k, 0:n2] -= W[0:k, 0:n2]
    else:
        W[0:m2, 0:k] = W[0:m2, 0:k] @ Top
        A1[0:m2, 0:k] -= W[0:m2, 0:k]
    _, A2 = pamm("A", side, storev, direct, m2, n2, k, l, A1, A2, V, W)
    return A1, A2


def pentagonal(side, direct, storev, k, l, V):
    """V as the dense matrix xTPRFB reads: zeros outside V1 and V2's
    trapezoid (V is m x k / n x k columnwise, k x m / k x n rowwise)."""
    V = _f64(V)
    col, fwd = storev.upper() == "C", direct.upper() == "F"
    if col:
        p = V.shape[0]
        for i in range(l):
            r = p - l + i if fwd else i
            for j in range(k):
                if (fwd and j < i) or (not fwd and j > k - l + i):
                    V[r, j] = 0.0
    else:
        p = V.shape[1]
        for j in range(l):
            c = p - l + j if fwd else j
            for i in range(k):
                if (fwd and i < j) or (not fwd and i > k - l + j):
                    V[i, c] = 0.0
    return V


def tprfb_dense(side, trans, direct, storev, l, V, T, A, B):
    """The closed form of xTPRFB (H = I - W T W^T or I - W^T T W with W
    the identity stacked with the pentagonal V) applied to [A; B], [B; A],
    [A B] or [B A]. Returns (A, B) in FP64."""
    side, trans, direct, storev = side.upper(), trans.upper(), direct.upper(), storev.upper()
    A, B = _f64(A), _f64(B)
    k = T.shape[0]
    Vd = pentagonal(side, direct, storev, k, l, V)
    Td = np.triu(_f64(T)) if direct == "F" else np.tril(_f64(T))
    I = np.eye(k)
    if storev == "C":
        Wm = np.vstack([I, Vd]) if direct == "F" else np.vstack([Vd, I])
        H = np.eye(Wm.shape[0]) - Wm @ Td @ Wm.T
    else:
        Wm = np.hstack([I, Vd]) if direct == "F" else np.hstack([Vd, I])
        H = np.eye(Wm.shape[1]) - Wm.T @ Td @ Wm
    Hop = H if trans == "N" else H.T
    if side == "L":
        C = np.vstack([A, B]) if direct == "F" else np.vstack([B, A])
        C = Hop @ C
        if direct == "F":
            return C[:k], C[k:]
        return C[C.shape[0] - k:], C[:C.shape[0] - k]
    C = np.hstack([A, B]) if direct == "F" else np.hstack([B, A])
    C = C @ Hop
    if direct == "F":
        return C[:, :k], C[:, k:]
    return C[:, C.shape[1] - k:], C[:, :C.shape[1] - k]
