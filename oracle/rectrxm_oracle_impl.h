/*
 * Type-generic body of the CPU oracle for NextLA.jl's recursive TRSM/TRMM.
 * TEST INFRASTRUCTURE ONLY: included twice by rectrxm_oracle.c with
 *   OT  = element type (double / float)
 *   OSF = symbol suffix (d / s)
 * Every routine restates the reference's arithmetic in the reference's order,
 * so results agree with the Julia code to rounding of the same operations.
 */

#define OCAT2(a, b) a##_##b
#define OCAT(a, b) OCAT2(a, b)
#define OFN(name) OCAT(name, OSF)

/* A strided matrix view: element (i, j) (0-based) lives at p[i*rs + j*cs].
 * Transpose(A) is the same storage with rs/cs swapped, exactly as Julia's
 * lazy Transpose wrapper used at src/rectrxm.jl:56-59. */
typedef struct {
    OT *p;
    long rs, cs;
    long nr, nc;
} OFN(oview);

static inline OT *OFN(oat)(OFN(oview) v, long i, long j) { return v.p + i * v.rs + j * v.cs; }

static OFN(oview) OFN(osub)(OFN(oview) v, long r0, long nr, long c0, long nc) {
    OFN(oview) s = v;
    s.p = v.p + r0 * v.rs + c0 * v.cs;
    s.nr = nr;
    s.nc = nc;
    return s;
}

static OFN(oview) OFN(otrans)(OFN(oview) v) {
    OFN(oview) t = v;
    t.rs = v.cs;
    t.cs = v.rs;
    t.nr = v.nc;
    t.nc = v.nr;
    return t;
}

/* matmul! kernel, src/matmul.jl:5-66.  out[I,J] += alpha * outval, where outval
 * is accumulated tile by tile: each 32-wide K tile forms its own partial sum
 * tmp (src/matmul.jl:50-54) which is then added to outval.  As on KA's CPU
 * backend, each 32 x 32 output tile (one work-group of the reference) is one
 * unit of work, spread over threads; the per-element operation order is the
 * reference's. */
static void OFN(omatmul)(OFN(oview) out, OFN(oview) in1, OFN(oview) in2, long N, long R, long M,
                         double alpha, int nthreads) {
    enum { TILE = 32 };
    long tn = (N + TILE - 1) / TILE, tm = (M + TILE - 1) / TILE;
#pragma omp parallel for schedule(dynamic, 1) num_threads(nthreads > 0 ? nthreads : 1)
    for (long tile = 0; tile < tn * tm; ++tile) {
        long I0 = (tile % tn) * TILE, J0 = (tile / tn) * TILE;
        long ni = N - I0 < TILE ? N - I0 : TILE, nj = M - J0 < TILE ? M - J0 : TILE;
        OT acc[TILE][TILE], tmp[TILE][TILE], a[TILE][TILE]; /* [j][i], [j][i], [k][i] */
        for (long j = 0; j < nj; ++j)
            for (long i = 0; i < ni; ++i) acc[j][i] = (OT)0;
        for (long t0 = 0; t0 < R; t0 += TILE) {
            long nk = R - t0 < TILE ? R - t0 : TILE;
            for (long k = 0; k < nk; ++k)
                for (long i = 0; i < ni; ++i) a[k][i] = *OFN(oat)(in1, I0 + i, t0 + k);
            for (long j = 0; j < nj; ++j) {
                for (long i = 0; i < ni; ++i) tmp[j][i] = (OT)0;
                for (long k = 0; k < nk; ++k) {
                    OT y = *OFN(oat)(in2, t0 + k, J0 + j);
                    for (long i = 0; i < ni; ++i) tmp[j][i] += a[k][i] * y;
                }
                for (long i = 0; i < ni; ++i) acc[j][i] += tmp[j][i];
            }
        }
        /* alpha arrives as a Float64 (src/matmul.jl:6) and is converted by the
         * product with eltype(output). */
        for (long j = 0; j < nj; ++j)
            for (long i = 0; i < ni; ++i) *OFN(oat)(out, I0 + i, J0 + j) += (OT)alpha * acc[j][i];
    }
}

/* GEMM_ADD!(A, B, C): C += A*B with N = size(A,1), R = size(B,1), M = size(B,2)
 * (src/matmul.jl:69-73). */
static void OFN(ogemm_add)(OFN(oview) A, OFN(oview) B, OFN(oview) C, int nthreads) {
    OFN(omatmul)(C, A, B, A.nr, B.nr, B.nc, 1.0, nthreads);
}

/* GEMM_SUB!(A, B, C): A -= B*C with N = size(A,1), R = size(C,1), M = size(A,2)
 * (src/matmul.jl:76-80). */
static void OFN(ogemm_sub)(OFN(oview) A, OFN(oview) B, OFN(oview) C, int nthreads) {
    OFN(omatmul)(A, B, C, A.nr, C.nr, A.nc, -1.0, nthreads);
}

/* lower_left_kernel, src/trsm.jl:5-33, called with Transpose(A) (src/trsm.jl:128-132).
 * Per RHS column: pre-divide b by the diagonal (:16-17), then for i = 1..n every
 * row > i subtracts (A[row,i]/diag[row]) * b[i] (:21-27). */
static void OFN(oleft_lower_trsm)(OFN(oview) A, OFN(oview) B, int nthreads) {
    long n = B.nr, m = B.nc;
    OFN(oview) At = OFN(otrans)(A);
#pragma omp parallel for schedule(static) num_threads(nthreads > 0 ? nthreads : 1)
    for (long col = 0; col < m; ++col) {
        OT *d = (OT *)malloc(sizeof(OT) * (size_t)n);
        OT *bc = (OT *)malloc(sizeof(OT) * (size_t)n);
        for (long r = 0; r < n; ++r) {
            d[r] = *OFN(oat)(At, r, r);
            bc[r] = *OFN(oat)(B, r, col) / d[r];
        }
        for (long i = 0; i < n; ++i)
            for (long r = i + 1; r < n; ++r) {
                OT acol = *OFN(oat)(At, i, r) / d[r];
                bc[r] -= acol * bc[i];
            }
        for (long r = 0; r < n; ++r) *OFN(oat)(B, r, col) = bc[r];
        free(d);
        free(bc);
    }
}

/* upper_left_kernel, src/trsm.jl:36-64: backward sweep i = n..1, rows < i. */
static void OFN(oleft_upper_trsm)(OFN(oview) A, OFN(oview) B, int nthreads) {
    long n = B.nr, m = B.nc;
#pragma omp parallel for schedule(static) num_threads(nthreads > 0 ? nthreads : 1)
    for (long col = 0; col < m; ++col) {
        OT *d = (OT *)malloc(sizeof(OT) * (size_t)n);
        OT *bc = (OT *)malloc(sizeof(OT) * (size_t)n);
        for (long r = 0; r < n; ++r) {
            d[r] = *OFN(oat)(A, r, r);
            bc[r] = *OFN(oat)(B, r, col) / d[r];
        }
        for (long i = n - 1; i >= 0; --i)
            for (long r = 0; r < i; ++r) {
                OT acol = *OFN(oat)(A, r, i) / d[r];
                bc[r] -= acol * bc[i];
            }
        for (long r = 0; r < n; ++r) *OFN(oat)(B, r, col) = bc[r];
        free(d);
        free(bc);
    }
}

/* right_lower_kernel, src/trsm.jl:67-95: X*A = B with A lower; one B row per
 * work-group, backward over columns, A[i,col]/diag[col]. */
static void OFN(oright_lower_trsm)(OFN(oview) A, OFN(oview) B, int nthreads) {
    long n = B.nc, m = B.nr;
#pragma omp parallel for schedule(static) num_threads(nthreads > 0 ? nthreads : 1)
    for (long row = 0; row < m; ++row) {
        OT *d = (OT *)malloc(sizeof(OT) * (size_t)n);
        OT *br = (OT *)malloc(sizeof(OT) * (size_t)n);
        for (long c = 0; c < n; ++c) {
            d[c] = *OFN(oat)(A, c, c);
            br[c] = *OFN(oat)(B, row, c) / d[c];
        }
        for (long i = n - 1; i >= 0; --i)
            for (long c = 0; c < i; ++c) {
                OT arow = *OFN(oat)(A, i, c) / d[c];
                br[c] -= br[i] * arow;
            }
        for (long c = 0; c < n; ++c) *OFN(oat)(B, row, c) = br[c];
        free(d);
        free(br);
    }
}

/* right_upper_kernel, src/trsm.jl:98-126, called with Transpose(A) (:146-150):
 * forward over columns, At[col,i] = A[i,col]. */
static void OFN(oright_upper_trsm)(OFN(oview) A, OFN(oview) B, int nthreads) {
    long n = B.nc, m = B.nr;
    OFN(oview) At = OFN(otrans)(A);
#pragma omp parallel for schedule(static) num_threads(nthreads > 0 ? nthreads : 1)
    for (long row = 0; row < m; ++row) {
        OT *d = (OT *)malloc(sizeof(OT) * (size_t)n);
        OT *br = (OT *)malloc(sizeof(OT) * (size_t)n);
        for (long c = 0; c < n; ++c) {
            d[c] = *OFN(oat)(At, c, c);
            br[c] = *OFN(oat)(B, row, c) / d[c];
        }
        for (long i = 0; i < n; ++i)
            for (long c = i + 1; c < n; ++c) {
                OT arow = *OFN(oat)(At, c, i) / d[c];
                br[c] -= br[i] * arow;
            }
        for (long c = 0; c < n; ++c) *OFN(oat)(B, row, c) = br[c];
        free(d);
        free(br);
    }
}

/* TRMM leaves, src/trmm.jl:43-312.  Each output element is a dot product over
 * the triangle of the (<=16) diagonal block, summed in ascending k (:91, :164,
 * :233, :296); all reads precede all writes (barrier at :98-104).
 * Left side: output column j depends only on column j of B, so it is staged
 * per column.  Right side: output row i depends only on row i of B. */
static void OFN(otrmm_leaf)(int left, int lower, OFN(oview) A, OFN(oview) B, int nthreads) {
    long nr = B.nr, nc = B.nc, N = A.nr;
    if (left) {
#pragma omp parallel for schedule(static) num_threads(nthreads > 0 ? nthreads : 1)
        for (long j = 0; j < nc; ++j) {
            OT *out = (OT *)malloc(sizeof(OT) * (size_t)(nr > 0 ? nr : 1));
            for (long i = 0; i < nr; ++i) {
                OT s = (OT)0;
                long k0 = lower ? 0 : i, k1 = lower ? i : N - 1; /* :91 k in 1:i ; :164 k in i:N */
                for (long k = k0; k <= k1; ++k) s += *OFN(oat)(A, i, k) * *OFN(oat)(B, k, j);
                out[i] = s;
            }
            for (long i = 0; i < nr; ++i) *OFN(oat)(B, i, j) = out[i];
            free(out);
        }
    } else {
#pragma omp parallel for schedule(static) num_threads(nthreads > 0 ? nthreads : 1)
        for (long i = 0; i < nr; ++i) {
            OT *out = (OT *)malloc(sizeof(OT) * (size_t)(nc > 0 ? nc : 1));
            for (long j = 0; j < nc; ++j) {
                OT s = (OT)0;
                long k0 = lower ? j : 0, k1 = lower ? N - 1 : j; /* :233 k in j:N ; :296 k in 1:j */
                for (long k = k0; k <= k1; ++k) s += *OFN(oat)(B, i, k) * *OFN(oat)(A, k, j);
                out[j] = s;
            }
            for (long j = 0; j < nc; ++j) *OFN(oat)(B, i, j) = out[j];
            free(out);
        }
    }
}

/* Leaf dispatch, src/rectrxm.jl:103-125. */
static void OFN(oleaf)(char func, char side, char uplo, OFN(oview) A, OFN(oview) B, int nthreads) {
    if (func == 'S') {
        if (side == 'L' && uplo == 'L')
            OFN(oleft_lower_trsm)(A, B, nthreads);
        else if (side == 'L' && uplo == 'U')
            OFN(oleft_upper_trsm)(A, B, nthreads);
        else if (side == 'R' && uplo == 'L')
            OFN(oright_lower_trsm)(A, B, nthreads);
        else
            OFN(oright_upper_trsm)(A, B, nthreads);
    } else {
        OFN(otrmm_leaf)(side == 'L', uplo == 'L', A, B, nthreads);
    }
}

/* unified_rec, src/rectrxm.jl:101-198. */
static void OFN(ounified_rec)(char func, char side, char uplo, OFN(oview) A, long n, OFN(oview) B,
                              long threshold, int nthreads) {
    if (n <= threshold) { /* :103-125 */
        OFN(oleaf)(func, side, uplo, A, B, nthreads);
        return;
    }
    /* :129-134 split rule: n/2 for a power of two, else the largest power of two below n */
    long mid;
    if ((n & (n - 1)) == 0)
        mid = n / 2;
    else {
        mid = 1;
        while (mid * 2 < n) mid *= 2;
    }
    long rem = n - mid;
    OFN(oview) A11 = OFN(osub)(A, 0, mid, 0, mid);          /* :137 */
    OFN(oview) A22 = OFN(osub)(A, mid, rem, mid, rem);      /* :138 */
    OFN(oview) A21 = OFN(osub)(A, mid, rem, 0, mid);        /* :139 */
    OFN(oview) A12 = OFN(osub)(A, 0, mid, mid, rem);        /* :140 */
    OFN(oview) B1, B2;
    if (side == 'L') { /* :143-149 */
        B1 = OFN(osub)(B, 0, mid, 0, B.nc);
        B2 = OFN(osub)(B, mid, rem, 0, B.nc);
    } else {
        B1 = OFN(osub)(B, 0, B.nr, 0, mid);
        B2 = OFN(osub)(B, 0, B.nr, mid, rem);
    }
    int forward = (side == 'L' && uplo == 'L' && func == 'S') || (side == 'R' && uplo == 'U' && func == 'S') ||
                  (side == 'L' && uplo == 'U' && func == 'M') || (side == 'R' && uplo == 'L' && func == 'M');
    if (forward) { /* :153-176 */
        OFN(ounified_rec)(func, side, uplo, A11, mid, B1, threshold, nthreads);
        if (side == 'L') {
            if (func == 'S')
                OFN(ogemm_sub)(B2, A21, B1, nthreads); /* :164 B2 -= A21*B1 */
            else
                OFN(ogemm_add)(A12, B2, B1, nthreads); /* :166 B1 += A12*B2 */
        } else {
            if (func == 'S')
                OFN(ogemm_sub)(B2, B1, A12, nthreads); /* :170 B2 -= B1*A12 */
            else
                OFN(ogemm_add)(B2, A21, B1, nthreads); /* :172 B1 += B2*A21 */
        }
        OFN(ounified_rec)(func, side, uplo, A22, rem, B2, threshold, nthreads);
    } else { /* :177-196 */
        OFN(ounified_rec)(func, side, uplo, A22, rem, B2, threshold, nthreads);
        if (side == 'L') {
            if (func == 'S')
                OFN(ogemm_sub)(B1, A12, B2, nthreads); /* :184 B1 -= A12*B2 */
            else
                OFN(ogemm_add)(A21, B1, B2, nthreads); /* :186 B2 += A21*B1 */
        } else {
            if (func == 'S')
                OFN(ogemm_sub)(B1, B2, A21, nthreads); /* :190 B1 -= B2*A21 */
            else
                OFN(ogemm_add)(B1, A12, B2, nthreads); /* :192 B2 += B1*A12 */
        }
        OFN(ounified_rec)(func, side, uplo, A11, mid, B1, threshold, nthreads);
    }
}

/* unified_rectrxm!, src/rectrxm.jl:43-76.  A is n x n column-major (lda),
 * B is n x m (side 'L') or m x n (side 'R') column-major (ldb). */
int OFN(oracle_rectrxm)(char side, char uplo, char trans, char func, double alpha, long n, long m,
                        const OT *Ap, long lda, OT *Bp, long ldb, int nthreads) {
    long threshold = 16; /* :52 */
    OFN(oview) A = {(OT *)Ap, 1, lda, n, n};
    OFN(oview) B;
    if (side == 'L') {
        B.p = Bp; B.rs = 1; B.cs = ldb; B.nr = n; B.nc = m;
    } else {
        B.p = Bp; B.rs = 1; B.cs = ldb; B.nr = m; B.nc = n;
    }
    if (trans == 'T' || trans == 'C') { /* :56-59; 'C' equals 'T' for real element types */
        A = OFN(otrans)(A);
        uplo = (uplo == 'L') ? 'U' : 'L';
    }
    if (func == 'S') { /* :62-65 */
        threshold = 256;
        for (long j = 0; j < B.nc; ++j)
            for (long i = 0; i < B.nr; ++i) *OFN(oat)(B, i, j) = (OT)alpha * *OFN(oat)(B, i, j);
    }
    OFN(ounified_rec)(func, side, uplo, A, n, B, threshold, nthreads); /* :68 */
    if (func == 'M') { /* :71-73 */
        for (long j = 0; j < B.nc; ++j)
            for (long i = 0; i < B.nr; ++i) *OFN(oat)(B, i, j) = (OT)alpha * *OFN(oat)(B, i, j);
    }
    return 0;
}

/* Same recursion with caller-chosen thresholds; used to exercise the
 * non-power-of-two split branch (src/rectrxm.jl:132) and the GEMM updates at
 * sizes the reference's own thresholds would send straight to a leaf. */
int OFN(oracle_rectrxm_thr)(char side, char uplo, char trans, char func, double alpha, long n, long m,
                            const OT *Ap, long lda, OT *Bp, long ldb, long threshold, int nthreads) {
    OFN(oview) A = {(OT *)Ap, 1, lda, n, n};
    OFN(oview) B = {Bp, 1, ldb, side == 'L' ? n : m, side == 'L' ? m : n};
    if (trans == 'T' || trans == 'C') {
        A = OFN(otrans)(A);
        uplo = (uplo == 'L') ? 'U' : 'L';
    }
    if (func == 'S')
        for (long j = 0; j < B.nc; ++j)
            for (long i = 0; i < B.nr; ++i) *OFN(oat)(B, i, j) = (OT)alpha * *OFN(oat)(B, i, j);
    OFN(ounified_rec)(func, side, uplo, A, n, B, threshold, nthreads);
    if (func == 'M')
        for (long j = 0; j < B.nc; ++j)
            for (long i = 0; i < B.nr; ++i) *OFN(oat)(B, i, j) = (OT)alpha * *OFN(oat)(B, i, j);
    return 0;
}

/* GEMM_ADD!(X, Y, C): C += X*Y ; X is N x R, Y is R x M, C is N x M. */
void OFN(oracle_gemm_add)(long N, long R, long M, const OT *X, long ldx, const OT *Y, long ldy, OT *C, long ldc,
                          int nthreads) {
    OFN(oview) vx = {(OT *)X, 1, ldx, N, R}, vy = {(OT *)Y, 1, ldy, R, M}, vc = {C, 1, ldc, N, M};
    OFN(ogemm_add)(vx, vy, vc, nthreads);
}

/* GEMM_SUB!(C, X, Y): C -= X*Y ; C is N x M, X is N x R, Y is R x M. */
void OFN(oracle_gemm_sub)(long N, long R, long M, OT *C, long ldc, const OT *X, long ldx, const OT *Y, long ldy,
                          int nthreads) {
    OFN(oview) vx = {(OT *)X, 1, ldx, N, R}, vy = {(OT *)Y, 1, ldy, R, M}, vc = {C, 1, ldc, N, M};
    OFN(ogemm_sub)(vc, vx, vy, nthreads);
}

/* The eight exported leaf functions (LeftLowerTRSM! ... RightUpperTRMM!), one
 * kernel launch each in the reference (src/trsm.jl:128-150, src/trmm.jl:332-393),
 * applied to an n x n block A and B (n x m left, m x n right). */
void OFN(oracle_leaf)(char func, char side, char uplo, long n, long m, const OT *Ap, long lda, OT *Bp, long ldb,
                      int nthreads) {
    OFN(oview) A = {(OT *)Ap, 1, lda, n, n};
    OFN(oview) B = {Bp, 1, ldb, side == 'L' ? n : m, side == 'L' ? m : n};
    OFN(oleaf)(func, side, uplo, A, B, nthreads);
}

#undef OFN
#undef OCAT
#undef OCAT2
