/*
 * nla.h — C ABI of libnlahip, the MI355X (gfx950) implementation of NextLA.jl's
 * recursive triangular solve / multiply hot path.
 *
 * Conventions shared by every entry point
 *   - Matrices are column-major device pointers with a leading dimension
 *     (element (i, j) at p[i + j*ld], 0-based), as Julia's Matrix / ROCMatrix.
 *   - n is the order of the triangular matrix A; m is the other dimension of B:
 *     B is n x m for side 'L' and m x n for side 'R'.
 *   - Only the `uplo` triangle of A is read; the diagonal is non-unit.
 *   - trans: 'N', 'T' or 'C' ('C' equals 'T' for the real element types).
 *   - dtype: NLA_F32 or NLA_F64; A and B share it.
 *   - stream: a hipStream_t passed as void* (NULL = the null stream).  Work is
 *     enqueued asynchronously; nothing synchronizes inside the plain path, and
 *     results are visible after the caller synchronizes the stream.
 *   - Return value: 0 on success; -k when argument k (1-based) is invalid,
 *     checked before any device work, so B is untouched on error; a positive
 *     NLA_ERR_* code when the HIP runtime reports a failure.
 *   - Device memory: the plain TRSM path keeps one buffer per (device,
 *     stream) for the inverses of its diagonal 64-blocks (n*64 elements),
 *     allocated in stream order the first time a stream needs it and grown
 *     when a larger n arrives; it is reused by later calls on that stream and
 *     held until the process ends.  TRMM, the GEMM entry points and the
 *     leaves allocate nothing.  The batched and mixed-precision forms work in
 *     a caller-provided workspace.
 */
#ifndef NLA_H
#define NLA_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum nla_dtype {
    NLA_F32 = 0,
    NLA_F64 = 1,
};

enum nla_status {
    NLA_OK = 0,
    NLA_ERR_HIP = 1,          /* a HIP runtime call failed */
    NLA_ERR_NO_DEVICE = 2,    /* no gfx950 device visible */
    NLA_ERR_UNSUPPORTED = 3,  /* combination not supported by this build */
    NLA_ERR_WORKSPACE = 4,    /* workspace too small / budget exceeded */
    NLA_ERR_NOT_CONVERGED = 5 /* iterative refinement did not reach the target */
};

/* Library / device information.  Returns the ABI version (>= 1). */
int32_t nla_abi_version(void);

/* Name of the last error on the calling thread (static string, never NULL). */
const char *nla_last_error(void);

/*
 * unified_rectrxm!(side, uplo, transpose, alpha, func, A, B)
 * Replaces src/rectrxm.jl:43-76 (with the recursion src/rectrxm.jl:101-198).
 *   func 'S' (TRSM): solve op(A)·X = alpha·B (side 'L') or X·op(A) = alpha·B
 *                    (side 'R'); X overwrites B.
 *   func 'M' (TRMM): B <- alpha·op(A)·B (side 'L') or alpha·B·op(A) (side 'R').
 * Argument positions for error codes: 1 side, 2 uplo, 3 trans, 4 func,
 * 5 dtype, 6 alpha, 7 n, 8 m, 9 A, 10 lda, 11 B, 12 ldb, 13 stream.
 */
int32_t nla_rectrxm(char side, char uplo, char trans, char func, int32_t dtype, double alpha, int64_t n,
                    int64_t m, const void *A, int64_t lda, void *B, int64_t ldb, void *stream);

/*
 * nla_rectrxm without allocation (no counterpart in the reference; for
 * callers that cannot allocate device memory: code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc).  Same arguments,
 * result and argument checks as nla_rectrxm.  The TRSM's scratch (the
 * inverses of A's diagonal 64-blocks and a flag per block, at most
 * NLA_NOALLOC_SCRATCH_BYTES: enough for FP64 n up to 32768, FP32 up to 65536) lives in a
 * device array of the library that each device holds until the process
 * ends; a larger n returns NLA_ERR_WORKSPACE before any device work (B
 * untouched).  The call makes no hipMalloc / hipFree.  Calls that overlap on
 * one device (two streams or host threads at once) share that array and
 * race; serialize them.
 */
#define NLA_NOALLOC_SCRATCH_BYTES ((int64_t)16 * 1024 * 1024 + 64 * 1024)

int32_t nla_rectrxm_noalloc(char side, char uplo, char trans, char func, int32_t dtype, double alpha, int64_t n,
                            int64_t m, const void *A, int64_t lda, void *B, int64_t ldb, void *stream);

/*
 * GEMM_ADD!(X, Y, C): C += X·Y, X is N x R, Y is R x M, C is N x M.
 * Replaces src/matmul.jl:69-73 (kernel matmul!, src/matmul.jl:5-66).
 */
int32_t nla_gemm_add(int32_t dtype, int64_t N, int64_t R, int64_t M, const void *X, int64_t ldx, const void *Y,
                     int64_t ldy, void *C, int64_t ldc, void *stream);

/*
 * GEMM_SUB!(C, X, Y): C -= X·Y, C is N x M, X is N x R, Y is R x M.
 * Replaces src/matmul.jl:76-80.  C comes first, as in the reference.
 */
int32_t nla_gemm_sub(int32_t dtype, int64_t N, int64_t R, int64_t M, void *C, int64_t ldc, const void *X,
                     int64_t ldx, const void *Y, int64_t ldy, void *stream);

/*
 * General GEMM update used by the recursion: C = beta·C + alpha·op(X)·op(Y),
 * op(X) is M x K, op(Y) is K x N; transx/transy in {'N','T'}.
 * No counterpart of its own in the reference: it is matmul! (src/matmul.jl:5-66)
 * with the transA/transB arguments that kernel declares and the alpha scaling
 * of src/rectrxm.jl:64,72 folded in.
 */
int32_t nla_gemm(int32_t dtype, char transx, char transy, int64_t M, int64_t N, int64_t K, double alpha,
                 const void *X, int64_t ldx, const void *Y, int64_t ldy, double beta, void *C, int64_t ldc,
                 void *stream);

/*
 * Diagonal-block leaves: LeftLowerTRSM!, LeftUpperTRSM!, RightLowerTRSM!,
 * RightUpperTRSM! (src/trsm.jl:128-150) for func 'S', and LeftLowerTRMM!,
 * LeftUpperTRMM!, RightLowerTRMM!, RightUpperTRMM! (src/trmm.jl:332-393) for
 * func 'M'.  Solves A·X = B / X·A = B or forms A·B / B·A in place, with A
 * n x n (uplo triangle, no transpose, alpha = 1).  The reference kernels are
 * limited to n <= 1024 (TRSM) and n <= 16 (TRMM); here any n is accepted and
 * blocks larger than one leaf go through the recursion.
 * Argument positions: 1 func, 2 side, 3 uplo, 4 dtype, 5 n, 6 m, 7 A, 8 lda,
 * 9 B, 10 ldb, 11 stream.
 */
int32_t nla_leaf(char func, char side, char uplo, int32_t dtype, int64_t n, int64_t m, const void *A, int64_t lda,
                 void *B, int64_t ldb, void *stream);

/*
 * Batched solves/multiplies of many independent systems (no counterpart in
 * the reference, which makes one unified_rectrxm! call per system).  All
 * systems of a batch share side, uplo, trans, func and dtype; each has its
 * own n (<= 65536), m, A, B and alpha.  A system of up to 1024 rows is solved
 * as one fused diagonal panel by a persistent kernel that pulls (system,
 * 16-column strip) tasks, longest first, from a device queue; a larger one
 * in 512-row stages, each a panel launch plus one grouped GEMM over all
 * systems, so a plan takes a fixed number of launches.
 *
 *   nla_batch_workspace_bytes  device workspace the plan needs (descriptors,
 *                              task queue and, for TRSM, the inverted 64x64
 *                              diagonal blocks of every system).
 *   nla_batch_plan_create      validates the items and uploads the plan into
 *                              the caller's workspace, ordered on `stream`;
 *                              the plan keeps pointers, so A_i and B_i must
 *                              stay allocated.
 *   nla_batch_plan_execute     enqueues the batch on `stream`: kernels only,
 *                              no allocation, no memset, no host copy, so it
 *                              can be captured into a hipGraph; with
 *                              use_graph != 0 the library captures it itself
 *                              on the first call and replays the graph after.
 *   nla_batch_plan_destroy     releases the host handle (and its graph).
 * Argument positions for nla_batch_plan_create: 1 side, 2 uplo, 3 trans,
 * 4 func, 5 dtype, 6 items, 7 count, 8 workspace, 9 workspace_bytes,
 * 10 stream, 11 plan.
 */
typedef struct nla_batch_item {
    int64_t n;       /* order of A_i (<= 65536) */
    int64_t m;       /* other dimension of B_i */
    const void *A;   /* column-major, leading dimension lda */
    int64_t lda;
    void *B;         /* column-major, leading dimension ldb */
    int64_t ldb;
    double alpha;
} nla_batch_item;

typedef struct nla_batch_plan_s *nla_batch_plan;

int64_t nla_batch_workspace_bytes(char func, int32_t dtype, const nla_batch_item *items, int64_t count);

int32_t nla_batch_plan_create(char side, char uplo, char trans, char func, int32_t dtype, const nla_batch_item *items,
                              int64_t count, void *workspace, int64_t workspace_bytes, void *stream,
                              nla_batch_plan *plan);

int32_t nla_batch_plan_execute(nla_batch_plan plan, void *stream, int32_t use_graph);

int32_t nla_batch_plan_destroy(nla_batch_plan plan);

/*
 * The task graph of a batch plan, for host-side checking (no device work):
 * one record per task, in launch order (launch 0 is the 64-block inversion
 * for func 'S'), in canonical coordinates (rows: the order of op(A); side
 * 'R' systems are described on their transpose, columns = the m systems'
 * right-hand sides).
 *   kind 0: invert the 64-block of item `system` at rows r0..r0+rows;
 *   kind 1: panel strip: rows r0..r0+rows, columns c0..c0+cols of B are
 *           solved (TRSM: X = alpha inv(M_rr) B) or multiplied
 *           (TRMM: B = alpha tri(M_rr) B) with the diagonal block M_rr;
 *   kind 2: GEMM tile: B(rows, cols) = beta B(rows, cols)
 *           + alpha M(rows, k0..k0+kn) B(k0..k0+kn, cols).
 * Tasks of one launch run concurrently in any order; launches run in order.
 * Returns the number of records (writing at most `capacity`), or -k for an
 * invalid argument k (the checks of nla_batch_plan_create).
 */
typedef struct nla_batch_task_desc {
    int32_t launch, kind;
    int64_t system;
    int64_t r0, rows, c0, cols, k0, kn;
    double alpha, beta;
} nla_batch_task_desc;

int64_t nla_batch_plan_describe(char side, char uplo, char trans, char func, int32_t dtype, const nla_batch_item *items,
                                int64_t count, nla_batch_task_desc *out, int64_t capacity);

/*
 * BLAS-style xTRSM / xTRMM with the diagonal argument: the same operation as
 * nla_rectrxm with func 'S' / 'M', plus diag:
 *   'N'  non-unit diagonal (as nla_rectrxm);
 *   'U'  unit diagonal: A's diagonal is taken as 1 and never read.
 * Replaces the reference's trsm(side, uplo, transa, diag, alpha, A, B) and
 * trmm(...) wrappers (src/trsm.jl:186-204, src/trmm.jl:430-447), which accept
 * transa and diag but ignore them, and gives getrf2!'s unit-lower solve of
 * the recursive LU (src/lu.jl:277) a GPU path.
 * Argument positions: 1 side, 2 uplo, 3 transa, 4 diag, 5 dtype, 6 alpha,
 * 7 n, 8 m, 9 A, 10 lda, 11 B, 12 ldb, 13 stream.
 */
int32_t nla_trsm(char side, char uplo, char transa, char diag, int32_t dtype, double alpha, int64_t n, int64_t m,
                 const void *A, int64_t lda, void *B, int64_t ldb, void *stream);

int32_t nla_trmm(char side, char uplo, char transa, char diag, int32_t dtype, double alpha, int64_t n, int64_t m,
                 const void *A, int64_t lda, void *B, int64_t ldb, void *stream);

/*
 * Recursive LU factorization with partial pivoting, A = P·L·U, in place
 * (replaces getrf2!(A, ipiv, info), src/lu.jl:185-300: LAPACK dgetrf2's
 * recursion).  Splits the columns at n1 = min(m,n)/2, factors the left
 * block, applies its interchanges to the right block, solves
 * A12 = L11^-1·A12 with the unit-lower TRSM of this library (nla_trsm,
 * diag 'U'), updates A22 -= A21·A12 (nla_gemm), factors A22 and applies its
 * interchanges to the left block; a single column is one kernel (pivot =
 * first index of the largest |a|, swap, scale by the reciprocal).
 *   A     m x n column-major (leading dimension lda), FP32 or FP64;
 *   ipiv  device array of min(m,n) int64: row i was interchanged with row
 *         ipiv[i] (1-based, as the reference);
 *   info  device int64: 0, or the 1-based index of the first exactly zero
 *         pivot (the factorization still completes, as LAPACK's); -1 if a
 *         leaf spread over several work-groups (NLA_CFG_LU_SPLIT_ROWS) timed
 *         out waiting for the others, when the factors are not valid.
 * Enqueued on `stream`, no host synchronization.  Differences from the
 * reference: info is returned (the reference's Int argument cannot carry
 * it out), and the factorization does not stop at a zero pivot.
 * Argument positions: 1 dtype, 2 m, 3 n, 4 A, 5 lda, 6 ipiv, 7 info,
 * 8 stream.
 */
int32_t nla_getrf2(int32_t dtype, int64_t m, int64_t n, void *A, int64_t lda, int64_t *ipiv, int64_t *info,
                   void *stream);

/*
 * LAUUM: A := U·U^T (uplo 'U') or L^T·L (uplo 'L') in place, reading and
 * writing only the uplo triangle (replaces lauum!(uplo, n, A, ib),
 * src/lauum.jl:52-186; the block size ib is internal here).  Recursion on a
 * 2 x 2 split: the diagonal blocks recursively (64-row blocks in one
 * work-group each), the off-diagonal product into the diagonal block by the
 * GEMM update (through a scratch square, then added to the triangle), and
 * the off-diagonal block by this library's TRMM (nla_rectrxm, trans 'T').
 * The workspace (nla_lauum_workspace_bytes: (n/2 rounded up to 64)^2
 * elements) is the caller's; none is needed for n <= 64.
 * Argument positions: 1 uplo, 2 dtype, 3 n, 4 A, 5 lda, 6 workspace,
 * 7 workspace_bytes, 8 stream.
 */
int64_t nla_lauum_workspace_bytes(int32_t dtype, int64_t n);

int32_t nla_lauum(char uplo, int32_t dtype, int64_t n, void *A, int64_t lda, void *workspace, int64_t workspace_bytes,
                  void *stream);

/*
 * Block reflector application (replaces larfb!(side, trans, direct, storev,
 * m, n, k, V, ldv, T, C, work), src/larfb.jl:37-142, for forward,
 * columnwise reflectors, the only form the reference supports):
 * H = I - V·T·V^T with V unit lower (m x k for side 'L', n x k for 'R'; its
 * diagonal and upper part are not read) and T k x k upper triangular;
 * C (m x n) := H·C, H^T·C, C·H or C·H^T (trans 'N' applies H, 'T' / 'C'
 * its transpose).  The three triangular multiplies go through this
 * library's TRMM (nla_trmm, V1 with diag 'U') and the products with V2
 * through the GEMM update.  work: device block of at least (side 'L' ? n : m)
 * rows by k columns, leading dimension ldwork.
 * Argument positions: 1 side, 2 trans, 3 direct, 4 storev, 5 dtype, 6 m,
 * 7 n, 8 k, 9 V, 10 ldv, 11 T, 12 ldt, 13 C, 14 ldc, 15 work, 16 ldwork,
 * 17 stream.
 */
int32_t nla_larfb(char side, char trans, char direct, char storev, int32_t dtype, int64_t m, int64_t n, int64_t k,
                  const void *V, int64_t ldv, const void *T, int64_t ldt, void *C, int64_t ldc, void *work,
                  int64_t ldwork, void *stream);

/*
 * Triangular-pentagonal block reflectors of tile QR (tsmqr!/ttmqr!):
 * H = I - W·T·W^T with W = [I; V] (columnwise, forward; [V; I] backward) or
 * its rowwise transpose, V pentagonal: a rectangular block V1 and an l-row
 * (columnwise) or l-column (rowwise) block V2 whose l x l part is triangular
 * (upper for columnwise forward and rowwise backward, lower otherwise); real
 * types only.
 *
 * nla_pamm replaces pamm!(op, side, storev, direct, m, n, k, l, A1, A2, V, W),
 * src/pamm.jl:50-466:
 *   op 'W': W = A1 + op(V)·A2 (side 'L': W, A1 k x n, A2 m x n, V m x k
 *           columnwise or k x m rowwise) or W = A1 + A2·op(V) (side 'R':
 *           W, A1 m x k, A2 m x n, V n x k or k x n);
 *   op 'A': A2 -= op(V)·W or A2 -= W·op(V); W's l triangle rows (columns)
 *           are overwritten.  A1 is not read (may be NULL).
 * Every product with V2's triangle is this library's TRMM on that l x l
 * block in place; the rest are GEMM updates.  l <= k and l <= m (side 'L')
 * or n ('R'); m, n or k zero returns at once.
 * Argument positions: 1 op, 2 side, 3 storev, 4 direct, 5 dtype, 6 m, 7 n,
 * 8 k, 9 l, 10 A1, 11 lda1, 12 A2, 13 lda2, 14 V, 15 ldv, 16 W, 17 ldw,
 * 18 stream.
 */
int32_t nla_pamm(char op, char side, char storev, char direct, int32_t dtype, int64_t m, int64_t n, int64_t k,
                 int64_t l, const void *A1, int64_t lda1, void *A2, int64_t lda2, const void *V, int64_t ldv, void *W,
                 int64_t ldw, void *stream);

/*
 * nla_parfb replaces parfb!(side, trans, direct, storev, m1, n1, m2, n2, k,
 * l, A1, A2, V, T, work), src/parfb.jl:46-202 (LAPACK's tprfb): applies H
 * (trans 'N') or H^T ('T'/'C') to C = [A1; A2] (side 'L', direct 'F'; [A2; A1]
 * for 'B'; A1's first k of its m1 >= k rows, A2 m2 x n2 with n1 = n2) or
 * C = [A1 A2] (side 'R'; A1's first k of its n1 >= k columns, m1 = m2), in
 * place.  T is k x k, upper
 * (forward) or lower (backward), only that triangle read.  work: k x n2
 * (side 'L') or m2 x k ('R') device block.  Steps: nla_pamm 'W', the TRMM
 * by op(T), A1 -= work, nla_pamm 'A'.  An empty A1 or A2 returns at once
 * (as LAPACK; the reference would then apply T to a work block it never
 * wrote).
 * Argument positions: 1 side, 2 trans, 3 direct, 4 storev, 5 dtype, 6 m1,
 * 7 n1, 8 m2, 9 n2, 10 k, 11 l, 12 A1, 13 lda1, 14 A2, 15 lda2, 16 V, 17 ldv,
 * 18 T, 19 ldt, 20 work, 21 ldwork, 22 stream.
 */
int32_t nla_parfb(char side, char trans, char direct, char storev, int32_t dtype, int64_t m1, int64_t n1, int64_t m2,
                  int64_t n2, int64_t k, int64_t l, void *A1, int64_t lda1, void *A2, int64_t lda2, const void *V,
                  int64_t ldv, const void *T, int64_t ldt, void *work, int64_t ldwork, void *stream);

/*
 * Mixed-precision TRSM with FP64 iterative refinement (no counterpart in the
 * reference; the FP64 solution it converges to is that of unified_rectrxm!'s
 * TRSM, src/rectrxm.jl:43-76).  Solves
 * op(A)·X = alpha·B (side 'L') or X·op(A) = alpha·B (side 'R') for FP64 A, B:
 *   - correction solves run the recursion on an fp32 copy of A's triangle,
 *     with fp32 diagonal blocks and GEMM updates in `precision`:
 *     NLA_LP_BF16 / NLA_LP_FP16 round both operands (fp16: after a
 *     power-of-two scaling of each work-group's operand panels) and
 *     accumulate in fp32 on MFMA;
 *     NLA_LP_FP32 is the plain fp32 update; NLA_LP_FP16X2 splits each
 *     scaled operand into fp16 pairs hi + lo (about 22 significant bits)
 *     and forms hi*hi + hi*lo + lo*hi with three fp16 MFMA products, an
 *     update close to fp32's accuracy at the fp16 kernel's cost;
 *   - every step forms R = alpha·B - op(A)·X (or alpha·B - X·op(A)) to FP64
 *     accuracy (nla_residual: int8 modular products for the shapes
 *     NLA_CFG_RESIDUAL allows, which round the iterate X in B to >= 55 bits
 *     of each system's largest entry first; else the FP64 MFMA GEMM
 *     restricted to the triangle) and the normwise backward error
 *     of every system (a column of B for side 'L', a row for side 'R'):
 *     max_j ||r_j||_2 / (||A||_lb ||x_j||_2 + ||alpha·b_j||_2), where
 *     ||A||_lb is the largest column 2-norm of A's triangle, a lower bound of
 *     ||A||_2 (so the value bounds the error with the exact 2-norm from
 *     above); the loop stops once that is <= tol, otherwise X += the
 *     correction solve of R.  The whole-matrix form ||R||_F / (||A||_F
 *     ||X||_F + ||alpha·B||_F), never larger, is reported beside it;
 *   - a step that less than halves the backward error raises the update
 *     precision for the following steps (bf16 -> fp16 -> fp16 pairs ->
 *     fp32);
 *   - NLA_LP_AUTO: X0 from fp16-pair updates, then each correction in the
 *     precision (bf16, fp16, fp16 pairs or fp32) with the least predicted
 *     cost to reach tol, from the measured error and the contraction of the
 *     last step;
 *   - NLA_LP_FP64: X0 is the plain FP64 solve (nla_rectrxm, its scratch in
 *     the workspace), certified by the same residual and stopping test;
 *     corrections, if ever needed, in fp32.  No low-precision update: on
 *     MI355X this is the fastest certified route at the sizes measured (a
 *     solve and one residual), offered beside the refinement routes.
 * X overwrites B.  The call synchronizes `stream` once per step (the
 * stopping test reads two norms on the host).  NLA_ERR_NOT_CONVERGED: after
 * max_iter (<= NLA_MIXED_MAX_ITER) correction steps the error is still above
 * tol; B then holds the last iterate and info its error.  The workspace
 * (nla_mixed_workspace_bytes, 256-byte aligned) holds about
 * 4*n^2 + 20*n*m bytes, plus nla_residual_workspace_bytes(NLA_RES_INT8)
 * when the int8 residual is used (about 8.5*n^2 + 43*n*m at n = 16384).
 * Argument positions: 1 side, 2 uplo, 3 trans, 4 precision, 5 alpha, 6 n,
 * 7 m, 8 A, 9 lda, 10 B, 11 ldb, 12 tol, 13 max_iter, 14 workspace,
 * 15 workspace_bytes, 16 stream, 17 info.
 */
enum nla_precision {
    NLA_LP_BF16 = 0,
    NLA_LP_FP16 = 1,
    NLA_LP_FP32 = 2,
    NLA_LP_AUTO = 3,
    NLA_LP_FP64 = 4,
    NLA_LP_FP16X2 = 5
};

#define NLA_MIXED_MAX_ITER 30

typedef struct nla_mixed_info {
    int32_t iterations;     /* correction steps applied */
    int32_t precision;      /* update precision of the last step (nla_precision) */
    double backward_error;  /* per-system form (the stopping test) of the X returned in B */
    double history[NLA_MIXED_MAX_ITER + 1]; /* backward_error after 0, 1, ... steps */
    double backward_error_frobenius; /* whole-matrix (Frobenius) form of the same X */
    int32_t residual;       /* nla_residual_method of the last residual */
} nla_mixed_info;

int64_t nla_mixed_workspace_bytes(char side, int64_t n, int64_t m);

int32_t nla_rectrxm_mixed(char side, char uplo, char trans, int32_t precision, double alpha, int64_t n, int64_t m,
                          const double *A, int64_t lda, double *B, int64_t ldb, double tol, int32_t max_iter,
                          void *workspace, int64_t workspace_bytes, void *stream, nla_mixed_info *info);

/*
 * FP64-accurate residual of a triangular system (no counterpart in the
 * reference; the residual step of nla_rectrxm_mixed on its own):
 *   R = alpha·B - op(A)·X (side 'L': X, B, R are n x m) or
 *   R = alpha·B - X·op(A) (side 'R': m x n), FP64 throughout.
 *   NLA_RES_FP64: the FP64 MFMA GEMM restricted to A's triangle; X is read
 *                 only.
 *   NLA_RES_INT8: Ozaki's scheme II on the int8 MFMA.  op(A)'s rows are
 *                 rounded to bA and X's systems (columns for side 'L', rows
 *                 for 'R') to bX fixed-point bits of their largest entry
 *                 (bA, bX >= 55, from the moduli count the order needs), X
 *                 is overwritten by its rounded value, and R is the residual
 *                 of that X for the rounded A computed exactly (as
 *                 integers modulo up to 20 moduli <= 256, then combined by
 *                 the Chinese remainder theorem) and rounded once to FP64.
 *                 n <= 65536 (NLA_ERR_UNSUPPORTED above).  The call
 *                 synchronizes `stream` once to read its range check; when
 *                 an input is not finite or alpha·B lies outside the
 *                 moduli's range, R is recomputed by the FP64 method.
 * R may not alias B or X.  The workspace (nla_residual_workspace_bytes,
 * 256-byte aligned): about (T/2 + 2T)·n·m + T·n^2/2 bytes for int8 with
 * T ~ 17 moduli, none beyond 256 bytes for FP64.
 * Argument positions: 1 side, 2 uplo, 3 trans, 4 method, 5 alpha, 6 n, 7 m,
 * 8 A, 9 lda, 10 X, 11 ldx, 12 B, 13 ldb, 14 R, 15 ldr, 16 workspace,
 * 17 workspace_bytes, 18 stream.
 */
enum nla_residual_method { NLA_RES_FP64 = 0, NLA_RES_INT8 = 1 };

int64_t nla_residual_workspace_bytes(char side, int32_t method, int64_t n, int64_t m);

int32_t nla_residual(char side, char uplo, char trans, int32_t method, double alpha, int64_t n, int64_t m,
                     const double *A, int64_t lda, double *X, int64_t ldx, const double *B, int64_t ldb, double *R,
                     int64_t ldr, void *workspace, int64_t workspace_bytes, void *stream);

/*
 * Batched FP64-accurate TRSM by fp32 correction solves and FP64 iterative
 * refinement (no counterpart in the reference; each system converges to the
 * FP64 solution of its unified_rectrxm! TRSM, src/rectrxm.jl:43-76).  All
 * systems share side, uplo and trans; each has its own FP64 A_i, B_i, n_i
 * (<= 65536), m_i and alpha_i (nla_batch_item, func 'S').  Per step, for the
 * whole batch at once:
 *   - a batched fp32 solve (nla_batch_* on an fp32 copy of each A_i's
 *     triangle) of the residual scaled by a power of two;
 *   - X_i += correction (FP64) for every system not yet done;
 *   - R_i = alpha_i·B_i - op(A_i)·X_i from a batched FP64 TRMM, and its
 *     normwise backward error ||R_i||_F / (||A_i||_F ||X_i||_F +
 *     ||alpha_i·B_i||_F); a system is done (and its X_i frozen) once that is
 *     <= tol.
 * The loop ends when every system is done or after max_iter (<=
 * NLA_MIXED_MAX_ITER) corrections; the stream is synchronized once per step.
 * A system's result, error and step count do not depend on the rest of the
 * batch.  X_i overwrites B_i.  backward_error and iterations (host arrays of
 * count entries, either may be NULL) receive each system's final error and
 * its number of corrections.  NLA_ERR_NOT_CONVERGED: some system is above
 * tol after max_iter steps.  A_i's entries must lie in the fp32 range.
 *   nla_batch_mixed_workspace_bytes  device workspace of a plan (about
 *     4*n^2 + 20*n*m bytes per system plus the two batched plans);
 *   nla_batch_mixed_plan_create      validates and uploads descriptors on
 *     `stream` (A_i, B_i must stay allocated while the plan exists);
 *   nla_batch_mixed_plan_execute     runs the refinement; use_graph != 0
 *     replays the batched solve and product from hipGraphs;
 *   nla_batch_mixed_plan_destroy     releases the plan.
 * Argument positions for create: 1 side, 2 uplo, 3 trans, 4 items, 5 count,
 * 6 workspace, 7 workspace_bytes, 8 stream, 9 plan; for execute: 1 plan,
 * 2 tol, 3 max_iter, 4 stream, 5 use_graph, 6 backward_error, 7 iterations.
 */
typedef struct nla_batch_mixed_plan_s *nla_batch_mixed_plan;

int64_t nla_batch_mixed_workspace_bytes(char side, const nla_batch_item *items, int64_t count);

int32_t nla_batch_mixed_plan_create(char side, char uplo, char trans, const nla_batch_item *items, int64_t count,
                                    void *workspace, int64_t workspace_bytes, void *stream,
                                    nla_batch_mixed_plan *plan);

int32_t nla_batch_mixed_plan_execute(nla_batch_mixed_plan plan, double tol, int32_t max_iter, void *stream,
                                     int32_t use_graph, double *backward_error, int32_t *iterations);

int32_t nla_batch_mixed_plan_destroy(nla_batch_mixed_plan plan);

/*
 * Continuous-batching request server (no counterpart in the reference, which
 * serves one unified_rectrxm! call at a time, test/unified_rectrxm.jl:33).
 * A request is one unified_rectrxm!(side, uplo, trans, alpha, func, A, B)
 * call whose column-major A and B already live in HBM.  The server admits
 * requests against an HBM budget (B; A unless it is a resident factor, which
 * is charged once; the batched workspace a request needs), groups queued
 * requests of one signature (dtype, side, uplo, trans, func) into one
 * batched plan (nla_batch_*) on one of max_inflight streams, and reports
 * completions from nla_serve_step.  With fp64_trsm = NLA_SERVE_REFINED,
 * FP64 TRSM requests are served by batched refinement (nla_batch_mixed_*,
 * per-system backward error <= tol) on `workers` worker threads, each with
 * its own stream.  A request's result is the same alone and in any batch.
 *
 *   nla_serve_submit records an event on `stream` (the submitter's): the
 *     request runs after the work queued there so far (writing A or B).  A
 *     and B must stay allocated until the request's completion is reported.
 *     a_key > 0 names a resident factor: its first request registers A under
 *     the key (charged to the budget once); later requests may pass A = NULL.
 *     Resident factors no request refers to are evicted, least recently used
 *     first, when the budget needs room; nla_serve_evicted lists them.
 *   nla_serve_step retires finished batches, launches what the budget and
 *     free streams allow (never blocks on the device) and hands out up to
 *     `capacity` completion records (a status other than NLA_OK fails that
 *     request only; NLA_ERR_WORKSPACE: it can never fit the budget).
 *   nla_serve_drain steps until nothing is queued or running; the
 *     completions stay for nla_serve_step to hand out.
 *   nla_serve_destroy drains, stops the workers, releases the streams.
 * One host thread at a time may call a server's functions.
 */
enum nla_serve_fp64 { NLA_SERVE_PLAIN = 0, NLA_SERVE_REFINED = 1 };

typedef struct nla_serve_request {
    char side, uplo, trans, func; /* as nla_rectrxm */
    int32_t dtype;
    double alpha;
    int64_t n, m;
    const void *A; /* may be NULL for a registered resident factor */
    int64_t lda;
    void *B;
    int64_t ldb;
    int64_t a_key; /* > 0: resident factor key; 0: none */
} nla_serve_request;

typedef struct nla_serve_done {
    int64_t ticket;
    int32_t status;        /* NLA_OK or the request's error */
    int32_t iterations;    /* refined requests: correction steps */
    double backward_error; /* refined requests: per-system backward error */
} nla_serve_done;

typedef struct nla_serve_stats_t {
    int64_t budget_bytes, used_bytes, queued, running, completed_pending, resident;
} nla_serve_stats_t;

typedef struct nla_server_s *nla_server;

int32_t nla_serve_create(int64_t budget_bytes, int32_t max_inflight, int32_t max_batch, int32_t fp64_trsm,
                         int32_t workers, int32_t use_graph, nla_server *out);

/* refined serving: target, step limit, and the update precision of requests
 * above the batched size limit (defaults 1e-14, 10, NLA_LP_FP16) */
int32_t nla_serve_set_refinement(nla_server server, double tol, int32_t max_iter, int32_t precision);

int32_t nla_serve_submit(nla_server server, const nla_serve_request *request, void *stream, int64_t *ticket);

int32_t nla_serve_step(nla_server server, nla_serve_done *done, int64_t capacity, int64_t *ndone);

int32_t nla_serve_drain(nla_server server);

int32_t nla_serve_stats(nla_server server, nla_serve_stats_t *stats);

int32_t nla_serve_evicted(nla_server server, int64_t *keys, int64_t capacity, int64_t *count);

int32_t nla_serve_destroy(nla_server server);

/*
 * Process-wide choice of the methods used below the recursion (no
 * counterpart in the reference, whose thresholds are constants,
 * src/rectrxm.jl:52,63).  Defaults come from the environment
 * (NLA_LEAF_MODE=subst, NLA_PANEL=0) and can be changed between calls:
 *   NLA_CFG_LEAF_MODE: 1 = inverted 64-blocks applied on MFMA (default),
 *                      0 = register/LDS substitution leaf;
 *   NLA_CFG_PANEL:     1 = diagonal blocks up to the panel size solved by
 *                      one fused panel launch (default, needs leaf mode 1),
 *                      0 = recurse to 64-leaves;
 *   NLA_CFG_PANEL_ROWS_F32 / _F64: panel size for that dtype, a multiple of
 *                      64 in [64, 1024], or 0 (default): by layout, 512
 *                      where the panel's rows are contiguous (left side
 *                      with trans 'N', right side with 'T'), else 256;
 *                      above 512 rows FP64 panels use 16 columns per group;
 *   NLA_CFG_PANEL_STRIP_F32 / _F64: right-hand-side columns per panel
 *                      work-group, 16 (default) or 32;
 *   NLA_CFG_PANEL_WAVES: waves per panel work-group (panels up to 512 rows):
 *                      4, or 8, where two waves share each 16-row output
 *                      tile and split its k range; 0 (default) = 4;
 *   NLA_CFG_BATCH_CLASSES: size classes new batch plans may use, bit c for
 *                      systems up to 128 << c rows (c = 0..3; the 1024 class
 *                      is always on; default 0x1F: all, with bit 4).  A system goes to the
 *                      smallest enabled class that holds it; smaller classes
 *                      use less LDS per work-group, so more of them run per CU.
 *                      Bit 4 (0x10): systems above 512 rows are solved in
 *                      512-row stages (panels of the 512 class plus a grouped
 *                      GEMM) instead of as one 1024-row panel.
 *   NLA_CFG_LU_SPLIT_ROWS: nla_getrf2 leaves (blocks of up to 8 columns) of at
 *                      least this many rows are held in registers: one
 *                      work-group up to 4096 rows, above that one per 2048
 *                      rows up to 131072 rows (default 1: every leaf up to
 *                      that size; 0 = the one-work-group leaf that works in
 *                      global memory, for every size).
 *   NLA_CFG_RESIDUAL:  residual of nla_rectrxm_mixed: NLA_RES_INT8 (default)
 *                      for n in [2048, 65536] and m >= 64, FP64 otherwise;
 *                      NLA_RES_FP64 for every size.  Read when the
 *                      workspace size is queried and when the solve starts
 *                      (the int8 residual needs the larger workspace).
 *   NLA_CFG_LU_SPIN_LIMIT: polls of a multi-work-group LU leaf's bounded
 *                      grid-barrier wait before it gives up (0 = default,
 *                      about a second).  A leaf whose wait runs out (its
 *                      work-groups not all resident, e.g. beside another
 *                      persistent grid) is redone from a copy saved before
 *                      it by the one-work-group leaf, so nla_getrf2's result
 *                      does not depend on it; a small value forces that path.
 *   NLA_CFG_OZ_VARIANT: schedule of the int8 residual's modular products
 *                      (0, default: per-XCD counters over super-tiles of
 *                      4 x 8 tiles, a 4-stage LDS ring of operand copies;
 *                      1: tiles in row order from one counter, 4 stages;
 *                      2: row order, 3 stages; 3: 8 x 4 super-tiles, 4
 *                      stages); results are identical, only the speed
 *                      differs.
 *   NLA_CFG_BATCH_ORDER: order of the tasks within every launch of new batch
 *                      plans (0, default: longest first for load balance;
 *                      k > 0: a pseudo-random order from seed k, for testing
 *                      that results do not depend on it).
 * Returns the previous value, or -1 for an unknown key or invalid value.
 */
enum nla_config_key {
    NLA_CFG_LEAF_MODE = 0,
    NLA_CFG_PANEL = 1,
    NLA_CFG_PANEL_ROWS_F32 = 2,
    NLA_CFG_PANEL_ROWS_F64 = 3,
    NLA_CFG_PANEL_STRIP_F32 = 4,
    NLA_CFG_PANEL_STRIP_F64 = 5,
    NLA_CFG_BATCH_CLASSES = 6,
    NLA_CFG_PANEL_WAVES = 7,
    NLA_CFG_LU_SPLIT_ROWS = 8,
    NLA_CFG_RESIDUAL = 9,
    NLA_CFG_LU_SPIN_LIMIT = 10,
    NLA_CFG_OZ_VARIANT = 11,
    NLA_CFG_BATCH_ORDER = 12
};

int32_t nla_set_config(int32_t key, int32_t value);

/*
 * Kernel timing (instrumentation; no counterpart in the reference, which has
 * no timers).  While enabled, every kernel the library launches is bracketed
 * by HIP events on its stream and tagged with its algorithmic work.
 *   nla_prof_enable(1) clears the records and starts recording; (0) stops.
 *   nla_prof_summary(kind, ...) waits for the recorded events and returns the
 *   launch count, summed kernel time (ms), algorithmic flops and bytes of one
 *   kernel kind: NLA_KIND_GEMM (fp32/fp64 GEMM updates), NLA_KIND_LEAF
 *   (diagonal-block leaves and panels), NLA_KIND_LP_GEMM (bf16/fp16 updates of
 *   the mixed-precision solve), NLA_KIND_RESIDUAL (its residuals: the FP64
 *   product, or the int8 method's conversions and reconstruction),
 *   NLA_KIND_RESIDUAL_GEMM (the int8 method's modular products; flops are
 *   its int8 operations, T·n^2·m for T moduli) or NLA_KIND_LP_PAIR_GEMM (the
 *   fp16-pair updates; flops are the update's 2·M·N·K, which the kernel
 *   forms with three fp16 products).
 */
enum nla_kernel_kind {
    NLA_KIND_GEMM = 0,
    NLA_KIND_LEAF = 1,
    NLA_KIND_LP_GEMM = 2,
    NLA_KIND_RESIDUAL = 3,
    NLA_KIND_RESIDUAL_GEMM = 4,
    NLA_KIND_LP_PAIR_GEMM = 5
};

int32_t nla_prof_enable(int32_t on);

int32_t nla_prof_summary(int32_t kind, int64_t *launches, double *ms, double *flops, double *bytes);

#ifdef __cplusplus
}
#endif

#endif /* NLA_H */
