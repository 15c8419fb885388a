// Host-side declarations shared by the launchers and the C-ABI layer.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace nla {

// Largest diagonal block a single leaf launch handles.
constexpr int LEAF_MAX_NB = 64;

// C = beta*C + alpha*op(X)*op(Y); op(X) is M x K, op(Y) is K x N.
// tri: 0 = dense operands; 1/2 = op(X) restricted to its lower/upper
// triangle (k <= r / k >= r); 3/4 = op(Y) restricted to its lower/upper
// triangle (c <= k / c >= k); k-steps outside the triangle are skipped
// (general tile shapes take tri for FP64 only).  leaf != 0 selects the tile
// shape used for diagonal-block leaves: one leaf (<= 64) along M (leaf = 1,
// left side) or N (leaf = 2, right side) and along K.  C may then alias the
// operand it overwrites, because one work-group covers the whole leaf
// dimension of the strip it reads and writes.
template <typename T>
struct GemmArgs {
    int64_t M, N, K;
    bool tx, ty;
    T alpha, beta;
    const T *X;
    int64_t ldx;
    const T *Y;
    int64_t ldy;
    T *C;
    int64_t ldc;
    int tri = 0;
    int leaf = 0;  // 0: general; 1: left leaf (M <= 64); 2: right leaf (N <= 64)
    // General shapes with tri != 0: a zeroed device counter through which a
    // persistent grid takes the tiles longest first (nullptr: one
    // work-group per tile).
    int *queue = nullptr;
};

// Canonical leaf: L(i,j) = A[i*ars + j*acs] (lower, nb x nb),
// Bc(i,c) = B[i*brs + c*bcs] (nb x m).  TRSM: L X = scale*Bc.
// TRMM: Bc <- scale*(L Bc).
template <typename T>
struct LeafArgs {
    bool trmm;
    int nb;
    int64_t m;
    const T *A;
    int64_t ars, acs;
    T *B;
    int64_t brs, bcs;
    T scale;
};

// Largest diagonal block the fused panel kernel solves in one launch, and the
// default RHS columns per work-group (16 or 32 per launch).
constexpr int PANEL_MAX = 1024;
constexpr int PANEL_STRIP = 16;

// Fused diagonal panel: M(i,j) = M[i*mrs + j*mcs] (nbig x nbig, lower or
// upper), Bc(i,c) = B[i*brs + c*bcs] (nbig x m), inverse 64-blocks of M's
// diagonal at W + k*64*64 with element (i,j) at [i*wrs + j*wcs] (TRSM only).
template <typename T>
struct PanelArgs {
    bool trmm, lower;
    int nbig;
    int64_t m;
    const T *M;
    int64_t mrs, mcs;
    T *B;
    int64_t brs, bcs;
    const T *W;
    int64_t wrs, wcs;
    T scale;
    // TRSM: per 64-block flags from launch_trtri_blocks (nullable): a flagged
    // block's product with its inverse is followed by one refinement step
    const int *F = nullptr;
    int strip = PANEL_STRIP;  // RHS columns per work-group: 16 or 32
    bool unit = false;        // unit diagonal (taken as 1, not read)
    int waves = 4;            // waves per work-group: 4, or 8 splitting each phase's k range (<= 512 rows)
    // TRSM, fp32: *xmax = max(*xmax, max |X|) over the panel's solution (a
    // float bit pattern), the scale source of the updates that read it
    unsigned *xmax = nullptr;
};

template <typename T>
hipError_t launch_panel(const PanelArgs<T> &a, hipStream_t s);

// Largest system the persistent batched solver takes as a single panel;
// larger systems (up to BATCH_N_MAX) are solved in stages of BATCH_STAGE
// rows: a batched panel launch per stage, then a grouped GEMM that applies
// the stage's block to the rest of its system.
constexpr int BATCH_MAX = 1024;
// Class-mask bit: systems above 512 rows are solved in 512-row stages instead
// of as one 1024-row panel.
constexpr int BATCH_STAGE_ABOVE_512 = 0x10;
constexpr int BATCH_STAGE = 512;
constexpr int BATCH_N_MAX = 65536;

// One GEMM of a grouped launch: C = beta*C + alpha*op(X)*op(Y) with the
// operand orientation (TX, TY) shared by the launch; tiles_m row tiles.
template <typename T>
struct GroupGemm {
    int64_t M, N, K;
    const T *X;
    int64_t ldx;
    const T *Y;
    int64_t ldy;
    T *C;
    int64_t ldc;
    T alpha, beta;
    int tiles_m;
};

// Tile of the grouped GEMM: GROUP_TILE_R along a system's rows and
// group_tile_c(element size) along its right-hand sides (FP64 32: most staged
// systems have few; FP32 64, its 32 x 32 MFMA blocks filling four waves), so
// a GEMM and its transpose (side 'R') give the same tile count.  rows_m: the
// GEMM's M runs along the system's rows (side 'L'; tiles R x C), else along
// the right-hand sides (side 'R'; tiles C x R).
// FP64 GEMMs of at most GROUP_NARROW right-hand sides (half the staged
// systems) take tiles GROUP_NARROW wide, in a launch of their own.
constexpr int GROUP_TILE_R = 64, GROUP_NARROW = 16;
inline bool group_narrow(int64_t rhs, int elem_bytes) { return elem_bytes == 8 && rhs <= GROUP_NARROW; }
inline int group_tile_c(int elem_bytes, bool narrow) { return elem_bytes == 8 ? (narrow ? GROUP_NARROW : 32) : 64; }
inline int group_tile_m(bool rows_m, int elem_bytes, bool narrow) {
    return rows_m ? GROUP_TILE_R : group_tile_c(elem_bytes, narrow);
}
inline int group_tile_n(bool rows_m, int elem_bytes, bool narrow) {
    return rows_m ? group_tile_c(elem_bytes, narrow) : GROUP_TILE_R;
}

// Persistent grouped GEMM over tasks (gemm index, tile index) taken from a
// counter (cleared by a kernel first); vec: every X and Y of the launch has
// 16-byte aligned columns and leading dimensions.
template <typename T>
hipError_t launch_grouped_gemm(const GroupGemm<T> *g, const int2 *tasks, int ntasks, int *counter, bool tx, bool ty,
                               bool vec, bool rows_m, bool narrow, int groups, hipStream_t s);
template <typename T>
const void *grouped_gemm_ptr(bool tx, bool ty, bool vec, bool rows_m, bool narrow);

// One system of a batch in the panel's canonical form (see PanelArgs).
template <typename T>
struct BatchSys {
    int n, lower;
    int64_t m;
    const T *M;
    int64_t mrs, mcs;
    T *B;
    int64_t brs, bcs;
    const T *W;  // this system's inverse 64-blocks (TRSM)
    int64_t wrs, wcs;
    T scale;
    const int *F;  // their ill-conditioning flags (TRSM)
};

struct BatchTask {
    int sys, strip;
};

// RHS columns per batched task.
template <typename T>
struct BatchStrip {
    static constexpr int W = 16;
};

// Size classes of the persistent batched solver: the kernel for class NBIG
// takes systems of up to NBIG rows and sizes its LDS strip for NBIG.
constexpr int BATCH_CLASSES[4] = {128, 256, 512, 1024};

// Classes a new batch plan may use: bit c enables BATCH_CLASSES[c] (the
// largest is always used); set with nla_set_config(NLA_CFG_BATCH_CLASSES).
int batch_class_mask();
// Task order of new batch plans: 0 longest first; k > 0 a pseudo-random order
// within every launch from seed k (nla_set_config(NLA_CFG_BATCH_ORDER)).
int batch_task_order();
int lu_split_rows();
// Polls of a split LU leaf's bounded grid-barrier wait before it gives up
// and leaves the block to the recovery launch (0 = the default, about a
// second); set with nla_set_config(NLA_CFG_LU_SPIN_LIMIT).
int lu_spin_limit();

// Kernel entry of the persistent batched solver for one size class (for
// occupancy queries); nullptr for an unknown class.
template <typename T>
const void *batch_kernel_ptr(bool trmm, int nbig);

// Tasks [0, ntasks) of one size class; *counter is cleared by a kernel first.
template <typename T>
hipError_t launch_batch(const BatchSys<T> *sys, const BatchTask *tasks, int ntasks, int *counter, bool trmm,
                        int nbig, int groups, hipStream_t s);

// Inverse 64-blocks of every system of a batch (TRSM): one work-group per
// (system, block) pair listed in blk (system index, block index).
template <typename T>
hipError_t launch_batch_trtri(const BatchSys<T> *sys, const int2 *blk, int nblk, hipStream_t s);

template <typename T>
hipError_t launch_gemm(const GemmArgs<T> &a, hipStream_t s);

template <typename T>
hipError_t launch_leaf(const LeafArgs<T> &a, hipStream_t s);

// Inverses of the diagonal blocks of an n x n triangular view
// M(i,j) = A[i*rs + j*cs] (lower if lower, else upper): block b covers rows
// and columns [b*nb, min(n, (b+1)*nb)); W receives inv(M_bb) as an nb x nb
// column-major block at W + b*nb*nb (zero outside the triangle and padded
// with the identity past n).  unit: the diagonal is taken as 1, not read.
template <typename T>
hipError_t launch_trtri_blocks(int64_t n, int nb, const T *A, int64_t rs, int64_t cs, bool lower, T *W,
                               hipStream_t s, bool unit = false, int *flags = nullptr);
// nb = 64 with flags: flags[b] = 1 when block b's 1-norm condition number
// ||M_bb||_1 ||inv(M_bb)||_1 exceeds LEAF_REFINE_KAPPA (or is not finite),
// else 0.  Applying the inverse of such a block loses accuracy in
// proportion to its condition number; the panels refine those blocks.
constexpr double LEAF_REFINE_KAPPA = 256.0;

template <typename T>
hipError_t launch_zero(int64_t rows, int64_t cols, T *B, int64_t ldb, hipStream_t s);

// ---- plain path entry helpers (nla_api.cpp) ---------------------------------

// Bytes of the diagonal-inverse workspace of a plain call (TRSM only).
int64_t rectrxm_workspace_bytes(char func, int32_t dtype, int64_t n);
// nla_rectrxm's argument checks (0 or -k; characters upper-cased in place).
int32_t rectrxm_check(char &side, char &uplo, char &trans, char &func, int32_t dtype, int64_t n, int64_t m,
                      const void *A, int64_t lda, const void *B, int64_t ldb);
// nla_rectrxm with the caller's workspace ws of ws_bytes bytes (checked
// against rectrxm_workspace_bytes: NLA_ERR_WORKSPACE when too small); never
// allocates.
int32_t rectrxm_in_workspace(char side, char uplo, char trans, char func, int32_t dtype, double alpha, int64_t n,
                             int64_t m, const void *A, int64_t lda, void *B, int64_t ldb, hipStream_t s, void *ws,
                             int64_t ws_bytes);
int32_t report_status(int32_t code, const char *msg);
int32_t report_hip(hipError_t e, const char *where);
int32_t report_bad_arg(int k, const char *what);

// ---- mixed precision ------------------------------------------------------

// Where the power-of-two scale of one low-precision GEMM operand comes from
// (fp16 and fp16-pair updates; bf16 has fp32's range and is not scaled).
// Each work-group scales its own operand panel (its BM rows of op(X), its BN
// columns of op(Y), over the whole k range) from the largest magnitude the
// source records for that panel:
//   tab == nullptr: no scaling;
//   ld > 0: a table of 64 x 64 tile maxima of a stored matrix, tile (i, j) at
//     tab[i + j*ld]; operand element (o, k) is stored at (r0 + o, c0 + k),
//     or at (r0 + k, c0 + o) with trans (r0, c0 multiples of 64);
//   ld == 0: slots tab[r0 .. r0 + c0 - 1], one maximum for the whole operand.
// Maxima are float bit patterns (non-negative floats order like them).
struct LpScale {
    const unsigned *tab = nullptr;
    int64_t ld = 0;
    int64_t r0 = 0, c0 = 0;
    bool trans = false;
};

// Low-precision GEMM update C = beta*C + alpha*op(X)*op(Y) on fp32 storage:
// op(X) and op(Y) are scaled by powers of two chosen from their scale
// sources, rounded to bf16 (lp = 0), fp16 (lp = 1) or fp16 pairs (lp = 2),
// multiplied on MFMA with fp32 accumulation and unscaled on the way out.
// X, Y need 16-byte aligned columns (rows when transposed) and
// ldx, ldy % 4 == 0.
struct LpGemmArgs {
    int64_t M, N, K;
    bool tx, ty;
    float alpha, beta;
    const float *X;
    int64_t ldx;
    const float *Y;
    int64_t ldy;
    float *C;
    int64_t ldc;
    int lp;
    LpScale scx, scy;
};
hipError_t launch_lp_gemm(const LpGemmArgs &a, hipStream_t s);
// tab[i + j*nt] = max |A32(r, c)| over tile (i, j) (64 x 64) of the stored
// triangle (lower: r >= c, upper: r <= c), as a float bit pattern; 0 for
// tiles outside the triangle.  nt = ceil(n / 64).
hipError_t launch_tile_absmax(int64_t n, const float *A32, int64_t ld32, bool lower, unsigned *tab, int64_t nt,
                              hipStream_t s);
hipError_t launch_fill_u32(unsigned *p, int64_t n, unsigned v, hipStream_t s);
// Column-major elementwise steps of the refinement loop.
hipError_t launch_tri_to_f32(int64_t n, const double *A, int64_t lda, bool lower, float *A32, int64_t ld32,
                             hipStream_t s);
hipError_t launch_scale_copy(int64_t rows, int64_t cols, double alpha, const double *B, int64_t ldb, double *S,
                             int64_t lds, hipStream_t s);
hipError_t launch_down(int64_t rows, int64_t cols, double scale, const double *R, int64_t ldr, float *D,
                       int64_t ldd, hipStream_t s);
hipError_t launch_up(int64_t rows, int64_t cols, double sigma, const float *D, int64_t ldd, double beta, double *X,
                     int64_t ldx, hipStream_t s);
// out[0] = sum of squares, out[1] = max |.| of a column-major matrix (tri:
// 0 whole, 1 lower triangle, 2 upper triangle); part holds NORM_PARTS_MAX
// pairs of partial results.
constexpr int NORM_PARTS_MAX = 1024;
hipError_t launch_norms(int64_t rows, int64_t cols, const double *P, int64_t ld, int tri, double *part,
                        double *out, hipStream_t s);
// Per-system norms of the stopping test (a system is a column of the
// matrix when by_col, else a row).  launch_sys_sumsq writes, per (chunk c,
// system j), the sum of squares and the absolute maximum over chunk c of the
// reduction dimension to part[2*(c*nsys + j) + {0, 1}]; nch chunks (at most
// SYS_CHUNKS_MAX; sys_chunks() picks a count).  tri (by_col only): 1 / 2
// restricts column c to rows >= c / <= c.
constexpr int SYS_CHUNKS_MAX = 64;
int sys_chunks(int64_t nsys, int64_t len);
hipError_t launch_sys_sumsq(int64_t rows, int64_t cols, const double *P, int64_t ld, bool by_col, int tri, int nch,
                            double *part, hipStream_t s);
// sys2[j] (if not null) = system j's sum of squares; out[0] = total sum of
// squares, out[1] = max |.|, out[2] = largest system 2-norm.
hipError_t launch_sys_fold(int nch, int64_t nsys, const double *part, double *sys2, double *out, hipStream_t s);
// out[0] = ||R||_F^2, out[1] = max |R|, out[2] = ||X||_F^2, out[3] =
// max_j ||r_j|| / (*anorm_lb ||x_j|| + sqrt(s2[j])).
hipError_t launch_sys_eta(int nch, int64_t nsys, const double *pr, const double *px, const double *s2,
                          const double *anorm_lb, double *out, hipStream_t s);

// ---- FP64-accurate residual on the int8 MFMA (nla_ozaki.hip) -------------

// Pairwise coprime moduli, largest first; a residual uses the first T.
constexpr int OZ_TMAX = 20;
constexpr int OZ_MODULI[OZ_TMAX] = {256, 255, 253, 251, 247, 241, 239, 233, 229, 227,
                                    223, 217, 211, 199, 197, 193, 191, 181, 179, 173};
// Largest order: int32 sums of K products of residues (|r| <= 128) stay
// exact for K <= 2^17; 2^16 also keeps the grids of every launch in range.
constexpr int64_t OZ_N_MAX = 65536;
// Fewest bits kept of A's rows and X's columns (rounding error 2^-55 of a
// row's / column's largest entry).
constexpr int OZ_BITS_MIN = 55;

// Constants of the modular products and the reconstruction (kernel argument).
struct OzConst {
    int T;
    int p[OZ_TMAX];
    int y[OZ_TMAX];  // (M / p_t)^-1 mod p_t
    int rc[OZ_TMAX][3];  // centred residues of 2^16, 2^32, 2^48 modulo p_t
    int rhi[OZ_TMAX];    // centred residues lie in [rhi - p + 1, rhi], rhi = floor((p - 1) / 2)
    double invpd[OZ_TMAX];
    // w_t = y_t (M / p_t) (the CRT basis: w_t = 1 mod p_t, 0 mod the
    // others) and M, as four base-2^40 limbs each (exact in double)
    double w[OZ_TMAX][4];
    double ml[4];
    double M;     // product of the T moduli (rounded)
    double invm;  // 1 / M (rounded)
    double smax;  // largest |S_int| the moduli cover (M / 8)
};

// Schedule of the modular-product kernel (NLA_CFG_OZ_VARIANT, 0..3).
int oz_variant();

// Moduli count and bit split for an order-n residual, and the residue planes'
// sizes (256-padded orders npad, mpad).  T = 0: n outside (0, OZ_N_MAX].
struct OzPlan {
    int T, bA, bX;
    int64_t npad, mpad, aplane, xplane, cplane;
    OzConst C;
};
OzPlan oz_plan(int64_t n, int64_t m);

// Workspace offsets (bytes from the workspace start) of an OzPlan.
struct OzLayout {
    int64_t a, x, c, ea, fx, mxa, mxx, flag, queue, total;
};
OzLayout oz_layout(const OzPlan &P, int64_t base);

// Once per solve: A's row scales and residue planes.  Ahat(i,k) =
// A[i*ars + k*acs] (the triangle of op(A), or of its transpose for side R,
// lower or upper).  Clears the flag word; a non-finite A sets it.
hipError_t oz_prepare(const OzPlan &P, const OzLayout &L, char *ws, const double *A, int64_t ars, int64_t acs,
                      bool lower, int64_t n, hipStream_t s);
// The three phases of R = S - Ahat * Xhat (out_t = false: R n x m; out_t =
// true: R m x n, its transpose, for side R), Xhat(k,j) = X[j*xso + k*xsk]:
//   oz_convert_x: X's system scales and residue planes; X is replaced by its
//     value rounded to bX bits per system (a non-finite X sets the flag);
//   oz_products: the T modular products (int8 MFMA);
//   oz_reconstruct: R from S and the product residues; an S that is not
//     finite or exceeds the moduli's range sets the flag word (R is then
//     not valid).
hipError_t oz_convert_x(const OzPlan &P, const OzLayout &L, char *ws, double *X, int64_t xso, int64_t xsk,
                        int64_t n, int64_t m, hipStream_t s);
hipError_t oz_products(const OzPlan &P, const OzLayout &L, char *ws, bool lower, bool out_t, hipStream_t s);
hipError_t oz_reconstruct(const OzPlan &P, const OzLayout &L, char *ws, bool out_t, int64_t n, int64_t m,
                          const double *S, int64_t lds, double *R, int64_t ldr, hipStream_t s);

// ---- batched refinement ----------------------------------------------------

// One FP64 system of a refined batch and its work arrays (column-major; the
// B-shaped ones rows x cols with leading dimension ldw).  Tasks cover bcb
// columns of the B-shaped arrays or acb columns of A.
struct MixSys {
    const double *A;  // stored triangle (lower or upper) read
    int64_t lda;
    float *A32;  // fp32 copy of that triangle
    int64_t lda32;
    double *B;  // right-hand side in, solution out
    int64_t ldb;
    double *S, *T;  // alpha*B; op(A)*X, then the residual
    float *D;       // scaled residual, then the correction
    int64_t ldw;
    int64_t rows, cols, n;
    double alpha;
    int lower;
    int bcb, acb;
    int bt0, bnt, at0, ant;  // task ranges [bt0, bt0 + bnt), [at0, at0 + ant)
};

// Per-system state of the refinement loop.
struct MixState {
    double anorm, snorm, sigma, bwd;
    int done, pad;
};

hipError_t launch_mix_prep_a(const MixSys *sys, const int2 *atasks, int nat, double *apart, hipStream_t s);
hipError_t launch_mix_prep_b(const MixSys *sys, const int2 *btasks, int nbt, double *bpart, hipStream_t s);
hipError_t launch_mix_down(const MixSys *sys, const int2 *btasks, int nbt, const MixState *st, bool from_t,
                           hipStream_t s);
hipError_t launch_mix_up(const MixSys *sys, const int2 *btasks, int nbt, const MixState *st, bool first,
                         hipStream_t s);
hipError_t launch_mix_resid(const MixSys *sys, const int2 *btasks, int nbt, double *bpart, hipStream_t s);
hipError_t launch_mix_final(const MixSys *sys, int64_t count, const double *apart, const double *bpart,
                            MixState *st, bool init, double tol, hipStream_t s);

}  // namespace nla
