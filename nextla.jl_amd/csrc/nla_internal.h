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
// triangle (c <= k / c >= k).  leaf != 0 selects the tile shape used for
// diagonal-block leaves: one leaf (<= 64) along M (leaf = 1, left side) or N
// (leaf = 2, right side) and along K.  C may then alias the operand it
// overwrites, because one work-group covers the whole leaf dimension of the
// strip it reads and writes.
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
// RHS columns per work-group.
constexpr int PANEL_MAX = 512;
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
};

template <typename T>
hipError_t launch_panel(const PanelArgs<T> &a, hipStream_t s);

// Largest system the persistent batched solver takes (one panel per system).
constexpr int BATCH_MAX = 1024;

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
};

struct BatchTask {
    int sys, strip;
};

// RHS columns per batched task.
template <typename T>
struct BatchStrip {
    static constexpr int W = 16;
};

// Kernel entry of the persistent batched solver (for occupancy queries).
template <typename T>
const void *batch_kernel_ptr(bool trmm);

template <typename T>
hipError_t launch_batch(const BatchSys<T> *sys, const BatchTask *tasks, int ntasks, int *counter, bool trmm,
                        int groups, hipStream_t s);

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
// with the identity past n).
template <typename T>
hipError_t launch_trtri_blocks(int64_t n, int nb, const T *A, int64_t rs, int64_t cs, bool lower, T *W,
                               hipStream_t s);

template <typename T>
hipError_t launch_zero(int64_t rows, int64_t cols, T *B, int64_t ldb, hipStream_t s);

}  // namespace nla
