// Diagonal-block leaves of the recursive triangular solve / multiply.
//
// Replaces the eight reference leaf kernels: lower_left_kernel,
// upper_left_kernel, right_lower_kernel, right_upper_kernel (src/trsm.jl:5-126)
// and Left/Right{Lower,Upper}TRMM_kernel! (src/trmm.jl:43-312).
//
// All eight (side, uplo) forms, with or without transpose, are reduced to one
// canonical problem on a lower-triangular nb x nb block L:
//   TRSM:  L * X = pre * Bc        (forward substitution)
//   TRMM:  Bc <- post * (L * Bc)
// with L(i, j) = A[i*ars + j*acs] and Bc(i, c) = B[i*brs + c*bcs].  Upper
// triangles are made lower by the reversal permutation (negative strides from
// the last diagonal element), and right-side problems are solved on the
// transpose (X*op(A) = B  <=>  op(A)^T X^T = B^T), so the caller only picks
// pointers and signed strides.
//
// Each lane owns one right-hand-side column and keeps all NB of its entries in
// registers; the block of L sits in LDS and is read with wave-uniform
// addresses (one broadcast LDS read serves all 64 lanes), so the sweep has no
// barriers at all (the reference takes one barrier per row, src/trsm.jl:21-27).
// For the left side the RHS tile is contiguous along i in HBM, so it is staged
// through LDS to turn 64 strided column reads into coalesced row reads.
#include "nla_device.h"
#include "nla_internal.h"

namespace nla {

constexpr int LEAF_COLS = 64;  // RHS columns per work-group (one per lane)

template <typename T, int NB, bool STAGE, bool TRMM>
__global__ __launch_bounds__(LEAF_COLS) void leaf_kernel(int nb, int64_t m, const T *__restrict__ A, int64_t ars,
                                                         int64_t acs, T *__restrict__ B, int64_t brs, int64_t bcs,
                                                         T scale) {
    // TRSM reads L by columns (column j after x[j] is final), TRMM by rows
    // (row i is a dot product over j <= i); store L in the matching order.
    __shared__ __attribute__((aligned(16))) T sl[NB * NB];
    __shared__ T srinv[NB];
    __shared__ T sb[STAGE ? LEAF_COLS * (NB + 1) : 1];

    const int t = threadIdx.x;
    const int64_t c0 = (int64_t)blockIdx.x * LEAF_COLS;

    // Stage L, padded to NB with an identity block.
    for (int e = t; e < NB * NB; e += LEAF_COLS) {
        int i, j;
        if (ars == 1 || ars == -1) {
            i = e % NB;
            j = e / NB;
        } else {
            j = e % NB;
            i = e / NB;
        }
        T v;
        if (i < nb && j < nb)
            v = (j <= i) ? A[i * ars + j * acs] : T(0);
        else
            v = (i == j) ? T(1) : T(0);
        if (TRMM)
            sl[i * NB + j] = v;
        else
            sl[j * NB + i] = v;
    }
    if (!TRMM) {
        for (int i = t; i < NB; i += LEAF_COLS) srinv[i] = (i < nb) ? T(1) / A[i * ars + i * acs] : T(1);
    }

    // Load this lane's column of Bc.
    T x[NB];
    const int64_t col = c0 + t;
    if (STAGE) {
        for (int e = t; e < NB * LEAF_COLS; e += LEAF_COLS) {
            int i = e % NB, c = e / NB;
            int64_t gc = c0 + c;
            sb[c * (NB + 1) + i] = (i < nb && gc < m) ? B[i * brs + gc * bcs] : T(0);
        }
        __syncthreads();
#pragma unroll
        for (int i = 0; i < NB; ++i) x[i] = sb[t * (NB + 1) + i];
    } else {
        __syncthreads();
#pragma unroll
        for (int i = 0; i < NB; ++i) x[i] = (i < nb && col < m) ? B[i * brs + col * bcs] : T(0);
    }

    if (!TRMM) {
        // alpha of unified_rectrxm! (src/rectrxm.jl:64) applied on first touch.
        if (scale != T(1)) {
#pragma unroll
            for (int i = 0; i < NB; ++i) x[i] *= scale;
        }
        // Forward substitution, column oriented: once x[j] is final, remove
        // its contribution from every later row.
#pragma unroll
        for (int j = 0; j < NB; ++j) {
            x[j] = x[j] * srinv[j];
#pragma unroll
            for (int r = j + 1; r < NB; ++r) x[r] = x[r] - sl[j * NB + r] * x[j];
        }
    } else {
        // Triangular product in place: row i needs the original x[0..i], so
        // rows are formed from the bottom up.
#pragma unroll
        for (int i = NB - 1; i >= 0; --i) {
            T s = T(0);
#pragma unroll
            for (int j = 0; j <= i; ++j) s += sl[i * NB + j] * x[j];
            x[i] = scale * s;
        }
    }

    if (STAGE) {
#pragma unroll
        for (int i = 0; i < NB; ++i) sb[t * (NB + 1) + i] = x[i];
        __syncthreads();
        for (int e = t; e < NB * LEAF_COLS; e += LEAF_COLS) {
            int i = e % NB, c = e / NB;
            int64_t gc = c0 + c;
            if (i < nb && gc < m) B[i * brs + gc * bcs] = sb[c * (NB + 1) + i];
        }
    } else {
        if (col < m) {
#pragma unroll
            for (int i = 0; i < NB; ++i)
                if (i < nb) B[i * brs + col * bcs] = x[i];
        }
    }
}

template <typename T, int NB, bool STAGE>
static hipError_t launch_leaf_nb(const LeafArgs<T> &a, hipStream_t s) {
    int64_t groups = (a.m + LEAF_COLS - 1) / LEAF_COLS;
    if (a.trmm)
        hipLaunchKernelGGL((leaf_kernel<T, NB, STAGE, true>), dim3((unsigned)groups), dim3(LEAF_COLS), 0, s, a.nb,
                           a.m, a.A, a.ars, a.acs, a.B, a.brs, a.bcs, a.scale);
    else
        hipLaunchKernelGGL((leaf_kernel<T, NB, STAGE, false>), dim3((unsigned)groups), dim3(LEAF_COLS), 0, s, a.nb,
                           a.m, a.A, a.ars, a.acs, a.B, a.brs, a.bcs, a.scale);
    return hipGetLastError();
}

template <typename T, bool STAGE>
static hipError_t launch_leaf_stage(const LeafArgs<T> &a, hipStream_t s) {
    if (a.nb <= 16) return launch_leaf_nb<T, 16, STAGE>(a, s);
    if (a.nb <= 32) return launch_leaf_nb<T, 32, STAGE>(a, s);
    return launch_leaf_nb<T, 64, STAGE>(a, s);
}

template <typename T>
hipError_t launch_leaf(const LeafArgs<T> &a, hipStream_t s) {
    if (a.m <= 0 || a.nb <= 0) return hipSuccess;
    if (a.nb > LEAF_MAX_NB) return hipErrorInvalidValue;
    // Stage through LDS when Bc is contiguous along i (left side).
    if (a.brs == 1 || a.brs == -1) return launch_leaf_stage<T, true>(a, s);
    return launch_leaf_stage<T, false>(a, s);
}

template hipError_t launch_leaf<float>(const LeafArgs<float> &, hipStream_t);
template hipError_t launch_leaf<double>(const LeafArgs<double> &, hipStream_t);

// Fill an m-by-n column-major block with zeros (alpha == 0 semantics of the
// BLAS: B is set to zero without being read).
template <typename T>
__global__ void zero_kernel(int64_t rows, int64_t cols, T *__restrict__ B, int64_t ldb) {
    int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    int64_t total = rows * cols;
    for (; idx < total; idx += (int64_t)gridDim.x * blockDim.x) {
        int64_t i = idx % rows, j = idx / rows;
        B[i + j * ldb] = T(0);
    }
}

template <typename T>
hipError_t launch_zero(int64_t rows, int64_t cols, T *B, int64_t ldb, hipStream_t s) {
    if (rows <= 0 || cols <= 0) return hipSuccess;
    int64_t total = rows * cols;
    int64_t blocks = (total + 255) / 256;
    if (blocks > 65536) blocks = 65536;
    hipLaunchKernelGGL(zero_kernel<T>, dim3((unsigned)blocks), dim3(256), 0, s, rows, cols, B, ldb);
    return hipGetLastError();
}

template hipError_t launch_zero<float>(int64_t, int64_t, float *, int64_t, hipStream_t);
template hipError_t launch_zero<double>(int64_t, int64_t, double *, int64_t, hipStream_t);

}  // namespace nla
