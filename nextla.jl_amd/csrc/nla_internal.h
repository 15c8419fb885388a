// Host-side declarations shared by the launchers and the C-ABI layer.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace nla {

// Largest diagonal block a single leaf launch handles.
constexpr int LEAF_MAX_NB = 64;

// C = beta*C + alpha*op(X)*op(Y); op(X) is M x K, op(Y) is K x N.
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

template <typename T>
hipError_t launch_gemm(const GemmArgs<T> &a, hipStream_t s);

template <typename T>
hipError_t launch_leaf(const LeafArgs<T> &a, hipStream_t s);

template <typename T>
hipError_t launch_zero(int64_t rows, int64_t cols, T *B, int64_t ldb, hipStream_t s);

}  // namespace nla
