// LAUUM: A := U·U^T (uplo 'U') or L^T·L (uplo 'L') in place, only the uplo
// triangle read and written (nla_lauum in include/nla.h; the reference's
// lauum!, src/lauum.jl:52-186 with lauu2!, src/lauu2.jl:48-107).
//
// Recursion on a 2 x 2 split (n1 a multiple of 64), lower case:
//   L11 := L11^T L11             (recursively)
//   L11 += L21^T L21             lower triangle only: the GEMM update into a
//                                scratch square, then a masked add
//   L21 := L22^T L21             TRMM of the hot path (nla_rectrxm 'L','L','T')
//   L22 := L22^T L22             (recursively)
// and the mirror image for upper (U11 += U12 U12^T; U12 := U12 U22^T).
// Diagonal blocks of up to 128 rows are one work-group each in LDS.
#include <hip/hip_runtime.h>

#include "../../include/nla.h"
#include "nla_internal.h"

namespace nla {

int32_t report_hip(hipError_t e, const char *where);
int32_t report_bad_arg(int k, const char *what);

namespace {

constexpr int LAUUM_LEAF = 128;  // largest diagonal block done in one work-group

// One diagonal block of nb <= NBL rows: lower R(i,j) = sum_{k>=i} L(k,i) L(k,j)
// for i >= j; upper R(i,j) = sum_{k>=j} U(i,k) U(j,k) for i <= j.  The
// triangle goes to LDS with zeros elsewhere, so every sum can run over all k
// from the block's first row; each thread owns a 4 x 4 block of R (the
// threads whose block lies outside the triangle only help load).  NBL = 64
// (256 threads) or 128 (1024 threads, 131 KB of LDS in FP64).  Both
// triangles are stored as rows of the factor the thread's index runs over
// (L(., i) for lower, U(i, .) for upper), 4 rows to a block with a stride
// of 4 NBL + 1: the 32 blocks one wave reads at a time start in 32
// different banks.
constexpr int LAUUM_RB = 4;

template <typename T, int NBL>
__global__ __launch_bounds__((NBL / 4) * (NBL / 4)) void lauum_leaf_kernel(int nb, T *__restrict__ A, int64_t lda,
                                                                        int lower) {
    constexpr int NB = NBL / LAUUM_RB, NT = NB * NB, BS = LAUUM_RB * NBL + 1;
    __shared__ T s[NB * BS];
    auto at = [](int i, int k) { return (i / LAUUM_RB) * BS + (i % LAUUM_RB) * NBL + k; };
    const int t = threadIdx.x;
    for (int e = t; e < NBL * NBL; e += NT) {
        const int r = e % NBL, c = e / NBL;
        const bool in = r < nb && c < nb && (lower ? r >= c : r <= c);
        s[lower ? at(c, r) : at(r, c)] = in ? A[r + (int64_t)c * lda] : T(0);
    }
    __syncthreads();
    const int bi = t % NB, bj = t / NB;  // rows 4 bi.., columns 4 bj..
    const int i0 = bi * LAUUM_RB, j0 = bj * LAUUM_RB;
    if ((lower ? bi < bj : bi > bj) || i0 >= nb || j0 >= nb) return;
    T acc[LAUUM_RB][LAUUM_RB];
#pragma unroll
    for (int a = 0; a < LAUUM_RB; ++a)
#pragma unroll
        for (int b = 0; b < LAUUM_RB; ++b) acc[a][b] = T(0);
    // x[a] = L(k, i0 + a) (lower) or U(i0 + a, k) (upper), zero for k below
    // the block's first row / column; y likewise for j0
    const int k0 = lower ? i0 : j0;
    for (int k = k0; k < nb; ++k) {
        T x[LAUUM_RB], y[LAUUM_RB];
#pragma unroll
        for (int a = 0; a < LAUUM_RB; ++a) {
            x[a] = s[at(i0 + a, k)];
            y[a] = s[at(j0 + a, k)];
        }
#pragma unroll
        for (int a = 0; a < LAUUM_RB; ++a)
#pragma unroll
            for (int b = 0; b < LAUUM_RB; ++b) acc[a][b] += x[a] * y[b];
    }
#pragma unroll
    for (int a = 0; a < LAUUM_RB; ++a)
#pragma unroll
        for (int b = 0; b < LAUUM_RB; ++b) {
            const int i = i0 + a, j = j0 + b;
            if (i < nb && j < nb && (lower ? i >= j : i <= j)) A[i + (int64_t)j * lda] = acc[a][b];
        }
}

// C(i,j) += W(i,j) on C's lower (or upper) triangle; W n x n, leading dim n.
template <typename T>
__global__ void tri_add_kernel(int64_t n, const T *__restrict__ W, T *__restrict__ C, int64_t ldc, int lower) {
    for (int64_t j = blockIdx.y; j < n; j += gridDim.y)
        for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
            if (lower ? i >= j : i <= j) C[i + j * ldc] += W[i + j * n];
}

template <typename T>
hipError_t launch_tri_add(int64_t n, const T *W, T *C, int64_t ldc, bool lower, hipStream_t s) {
    const dim3 grid((unsigned)((n + 255) / 256), (unsigned)(n < 65535 ? n : 65535));
    hipLaunchKernelGGL(tri_add_kernel<T>, grid, dim3(256), 0, s, n, W, C, ldc, lower ? 1 : 0);
    return hipGetLastError();
}

template <typename T>
struct Lauum {
    int32_t dtype;
    int64_t lda;
    bool lower;
    T *W;  // scratch, at least (n/2 rounded up to 64)^2 elements
    hipStream_t s;
    int32_t st = NLA_OK;

    void rec(T *A, int64_t n) {
        if (st != NLA_OK || n <= 0) return;
        if (n <= LAUUM_LEAF) {
            if (n <= 64)
                hipLaunchKernelGGL((lauum_leaf_kernel<T, 64>), dim3(1), dim3(256), 0, s, (int)n, A, lda, lower ? 1 : 0);
            else
                hipLaunchKernelGGL((lauum_leaf_kernel<T, 128>), dim3(1), dim3(1024), 0, s, (int)n, A, lda,
                                   lower ? 1 : 0);
            hipError_t e = hipGetLastError();
            if (e != hipSuccess) st = report_hip(e, "lauum leaf");
            return;
        }
        const int64_t n1 = ((n / 2 + 63) / 64) * 64, n2 = n - n1;
        T *A11 = A, *A22 = A + n1 + n1 * lda;
        rec(A11, n1);
        if (st != NLA_OK) return;
        if (lower) {
            T *A21 = A + n1;
            // W = L21^T L21 (n1 x n1), then its lower triangle into L11
            st = nla_gemm(dtype, 'T', 'N', n1, n1, n2, 1.0, A21, lda, A21, lda, 0.0, W, n1, s);
            if (st != NLA_OK) return;
            hipError_t e = launch_tri_add<T>(n1, W, A11, lda, true, s);
            if (e != hipSuccess) {
                st = report_hip(e, "lauum add");
                return;
            }
            // L21 := L22^T L21
            st = nla_rectrxm('L', 'L', 'T', 'M', dtype, 1.0, n2, n1, A22, lda, A21, lda, s);
        } else {
            T *A12 = A + n1 * lda;
            // W = U12 U12^T (n1 x n1), then its upper triangle into U11
            st = nla_gemm(dtype, 'N', 'T', n1, n1, n2, 1.0, A12, lda, A12, lda, 0.0, W, n1, s);
            if (st != NLA_OK) return;
            hipError_t e = launch_tri_add<T>(n1, W, A11, lda, false, s);
            if (e != hipSuccess) {
                st = report_hip(e, "lauum add");
                return;
            }
            // U12 := U12 U22^T
            st = nla_rectrxm('R', 'U', 'T', 'M', dtype, 1.0, n2, n1, A22, lda, A12, lda, s);
        }
        if (st != NLA_OK) return;
        rec(A22, n2);
    }
};

}  // namespace
}  // namespace nla

using namespace nla;

extern "C" {

int64_t nla_lauum_workspace_bytes(int32_t dtype, int64_t n) {
    if ((dtype != NLA_F32 && dtype != NLA_F64) || n < 0) return -1;
    const int64_t h = ((n / 2 + 63) / 64) * 64;
    return (n > LAUUM_LEAF ? h * h : 1) * (dtype == NLA_F64 ? 8 : 4);
}

int32_t nla_lauum(char uplo, int32_t dtype, int64_t n, void *A, int64_t lda, void *workspace,
                  int64_t workspace_bytes, void *stream) {
    uplo = (uplo == 'l') ? 'L' : (uplo == 'u') ? 'U' : uplo;
    if (uplo != 'L' && uplo != 'U') return report_bad_arg(1, "uplo");
    if (dtype != NLA_F32 && dtype != NLA_F64) return report_bad_arg(2, "dtype");
    if (n < 0) return report_bad_arg(3, "n");
    if (n > 0 && A == nullptr) return report_bad_arg(4, "A");
    if (lda < (n > 1 ? n : 1)) return report_bad_arg(5, "lda");
    const int64_t need = nla_lauum_workspace_bytes(dtype, n);
    if (n > LAUUM_LEAF && (workspace == nullptr || workspace_bytes < need)) return report_bad_arg(7, "workspace_bytes");
    if (n == 0) return NLA_OK;
    hipStream_t s = (hipStream_t)stream;
    if (dtype == NLA_F64) {
        Lauum<double> r{dtype, lda, uplo == 'L', (double *)workspace, s};
        r.rec((double *)A, n);
        return r.st;
    }
    Lauum<float> r{dtype, lda, uplo == 'L', (float *)workspace, s};
    r.rec((float *)A, n);
    return r.st;
}

}  // extern "C"
