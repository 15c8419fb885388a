// Block reflector application (nla_larfb in include/nla.h): the reference's
// larfb! (src/larfb.jl:37-142) for forward, columnwise reflectors (the only
// form it supports), H = I - V T V^T with V unit lower (m x k for side 'L',
// n x k for side 'R') and T k x k upper triangular.  Each triangular
// multiply goes through this library's TRMM (nla_trmm: V1 with a unit
// diagonal that is never read, T upper) and each product with V2 through
// the GEMM update; the transposed copies in and out of the work block are
// two small elementwise kernels.
#include <hip/hip_runtime.h>

#include "../../include/nla.h"
#include "nla_internal.h"

namespace nla {

int32_t report_hip(hipError_t e, const char *where);
int32_t report_bad_arg(int k, const char *what);

namespace {

// W(j, i) = C(i, j) for C r x c (W c x r); tr = false: W(i, j) = C(i, j).
template <typename T>
__global__ void copy_block_kernel(int64_t r, int64_t c, const T *__restrict__ C, int64_t ldc, T *__restrict__ W,
                                  int64_t ldw, int tr) {
    for (int64_t j = blockIdx.y; j < c; j += gridDim.y)
        for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < r; i += (int64_t)gridDim.x * blockDim.x) {
            const T v = C[i + j * ldc];
            if (tr) W[j + i * ldw] = v;
            else W[i + j * ldw] = v;
        }
}

// C(i, j) -= W(j, i) (tr) or W(i, j), C r x c.
template <typename T>
__global__ void sub_block_kernel(int64_t r, int64_t c, T *__restrict__ C, int64_t ldc, const T *__restrict__ W,
                                 int64_t ldw, int tr) {
    for (int64_t j = blockIdx.y; j < c; j += gridDim.y)
        for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < r; i += (int64_t)gridDim.x * blockDim.x)
            C[i + j * ldc] -= tr ? W[j + i * ldw] : W[i + j * ldw];
}

inline dim3 grid2(int64_t r, int64_t c) {
    return dim3((unsigned)((r + 255) / 256 < 64 ? (r + 255) / 256 : 64), (unsigned)(c < 65535 ? c : 65535));
}

// copy (sub = false) or subtract (sub = true) between the r x c block of C
// and the work block, transposed when tr
int32_t block_op(bool sub, int32_t dtype, int64_t r, int64_t c, void *C, int64_t ldc, void *W, int64_t ldw, bool tr,
                 hipStream_t s) {
    const int t = tr ? 1 : 0;
    if (dtype == NLA_F64) {
        if (sub)
            hipLaunchKernelGGL(sub_block_kernel<double>, grid2(r, c), dim3(256), 0, s, r, c, (double *)C, ldc,
                               (const double *)W, ldw, t);
        else
            hipLaunchKernelGGL(copy_block_kernel<double>, grid2(r, c), dim3(256), 0, s, r, c, (const double *)C, ldc,
                               (double *)W, ldw, t);
    } else {
        if (sub)
            hipLaunchKernelGGL(sub_block_kernel<float>, grid2(r, c), dim3(256), 0, s, r, c, (float *)C, ldc,
                               (const float *)W, ldw, t);
        else
            hipLaunchKernelGGL(copy_block_kernel<float>, grid2(r, c), dim3(256), 0, s, r, c, (const float *)C, ldc,
                               (float *)W, ldw, t);
    }
    return report_hip(hipGetLastError(), sub ? "larfb subtract" : "larfb copy");
}

}  // namespace
}  // namespace nla

using namespace nla;

extern "C" int32_t nla_larfb(char side, char trans, char direct, char storev, int32_t dtype, int64_t m, int64_t n,
                             int64_t k, const void *V, int64_t ldv, const void *Tm, int64_t ldt, void *C, int64_t ldc,
                             void *work, int64_t ldwork, void *stream) {
    auto up = [](char c) { return (c >= 'a' && c <= 'z') ? (char)(c - 'a' + 'A') : c; };
    side = up(side); trans = up(trans); direct = up(direct); storev = up(storev);
    if (side != 'L' && side != 'R') return report_bad_arg(1, "side");
    if (trans != 'N' && trans != 'T' && trans != 'C') return report_bad_arg(2, "trans");
    if (direct != 'F') return report_bad_arg(3, "direct (only 'F', as the reference)");
    if (storev != 'C') return report_bad_arg(4, "storev (only 'C', as the reference)");
    if (dtype != NLA_F32 && dtype != NLA_F64) return report_bad_arg(5, "dtype");
    if (m < 0) return report_bad_arg(6, "m");
    if (n < 0) return report_bad_arg(7, "n");
    if (k < 0) return report_bad_arg(8, "k");
    if (m == 0 || n == 0 || k == 0) return NLA_OK;  // src/larfb.jl:38-40
    const int64_t nv = side == 'L' ? m : n;       // rows of V
    if (k > nv) return report_bad_arg(8, "k (more reflectors than rows of V)");
    if (!V || ldv < nv) return report_bad_arg(10, "ldv");
    if (!Tm || ldt < k) return report_bad_arg(12, "ldt");
    if (!C || ldc < m) return report_bad_arg(14, "ldc");
    const int64_t wr = side == 'L' ? n : m;  // W is wr x k
    if (!work || ldwork < wr) return report_bad_arg(16, "ldwork");
    hipStream_t s = (hipStream_t)stream;
    const int64_t es = dtype == NLA_F64 ? 8 : 4;
    const char *Vb = (const char *)V;
    char *Cb = (char *)C;
    const void *V2 = Vb + k * es;  // rows k.. of V
    int32_t st;
#define NLA_CK(expr)                       \
    do {                                   \
        st = (expr);                       \
        if (st != NLA_OK) return st;       \
    } while (0)
    // H = I - V T V^T; with trans 'T'/'C' the transpose H^T = I - V T^T V^T
    const char tT = (trans == 'N') ? 'T' : 'N';  // op on T when W holds (.)^T rows (left side)
    if (side == 'L') {
        // W = C1^T (n x k); W = W V1; W += C2^T V2; W = W op(T); C2 -= V2 W^T; W = W V1^T; C1 -= W^T
        NLA_CK(block_op(false, dtype, k, n, C, ldc, work, ldwork, true, s));
        NLA_CK(nla_trmm('R', 'L', 'N', 'U', dtype, 1.0, k, n, V, ldv, work, ldwork, s));
        if (m > k)
            NLA_CK(nla_gemm(dtype, 'T', 'N', n, k, m - k, 1.0, Cb + k * es, ldc, V2, ldv, 1.0, work, ldwork, s));
        NLA_CK(nla_trmm('R', 'U', tT, 'N', dtype, 1.0, k, n, Tm, ldt, work, ldwork, s));
        if (m > k)
            NLA_CK(nla_gemm(dtype, 'N', 'T', m - k, n, k, -1.0, V2, ldv, work, ldwork, 1.0, Cb + k * es, ldc, s));
        NLA_CK(nla_trmm('R', 'L', 'T', 'U', dtype, 1.0, k, n, V, ldv, work, ldwork, s));
        NLA_CK(block_op(true, dtype, k, n, C, ldc, work, ldwork, true, s));
    } else {
        // W = C1 (m x k); W = W V1; W += C2 V2; W = W op(T); C2 -= W V2^T; W = W V1^T; C1 -= W
        NLA_CK(block_op(false, dtype, m, k, C, ldc, work, ldwork, false, s));
        NLA_CK(nla_trmm('R', 'L', 'N', 'U', dtype, 1.0, k, m, V, ldv, work, ldwork, s));
        if (n > k)
            NLA_CK(nla_gemm(dtype, 'N', 'N', m, k, n - k, 1.0, Cb + k * ldc * es, ldc, V2, ldv, 1.0, work, ldwork,
                            s));
        NLA_CK(nla_trmm('R', 'U', trans == 'N' ? 'N' : 'T', 'N', dtype, 1.0, k, m, Tm, ldt, work, ldwork, s));
        if (n > k)
            NLA_CK(nla_gemm(dtype, 'N', 'T', m, n - k, k, -1.0, work, ldwork, V2, ldv, 1.0, Cb + k * ldc * es, ldc,
                            s));
        NLA_CK(nla_trmm('R', 'L', 'T', 'U', dtype, 1.0, k, m, V, ldv, work, ldwork, s));
        NLA_CK(block_op(true, dtype, m, k, C, ldc, work, ldwork, false, s));
    }
#undef NLA_CK
    return NLA_OK;
}
