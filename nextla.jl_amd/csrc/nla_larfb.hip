// Block reflector application (nla_larfb in include/nla.h): the reference's
// larfb! (src/larfb.jl:37-142) for forward, columnwise reflectors (the only
// form it supports), H = I - V T V^T with V unit lower (m x k for side 'L',
// n x k for side 'R') and T k x k upper triangular.  Each triangular
// multiply goes through this library's TRMM (nla_trmm: V1 with a unit
// diagonal that is never read, T upper) and each product with V2 through
// the GEMM update; the transposed copies in and out of the work block are
// two small elementwise kernels.
#include <hip/hip_runtime.h>

#include <algorithm>

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

// C(i, j) -= W(j, i) (tr) or W(i, j), C r x c; with add, += instead.
template <typename T>
__global__ void sub_block_kernel(int64_t r, int64_t c, T *__restrict__ C, int64_t ldc, const T *__restrict__ W,
                                 int64_t ldw, int tr, int add) {
    for (int64_t j = blockIdx.y; j < c; j += gridDim.y)
        for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < r; i += (int64_t)gridDim.x * blockDim.x) {
            const T w = tr ? W[j + i * ldw] : W[i + j * ldw];
            C[i + j * ldc] = add ? C[i + j * ldc] + w : C[i + j * ldc] - w;
        }
}

inline dim3 grid2(int64_t r, int64_t c) {
    return dim3((unsigned)((r + 255) / 256 < 64 ? (r + 255) / 256 : 64), (unsigned)(c < 65535 ? c : 65535));
}

// copy (sub = false: W <- C) or subtract (sub = true: C -= W; with add,
// C += W) between the r x c block of C and the work block, transposed when tr
int32_t block_op(bool sub, int32_t dtype, int64_t r, int64_t c, void *C, int64_t ldc, void *W, int64_t ldw, bool tr,
                 hipStream_t s, bool add = false) {
    if (r <= 0 || c <= 0) return NLA_OK;
    const int t = tr ? 1 : 0, a = add ? 1 : 0;
    if (dtype == NLA_F64) {
        if (sub)
            hipLaunchKernelGGL(sub_block_kernel<double>, grid2(r, c), dim3(256), 0, s, r, c, (double *)C, ldc,
                               (const double *)W, ldw, t, a);
        else
            hipLaunchKernelGGL(copy_block_kernel<double>, grid2(r, c), dim3(256), 0, s, r, c, (const double *)C, ldc,
                               (double *)W, ldw, t);
    } else {
        if (sub)
            hipLaunchKernelGGL(sub_block_kernel<float>, grid2(r, c), dim3(256), 0, s, r, c, (float *)C, ldc,
                               (const float *)W, ldw, t, a);
        else
            hipLaunchKernelGGL(copy_block_kernel<float>, grid2(r, c), dim3(256), 0, s, r, c, (const float *)C, ldc,
                               (float *)W, ldw, t);
    }
    return report_hip(hipGetLastError(), sub ? "block subtract" : "block copy");
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

// ---- pamm! / parfb!: triangular-pentagonal block reflectors ------------------
//
// The tile-QR kernels tsmqr!/ttmqr! apply H = I - W·T·W^T where W stacks an
// identity block over a pentagonal V (src/parfb.jl:1-45): V is a rectangular
// block V1 plus an l-row (columnwise storage) or l-column (rowwise) block V2
// that is triangular in its last (forward) or first (backward) l x l part.
// pamm! forms the work block W = A1 + op(V)·A2 ('W') or updates
// A2 -= op(V)·W ('A') (src/pamm.jl:86-466); each product with V2's triangle
// goes through this library's TRMM (non-unit diagonal, the triangle of that
// l x l block read in place) and each product with a rectangular part of V
// through the GEMM update.  parfb! (src/parfb.jl:46-202) is pamm 'W', a TRMM
// by op(T), A1 -= W, pamm 'A'.  Indices below are 0-based: the reference's
// mp = min(m-l+1, m) and kp = min(l+1, k) (forward) or mp = min(l+1, m) and
// kp = min(k-l+1, k) (backward) are mp0 = mp - 1 and kp0 = kp - 1.
namespace nla {
namespace {

struct Pamm {
    int32_t dtype;
    int64_t es;
    hipStream_t s;
    char *at(const void *p, int64_t ld, int64_t i, int64_t j) const {
        return (char *)p + (i + j * ld) * es;
    }
    // B (r x c at b) := op(tri(V)) B, the l = r triangle of V at v
    int32_t tri_left(char uplo, bool tr, const char *v, int64_t ldv, int64_t r, int64_t c, char *b, int64_t ldb) const {
        if (r <= 0 || c <= 0) return NLA_OK;
        return nla_trmm('L', uplo, tr ? 'T' : 'N', 'N', dtype, 1.0, r, c, v, ldv, b, ldb, s);
    }
    // B (r x c at b) := B op(tri(V)), the l = c triangle of V at v
    int32_t tri_right(char uplo, bool tr, const char *v, int64_t ldv, int64_t r, int64_t c, char *b, int64_t ldb) const {
        if (r <= 0 || c <= 0) return NLA_OK;
        return nla_trmm('R', uplo, tr ? 'T' : 'N', 'N', dtype, 1.0, c, r, v, ldv, b, ldb, s);
    }
    // C (M x N) := beta C + alpha op(X) op(Y), K the inner dimension (every
    // call below has K > 0: m, n > 0 after the quick return, and the
    // K = m - l, n - l, k - l products are taken only when positive)
    int32_t gemm(char tx, char ty, int64_t M, int64_t N, int64_t K, double alpha, const char *x, int64_t ldx,
                 const char *y, int64_t ldy, double beta, char *c, int64_t ldc) const {
        if (M <= 0 || N <= 0 || K <= 0) return NLA_OK;
        return nla_gemm(dtype, tx, ty, M, N, K, alpha, x, ldx, y, ldy, beta, c, ldc, s);
    }
    int32_t copy(int64_t r, int64_t c, const char *src, int64_t lds, char *dst, int64_t ldd) const {
        return block_op(false, dtype, r, c, (void *)src, lds, dst, ldd, false, s);
    }
    int32_t sub(int64_t r, int64_t c, char *dst, int64_t ldd, const char *w, int64_t ldw) const {
        return block_op(true, dtype, r, c, dst, ldd, (void *)w, ldw, false, s);
    }
    int32_t add(int64_t r, int64_t c, char *dst, int64_t ldd, const char *w, int64_t ldw) const {
        return block_op(true, dtype, r, c, dst, ldd, (void *)w, ldw, false, s, true);
    }
};

#define NLA_CK(expr)                 \
    do {                             \
        const int32_t st_ = (expr);  \
        if (st_ != NLA_OK) return st_; \
    } while (0)

// W = A1 + op(V) A2 (left: W, A1 k x n, A2 m x n) or A1 + A2 op(V) (right:
// W, A1 m x k, A2 m x n); src/pamm.jl:101-324.
int32_t pamm_w(const Pamm &P, bool left, bool col, bool fwd, int64_t m, int64_t n, int64_t k, int64_t l,
               const void *A1, int64_t la1, const void *A2, int64_t la2, const void *V, int64_t lv, void *W,
               int64_t lw) {
    const int64_t kp0 = fwd ? std::min(l, k - 1) : std::min(k - l, k - 1);
    if (left) {
        const int64_t mp0 = fwd ? std::min(m - l, m - 1) : std::min(l, m - 1);
        if (fwd) {
            if (l > 0) {
                NLA_CK(P.copy(l, n, P.at(A2, la2, m - l, 0), la2, P.at(W, lw, 0, 0), lw));
                if (col) {
                    NLA_CK(P.tri_left('U', true, P.at(V, lv, mp0, 0), lv, l, n, P.at(W, lw, 0, 0), lw));
                    if (m > l) NLA_CK(P.gemm('T', 'N', l, n, m - l, 1.0, P.at(V, lv, 0, 0), lv, P.at(A2, la2, 0, 0), la2,
                                             1.0, P.at(W, lw, 0, 0), lw));
                } else {
                    NLA_CK(P.tri_left('L', false, P.at(V, lv, 0, mp0), lv, l, n, P.at(W, lw, 0, 0), lw));
                    if (m > l) NLA_CK(P.gemm('N', 'N', l, n, m - l, 1.0, P.at(V, lv, 0, 0), lv, P.at(A2, la2, 0, 0), la2,
                                             1.0, P.at(W, lw, 0, 0), lw));
                }
            }
            if (k > l) {
                if (col) NLA_CK(P.gemm('T', 'N', k - l, n, m, 1.0, P.at(V, lv, 0, kp0), lv, P.at(A2, la2, 0, 0), la2, 0.0,
                                       P.at(W, lw, kp0, 0), lw));
                else NLA_CK(P.gemm('N', 'N', k - l, n, m, 1.0, P.at(V, lv, kp0, 0), lv, P.at(A2, la2, 0, 0), la2, 0.0,
                                   P.at(W, lw, kp0, 0), lw));
            }
        } else {
            if (l > 0) {
                NLA_CK(P.copy(l, n, P.at(A2, la2, 0, 0), la2, P.at(W, lw, k - l, 0), lw));
                if (col) {
                    NLA_CK(P.tri_left('L', true, P.at(V, lv, 0, kp0), lv, l, n, P.at(W, lw, kp0, 0), lw));
                    if (m > l) NLA_CK(P.gemm('T', 'N', l, n, m - l, 1.0, P.at(V, lv, mp0, kp0), lv, P.at(A2, la2, mp0, 0),
                                             la2, 1.0, P.at(W, lw, kp0, 0), lw));
                } else {
                    NLA_CK(P.tri_left('U', false, P.at(V, lv, kp0, 0), lv, l, n, P.at(W, lw, kp0, 0), lw));
                    if (m > l) NLA_CK(P.gemm('N', 'N', l, n, m - l, 1.0, P.at(V, lv, kp0, mp0), lv, P.at(A2, la2, mp0, 0),
                                             la2, 1.0, P.at(W, lw, kp0, 0), lw));
                }
            }
            if (k > l) {
                if (col) NLA_CK(P.gemm('T', 'N', k - l, n, m, 1.0, P.at(V, lv, 0, 0), lv, P.at(A2, la2, 0, 0), la2, 0.0,
                                       P.at(W, lw, 0, 0), lw));
                else NLA_CK(P.gemm('N', 'N', k - l, n, m, 1.0, P.at(V, lv, 0, 0), lv, P.at(A2, la2, 0, 0), la2, 0.0,
                                   P.at(W, lw, 0, 0), lw));
            }
        }
        return P.add(k, n, P.at(W, lw, 0, 0), lw, P.at(A1, la1, 0, 0), la1);
    }
    const int64_t np0 = fwd ? std::min(n - l, n - 1) : std::min(l, n - 1);
    if (fwd) {
        if (l > 0) {
            NLA_CK(P.copy(m, l, P.at(A2, la2, 0, n - l), la2, P.at(W, lw, 0, 0), lw));
            if (col) {
                NLA_CK(P.tri_right('U', false, P.at(V, lv, np0, 0), lv, m, l, P.at(W, lw, 0, 0), lw));
                if (n > l) NLA_CK(P.gemm('N', 'N', m, l, n - l, 1.0, P.at(A2, la2, 0, 0), la2, P.at(V, lv, 0, 0), lv,
                                         1.0, P.at(W, lw, 0, 0), lw));
            } else {
                NLA_CK(P.tri_right('L', true, P.at(V, lv, 0, np0), lv, m, l, P.at(W, lw, 0, 0), lw));
                if (n > l) NLA_CK(P.gemm('N', 'T', m, l, n - l, 1.0, P.at(A2, la2, 0, 0), la2, P.at(V, lv, 0, 0), lv,
                                         1.0, P.at(W, lw, 0, 0), lw));
            }
        }
        if (k > l) {
            if (col) NLA_CK(P.gemm('N', 'N', m, k - l, n, 1.0, P.at(A2, la2, 0, 0), la2, P.at(V, lv, 0, kp0), lv, 0.0,
                                   P.at(W, lw, 0, kp0), lw));
            else NLA_CK(P.gemm('N', 'T', m, k - l, n, 1.0, P.at(A2, la2, 0, 0), la2, P.at(V, lv, kp0, 0), lv, 0.0,
                               P.at(W, lw, 0, kp0), lw));
        }
    } else {
        if (l > 0) {
            NLA_CK(P.copy(m, l, P.at(A2, la2, 0, 0), la2, P.at(W, lw, 0, k - l), lw));
            if (col) {
                NLA_CK(P.tri_right('L', false, P.at(V, lv, 0, kp0), lv, m, l, P.at(W, lw, 0, kp0), lw));
                if (n > l) NLA_CK(P.gemm('N', 'N', m, l, n - l, 1.0, P.at(A2, la2, 0, np0), la2, P.at(V, lv, np0, kp0),
                                         lv, 1.0, P.at(W, lw, 0, kp0), lw));
            } else {
                NLA_CK(P.tri_right('U', true, P.at(V, lv, kp0, 0), lv, m, l, P.at(W, lw, 0, kp0), lw));
                if (n > l) NLA_CK(P.gemm('N', 'T', m, l, n - l, 1.0, P.at(A2, la2, 0, np0), la2, P.at(V, lv, kp0, np0),
                                         lv, 1.0, P.at(W, lw, 0, kp0), lw));
            }
        }
        if (k > l) {
            if (col) NLA_CK(P.gemm('N', 'N', m, k - l, n, 1.0, P.at(A2, la2, 0, 0), la2, P.at(V, lv, 0, 0), lv, 0.0,
                                   P.at(W, lw, 0, 0), lw));
            else NLA_CK(P.gemm('N', 'T', m, k - l, n, 1.0, P.at(A2, la2, 0, 0), la2, P.at(V, lv, 0, 0), lv, 0.0,
                               P.at(W, lw, 0, 0), lw));
        }
    }
    return P.add(m, k, P.at(W, lw, 0, 0), lw, P.at(A1, la1, 0, 0), la1);
}

// A2 -= op(V) W (left) or W op(V) (right); W's triangle rows / columns are
// overwritten by the TRMM; src/pamm.jl:327-466.
int32_t pamm_a(const Pamm &P, bool left, bool col, bool fwd, int64_t m, int64_t n, int64_t k, int64_t l, void *A2,
               int64_t la2, const void *V, int64_t lv, void *W, int64_t lw) {
    const int64_t kp0 = fwd ? std::min(l, k - 1) : std::min(k - l, k - 1);
    if (left) {
        const int64_t mp0 = fwd ? std::min(m - l, m - 1) : std::min(l, m - 1);
        if (fwd) {
            if (m > l) {
                if (col) NLA_CK(P.gemm('N', 'N', m - l, n, k, -1.0, P.at(V, lv, 0, 0), lv, P.at(W, lw, 0, 0), lw, 1.0,
                                       P.at(A2, la2, 0, 0), la2));
                else NLA_CK(P.gemm('T', 'N', m - l, n, k, -1.0, P.at(V, lv, 0, 0), lv, P.at(W, lw, 0, 0), lw, 1.0,
                                   P.at(A2, la2, 0, 0), la2));
            }
            if (l > 0) {
                if (k > l) {
                    if (col) NLA_CK(P.gemm('N', 'N', l, n, k - l, -1.0, P.at(V, lv, mp0, kp0), lv, P.at(W, lw, kp0, 0), lw,
                                           1.0, P.at(A2, la2, mp0, 0), la2));
                    else NLA_CK(P.gemm('T', 'N', l, n, k - l, -1.0, P.at(V, lv, kp0, mp0), lv, P.at(W, lw, kp0, 0), lw,
                                       1.0, P.at(A2, la2, mp0, 0), la2));
                }
                if (col) NLA_CK(P.tri_left('U', false, P.at(V, lv, mp0, 0), lv, l, n, P.at(W, lw, 0, 0), lw));
                else NLA_CK(P.tri_left('L', true, P.at(V, lv, 0, mp0), lv, l, n, P.at(W, lw, 0, 0), lw));
                NLA_CK(P.sub(l, n, P.at(A2, la2, m - l, 0), la2, P.at(W, lw, 0, 0), lw));
            }
        } else {
            if (m > l) {
                if (col) NLA_CK(P.gemm('N', 'N', m - l, n, k, -1.0, P.at(V, lv, mp0, 0), lv, P.at(W, lw, 0, 0), lw, 1.0,
                                       P.at(A2, la2, mp0, 0), la2));
                else NLA_CK(P.gemm('T', 'N', m - l, n, k, -1.0, P.at(V, lv, 0, mp0), lv, P.at(W, lw, 0, 0), lw, 1.0,
                                   P.at(A2, la2, mp0, 0), la2));
            }
            if (l > 0) {
                if (k > l) {
                    if (col) NLA_CK(P.gemm('N', 'N', l, n, k - l, -1.0, P.at(V, lv, 0, 0), lv, P.at(W, lw, 0, 0), lw, 1.0,
                                           P.at(A2, la2, 0, 0), la2));
                    else NLA_CK(P.gemm('T', 'N', l, n, k - l, -1.0, P.at(V, lv, 0, 0), lv, P.at(W, lw, 0, 0), lw, 1.0,
                                       P.at(A2, la2, 0, 0), la2));
                }
                if (col) NLA_CK(P.tri_left('L', false, P.at(V, lv, 0, kp0), lv, l, n, P.at(W, lw, kp0, 0), lw));
                else NLA_CK(P.tri_left('U', true, P.at(V, lv, kp0, 0), lv, l, n, P.at(W, lw, kp0, 0), lw));
                NLA_CK(P.sub(l, n, P.at(A2, la2, 0, 0), la2, P.at(W, lw, k - l, 0), lw));
            }
        }
        return NLA_OK;
    }
    const int64_t np0 = fwd ? std::min(n - l, n - 1) : std::min(l, n - 1);
    if (fwd) {
        if (n > l) {
            if (col) NLA_CK(P.gemm('N', 'T', m, n - l, k, -1.0, P.at(W, lw, 0, 0), lw, P.at(V, lv, 0, 0), lv, 1.0,
                                   P.at(A2, la2, 0, 0), la2));
            else NLA_CK(P.gemm('N', 'N', m, n - l, k, -1.0, P.at(W, lw, 0, 0), lw, P.at(V, lv, 0, 0), lv, 1.0,
                               P.at(A2, la2, 0, 0), la2));
        }
        if (l > 0) {
            if (k > l) {
                if (col) NLA_CK(P.gemm('N', 'T', m, l, k - l, -1.0, P.at(W, lw, 0, kp0), lw, P.at(V, lv, np0, kp0), lv,
                                       1.0, P.at(A2, la2, 0, np0), la2));
                else NLA_CK(P.gemm('N', 'N', m, l, k - l, -1.0, P.at(W, lw, 0, kp0), lw, P.at(V, lv, kp0, np0), lv,
                                   1.0, P.at(A2, la2, 0, np0), la2));
            }
            if (col) NLA_CK(P.tri_right('U', true, P.at(V, lv, np0, 0), lv, m, l, P.at(W, lw, 0, 0), lw));
            else NLA_CK(P.tri_right('L', false, P.at(V, lv, 0, np0), lv, m, l, P.at(W, lw, 0, 0), lw));
            NLA_CK(P.sub(m, l, P.at(A2, la2, 0, n - l), la2, P.at(W, lw, 0, 0), lw));
        }
    } else {
        if (n > l) {
            if (col) NLA_CK(P.gemm('N', 'T', m, n - l, k, -1.0, P.at(W, lw, 0, 0), lw, P.at(V, lv, np0, 0), lv, 1.0,
                                   P.at(A2, la2, 0, np0), la2));
            else NLA_CK(P.gemm('N', 'N', m, n - l, k, -1.0, P.at(W, lw, 0, 0), lw, P.at(V, lv, 0, np0), lv, 1.0,
                               P.at(A2, la2, 0, np0), la2));
        }
        if (l > 0) {
            if (k > l) {
                if (col) NLA_CK(P.gemm('N', 'T', m, l, k - l, -1.0, P.at(W, lw, 0, 0), lw, P.at(V, lv, 0, 0), lv, 1.0,
                                       P.at(A2, la2, 0, 0), la2));
                else NLA_CK(P.gemm('N', 'N', m, l, k - l, -1.0, P.at(W, lw, 0, 0), lw, P.at(V, lv, 0, 0), lv, 1.0,
                                   P.at(A2, la2, 0, 0), la2));
            }
            if (col) NLA_CK(P.tri_right('L', true, P.at(V, lv, 0, kp0), lv, m, l, P.at(W, lw, 0, kp0), lw));
            else NLA_CK(P.tri_right('U', false, P.at(V, lv, kp0, 0), lv, m, l, P.at(W, lw, 0, kp0), lw));
            NLA_CK(P.sub(m, l, P.at(A2, la2, 0, 0), la2, P.at(W, lw, 0, k - l), lw));
        }
    }
    return NLA_OK;
}

char upc(char c) { return (c >= 'a' && c <= 'z') ? (char)(c - 'a' + 'A') : c; }

// Leading-dimension checks shared by pamm and parfb (sizes of the
// reference's matrices: V m x k / k x m (left) or n x k / k x n (right)).
int32_t pamm_check_v(bool left, bool col, int64_t m, int64_t n, int64_t k, const void *V, int64_t ldv, int pos) {
    const int64_t vr = col ? (left ? m : n) : k;
    if (!V || ldv < std::max<int64_t>(1, vr)) return report_bad_arg(pos, "ldv");
    return NLA_OK;
}

}  // namespace
}  // namespace nla

extern "C" int32_t nla_pamm(char op, char side, char storev, char direct, int32_t dtype, int64_t m, int64_t n,
                            int64_t k, int64_t l, const void *A1, int64_t lda1, void *A2, int64_t lda2, const void *V,
                            int64_t ldv, void *W, int64_t ldw, void *stream) {
    op = upc(op); side = upc(side); storev = upc(storev); direct = upc(direct);
    if (op != 'W' && op != 'A') return report_bad_arg(1, "op");
    if (side != 'L' && side != 'R') return report_bad_arg(2, "side");
    if (storev != 'C' && storev != 'R') return report_bad_arg(3, "storev");
    if (direct != 'F' && direct != 'B') return report_bad_arg(4, "direct");
    if (dtype != NLA_F32 && dtype != NLA_F64) return report_bad_arg(5, "dtype");
    if (m < 0) return report_bad_arg(6, "m");
    if (n < 0) return report_bad_arg(7, "n");
    if (k < 0) return report_bad_arg(8, "k");
    if (l < 0) return report_bad_arg(9, "l");
    if (m == 0 || n == 0 || k == 0) return NLA_OK;  // src/pamm.jl:83-85
    const bool left = side == 'L', col = storev == 'C', fwd = direct == 'F';
    if (l > k || l > (left ? m : n)) return report_bad_arg(9, "l (the triangle must fit V)");
    // left: W, A1 k x n, A2 m x n; right: W, A1 m x k, A2 m x n
    const int64_t wr = left ? k : m;
    if (op == 'W' && (!A1 || lda1 < std::max<int64_t>(1, wr))) return report_bad_arg(11, "lda1");
    if (!A2 || lda2 < std::max<int64_t>(1, m)) return report_bad_arg(13, "lda2");
    int32_t st = pamm_check_v(left, col, m, n, k, V, ldv, 15);
    if (st != NLA_OK) return st;
    if (!W || ldw < std::max<int64_t>(1, wr)) return report_bad_arg(17, "ldw");
    Pamm P{dtype, dtype == NLA_F64 ? 8 : 4, (hipStream_t)stream};
    if (op == 'W') return pamm_w(P, left, col, fwd, m, n, k, l, A1, lda1, A2, lda2, V, ldv, W, ldw);
    return pamm_a(P, left, col, fwd, m, n, k, l, A2, lda2, V, ldv, W, ldw);
}

extern "C" int32_t nla_parfb(char side, char trans, char direct, char storev, int32_t dtype, int64_t m1, int64_t n1,
                             int64_t m2, int64_t n2, int64_t k, int64_t l, void *A1, int64_t lda1, void *A2,
                             int64_t lda2, const void *V, int64_t ldv, const void *Tm, int64_t ldt, void *work,
                             int64_t ldwork, void *stream) {
    side = upc(side); trans = upc(trans); direct = upc(direct); storev = upc(storev);
    if (side != 'L' && side != 'R') return report_bad_arg(1, "side");
    if (trans != 'N' && trans != 'T' && trans != 'C') return report_bad_arg(2, "trans");
    if (direct != 'F' && direct != 'B') return report_bad_arg(3, "direct");
    if (storev != 'C' && storev != 'R') return report_bad_arg(4, "storev");
    if (dtype != NLA_F32 && dtype != NLA_F64) return report_bad_arg(5, "dtype");
    const bool left = side == 'L', col = storev == 'C', fwd = direct == 'F';
    if (m1 < 0) return report_bad_arg(6, "m1");
    if (n1 < 0) return report_bad_arg(7, "n1");
    if (m2 < 0 || (!left && m1 != m2)) return report_bad_arg(8, "m2");
    if (n2 < 0 || (left && n1 != n2)) return report_bad_arg(9, "n2");
    if (k < 0) return report_bad_arg(10, "k");
    if (l < 0 || l > k) return report_bad_arg(11, "l");
    // src/parfb.jl:100-102, and an empty A2 (LAPACK's tprfb returns there
    // too; the reference would apply T to a work block it never wrote)
    if (m1 == 0 || n1 == 0 || m2 == 0 || n2 == 0 || k == 0) return NLA_OK;
    // only A1's first k rows (left) or columns (right) are transformed, as
    // the reference's loops over 1:k (tsmqr! passes taller blocks)
    if (left ? (m1 < k) : (n1 < k)) return report_bad_arg(left ? 6 : 7, "A1 needs k rows (left) or columns (right)");
    if (l > (left ? m2 : n2)) return report_bad_arg(11, "l (the triangle must fit V)");
    if (!A1 || lda1 < std::max<int64_t>(1, m1)) return report_bad_arg(13, "lda1");
    if (!A2 || lda2 < std::max<int64_t>(1, m2)) return report_bad_arg(15, "lda2");
    int32_t st = pamm_check_v(left, col, m2, n2, k, V, ldv, 17);
    if (st != NLA_OK) return st;
    if (!Tm || ldt < k) return report_bad_arg(19, "ldt");
    const int64_t wr = left ? k : m2;  // work: k x n2 (left) or m2 x k (right)
    if (!work || ldwork < std::max<int64_t>(1, wr)) return report_bad_arg(21, "ldwork");
    Pamm P{dtype, dtype == NLA_F64 ? 8 : 4, (hipStream_t)stream};
    NLA_CK(pamm_w(P, left, col, fwd, m2, n2, k, l, A1, lda1, A2, lda2, V, ldv, work, ldwork));
    const char tu = fwd ? 'U' : 'L';  // T upper (forward) or lower (backward)
    const bool tt = trans != 'N';
    if (left) {
        NLA_CK(P.tri_left(tu, tt, (const char *)Tm, ldt, k, n2, (char *)work, ldwork));
        NLA_CK(P.sub(k, n2, (char *)A1, lda1, (const char *)work, ldwork));
    } else {
        NLA_CK(P.tri_right(tu, tt, (const char *)Tm, ldt, m2, k, (char *)work, ldwork));
        NLA_CK(P.sub(m2, k, (char *)A1, lda1, (const char *)work, ldwork));
    }
    return pamm_a(P, left, col, fwd, m2, n2, k, l, A2, lda2, V, ldv, work, ldwork);
}
#undef NLA_CK
