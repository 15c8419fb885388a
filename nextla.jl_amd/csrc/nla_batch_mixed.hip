// Elementwise and reduction kernels of the batched refinement loop
// (nla_batch_mixed_* in include/nla.h).
//
// Work is cut into tasks of whole columns of one system (int2 {system,
// first column}); a system's column count per task depends only on its own
// shape, every thread walks its elements in a fixed order and the per-task
// partial sums are combined in task order by one thread per system, so every
// quantity a system's refinement decides on (norms, backward error, scale)
// is the same whatever else is in the batch.
#include <hip/hip_runtime.h>

#include "nla_internal.h"

namespace nla {

namespace {

constexpr int MIX_THREADS = 256;

// Thread -> (row phase, column phase) for a column block of `rows` rows: a
// system with fewer than 256 rows takes several columns per pass.
struct ColWalk {
    int rp, cpi, r0, c0;
    __device__ explicit ColWalk(int64_t rows) {
        int lg = 0;
        while (lg < 8 && ((int64_t)1 << lg) < rows) ++lg;
        rp = 1 << lg;
        cpi = MIX_THREADS / rp;
        r0 = (int)threadIdx.x & (rp - 1);
        c0 = (int)threadIdx.x >> lg;
    }
};

__device__ inline double block_sum(double v, double *sh) {
    sh[threadIdx.x] = v;
    __syncthreads();
    for (int w = MIX_THREADS / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
        __syncthreads();
    }
    const double r = sh[0];
    __syncthreads();
    return r;
}

__device__ inline double block_max(double v, double *sh) {
    sh[threadIdx.x] = v;
    __syncthreads();
    for (int w = MIX_THREADS / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) sh[threadIdx.x] = fmax(sh[threadIdx.x], sh[threadIdx.x + w]);
        __syncthreads();
    }
    const double r = sh[0];
    __syncthreads();
    return r;
}

// Power of two sigma with m * sigma in [0.5, 1) (1 for 0, inf or NaN).
__device__ inline double pow2_scale(double m) {
    if (!(m > 0.0) || !isfinite(m)) return 1.0;
    int e;
    (void)frexp(m, &e);
    return ldexp(1.0, -e);
}

// Loop over the elements (r, c) of a B-shaped task.
#define MIX_B_TASK                                                                         \
    const int2 tk = tasks[blockIdx.x];                                                     \
    const MixSys &S = sys[tk.x];                                                           \
    const int64_t c_end = tk.y + S.bcb < S.cols ? tk.y + S.bcb : S.cols;                   \
    const ColWalk cw(S.rows);                                                              \
    for (int64_t c = tk.y + cw.c0; c < c_end; c += cw.cpi)                                 \
        for (int64_t r = cw.r0; r < S.rows; r += cw.rp)

__global__ __launch_bounds__(MIX_THREADS) void mix_prep_a_kernel(const MixSys *__restrict__ sys,
                                                                 const int2 *__restrict__ tasks,
                                                                 double *__restrict__ part) {
    __shared__ double sh[MIX_THREADS];
    const int2 tk = tasks[blockIdx.x];
    const MixSys &S = sys[tk.x];
    const int64_t n = S.n, c_end = tk.y + S.acb < n ? tk.y + S.acb : n;
    const ColWalk cw(n);
    double ss = 0.0;
    for (int64_t c = tk.y + cw.c0; c < c_end; c += cw.cpi) {
        const int64_t lo = S.lower ? c : 0, hi = S.lower ? n : c + 1;
        for (int64_t r = lo + cw.r0; r < hi; r += cw.rp) {
            const double a = S.A[r + c * S.lda];
            S.A32[r + c * S.lda32] = (float)a;
            ss += a * a;
        }
    }
    ss = block_sum(ss, sh);
    if (threadIdx.x == 0) part[blockIdx.x] = ss;
}

// S = alpha * B; partial sum of squares and maximum of |S|.
__global__ __launch_bounds__(MIX_THREADS) void mix_prep_b_kernel(const MixSys *__restrict__ sys,
                                                                 const int2 *__restrict__ tasks,
                                                                 double *__restrict__ part) {
    __shared__ double sh[MIX_THREADS];
    double ss = 0.0, mx = 0.0;
    MIX_B_TASK {
        const double v = S.alpha * S.B[r + c * S.ldb];
        S.S[r + c * S.ldw] = v;
        ss += v * v;
        mx = fmax(mx, fabs(v));
    }
    ss = block_sum(ss, sh);
    mx = block_max(mx, sh);
    if (threadIdx.x == 0) {
        part[3 * blockIdx.x] = ss;
        part[3 * blockIdx.x + 1] = mx;
        part[3 * blockIdx.x + 2] = 0.0;
    }
}

// D = fp32(sigma * (from_t ? T : S)).
__global__ __launch_bounds__(MIX_THREADS) void mix_down_kernel(const MixSys *__restrict__ sys,
                                                               const int2 *__restrict__ tasks,
                                                               const MixState *__restrict__ st, int from_t) {
    const double sg = st[tasks[blockIdx.x].x].sigma;
    MIX_B_TASK {
        const double v = from_t ? S.T[r + c * S.ldw] : S.S[r + c * S.ldw];
        S.D[r + c * S.ldw] = (float)(sg * v);
    }
}

// X (in B) = D / sigma on the first pass, else X += D / sigma while the system
// is not done; then T = X.
__global__ __launch_bounds__(MIX_THREADS) void mix_up_kernel(const MixSys *__restrict__ sys,
                                                             const int2 *__restrict__ tasks,
                                                             const MixState *__restrict__ st, int first) {
    const MixState &q = st[tasks[blockIdx.x].x];
    const double inv = 1.0 / q.sigma;  // a power of two: exact
    const bool add = !first && !q.done;
    MIX_B_TASK {
        double x;
        if (first) {
            x = (double)S.D[r + c * S.ldw] * inv;
            S.B[r + c * S.ldb] = x;
        } else if (add) {
            x = S.B[r + c * S.ldb] + (double)S.D[r + c * S.ldw] * inv;
            S.B[r + c * S.ldb] = x;
        } else {
            x = S.B[r + c * S.ldb];
        }
        S.T[r + c * S.ldw] = x;
    }
}

// T = S - T (T held op(A) X): the residual; partial sums of its squares, its
// maximum and the squares of X.
__global__ __launch_bounds__(MIX_THREADS) void mix_resid_kernel(const MixSys *__restrict__ sys,
                                                                const int2 *__restrict__ tasks,
                                                                double *__restrict__ part) {
    __shared__ double sh[MIX_THREADS];
    double rr = 0.0, mx = 0.0, xx = 0.0;
    MIX_B_TASK {
        const double v = S.S[r + c * S.ldw] - S.T[r + c * S.ldw];
        S.T[r + c * S.ldw] = v;
        rr += v * v;
        mx = fmax(mx, fabs(v));
        const double x = S.B[r + c * S.ldb];
        xx += x * x;
    }
    rr = block_sum(rr, sh);
    mx = block_max(mx, sh);
    xx = block_sum(xx, sh);
    if (threadIdx.x == 0) {
        part[3 * blockIdx.x] = rr;
        part[3 * blockIdx.x + 1] = mx;
        part[3 * blockIdx.x + 2] = xx;
    }
}

// One thread per system: combine its task partials in task order.
//   init: ||A||_F, ||alpha B||_F and the scale of the first solve;
//   step: the backward error of the current X, the done flag (set once,
//         kept) and the scale of the next correction.
__global__ void mix_final_kernel(const MixSys *__restrict__ sys, int64_t count, const double *__restrict__ apart,
                                 const double *__restrict__ bpart, MixState *__restrict__ st, int init, double tol) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
    const MixSys &S = sys[i];
    MixState &q = st[i];
    double s0 = 0.0, mx = 0.0, s2 = 0.0;
    for (int t = S.bt0; t < S.bt0 + S.bnt; ++t) {
        s0 += bpart[3 * t];
        mx = fmax(mx, bpart[3 * t + 1]);
        s2 += bpart[3 * t + 2];
    }
    if (init) {
        double aa = 0.0;
        for (int t = S.at0; t < S.at0 + S.ant; ++t) aa += apart[t];
        q.anorm = sqrt(aa);
        q.snorm = sqrt(s0);
        q.sigma = pow2_scale(mx);
        q.bwd = S.bnt == 0 ? 0.0 : __builtin_inf();
        q.done = S.bnt == 0;
        return;
    }
    q.sigma = pow2_scale(mx);
    if (q.done) return;
    const double den = q.anorm * sqrt(s2) + q.snorm;
    const double b = s0 == 0.0 ? 0.0 : sqrt(s0) / den;
    q.bwd = b;
    q.done = b <= tol;
}

}  // namespace

#define MIX_LAUNCH(kernel, ntasks, ...)                                                            \
    do {                                                                                           \
        if ((ntasks) <= 0) return hipSuccess;                                                      \
        hipLaunchKernelGGL(kernel, dim3((unsigned)(ntasks)), dim3(MIX_THREADS), 0, s, __VA_ARGS__); \
        return hipGetLastError();                                                                  \
    } while (0)

hipError_t launch_mix_prep_a(const MixSys *sys, const int2 *atasks, int nat, double *apart, hipStream_t s) {
    MIX_LAUNCH(mix_prep_a_kernel, nat, sys, atasks, apart);
}

hipError_t launch_mix_prep_b(const MixSys *sys, const int2 *btasks, int nbt, double *bpart, hipStream_t s) {
    MIX_LAUNCH(mix_prep_b_kernel, nbt, sys, btasks, bpart);
}

hipError_t launch_mix_down(const MixSys *sys, const int2 *btasks, int nbt, const MixState *st, bool from_t,
                           hipStream_t s) {
    MIX_LAUNCH(mix_down_kernel, nbt, sys, btasks, st, from_t ? 1 : 0);
}

hipError_t launch_mix_up(const MixSys *sys, const int2 *btasks, int nbt, const MixState *st, bool first,
                         hipStream_t s) {
    MIX_LAUNCH(mix_up_kernel, nbt, sys, btasks, st, first ? 1 : 0);
}

hipError_t launch_mix_resid(const MixSys *sys, const int2 *btasks, int nbt, double *bpart, hipStream_t s) {
    MIX_LAUNCH(mix_resid_kernel, nbt, sys, btasks, bpart);
}

hipError_t launch_mix_final(const MixSys *sys, int64_t count, const double *apart, const double *bpart,
                            MixState *st, bool init, double tol, hipStream_t s) {
    if (count <= 0) return hipSuccess;
    hipLaunchKernelGGL(mix_final_kernel, dim3((unsigned)((count + 255) / 256)), dim3(256), 0, s, sys, count, apart,
                       bpart, st, init ? 1 : 0, tol);
    return hipGetLastError();
}

}  // namespace nla
