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
// The RHS tile is staged through LDS so that HBM is always walked along its
// contiguous direction (down columns for the left side, along rows for the
// right side) whatever the lane-to-column mapping.
#include "nla_device.h"
#include "nla_internal.h"

namespace nla {

constexpr int LEAF_COLS = 64;  // RHS columns per work-group (one per lane)

// ICONTIG: Bc is contiguous along i (left side) rather than along c (right
// side); the RHS tile is staged through LDS either way, walking the
// contiguous direction so every wave-wide load and store is coalesced.
template <typename T, int NB, bool ICONTIG, bool TRMM>
__global__ __launch_bounds__(LEAF_COLS) void leaf_kernel(int nb, int64_t m, const T *__restrict__ A, int64_t ars,
                                                         int64_t acs, T *__restrict__ B, int64_t brs, int64_t bcs,
                                                         T scale) {
    // TRSM reads L by columns (column j after x[j] is final), TRMM by rows
    // (row i is a dot product over j <= i); store L in the matching order.
    __shared__ __attribute__((aligned(16))) T sl[NB * NB];
    __shared__ T srinv[NB];
    __shared__ T sb[LEAF_COLS * (NB + 1)];

    const int t = threadIdx.x;
    const int64_t c0 = (int64_t)blockIdx.x * LEAF_COLS;

    // Stage L, padded to NB with an identity block.  Loads are issued in
    // batches of SB before any LDS write so HBM/L2 latency overlaps.
    constexpr int SB = 16;
    const bool arow = (ars == 1 || ars == -1);  // walk the contiguous direction
#pragma unroll 1
    for (int e0 = 0; e0 < NB * NB; e0 += SB * LEAF_COLS) {
        T v[SB];
#pragma unroll
        for (int u = 0; u < SB; ++u) {
            int e = e0 + u * LEAF_COLS + t;
            int i = arow ? e % NB : e / NB;
            int j = arow ? e / NB : e % NB;
            bool in = e < NB * NB;
            if (in && i < nb && j < nb)
                v[u] = (j <= i) ? A[i * ars + j * acs] : T(0);
            else
                v[u] = (i == j) ? T(1) : T(0);
        }
#pragma unroll
        for (int u = 0; u < SB; ++u) {
            int e = e0 + u * LEAF_COLS + t;
            int i = arow ? e % NB : e / NB;
            int j = arow ? e / NB : e % NB;
            if (e < NB * NB) {
                if (TRMM)
                    sl[i * NB + j] = v[u];
                else
                    sl[j * NB + i] = v[u];
            }
        }
    }
    if (!TRMM) {
        for (int i = t; i < NB; i += LEAF_COLS) srinv[i] = (i < nb) ? T(1) / A[i * ars + i * acs] : T(1);
    }

    // Load this lane's column of Bc.
    T x[NB];
#pragma unroll 1
    for (int e0 = 0; e0 < NB * LEAF_COLS; e0 += SB * LEAF_COLS) {
        T v[SB];
#pragma unroll
        for (int u = 0; u < SB; ++u) {
            int e = e0 + u * LEAF_COLS + t;
            int i = ICONTIG ? e % NB : e / LEAF_COLS;
            int c = ICONTIG ? e / NB : e % LEAF_COLS;
            int64_t gc = c0 + c;
            v[u] = (i < nb && gc < m) ? B[i * brs + gc * bcs] : T(0);
        }
#pragma unroll
        for (int u = 0; u < SB; ++u) {
            int e = e0 + u * LEAF_COLS + t;
            int i = ICONTIG ? e % NB : e / LEAF_COLS;
            int c = ICONTIG ? e / NB : e % LEAF_COLS;
            sb[c * (NB + 1) + i] = v[u];
        }
    }
    __syncthreads();
#pragma unroll
    for (int i = 0; i < NB; ++i) x[i] = sb[t * (NB + 1) + i];

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

#pragma unroll
    for (int i = 0; i < NB; ++i) sb[t * (NB + 1) + i] = x[i];
    __syncthreads();
#pragma unroll 16
    for (int e = t; e < NB * LEAF_COLS; e += LEAF_COLS) {
        int i = ICONTIG ? e % NB : e / LEAF_COLS;
        int c = ICONTIG ? e / NB : e % LEAF_COLS;
        int64_t gc = c0 + c;
        if (i < nb && gc < m) B[i * brs + gc * bcs] = sb[c * (NB + 1) + i];
    }
}

// Inverse of one 64 x 64 triangular block with 256 threads, in three stages
// (the blocked scheme of LAPACK's trtri), in place in LDS:
//   1. the four 16 x 16 diagonal sub-blocks by substitution on the columns of
//      the identity (64 lanes, one column each, 16-step chains), written over
//      the sub-blocks once every lane has read them;
//   2. 32 x 32 inverses: inv([A 0; C B]) = [inv A, 0; -inv B * C * inv A, inv B],
//      C * inv A through the scratch st, the result written over C;
//   3. the same step from 32 to 64.
// sl: the canonical lower block (column-major, 64 x 64), which becomes its
// inverse; st: 32 x 32 scratch.  The parts above the diagonal stay zero, so
// the products run over full, unrolled k ranges and the extra terms add
// exact zeros.  Upper blocks are inverted as J U J.  40 KB of LDS in FP64
// (four work-groups per CU; with a separate inverse it was 72 KB, two).
template <typename T>
__device__ __forceinline__ void trtri64_block(int64_t n, int64_t blk, const T *__restrict__ A, int64_t rs, int64_t cs, bool lower,
                              bool unit, T *__restrict__ W, T *sl, T *st, int *__restrict__ F) {
    constexpr int NB = 64, NT = 256;
    const int t = threadIdx.x;
    const int64_t off = blk * NB;
    const int nb = (int)(n - off < NB ? n - off : NB);
    const T *base = lower ? A + off * (rs + cs) : A + (off + nb - 1) * (rs + cs);
    const int64_t ars = lower ? rs : -rs, acs = lower ? cs : -cs;
    const bool arow = (ars == 1 || ars == -1);
    // stage L (identity padding past nb).  All 16 loads of a thread are
    // issued before the first LDS store.
    constexpr int PER = NB * NB / NT;
    T v[PER];
#pragma unroll
    for (int u = 0; u < PER; ++u) {
        const int e = t + u * NT;
        const int i = arow ? e % NB : e / NB;
        const int j = arow ? e / NB : e % NB;
        // a unit diagonal is taken as 1 and not read
        const bool ok = i < nb && j < nb && j <= i && !(unit && i == j);
        v[u] = *(ok ? base + i * ars + j * acs : base);
        v[u] = ok ? v[u] : (i == j ? T(1) : T(0));
    }
#pragma unroll
    for (int u = 0; u < PER; ++u) {
        const int e = t + u * NT;
        const int i = arow ? e % NB : e / NB;
        const int j = arow ? e / NB : e % NB;
        sl[j * NB + i] = v[u];
    }
    __syncthreads();
    // ||L||_1 (largest column sum) while L is still there: wave 0, lane = column
    // (lane t walks its column from row t on, so that the 64 lanes of a read
    // fall in different banks rather than all in one)
    T nl = T(0);
    if (F != nullptr && t < NB) {
#pragma unroll 8
        for (int i = 0; i < NB; ++i) {
            const T a = sl[t * NB + ((i + t) & (NB - 1))];
            nl += a < T(0) ? -a : a;
        }
        for (int o = 32; o > 0; o >>= 1) {
            const T other = __shfl_xor(nl, o);
            nl = other > nl ? other : nl;
        }
    }
    // 1. sub-block b, column c: y = inv(D_b) e_c
    T x[16];
    const int b1 = t >> 4, c1 = t & 15, o1 = 16 * b1;
    if (t < 64) {
#pragma unroll
        for (int i = 0; i < 16; ++i) x[i] = (i == c1) ? T(1) : T(0);
#pragma unroll
        for (int j = 0; j < 16; ++j) {
            x[j] = x[j] / sl[(o1 + j) * NB + o1 + j];
#pragma unroll
            for (int r = j + 1; r < 16; ++r) x[r] -= sl[(o1 + j) * NB + o1 + r] * x[j];
        }
    }
    __syncthreads();  // every lane has read its sub-block
    if (t < 64) {
#pragma unroll
        for (int i = 0; i < 16; ++i) sl[(o1 + c1) * NB + o1 + i] = x[i];
    }
    __syncthreads();
    // 2. two 32-blocks p: st_p = C_p * inv(D_2p) (16 x 16), then
    //    C_p <- -inv(D_2p+1) * st_p.  Thread t: block p = t>>7, row i,
    //    columns j and j+8.
    {
        const int p = t >> 7, q = t & 127, i = q & 15, j = q >> 4, o = 32 * p;
        T a0 = T(0), a1 = T(0);
#pragma unroll
        for (int k = 0; k < 16; ++k) {
            const T c = sl[(o + k) * NB + o + 16 + i];
            a0 += c * sl[(o + j) * NB + o + k];
            a1 += c * sl[(o + j + 8) * NB + o + k];
        }
        st[p * 256 + j * 16 + i] = a0;
        st[p * 256 + (j + 8) * 16 + i] = a1;
        __syncthreads();  // C_p read by every lane; st_p complete
        a0 = T(0);
        a1 = T(0);
#pragma unroll
        for (int k = 0; k < 16; ++k) {
            const T c = sl[(o + 16 + k) * NB + o + 16 + i];
            a0 += c * st[p * 256 + j * 16 + k];
            a1 += c * st[p * 256 + (j + 8) * 16 + k];
        }
        sl[(o + j) * NB + o + 16 + i] = -a0;
        sl[(o + j + 8) * NB + o + 16 + i] = -a1;
    }
    __syncthreads();
    // 3. st = Q * inv(P) (32 x 32, inv(P) lower), Q <- -inv(R) * st.
    //    Thread t: row i = t & 31, columns j0 + 8u.
    {
        const int i = t & 31, j0 = t >> 5;
        T a[4] = {T(0), T(0), T(0), T(0)};
#pragma unroll
        for (int k = 0; k < 32; ++k) {
            const T c = sl[k * NB + 32 + i];
#pragma unroll
            for (int u = 0; u < 4; ++u) a[u] += c * sl[(j0 + 8 * u) * NB + k];
        }
        // (step 2's reads of st ended at the barrier before this step)
#pragma unroll
        for (int u = 0; u < 4; ++u) st[(j0 + 8 * u) * 32 + i] = a[u];
        __syncthreads();  // st complete; Q read by every lane
#pragma unroll
        for (int u = 0; u < 4; ++u) a[u] = T(0);
#pragma unroll
        for (int k = 0; k < 32; ++k) {
            const T c = sl[(32 + k) * NB + 32 + i];
#pragma unroll
            for (int u = 0; u < 4; ++u) a[u] += c * st[(j0 + 8 * u) * 32 + k];
        }
#pragma unroll
        for (int u = 0; u < 4; ++u) sl[(j0 + 8 * u) * NB + 32 + i] = -a[u];
    }
    __syncthreads();
    // Condition of the block, kappa_1 = ||L||_1 ||inv(L)||_1 (both exact, the
    // inverse being in LDS): an explicitly inverted block applied as a
    // product has backward error growing with kappa * u, so a block above
    // LEAF_REFINE_KAPPA is flagged and the panels follow its product with
    // one step of refinement (X += inv(L) (B - L X)).
    if (F != nullptr && t < NB) {
        T ni = T(0);
#pragma unroll 8
        for (int i = 0; i < NB; ++i) {
            const T a = sl[t * NB + ((i + t) & (NB - 1))];
            ni += a < T(0) ? -a : a;
        }
        for (int o = 32; o > 0; o >>= 1) {
            const T other = __shfl_xor(ni, o);
            ni = other > ni ? other : ni;
        }
        if (t == 0) {
            const T kappa = nl * ni;
            F[blk] = (kappa > T(LEAF_REFINE_KAPPA) || !(kappa == kappa)) ? 1 : 0;
        }
    }
    // write: column-major 64 x 64; an upper block's inverse is J inv(L) J
    T *w = W + blk * NB * NB;
#pragma unroll 4
    for (int e = t; e < NB * NB; e += NT) {
        int i = e % NB, j = e / NB;  // element (i, j) of the stored inverse
        T val;
        if (lower || i >= nb || j >= nb)
            val = (i < nb && j < nb) || lower ? sl[j * NB + i] : (i == j ? T(1) : T(0));
        else
            val = sl[(nb - 1 - j) * NB + (nb - 1 - i)];
        w[e] = val;
    }
}

// Inverse of one NB x NB diagonal block per work-group, by forward
// substitution on the columns of the identity (one column per lane, the
// block in LDS).  Each column of the inverse is a backward-stable triangular
// solve; the inverses turn the diagonal-block leaves of the recursion into
// MFMA products.
template <typename T, int NB>
__device__ __forceinline__ void trtri_block(int64_t n, int64_t blk, const T *__restrict__ A, int64_t rs, int64_t cs, bool lower,
                            bool unit, T *__restrict__ W, T *sl, T *srinv) {
    const int t = threadIdx.x;
    const int64_t off = blk * NB;
    const int nb = (int)(n - off < NB ? n - off : NB);
    // canonical lower L: the block itself, or J*U*J for an upper block
    const T *base = lower ? A + off * (rs + cs) : A + (off + nb - 1) * (rs + cs);
    const int64_t ars = lower ? rs : -rs, acs = lower ? cs : -cs;
    const bool arow = (ars == 1 || ars == -1);
    constexpr int SB = 16;
#pragma unroll 1
    for (int e0 = 0; e0 < NB * NB; e0 += SB * LEAF_COLS) {
        T v[SB];
#pragma unroll
        for (int u = 0; u < SB; ++u) {
            int e = e0 + u * LEAF_COLS + t;
            int i = arow ? e % NB : e / NB;
            int j = arow ? e / NB : e % NB;
            if (e < NB * NB && i < nb && j < nb && !(unit && i == j))
                v[u] = (j <= i) ? base[i * ars + j * acs] : T(0);
            else
                v[u] = (i == j) ? T(1) : T(0);
        }
#pragma unroll
        for (int u = 0; u < SB; ++u) {
            int e = e0 + u * LEAF_COLS + t;
            int i = arow ? e % NB : e / NB;
            int j = arow ? e / NB : e % NB;
            if (e < NB * NB) sl[j * NB + i] = v[u];
        }
    }
    for (int i = t; i < NB; i += LEAF_COLS) srinv[i] = (i < nb && !unit) ? T(1) / base[i * ars + i * acs] : T(1);
    __syncthreads();
    if (t >= NB) return;
    T x[NB];
#pragma unroll
    for (int i = 0; i < NB; ++i) x[i] = (i == t) ? T(1) : T(0);
#pragma unroll
    for (int j = 0; j < NB; ++j) {
        x[j] = x[j] * srinv[j];
#pragma unroll
        for (int r = j + 1; r < NB; ++r) x[r] = x[r] - sl[j * NB + r] * x[j];
    }
    // column t of inv(L); for an upper block inv(U) = J inv(L) J
    T *w = W + blk * NB * NB;
    if (lower || t >= nb) {
#pragma unroll
        for (int i = 0; i < NB; ++i) w[i + t * NB] = x[i];
    } else {
#pragma unroll
        for (int i = 0; i < NB; ++i) {
            if (i < nb) w[(nb - 1 - i) + (nb - 1 - t) * NB] = x[i];
            else w[i + (nb - 1 - t) * NB] = T(0);
        }
    }
}

// Inverse of one NB x NB diagonal block per work-group, by forward
// substitution on the columns of the identity (one column per lane, the
// block in LDS).  Each column of the inverse is a backward-stable triangular
// solve; the inverses turn the diagonal-block leaves of the recursion into
// MFMA products.
template <typename T, int NB>
__global__ __launch_bounds__(LEAF_COLS) void trtri_kernel(int64_t n, const T *__restrict__ A, int64_t rs,
                                                          int64_t cs, int lower, int unit, T *__restrict__ W) {
    __shared__ __attribute__((aligned(16))) T sl[NB * NB];
    __shared__ T srinv[NB];
    trtri_block<T, NB>(n, blockIdx.x, A, rs, cs, lower != 0, unit != 0, W, sl, srinv);
}

// The 64-block inverse used by the recursion and the batched solver.
template <typename T>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4))) void trtri64_kernel(int64_t n, const T *__restrict__ A, int64_t rs, int64_t cs,
                                                      int lower, int unit, T *__restrict__ W, int *__restrict__ F) {
    __shared__ __attribute__((aligned(16))) T sl[64 * 64];
    __shared__ __attribute__((aligned(16))) T st[32 * 32];
    trtri64_block<T>(n, blockIdx.x, A, rs, cs, lower != 0, unit != 0, W, sl, st, F);
}

// Batched form: one work-group per (system, block) pair.
template <typename T>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4))) void batch_trtri_kernel(const BatchSys<T> *__restrict__ sys,
                                                         const int2 *__restrict__ blk) {
    __shared__ __attribute__((aligned(16))) T sl[64 * 64];
    __shared__ __attribute__((aligned(16))) T st[32 * 32];
    const int2 b = blk[blockIdx.x];
    const BatchSys<T> &S = sys[b.x];
    // the diagonal blocks of the effective matrix in its natural orientation:
    // for a right-side system M = op(A)^T, whose blocks are the transposed
    // blocks of op(A); the inverses are stored for op(A) (read transposed).
    const bool right = (S.wrs != 1);
    const int64_t rs = right ? S.mcs : S.mrs, cs = right ? S.mrs : S.mcs;
    const bool lower = right ? (S.lower == 0) : (S.lower != 0);
    trtri64_block<T>(S.n, b.y, S.M, rs, cs, lower, false, (T *)S.W, sl, st, (int *)S.F);
}

template <typename T>
hipError_t launch_batch_trtri(const BatchSys<T> *sys, const int2 *blk, int nblk, hipStream_t s) {
    if (nblk <= 0) return hipSuccess;
    hipLaunchKernelGGL(batch_trtri_kernel<T>, dim3(nblk), dim3(256), 0, s, sys, blk);
    return hipGetLastError();
}

template hipError_t launch_batch_trtri<float>(const BatchSys<float> *, const int2 *, int, hipStream_t);
template hipError_t launch_batch_trtri<double>(const BatchSys<double> *, const int2 *, int, hipStream_t);

template <typename T>
hipError_t launch_trtri_blocks(int64_t n, int nb, const T *A, int64_t rs, int64_t cs, bool lower, T *W,
                               hipStream_t s, bool unit, int *flags) {
    if (n <= 0) return hipSuccess;
    int64_t blocks = (n + nb - 1) / nb;
    const int lo = lower ? 1 : 0, un = unit ? 1 : 0;
    if (nb == 64)
        hipLaunchKernelGGL((trtri64_kernel<T>), dim3((unsigned)blocks), dim3(256), 0, s, n, A, rs, cs, lo, un, W,
                           flags);
    else if (nb == 32)
        hipLaunchKernelGGL((trtri_kernel<T, 32>), dim3((unsigned)blocks), dim3(LEAF_COLS), 0, s, n, A, rs, cs, lo, un,
                           W);
    else if (nb == 16)
        hipLaunchKernelGGL((trtri_kernel<T, 16>), dim3((unsigned)blocks), dim3(LEAF_COLS), 0, s, n, A, rs, cs, lo, un,
                           W);
    else
        return hipErrorInvalidValue;
    return hipGetLastError();
}

template hipError_t launch_trtri_blocks<float>(int64_t, int, const float *, int64_t, int64_t, bool, float *,
                                               hipStream_t, bool, int *);
template hipError_t launch_trtri_blocks<double>(int64_t, int, const double *, int64_t, int64_t, bool, double *,
                                                hipStream_t, bool, int *);

template <typename T, int NB, bool ICONTIG>
static hipError_t launch_leaf_nb(const LeafArgs<T> &a, hipStream_t s) {
    int64_t groups = (a.m + LEAF_COLS - 1) / LEAF_COLS;
    if (a.trmm)
        hipLaunchKernelGGL((leaf_kernel<T, NB, ICONTIG, true>), dim3((unsigned)groups), dim3(LEAF_COLS), 0, s, a.nb,
                           a.m, a.A, a.ars, a.acs, a.B, a.brs, a.bcs, a.scale);
    else
        hipLaunchKernelGGL((leaf_kernel<T, NB, ICONTIG, false>), dim3((unsigned)groups), dim3(LEAF_COLS), 0, s, a.nb,
                           a.m, a.A, a.ars, a.acs, a.B, a.brs, a.bcs, a.scale);
    return hipGetLastError();
}

template <typename T, bool ICONTIG>
static hipError_t launch_leaf_stage(const LeafArgs<T> &a, hipStream_t s) {
    if (a.nb <= 16) return launch_leaf_nb<T, 16, ICONTIG>(a, s);
    if (a.nb <= 32) return launch_leaf_nb<T, 32, ICONTIG>(a, s);
    return launch_leaf_nb<T, 64, ICONTIG>(a, s);
}

template <typename T>
hipError_t launch_leaf(const LeafArgs<T> &a, hipStream_t s) {
    if (a.m <= 0 || a.nb <= 0) return hipSuccess;
    if (a.nb > LEAF_MAX_NB) return hipErrorInvalidValue;
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
