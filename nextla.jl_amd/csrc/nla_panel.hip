// Fused diagonal-panel kernel: the bottom levels of the recursion in one
// launch.
//
// For a diagonal block of up to NBIG rows (a multiple of 64) of the effective
// triangular matrix M (lower or upper), every work-group owns a strip of W
// right-hand-side columns, keeps that strip of B in LDS for the whole block
// and runs the blocked algorithm the recursion would run below this size:
//   TRSM lower: for k: X_k = inv(M_kk) B_k ; B_i -= M_ik X_k (i > k)
//   TRSM upper: the same from the last 64-block backwards (i < k)
//   TRMM lower: from the last block up: B_i = tri(M_ii) B_i + sum_{k<i} M_ik B_k
//   TRMM upper: from the first block down, sum over k > i
// with inv(M_kk) from the batched diagonal inverses.  Each phase is a
// 64 x W x 64 MFMA product; no work-group depends on another (each owns its
// columns), so the panel needs one launch instead of the 2*(NBIG/64)-1
// launches of leaves and small-K updates it replaces.
//
// Replaces the bottom of unified_rec (src/rectrxm.jl:101-198) together with the
// leaves it calls (src/trsm.jl:128-150, src/trmm.jl:332-393).
#include "nla_device.h"
#include "nla_internal.h"

namespace nla {

#define NLA_LDS __attribute__((address_space(3)))

// MFMA output tile 64 x W per phase, split over 4 waves.
template <typename T>
struct PanelShape;
template <>
struct PanelShape<double> {  // 16x16 blocks: wave w owns row block w, both column blocks
    static constexpr int W = 32;
    static constexpr int RBW = 1, CBW = 2;  // blocks per wave (rows, cols)
    __device__ static inline int rb0(int wave) { return wave; }
    __device__ static inline int cb0(int) { return 0; }
};
template <>
struct PanelShape<float> {  // 32x32 blocks: wave w owns block (w%2, w/2)
    static constexpr int W = 64;
    static constexpr int RBW = 1, CBW = 1;
    __device__ static inline int rb0(int wave) { return wave & 1; }
    __device__ static inline int cb0(int wave) { return wave >> 1; }
};

template <typename T, int NBIG, bool TRMM>
__global__ __launch_bounds__(256) void panel_kernel(int nbig, int64_t m, const T *__restrict__ Mp, int64_t mrs,
                                                    int64_t mcs, int lower, T *__restrict__ Bp, int64_t brs,
                                                    int64_t bcs, const T *__restrict__ Wp, int64_t wrs,
                                                    int64_t wcs, T scale) {
    typedef Mfma<T> MF;
    typedef PanelShape<T> PS;
    constexpr int W = PS::W;
    constexpr int LDB = NBIG + (sizeof(T) == 8 ? 2 : 4);  // LDS row stride of one strip column
    constexpr int NKS = 64 / MF::KD;                      // k-steps per 64-deep phase
    __shared__ __attribute__((aligned(16))) T bs_raw[W * LDB];
    NLA_LDS T *bs = (NLA_LDS T *)bs_raw;

    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    const int64_t c0 = (int64_t)blockIdx.x * W;
    const int nsb = (nbig + 63) / 64;

    // ---- load the strip (coalesced along B's contiguous direction) ----
    {
        const bool rowc = (brs == 1 || brs == -1);
#pragma unroll 8
        for (int e = t; e < W * NBIG; e += 256) {
            int r = rowc ? e % NBIG : e / W;
            int c = rowc ? e / NBIG : e % W;
            int64_t gc = c0 + c;
            bool ok = r < nbig && gc < m;
            T v = *(ok ? Bp + r * brs + gc * bcs : Bp);
            bs[c * LDB + r] = ok ? (TRMM ? v : v * scale) : T(0);
        }
    }
    __syncthreads();

    const int lo = lane % MF::BD, lk = lane / MF::BD;
    const int rb = PS::rb0(wave), cbase = PS::cb0(wave);
    typename MF::acc_t acc[PS::CBW];

    // acc += op * Bs[rows of block kb]; op(i, k) = A[i*ars + k*acs] for
    // i < ni, k < nk (zero outside, and outside the triangle when tri != 0:
    // 1 keep k <= i, 2 keep k >= i).
    auto phase = [&](const T *A, int64_t ars, int64_t acs, int ni, int nk, int kb, int tri) {
        T af[NKS];
        const int i = rb * MF::BD + lo;
#pragma unroll
        for (int s = 0; s < NKS; ++s) {
            const int k = s * MF::KD + lk;
            bool ok = i < ni && k < nk;
            if (tri == 1) ok = ok && k <= i;
            if (tri == 2) ok = ok && k >= i;
            T v = *(ok ? A + i * ars + k * acs : A);
            af[s] = ok ? v : T(0);
        }
        const NLA_LDS T *src = bs + kb * 64;
#pragma unroll
        for (int s = 0; s < NKS; ++s) {
#pragma unroll
            for (int cb = 0; cb < PS::CBW; ++cb) {
                const int col = (cbase + cb) * MF::BD + lo;
                T bv = src[col * LDB + s * MF::KD + lk];
                acc[cb] = MF::mma(af[s], bv, acc[cb]);
            }
        }
    };
    auto zero_acc = [&]() {
#pragma unroll
        for (int cb = 0; cb < PS::CBW; ++cb) acc[cb] = MF::zero();
    };
    // Bs[block ib rows] = alpha * acc + beta * Bs[...] for this wave's tile
    auto write_block = [&](int ib, T alpha, T beta) {
        NLA_LDS T *dst = bs + ib * 64;
#pragma unroll
        for (int cb = 0; cb < PS::CBW; ++cb)
#pragma unroll
            for (int r = 0; r < MF::NR; ++r) {
                const int row = rb * MF::BD + MF::di(lane, r);
                const int col = (cbase + cb) * MF::BD + MF::dj(lane, r);
                NLA_LDS T *p = dst + col * LDB + row;
                T v = alpha * acc[cb][r];
                *p = (beta == T(0)) ? v : v + beta * *p;
            }
    };
    auto bsize = [&](int b) { return (b == nsb - 1) ? nbig - 64 * b : 64; };
    const T *Wb = Wp;

    if (!TRMM) {
        for (int step = 0; step < nsb; ++step) {
            const int k = lower ? step : nsb - 1 - step;
            // X_k = inv(M_kk) B_k
            zero_acc();
            phase(Wb + (int64_t)k * 64 * 64, wrs, wcs, bsize(k), bsize(k), k, 0);
            __syncthreads();
            write_block(k, T(1), T(0));
            __syncthreads();
            // B_i -= M_ik X_k for the blocks still to be solved
            const int i0 = lower ? k + 1 : 0, i1 = lower ? nsb : k;
            for (int i = i0; i < i1; ++i) {
                zero_acc();
                phase(Mp + (int64_t)(64 * i) * mrs + (int64_t)(64 * k) * mcs, mrs, mcs, bsize(i), bsize(k), k, 0);
                write_block(i, T(-1), T(1));
            }
            __syncthreads();
        }
    } else {
        for (int step = 0; step < nsb; ++step) {
            const int i = lower ? nsb - 1 - step : step;
            zero_acc();
            phase(Mp + (int64_t)(64 * i) * (mrs + mcs), mrs, mcs, bsize(i), bsize(i), i, lower ? 1 : 2);
            const int k0 = lower ? 0 : i + 1, k1 = lower ? i : nsb;
            for (int k = k0; k < k1; ++k)
                phase(Mp + (int64_t)(64 * i) * mrs + (int64_t)(64 * k) * mcs, mrs, mcs, bsize(i), bsize(k), k, 0);
            __syncthreads();  // every wave has read B_i
            write_block(i, T(1), T(0));
            __syncthreads();
        }
    }

    // ---- write the strip back ----
    {
        const bool rowc = (brs == 1 || brs == -1);
#pragma unroll 8
        for (int e = t; e < W * NBIG; e += 256) {
            int r = rowc ? e % NBIG : e / W;
            int c = rowc ? e / NBIG : e % W;
            int64_t gc = c0 + c;
            if (r < nbig && gc < m) Bp[r * brs + gc * bcs] = TRMM ? bs[c * LDB + r] * scale : bs[c * LDB + r];
        }
    }
}

template <typename T>
hipError_t launch_panel(const PanelArgs<T> &a, hipStream_t s) {
    if (a.m <= 0 || a.nbig <= 0) return hipSuccess;
    if (a.nbig > PANEL_MAX) return hipErrorInvalidValue;
    constexpr int W = PanelShape<T>::W;
    const int64_t groups = (a.m + W - 1) / W;
    if (a.trmm)
        hipLaunchKernelGGL((panel_kernel<T, PANEL_MAX, true>), dim3((unsigned)groups), dim3(256), 0, s, a.nbig, a.m,
                           a.M, a.mrs, a.mcs, a.lower ? 1 : 0, a.B, a.brs, a.bcs, a.W, a.wrs, a.wcs, a.scale);
    else
        hipLaunchKernelGGL((panel_kernel<T, PANEL_MAX, false>), dim3((unsigned)groups), dim3(256), 0, s, a.nbig,
                           a.m, a.M, a.mrs, a.mcs, a.lower ? 1 : 0, a.B, a.brs, a.bcs, a.W, a.wrs, a.wcs, a.scale);
    return hipGetLastError();
}

template hipError_t launch_panel<float>(const PanelArgs<float> &, hipStream_t);
template hipError_t launch_panel<double>(const PanelArgs<double> &, hipStream_t);

}  // namespace nla
