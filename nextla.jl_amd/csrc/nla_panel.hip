// Fused diagonal-panel solve: the bottom levels of the recursion in one
// launch, and the persistent batched solver built on the same routine.
//
// For a diagonal block of up to NBIG rows of the effective triangular matrix
// M (lower or upper), a work-group owns a strip of W right-hand-side columns,
// keeps that strip of B in LDS for the whole block and runs the blocked
// algorithm the recursion would run below this size:
//   TRSM lower: for k: X_k = inv(M_kk) B_k ; B_i -= M_ik X_k (i > k)
//   TRSM upper: the same from the last 64-block backwards (i < k)
//   TRMM lower: from the last block up: B_i = tri(M_ii) B_i + sum_{k<i} M_ik B_k
//   TRMM upper: from the first block down, sum over k > i
// with inv(M_kk) from the batched diagonal inverses.  Each phase is a
// 64 x W x 64 MFMA product whose M-operand fragments are fetched one phase
// ahead; no work-group depends on another (each owns its columns).
//
// Replaces the bottom of unified_rec (src/rectrxm.jl:101-198) together with the
// leaves it calls (src/trsm.jl:128-150, src/trmm.jl:332-393); the batched form
// has no counterpart in the reference (one call per system there).
#include "nla_device.h"
#include "nla_internal.h"

namespace nla {

#define NLA_LDS __attribute__((address_space(3)))

// The source of the M-operand elements outside a block (or its triangle):
// reading zeros there leaves nothing to mask when the fragment is used.
static __device__ __align__(16) const double panel_zero[2] = {0.0, 0.0};

// 16x16x4 f32 MFMA (2048 flops, 32 cycles): D[i][j] in lane j+16*(i/4),
// register i%4 (verified on gfx950 by a lane-layout probe).
struct Mfma16f {
    static constexpr int BD = 16, KD = 4, NR = 4;
    typedef f4_t acc_t;
    __device__ static inline acc_t zero() { return acc_t{0.f, 0.f, 0.f, 0.f}; }
    __device__ static inline acc_t mma(float a, float b, acc_t c) {
        return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
    }
    __device__ static inline int di(int l, int r) { return 4 * (l >> 4) + r; }
    __device__ static inline int dj(int l, int) { return l & 15; }
};

// Output tile of one phase: 64 rows x W columns over 4 waves.  With 16x16
// MFMA blocks and W = 16, wave w owns rows 16w..16w+15 of the 64-row block;
// narrow strips give many work-groups (m/16) for the column parallelism.
template <typename T>
struct Mfma16;
template <>
struct Mfma16<double> : Mfma<double> {};
template <>
struct Mfma16<float> : Mfma16f {};

template <typename T, int W_>
struct PanelShape {
    typedef Mfma16<T> MF;
    static constexpr int W = W_, CBW = W_ / 16;
    static_assert(W_ % 16 == 0, "strip width");
    __device__ static inline int rb0(int wave) { return wave; }
    __device__ static inline int cb0(int) { return 0; }
};

// A phase of the blocked algorithm for nsb 64-blocks.  kind: 0 = diagonal
// (TRSM: inv(M_kk) block; TRMM: masked M_ii), 1 = off-diagonal M_ik.
struct Phase {
    int i, k, kind;
    bool first, last;  // first / last phase of an output block group
};

// Phase (step, q) of the blocked algorithm, walked in order with phase_next
// (step s holds s + 1 phases for TRSM, nsb - s for TRMM).
__device__ inline Phase phase_at(int step, int q, int nsb, bool lower, bool trmm) {
    Phase ph;
    if (!trmm) {
        // Left-looking: output block i (top-down for lower, bottom-up for
        // upper) gathers its sources k (already solved) in registers, then
        // its diagonal phase subtracts them from B_i and applies inv(M_ii).
        // Step s has s source phases and one diagonal phase.
        const int i = lower ? step : nsb - 1 - step;
        ph.i = i;
        if (q < step) {
            ph.k = lower ? q : nsb - 1 - q;
            ph.kind = 1;
            ph.first = (q == 0);
            ph.last = false;
        } else {
            ph.k = i;
            ph.kind = 0;
            ph.first = (step == 0);
            ph.last = true;
        }
    } else {
        // output block i (bottom-up for lower): diagonal, then its sources
        const int i = lower ? nsb - 1 - step : step;
        ph.i = i;
        if (q == 0) {
            ph.k = i;
            ph.kind = 0;
        } else {
            ph.k = lower ? q - 1 : i + q;
            ph.kind = 1;
        }
        ph.first = (q == 0);
        ph.last = (q == nsb - step - 1);
    }
    return ph;
}

__device__ inline void phase_next(int &step, int &q, int nsb, bool trmm) {
    ++q;
    if (trmm ? (q >= nsb - step) : (q > step)) {
        ++step;
        q = 0;
    }
}

__device__ inline int phase_count(int nsb, bool trmm) {
    (void)trmm;
    return nsb * (nsb + 1) / 2;
}

// Solve / multiply one strip of W columns (starting at c0) of one diagonal
// panel; bs is the LDS strip (W x LDB).  NW waves: with 4, wave w computes
// rows 16w..16w+15 of every 64-row phase over its whole 64-deep k range;
// with 8, waves w and w+4 share those rows and take one half of the k range
// each (twice the waves to hide latency with the same work), their partial
// sums meeting in the 64 x W LDS scratch sc when a block is written.
template <typename T, int NBIG, int WS, bool TRMM, int NW = 4>
// Always inlined: with the single-panel kernel and the batched kernel both
// calling one instantiation, the compiler otherwise emits a real call (and
// the panel ran 50 % slower).
__device__ __forceinline__ void panel_strip(int nbig, int64_t m, int64_t c0, const T *__restrict__ Mp, int64_t mrs,
                                            int64_t mcs, bool lower, bool unit, T *__restrict__ Bp, int64_t brs,
                                            int64_t bcs,
                                            const T *__restrict__ Wp, int64_t wrs, int64_t wcs, T scale,
                                            const int *__restrict__ Fp, NLA_LDS T *bs, NLA_LDS T *sc = nullptr,
                                            unsigned *xmax = nullptr) {
    typedef PanelShape<T, WS> PS;
    typedef typename PS::MF MF;
    constexpr int W = PS::W;
    constexpr int LDB = NBIG + (sizeof(T) == 8 ? 2 : 4);
    constexpr int NKS = 64 / MF::KD;  // k-steps per 64-deep phase
    static_assert(NW == 4 || NW == 8, "4 or 8 waves");
    constexpr int NT = NW * 64, KSPLIT = NW / 4, NKW = NKS / KSPLIT;  // k-steps per wave

    const int t = threadIdx.x, lane = t & 63, wave = (t >> 6) & 3, kh = (t >> 6) >> 2;
    const int nsb = (nbig + 63) / 64;

    // ---- load the strip (coalesced along B's contiguous direction) ----
    // Each thread issues a batch of independent loads before it stores any
    // of them, so their latencies overlap; no division per element.
    //   rowc (unit row stride): thread t takes rows t, t+256, ... of every
    //   column; otherwise (unit column stride) it takes column t % W of rows
    //   t / W, t / W + 256 / W, ...
    const bool rowc = (brs == 1 || brs == -1);
    const int nrow = nsb * 64;  // rows past nbig up to the 64-boundary are zeroed
    static_assert(NT % W == 0, "strip width must divide the work-group");
    constexpr int RSTEP = NT / W;
    if (rowc) {
        for (int r = t; r < nrow; r += NT) {
            T v[W];
#pragma unroll
            for (int c = 0; c < W; ++c) {
                const bool ok = r < nbig && c0 + c < m;
                v[c] = *(ok ? Bp + r * brs + (c0 + c) * bcs : Bp);
                v[c] = ok ? (TRMM ? v[c] : v[c] * scale) : T(0);
            }
#pragma unroll
            for (int c = 0; c < W; ++c) bs[c * LDB + r] = v[c];
        }
    } else {
        const int c = t % W;
        const bool cok = c0 + c < m;
        constexpr int SL = 8;
        for (int r0 = t / W; r0 < nrow; r0 += RSTEP * SL) {
            T v[SL];
#pragma unroll
            for (int u = 0; u < SL; ++u) {
                const int r = r0 + u * RSTEP;
                const bool ok = cok && r < nbig;
                v[u] = *(ok ? Bp + r * brs + (c0 + c) * bcs : Bp);
                v[u] = ok ? (TRMM ? v[u] : v[u] * scale) : T(0);
            }
#pragma unroll
            for (int u = 0; u < SL; ++u) {
                const int r = r0 + u * RSTEP;
                if (r < nrow) bs[c * LDB + r] = v[u];
            }
        }
    }
    __syncthreads();

    const int lo = lane % MF::BD, lk = lane / MF::BD;
    const int rb = PS::rb0(wave), cbase = PS::cb0(wave);
    // two accumulator sets (even / odd k-steps) keep two independent MFMA
    // chains in flight per wave; they are summed when a block is written
    typename MF::acc_t acc[PS::CBW], acc2[PS::CBW];
    auto bsize = [&](int b) { return (b == nsb - 1) ? nbig - 64 * b : 64; };

    // M-operand fragment of phase ph: op(i, k) for this lane's row.  Elements
    // outside the block (or, for the TRMM diagonal, outside the triangle) are
    // read from panel_zero, so no instruction waits for the loads before the
    // fragment is used, one phase later.
    //
    // The lane's elements of a phase are op(i, kb + KD sw), sw < NKW, each
    // KD * acs past the last: one 64-bit add per load (a multiply per load
    // before, at a quarter of the rate).  Which of them lie inside the block
    // (and its triangle) is a prefix / suffix of sw, computed once per phase
    // as a bit mask.
    static_assert(MF::KD == 4, "four k per MFMA step");
    const int fi = rb * MF::BD + lo;            // the lane's row of a block
    const int kb = kh * NKW * MF::KD + lk;      // the lane's first k
    auto span = [](int c) { return c <= 0 ? 0u : (c >= NKW ? (NKW >= 32 ? ~0u : (1u << NKW) - 1u) : (1u << c) - 1u); };
    // global address space spelled out: a select between the block and
    // panel_zero otherwise degrades the loads to flat ones; its address is
    // taken once (the compiler would rebuild it before every select)
    typedef const __attribute__((address_space(1))) T GT;
    GT *zp = (GT *)panel_zero;
    asm volatile("" : "+v"(zp));
    // the lane's offset and the step between its k in M and W blocks
    const int64_t offM = (int64_t)fi * mrs + (int64_t)kb * mcs, stM = (int64_t)MF::KD * mcs;
    const int64_t offW = (int64_t)fi * wrs + (int64_t)kb * wcs, stW = (int64_t)MF::KD * wcs;
    auto fetch = [&](const Phase &ph, T *af) {
        const T *A;
        int64_t off, st;
        int tri = 0;
        if (ph.kind == 0 && !TRMM) {
            A = Wp + (int64_t)ph.k * 64 * 64;
            off = offW;
            st = stW;
        } else {
            A = Mp + (int64_t)(64 * ph.i) * mrs + (int64_t)(64 * ph.k) * mcs;
            off = offM;
            st = stM;
            if (ph.kind == 0) tri = lower ? 1 : 2;
        }
        const int ni = bsize(ph.i), nk = bsize(ph.k);
        // k < nk: the first ceil((nk - kb) / 4) steps; a unit diagonal is not
        // read (the strict triangle here, the block itself added when the
        // result is written)
        uint32_t bits = fi < ni ? span((nk - kb + 3) >> 2) : 0u;
        if (tri == 1) bits &= span((((unit ? fi - 1 : fi) - kb) >> 2) + 1);             // k <= i (k < i)
        if (tri == 2) bits &= ~span(((unit ? fi + 1 : fi) - kb + 3) >> 2);            // k >= i (k > i)
        GT *p = (GT *)(A + off);
        if (tri == 0 && ni == 64 && nk == 64) {  // a whole block (uniform): nothing to select
#pragma unroll
            for (int sw = 0; sw < NKW; ++sw) {
                af[sw] = *p;
                p += st;
            }
        } else {
#pragma unroll
            for (int sw = 0; sw < NKW; ++sw) {
                af[sw] = *(((bits >> sw) & 1u) ? p : zp);
                p += st;
            }
        }
    };
    auto compute = [&](const Phase &ph, const T *af) {
        const NLA_LDS T *src = bs + ph.k * 64;
        // all LDS operands of the phase are requested before the first MFMA
        T bv[NKW][PS::CBW];
#pragma unroll
        for (int sw = 0; sw < NKW; ++sw)
#pragma unroll
            for (int cb = 0; cb < PS::CBW; ++cb)
                bv[sw][cb] = src[((cbase + cb) * MF::BD + lo) * LDB + (kh * NKW + sw) * MF::KD + lk];
        // keep the scheduler from sinking the reads back next to their MFMAs
        __builtin_amdgcn_sched_barrier(0);
        __builtin_amdgcn_s_setprio(1);  // issue priority while in the MFMA cluster
#pragma unroll
        for (int sw = 0; sw < NKW; ++sw) {
            const T a = af[sw];
#pragma unroll
            for (int cb = 0; cb < PS::CBW; ++cb) {
                if (sw & 1) acc2[cb] = MF::mma(a, bv[sw][cb], acc2[cb]);
                else acc[cb] = MF::mma(a, bv[sw][cb], acc[cb]);
            }
        }
        __builtin_amdgcn_s_setprio(0);
    };
    auto write_block = [&](int ib, T alpha, T beta) {
        NLA_LDS T *dst = bs + ib * 64;
        if constexpr (KSPLIT == 2) {
            // the second k half hands its partial sums to the first
            if (kh == 1) {
#pragma unroll
                for (int cb = 0; cb < PS::CBW; ++cb)
#pragma unroll
                    for (int r = 0; r < MF::NR; ++r) {
                        const int row = rb * MF::BD + MF::di(lane, r);
                        const int col = (cbase + cb) * MF::BD + MF::dj(lane, r);
                        sc[col * 64 + row] = acc[cb][r] + acc2[cb][r];
                    }
            }
            __syncthreads();
            if (kh == 1) return;
#pragma unroll
            for (int cb = 0; cb < PS::CBW; ++cb)
#pragma unroll
                for (int r = 0; r < MF::NR; ++r) {
                    const int row = rb * MF::BD + MF::di(lane, r);
                    const int col = (cbase + cb) * MF::BD + MF::dj(lane, r);
                    NLA_LDS T *p = dst + col * LDB + row;
                    T v = alpha * ((acc[cb][r] + acc2[cb][r]) + sc[col * 64 + row]);
                    *p = (beta == T(0)) ? v : v + beta * *p;
                }
            return;
        }
#pragma unroll
        for (int cb = 0; cb < PS::CBW; ++cb)
#pragma unroll
            for (int r = 0; r < MF::NR; ++r) {
                const int row = rb * MF::BD + MF::di(lane, r);
                const int col = (cbase + cb) * MF::BD + MF::dj(lane, r);
                NLA_LDS T *p = dst + col * LDB + row;
                T v = alpha * (acc[cb][r] + acc2[cb][r]);
                *p = (beta == T(0)) ? v : v + beta * *p;
            }
    };
    auto zero_acc = [&]() {
#pragma unroll
        for (int cb = 0; cb < PS::CBW; ++cb) {
            acc[cb] = MF::zero();
            acc2[cb] = MF::zero();
        }
    };
    // An ill-conditioned diagonal block (flag set by the inverse kernel):
    // X_i = inv(M_ii) B_i followed by one refinement step X_i += inv(M_ii)
    // (B_i - M_ii X_i) (backward stable while kappa(M_ii) * u << 1).  A
    // rare path, in plain FMAs: thread t owns row t % 64 of the block and
    // columns t / 64 + NW * j; X goes through the 64 x W scratch sc.
    auto solve_refined = [&](int ib) {
        constexpr int NC = W / NW;
        const int r = t & 63, cg = t >> 6;
        const int nb = bsize(ib);
        const T *Wk = Wp + (int64_t)ib * 64 * 64;
        const T *Mk = Mp + (int64_t)(64 * ib) * (mrs + mcs);
        NLA_LDS T *blk = bs + ib * 64;
        T x[NC], v[NC];
        // X = inv(M_ii) B_i (the inverse is zero outside its triangle and
        // the identity past nb)
#pragma unroll
        for (int j = 0; j < NC; ++j) x[j] = T(0);
        for (int k = 0; k < 64; ++k) {
            const T w = Wk[r * wrs + k * wcs];
#pragma unroll
            for (int j = 0; j < NC; ++j) x[j] += w * blk[(cg + NW * j) * LDB + k];
        }
#pragma unroll
        for (int j = 0; j < NC; ++j) sc[(cg + NW * j) * 64 + r] = x[j];
        __syncthreads();
        // R = B_i - M_ii X
#pragma unroll
        for (int j = 0; j < NC; ++j) v[j] = unit ? sc[(cg + NW * j) * 64 + r] : T(0);
        if (r < nb) {
            const int k0 = lower ? 0 : (unit ? r + 1 : r);
            const int k1 = lower ? (unit ? r : r + 1) : nb;
            for (int k = k0; k < k1; ++k) {
                const T a = Mk[r * mrs + k * mcs];
#pragma unroll
                for (int j = 0; j < NC; ++j) v[j] += a * sc[(cg + NW * j) * 64 + k];
            }
        }
#pragma unroll
        for (int j = 0; j < NC; ++j) v[j] = blk[(cg + NW * j) * LDB + r] - v[j];
        __syncthreads();  // every thread has read B_i
#pragma unroll
        for (int j = 0; j < NC; ++j) blk[(cg + NW * j) * LDB + r] = v[j];
        __syncthreads();
        // X += inv(M_ii) R
        for (int k = 0; k < 64; ++k) {
            const T w = Wk[r * wrs + k * wcs];
#pragma unroll
            for (int j = 0; j < NC; ++j) x[j] += w * blk[(cg + NW * j) * LDB + k];
        }
        __syncthreads();  // every thread has read R
#pragma unroll
        for (int j = 0; j < NC; ++j) blk[(cg + NW * j) * LDB + r] = x[j];
        __syncthreads();  // X_i complete for the blocks that follow
    };
    // The accumulators are zero when an output block's first phase starts:
    // cleared once here and after every write that consumes them (a
    // conditional clear in the source phases cost a select per register).
    zero_acc();
    auto run = [&](const Phase &ph, const T *af) {
        if (!TRMM) {
            if (ph.kind == 1) {  // acc += M_ik X_k, kept in registers
                compute(ph, af);
            } else {  // B_i -= acc; X_i = inv(M_ii) B_i, in place
                if (!ph.first) write_block(ph.i, T(-1), T(1));
                zero_acc();
                __syncthreads();
                if (Fp != nullptr && Fp[ph.i] != 0) {
                    solve_refined(ph.i);
                    return;
                }
                compute(ph, af);
                __syncthreads();  // every wave has read B_i
                write_block(ph.i, T(1), T(0));
                zero_acc();
                __syncthreads();  // X_i complete for the blocks that follow
            }
            return;
        }
        compute(ph, af);
        if (ph.last) {
            __syncthreads();  // every wave has read B_i
            write_block(ph.i, T(1), unit ? T(1) : T(0));
            zero_acc();
            __syncthreads();
        }
    };

    // phases with the M-operand fragments of the next PD - 1 phases in
    // flight: a ring of PD register buffers, the loop body unrolled PD
    // times so every buffer index is a compile-time constant.  Fetches are
    // unconditional (past the last phase they repeat it): a fetch under a
    // branch made the compiler wait for it at the join, before the phase
    // that did not need it.
    constexpr int PD = 2;
    const int np = phase_count(nsb, TRMM);
    T fr[PD][NKW];
    Phase phr[PD];
    // the next phase to fetch, as (step, q) and its index
    int fstep = 0, fq = 0, fp = 0;
    auto fetch_next = [&](int slot) {
        phr[slot] = phase_at(fstep, fq, nsb, lower, TRMM);
        fetch(phr[slot], fr[slot]);
        if (fp + 1 < np) {  // past the last phase, repeat it
            phase_next(fstep, fq, nsb, TRMM);
            ++fp;
        }
    };
#pragma unroll
    for (int j = 0; j < PD - 1; ++j) fetch_next(j);
    const int np_full = np - np % PD;
    for (int p = 0; p < np_full; p += PD) {
#pragma unroll
        for (int u = 0; u < PD; ++u) {
            fetch_next((u + PD - 1) % PD);
            run(phr[u], fr[u]);
        }
    }
#pragma unroll
    for (int u = 0; u < PD - 1; ++u)
        if (np_full + u < np) run(phr[u], fr[u]);

    // ---- write the strip back ----
    T vmax = T(0);  // largest |X| written (fp32 TRSM with xmax)
    if (rowc) {
        for (int r = t; r < nbig; r += NT) {
#pragma unroll
            for (int c = 0; c < W; ++c) {
                const T v = bs[c * LDB + r];
                if (c0 + c < m) {
                    Bp[r * brs + (c0 + c) * bcs] = TRMM ? v * scale : v;
                    if (!TRMM) vmax = fabs(v) > vmax ? fabs(v) : vmax;
                }
            }
        }
    } else {
        const int c = t % W;
        if (c0 + c < m) {
            for (int r = t / W; r < nbig; r += RSTEP) {
                const T v = bs[c * LDB + r];
                Bp[r * brs + (c0 + c) * bcs] = TRMM ? v * scale : v;
                if (!TRMM) vmax = fabs(v) > vmax ? fabs(v) : vmax;
            }
        }
    }
    if constexpr (sizeof(T) == 4 && !TRMM) {
        if (xmax != nullptr) {
            // a NaN compares false above and never raises the maximum; the
            // refinement's own checks see it in the residual.  One global
            // atomic per work-group: every work-group of the launch targets
            // the same slot, and one per wave serialized there (+10 us per
            // launch).
            __shared__ unsigned s_max;
            unsigned b = __float_as_uint((float)vmax);
            for (int o = 32; o > 0; o >>= 1) {
                const unsigned u = __shfl_xor(b, o);
                b = u > b ? u : b;
            }
            if (t == 0) s_max = 0u;
            __syncthreads();
            if ((t & 63) == 0) atomicMax(&s_max, b);
            __syncthreads();
            if (t == 0 && s_max != 0u) atomicMax(xmax, s_max);
        }
    }
}

template <typename T, int NBIG, int WS, bool TRMM, int NW>
__global__ __launch_bounds__(NW * 64) void panel_kernel(int nbig, int64_t m, const T *__restrict__ Mp, int64_t mrs,
                                                        int64_t mcs, int lower, int unit, T *__restrict__ Bp,
                                                        int64_t brs, int64_t bcs, const T *__restrict__ Wp,
                                                        int64_t wrs, int64_t wcs, T scale,
                                                        const int *__restrict__ Fp, unsigned *xmax) {
    constexpr int LDB = NBIG + (sizeof(T) == 8 ? 2 : 4);
    __shared__ __attribute__((aligned(16))) T bs_raw[WS * LDB];
    // k-split partial sums (8 waves) and the refinement step's scratch
    __shared__ __attribute__((aligned(16))) T sc_raw[(NW == 8 || !TRMM) ? 64 * WS : 1];
    panel_strip<T, NBIG, WS, TRMM, NW>(nbig, m, (int64_t)blockIdx.x * WS, Mp, mrs, mcs, lower != 0, unit != 0, Bp,
                                       brs, bcs, Wp, wrs, wcs, scale, Fp, (NLA_LDS T *)bs_raw,
                                       (NLA_LDS T *)sc_raw, xmax);
}

template <typename T, int NBIG, int W, int NW>
static hipError_t launch_panel_w(const PanelArgs<T> &a, hipStream_t s) {
    const int64_t groups = (a.m + W - 1) / W;
    const int lo = a.lower ? 1 : 0, un = a.unit ? 1 : 0;
    if (a.trmm)
        hipLaunchKernelGGL((panel_kernel<T, NBIG, W, true, NW>), dim3((unsigned)groups), dim3(NW * 64), 0, s,
                           a.nbig, a.m, a.M, a.mrs, a.mcs, lo, un, a.B, a.brs, a.bcs, a.W, a.wrs, a.wcs, a.scale,
                           (const int *)nullptr, (unsigned *)nullptr);
    else
        hipLaunchKernelGGL((panel_kernel<T, NBIG, W, false, NW>), dim3((unsigned)groups), dim3(NW * 64), 0, s,
                           a.nbig, a.m, a.M, a.mrs, a.mcs, lo, un, a.B, a.brs, a.bcs, a.W, a.wrs, a.wcs, a.scale,
                           a.F, a.xmax);
    return hipGetLastError();
}

// Panels up to 512 rows use the 512-row LDS strip (4 or 8 waves); larger
// ones (up to PANEL_MAX) the 1024-row strip with 4 waves, whose FP64 form
// only fits LDS with 16 columns per work-group.
template <typename T>
hipError_t launch_panel(const PanelArgs<T> &a, hipStream_t s) {
    if (a.m <= 0 || a.nbig <= 0) return hipSuccess;
    if (a.nbig > PANEL_MAX) return hipErrorInvalidValue;
    if (a.strip != 16 && a.strip != 32) return hipErrorInvalidValue;
    if (a.waves != 4 && a.waves != 8) return hipErrorInvalidValue;
    if (a.nbig <= 512) {
        if (a.waves == 8)
            return a.strip == 32 ? launch_panel_w<T, 512, 32, 8>(a, s) : launch_panel_w<T, 512, 16, 8>(a, s);
        return a.strip == 32 ? launch_panel_w<T, 512, 32, 4>(a, s) : launch_panel_w<T, 512, 16, 4>(a, s);
    }
    if constexpr (sizeof(T) == 4) {
        if (a.strip == 32) return launch_panel_w<T, 1024, 32, 4>(a, s);
    }
    return launch_panel_w<T, 1024, 16, 4>(a, s);
}

template hipError_t launch_panel<float>(const PanelArgs<float> &, hipStream_t);
template hipError_t launch_panel<double>(const PanelArgs<double> &, hipStream_t);

// ---------------------------------------------------------------------------
// Persistent batched solver: work-groups pull (system, strip) tasks from a
// queue ordered by decreasing cost until it is empty.  Every system is one
// diagonal panel of up to NBIG rows; the kernel is instantiated per size
// class (NBIG = 128 ... 1024) so that the LDS strip, and with it the number
// of work-groups a CU holds, fits the largest system of the class.

template <typename T, bool TRMM, int NBIG>
__global__ __launch_bounds__(256) void batch_kernel(const BatchSys<T> *__restrict__ sys,
                                                    const BatchTask *__restrict__ tasks, int ntasks,
                                                    int *__restrict__ counter) {
    constexpr int WS = BatchStrip<T>::W;
    constexpr int LDB = NBIG + (sizeof(T) == 8 ? 2 : 4);
    __shared__ __attribute__((aligned(16))) T bs_raw[WS * LDB];
    __shared__ __attribute__((aligned(16))) T sc_raw[TRMM ? 1 : 64 * WS];  // refinement scratch
    __shared__ int s_idx;
    for (;;) {
        if (threadIdx.x == 0) s_idx = atomicAdd(counter, 1);
        __syncthreads();
        const int idx = s_idx;
        __syncthreads();
        if (idx >= ntasks) break;  // every wave of the group sees the same idx
        const BatchTask tk = tasks[idx];
        const BatchSys<T> &S = sys[tk.sys];
        panel_strip<T, NBIG, WS, TRMM>(S.n, S.m, (int64_t)tk.strip * WS, S.M, S.mrs, S.mcs, S.lower != 0, false, S.B,
                                       S.brs, S.bcs, S.W, S.wrs, S.wcs, S.scale, TRMM ? nullptr : S.F,
                                       (NLA_LDS T *)bs_raw, (NLA_LDS T *)sc_raw);
        __syncthreads();
    }
}

__global__ void zero_int_kernel(int *p) {
    if (threadIdx.x == 0 && blockIdx.x == 0) *p = 0;
}

template <typename T, bool TRMM>
static const void *batch_ptr_t(int nbig) {
    switch (nbig) {
        case 128: return (const void *)batch_kernel<T, TRMM, 128>;
        case 256: return (const void *)batch_kernel<T, TRMM, 256>;
        case 512: return (const void *)batch_kernel<T, TRMM, 512>;
        case 1024: return (const void *)batch_kernel<T, TRMM, 1024>;
        default: return nullptr;
    }
}

template <typename T>
const void *batch_kernel_ptr(bool trmm, int nbig) {
    return trmm ? batch_ptr_t<T, true>(nbig) : batch_ptr_t<T, false>(nbig);
}

template <typename T>
hipError_t launch_batch(const BatchSys<T> *sys, const BatchTask *tasks, int ntasks, int *counter, bool trmm,
                        int nbig, int groups, hipStream_t s) {
    if (ntasks <= 0) return hipSuccess;
    const void *k = batch_kernel_ptr<T>(trmm, nbig);
    if (!k) return hipErrorInvalidValue;
    hipLaunchKernelGGL(zero_int_kernel, dim3(1), dim3(64), 0, s, counter);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    void *args[] = {(void *)&sys, (void *)&tasks, (void *)&ntasks, (void *)&counter};
    return hipLaunchKernel(k, dim3(groups), dim3(256), args, 0, s);
}

template const void *batch_kernel_ptr<float>(bool, int);
template const void *batch_kernel_ptr<double>(bool, int);

template hipError_t launch_batch<float>(const BatchSys<float> *, const BatchTask *, int, int *, bool, int, int,
                                        hipStream_t);
template hipError_t launch_batch<double>(const BatchSys<double> *, const BatchTask *, int, int *, bool, int, int,
                                         hipStream_t);

}  // namespace nla
