// Kernels of the mixed-precision triangular solve with FP64 iterative
// refinement (nla_rectrxm_mixed in include/nla.h).
//
// The refinement loop needs, besides the plain recursion:
//   - a low-precision GEMM update for the correction solves: op(X), op(Y)
//     stored as fp32, scaled by a power of two per work-group operand panel
//     (fp16 kinds: from the largest magnitude recorded for the panel, so
//     fp16 neither overflows nor flushes its large entries), rounded to
//     bf16, fp16 or fp16 pairs on their way into LDS, multiplied on the
//     16x16x32 MFMA with fp32 accumulation, and unscaled in the epilogue;
//   - conversions between the FP64 iterate / residual and the fp32 solve
//     buffers, and the norms of the stopping test, as HBM-bound elementwise
//     kernels with deterministic (fixed-order) reductions.
// The FP64 residual itself is the MFMA GEMM of nla_gemm.hip restricted to
// the triangle of op(A).
#include <math.h>

#include "nla_device.h"
#include "nla_internal.h"

namespace nla {

#define NLA_LDS __attribute__((address_space(3)))

typedef __bf16 bf16x8_t __attribute__((ext_vector_type(8)));
typedef _Float16 f16x8_t __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x4_t __attribute__((ext_vector_type(4)));
typedef _Float16 f16x4_t __attribute__((ext_vector_type(4)));

template <int LP>
struct LpType;
template <>
struct LpType<0> {  // bf16: 8 significant bits, fp32's exponent range
    typedef __bf16 elem_t;
    typedef bf16x8_t op_t;
    typedef bf16x4_t quad_t;
    __device__ static inline f4_t mma(op_t a, op_t b, f4_t c) {
        return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
    }
};
template <>
struct LpType<1> {  // fp16: 11 significant bits, largest finite 65504
    typedef _Float16 elem_t;
    typedef f16x8_t op_t;
    typedef f16x4_t quad_t;
    __device__ static inline f4_t mma(op_t a, op_t b, f4_t c) {
        return __builtin_amdgcn_mfma_f32_16x16x32_f16(a, b, c, 0, 0, 0);
    }
};
// fp16 pairs: each scaled fp32 value v as hi + lo with hi = fp16(v) and
// lo = fp16(v - hi) (about 22 significant bits); a product takes the three
// terms hi*hi + hi*lo + lo*hi (each fp16 x fp16 product is exact in fp32; the
// dropped lo*lo term is below 2^-22 of it), three fp16 MFMAs instead of one.
template <>
struct LpType<2> : LpType<1> {};

template <int LP>
struct LpSplit {
    static constexpr bool value = false;
};
template <>
struct LpSplit<2> {
    static constexpr bool value = true;
};

// Scale that puts a block with absolute maximum `amax` (bit pattern of a
// non-negative float) just below 2^15: exact (a power of two), and the
// largest scaled entry stays below fp16's largest finite value.
__device__ inline float block_scale(unsigned bits) {
    float mx = __uint_as_float(bits);
    if (!(mx > 0.0f) || !(mx < INFINITY)) return 1.0f;
    int e;
    (void)frexpf(mx, &e);  // mx < 2^e
    int k = 15 - e;        // kept inside fp32's normal range
    k = k > 126 ? 126 : (k < -126 ? -126 : k);
    return ldexpf(1.0f, k);
}

// Stages a BO x 32 block (BO = 128 with 256 threads, 256 with 512) of an
// fp32 operand into LDS as BO rows of 32 k (64 bytes) in the low-precision
// type.  Row r sits in row slot
// r ^ ((r >> 2) & 3) (the low two bits flipped by the next two), and its
// 16-byte chunk c at chunk c ^ (2 * (bit 2 ^ bit 4 of r)) (lp_lds_off):
//   fragment reads (16 rows x one chunk per 16 lanes, 8 rows x 2 chunks per
//     b128 phase) meet 16 distinct 16-byte bank groups;
//   each 32-lane half of a tile store meets 32 distinct 8-byte bank pairs.
// Each thread moves 4 rows x 4 k (one 8-byte quad per row):
//   k-contiguous operand: quad t % 8 of rows t/8 + (NT/8) v (v = 0..3), four
//     16-byte loads along k (8 lanes cover 128 contiguous bytes of a row);
//   outer-contiguous operand: rows 4 g .. 4 g + 3 (g = 8 (t / 64) + t % 8),
//     quad (t / 8) % 8, four 16-byte loads along the rows (8 lanes cover
//     128 contiguous bytes of a column).
__device__ __forceinline__ int lp_lds_off(int row, int k) {  // element offset, k a multiple of 4
    const int slot = row ^ ((row >> 2) & 3);
    const int chunk = (k >> 3) ^ (((row >> 1) ^ (row >> 3)) & 2);
    return slot * 32 + ((chunk << 3) | (k & 7));
}

template <int LP, int BO, int BK, int NT, bool KCONTIG>
struct LpLoader {
    typedef typename LpType<LP>::elem_t E;
    typedef typename LpType<LP>::quad_t Q;
    static_assert(BK == 32 && ((BO == 128 && NT == 256) || (BO == 256 && NT == 512)),
                  "the LDS layout is laid out for BO x 32 tiles, 2 BO threads");
    static constexpr bool SPLIT = LpSplit<LP>::value;
    static constexpr int LDK = BK;
    static constexpr int TILE_ELEMS = BO * LDK;                     // one tile (hi parts for a split type)
    static constexpr int LDS_ELEMS = (SPLIT ? 2 : 1) * TILE_ELEMS;  // per stage; lo parts follow the hi tile

    float r[4][4];  // [row offset][k offset]
    uint32_t okbits;

    __device__ static inline int g_of(int t) { return KCONTIG ? t / 8 : 8 * (t / 64) + (t % 8); }
    __device__ static inline int q_of(int t) { return KCONTIG ? t % 8 : (t / 8) % 8; }
    // tile row of the thread's v-th row (r[v])
    __device__ static inline int row_of(int t, int v) { return KCONTIG ? g_of(t) + (NT / 8) * v : 4 * g_of(t) + v; }

    // Element (outer, k) at P[outer*so + k*sk]; (KCONTIG ? sk : so) == 1.
    template <bool CHECK>
    __device__ inline void load(const float *__restrict__ P, int64_t so, int64_t sk, int64_t o0, int64_t k0,
                                int64_t no, int64_t nk, int t) {
        if (CHECK) okbits = 0;
        const int k = 4 * q_of(t);
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            // KCONTIG: row v, 4 contiguous k; else: k + v, 4 contiguous rows
            const int64_t go = o0 + (KCONTIG ? row_of(t, v) : row_of(t, 0)), gk = k0 + k + (KCONTIG ? 0 : v);
            if (!CHECK) {
                f4_t x = *reinterpret_cast<const f4_t *>(P + go * so + gk * sk);
#pragma unroll
                for (int e = 0; e < 4; ++e) {
                    if (KCONTIG) r[v][e] = x[e];
                    else r[e][v] = x[e];
                }
            } else {
#pragma unroll
                for (int e = 0; e < 4; ++e) {
                    const int64_t eo = KCONTIG ? go : go + e, ek = KCONTIG ? gk + e : gk;
                    const bool ok = (eo < no) && (ek < nk);
                    const float x = *(ok ? P + eo * so + ek * sk : P);
                    if (KCONTIG) r[v][e] = x;
                    else r[e][v] = x;
                    okbits |= (ok ? 1u : 0u) << (v * 4 + e);
                }
            }
        }
    }

    // Zero the out-of-bounds elements of a CHECK load (bit order of load()).
    __device__ inline void apply_bounds() {
#pragma unroll
        for (int v = 0; v < 4; ++v)
#pragma unroll
            for (int e = 0; e < 4; ++e)
                if (!((okbits >> (v * 4 + e)) & 1u)) {
                    if (KCONTIG) r[v][e] = 0.0f;
                    else r[e][v] = 0.0f;
                }
    }

    __device__ inline void store(NLA_LDS E *lds, int t, float sc) const {
        const int k = 4 * q_of(t);
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            Q q, ql;
#pragma unroll
            for (int kk = 0; kk < 4; ++kk) {
                const float x = r[v][kk] * sc;
                q[kk] = (E)x;
                if (SPLIT) ql[kk] = (E)(x - (float)q[kk]);
            }
            *reinterpret_cast<NLA_LDS Q *>(lds + lp_lds_off(row_of(t, v), k)) = q;
            if (SPLIT) *reinterpret_cast<NLA_LDS Q *>(lds + TILE_ELEMS + lp_lds_off(row_of(t, v), k)) = ql;
        }
    }
};

// One BM x BN tile of C = beta*C + alpha*op(X)*op(Y) (fp32 C).  The MFMA
// computes C^T per 16x16 block (A operand from op(Y), B operand from op(X)),
// so the 16 lanes of a result register run down one column of C.
template <int LP, int BM, int BN, int BK, int WM, int WN, bool TX, bool TY, bool CHECK>
__device__ __forceinline__ void lp_tile(int64_t M, int64_t N, int64_t K, float alpha, const float *__restrict__ X,
                                        int64_t ldx, const float *__restrict__ Y, int64_t ldy, float beta, float *C,
                                        int64_t ldc, int64_t row0, int64_t col0, float sx, float sy,
                                        NLA_LDS typename LpType<LP>::elem_t *lds_xb,
                                        NLA_LDS typename LpType<LP>::elem_t *lds_yb) {
    typedef LpType<LP> LT;
    typedef typename LT::elem_t E;
    typedef typename LT::op_t OP;
    constexpr int NT = WM * WN * 64;
    constexpr int TWM = BM / WM, TWN = BN / WN;
    constexpr int RB = TWM / 16, CB = TWN / 16;
    static_assert(BK % 32 == 0, "BK must be a multiple of the MFMA depth");
    typedef LpLoader<LP, BM, BK, NT, TX> LX;   // rows of C; k-contiguous when TX
    typedef LpLoader<LP, BN, BK, NT, !TY> LY;  // columns of C; k-contiguous unless TY
    NLA_LDS E *lds_x[2] = {lds_xb, lds_xb + LX::LDS_ELEMS};
    NLA_LDS E *lds_y[2] = {lds_yb, lds_yb + LY::LDS_ELEMS};

    const int t = threadIdx.x;
    const int lane = t & 63, wave = t >> 6;
    const int wm = wave % WM, wn = wave / WM;
    const int64_t xso = TX ? ldx : 1, xsk = TX ? 1 : ldx;
    const int64_t yso = TY ? 1 : ldy, ysk = TY ? ldy : 1;

    f4_t acc[CB][RB];
#pragma unroll
    for (int cb = 0; cb < CB; ++cb)
#pragma unroll
        for (int rb = 0; rb < RB; ++rb) acc[cb][rb] = f4_t{0.0f, 0.0f, 0.0f, 0.0f};

    LX lx;
    LY ly;
    const int nkt = (int)((K + BK - 1) / BK);
    if (nkt > 0) {
        lx.template load<CHECK>(X, xso, xsk, row0, 0, M, K, t);
        ly.template load<CHECK>(Y, yso, ysk, col0, 0, N, K, t);
        if (CHECK) {
            lx.apply_bounds();
            ly.apply_bounds();
        }
        lx.store(lds_x[0], t, sx);
        ly.store(lds_y[0], t, sy);
    }
    __syncthreads();

    const int lo = lane & 15, lk = (lane >> 4) * 8;
    for (int kt = 0; kt < nkt; ++kt) {
        const int cur = kt & 1;
        const bool more = kt + 1 < nkt;
        if (more) {
            lx.template load<CHECK>(X, xso, xsk, row0, (int64_t)(kt + 1) * BK, M, K, t);
            ly.template load<CHECK>(Y, yso, ysk, col0, (int64_t)(kt + 1) * BK, N, K, t);
        }
        const NLA_LDS E *sx_ = lds_x[cur];
        const NLA_LDS E *sy_ = lds_y[cur];
        __builtin_amdgcn_s_setprio(1);  // issue priority while in the MFMA cluster
#pragma unroll
        for (int ks = 0; ks < BK; ks += 32) {
            OP a[CB], b[RB];
#pragma unroll
            for (int cb = 0; cb < CB; ++cb)
                a[cb] = *reinterpret_cast<const NLA_LDS OP *>(sy_ + lp_lds_off(wn * TWN + cb * 16 + lo, ks + lk));
#pragma unroll
            for (int rb = 0; rb < RB; ++rb)
                b[rb] = *reinterpret_cast<const NLA_LDS OP *>(sx_ + lp_lds_off(wm * TWM + rb * 16 + lo, ks + lk));
            if constexpr (LX::SPLIT) {
                // the small terms first: hi*lo + lo*hi, then hi*hi
                OP al[CB], bl[RB];
#pragma unroll
                for (int cb = 0; cb < CB; ++cb)
                    al[cb] = *reinterpret_cast<const NLA_LDS OP *>(sy_ + LY::TILE_ELEMS +
                                                                   lp_lds_off(wn * TWN + cb * 16 + lo, ks + lk));
#pragma unroll
                for (int rb = 0; rb < RB; ++rb)
                    bl[rb] = *reinterpret_cast<const NLA_LDS OP *>(sx_ + LX::TILE_ELEMS +
                                                                   lp_lds_off(wm * TWM + rb * 16 + lo, ks + lk));
#pragma unroll
                for (int cb = 0; cb < CB; ++cb)
#pragma unroll
                    for (int rb = 0; rb < RB; ++rb) {
                        acc[cb][rb] = LT::mma(a[cb], bl[rb], acc[cb][rb]);
                        acc[cb][rb] = LT::mma(al[cb], b[rb], acc[cb][rb]);
                    }
            }
#pragma unroll
            for (int cb = 0; cb < CB; ++cb)
#pragma unroll
                for (int rb = 0; rb < RB; ++rb) acc[cb][rb] = LT::mma(a[cb], b[rb], acc[cb][rb]);
        }
        __builtin_amdgcn_s_setprio(0);
        if (more) {
            if (CHECK) {
                lx.apply_bounds();
                ly.apply_bounds();
            }
            lx.store(lds_x[cur ^ 1], t, sx);
            ly.store(lds_y[cur ^ 1], t, sy);
        }
        __syncthreads();
    }

    // D[i][j] (lane j + 16*(i/4), register i%4) is C(row j, column i).
    // unscale in double: sx*sy may lie outside fp32's range
    const double f = (double)alpha / ((double)sx * (double)sy);
#pragma unroll
    for (int cb = 0; cb < CB; ++cb) {
        float cv[RB][4];
        if (beta != 0.0f) {
#pragma unroll
            for (int rb = 0; rb < RB; ++rb)
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    int64_t row = row0 + wm * TWM + rb * 16 + lo;
                    int64_t col = col0 + wn * TWN + cb * 16 + (lane >> 4) * 4 + r;
                    bool ok = !CHECK || (row < M && col < N);
                    cv[rb][r] = *(ok ? C + row + col * ldc : C);
                }
        }
#pragma unroll
        for (int rb = 0; rb < RB; ++rb)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                int64_t row = row0 + wm * TWM + rb * 16 + lo;
                int64_t col = col0 + wn * TWN + cb * 16 + (lane >> 4) * 4 + r;
                if (!CHECK || (row < M && col < N)) {
                    float v = (float)(f * (double)acc[cb][rb][r]);
                    C[row + col * ldc] = beta != 0.0f ? v + beta * cv[rb][r] : v;
                }
            }
    }
}

__device__ inline void lp_tile_coords(int G, int g, int tiles_m, int &tm, int &tn) {
    // XCD-aware renumbering as in the FP32/FP64 GEMM: work-group g runs on
    // XCD g % 8; each XCD walks a contiguous range of tiles in groups of 8
    // row tiles.
    const int NX = 8, GM = 8;
    int xcd = g % NX, idx = g / NX;
    int per = G / NX, rem = G % NX;
    int gp = (xcd < rem) ? xcd * (per + 1) + idx : rem * (per + 1) + (xcd - rem) * per + idx;
    int tiles_n = G / tiles_m;
    int group = GM * tiles_n;
    int gid = gp / group;
    int first = gid * GM;
    int gm = tiles_m - first < GM ? tiles_m - first : GM;
    int in = gp % group;
    tm = first + in % gm;
    tn = in / gm;
}

// Largest magnitude the scale source records for the operand panel with
// outer indices [o0, o0 + BO) (clipped to no) over k in [0, K), as a float
// bit pattern (0 when there is none); every thread of the work-group gets it.
template <int NT>
__device__ unsigned lp_panel_max(const LpScale &sc, int64_t o0, int BO, int64_t no, int64_t K, unsigned *red) {
    const int t = threadIdx.x;
    unsigned m = 0;
    if (sc.ld == 0) {
        for (int64_t i = t; i < sc.c0; i += NT) {
            const unsigned v = sc.tab[sc.r0 + i];
            m = v > m ? v : m;
        }
    } else {
        const int64_t o1 = (o0 + BO < no ? o0 + BO : no);
        int64_t rlo, rhi, clo, chi;  // stored element ranges [lo, hi)
        if (!sc.trans) {
            rlo = sc.r0 + o0; rhi = sc.r0 + o1; clo = sc.c0; chi = sc.c0 + K;
        } else {
            rlo = sc.r0; rhi = sc.r0 + K; clo = sc.c0 + o0; chi = sc.c0 + o1;
        }
        const int64_t tr0 = rlo >> 6, ntr = ((rhi + 63) >> 6) - tr0;
        const int64_t tc0 = clo >> 6, ntc = ((chi + 63) >> 6) - tc0;
        for (int64_t i = t; i < ntr * ntc; i += NT) {
            const int64_t tr = tr0 + i % ntr, tc = tc0 + i / ntr;
            const unsigned v = sc.tab[tr + tc * sc.ld];
            m = v > m ? v : m;
        }
    }
    for (int o = 32; o > 0; o >>= 1) {
        const unsigned v = __shfl_xor(m, o);
        m = v > m ? v : m;
    }
    if ((t & 63) == 0) red[t >> 6] = m;
    __syncthreads();
    m = 0;
#pragma unroll
    for (int w = 0; w < NT / 64; ++w) m = red[w] > m ? red[w] : m;
    __syncthreads();  // red is reused by the next call
    return m;
}

template <int LP, int BM, int BN, int BK, int WM, int WN, bool TX, bool TY>
__global__ __launch_bounds__(WM *WN * 64) void lp_gemm_kernel(int64_t M, int64_t N, int64_t K, float alpha,
                                                              const float *__restrict__ X, int64_t ldx,
                                                              const float *__restrict__ Y, int64_t ldy, float beta,
                                                              float *C, int64_t ldc, int tiles_m, LpScale scx,
                                                              LpScale scy) {
    typedef typename LpType<LP>::elem_t E;
    constexpr int NT = WM * WN * 64;
    typedef LpLoader<LP, BM, BK, NT, TX> LX;
    typedef LpLoader<LP, BN, BK, NT, !TY> LY;
    __shared__ __attribute__((aligned(16))) E lds_x[2 * LX::LDS_ELEMS];
    __shared__ __attribute__((aligned(16))) E lds_y[2 * LY::LDS_ELEMS];
    __shared__ unsigned red[NT / 64];
    int tm, tn;
    lp_tile_coords(gridDim.x, blockIdx.x, tiles_m, tm, tn);
    const int64_t row0 = (int64_t)tm * BM, col0 = (int64_t)tn * BN;
    // per-work-group scales: this tile's rows of op(X) and columns of op(Y)
    const float sx = scx.tab ? block_scale(lp_panel_max<NT>(scx, row0, BM, M, K, red)) : 1.0f;
    const float sy = scy.tab ? block_scale(lp_panel_max<NT>(scy, col0, BN, N, K, red)) : 1.0f;
    const bool interior = (row0 + BM <= M) && (col0 + BN <= N) && (K % BK == 0);
    if (interior)
        lp_tile<LP, BM, BN, BK, WM, WN, TX, TY, false>(M, N, K, alpha, X, ldx, Y, ldy, beta, C, ldc, row0, col0,
                                                        sx, sy, (NLA_LDS E *)lds_x, (NLA_LDS E *)lds_y);
    else
        lp_tile<LP, BM, BN, BK, WM, WN, TX, TY, true>(M, N, K, alpha, X, ldx, Y, ldy, beta, C, ldc, row0, col0,
                                                       sx, sy, (NLA_LDS E *)lds_x, (NLA_LDS E *)lds_y);
}

template <int LP, int BM, int BN, int BK, int WM, int WN, bool TX, bool TY>
static hipError_t lp_launch_cfg(const LpGemmArgs &a, hipStream_t s) {
    int tiles_m = (int)((a.M + BM - 1) / BM);
    int tiles_n = (int)((a.N + BN - 1) / BN);
    hipLaunchKernelGGL((lp_gemm_kernel<LP, BM, BN, BK, WM, WN, TX, TY>), dim3(tiles_m * tiles_n),
                       dim3(WM * WN * 64), 0, s, a.M, a.N, a.K, a.alpha, a.X, a.ldx, a.Y, a.ldy, a.beta, a.C,
                       a.ldc, tiles_m, a.scx, a.scy);
    return hipGetLastError();
}

template <int LP, int BM, int BN, int BK, int WM, int WN>
static hipError_t lp_launch_trans(const LpGemmArgs &a, hipStream_t s) {
    if (!a.tx && !a.ty) return lp_launch_cfg<LP, BM, BN, BK, WM, WN, false, false>(a, s);
    if (a.tx && !a.ty) return lp_launch_cfg<LP, BM, BN, BK, WM, WN, true, false>(a, s);
    if (!a.tx && a.ty) return lp_launch_cfg<LP, BM, BN, BK, WM, WN, false, true>(a, s);
    return lp_launch_cfg<LP, BM, BN, BK, WM, WN, true, true>(a, s);
}

hipError_t launch_lp_gemm(const LpGemmArgs &a, hipStream_t s) {
    if (a.M <= 0 || a.N <= 0) return hipSuccess;
    // Unchecked interior tiles issue 16-byte loads: every column (row, for a
    // transposed operand) must start 16-byte aligned.
    const bool vec = ((uintptr_t)a.X & 15) == 0 && ((uintptr_t)a.Y & 15) == 0 && a.ldx % 4 == 0 && a.ldy % 4 == 0;
    if (!vec) return hipErrorInvalidValue;
    // 256 x 256 tiles (eight waves of 64 x 128) halve the operand traffic of
    // the large updates; fp16 pairs keep 128 x 128 (two tiles per operand
    // would not fit the registers of the larger shape)
    const bool big = a.M >= 1024 && a.N >= 1024;
    if (a.lp == 0) return big ? lp_launch_trans<0, 256, 256, 32, 4, 2>(a, s) : lp_launch_trans<0, 128, 128, 32, 2, 2>(a, s);
    if (a.lp == 1) return big ? lp_launch_trans<1, 256, 256, 32, 4, 2>(a, s) : lp_launch_trans<1, 128, 128, 32, 2, 2>(a, s);
    if (a.lp == 2) return lp_launch_trans<2, 128, 128, 32, 2, 2>(a, s);
    return hipErrorInvalidValue;
}

// ---- elementwise and reduction kernels -----------------------------------

// Column-major rows x cols walks: threads along the rows (the unit-stride
// dimension), work-groups of the grid's y dimension along the columns.  A
// thread takes four rows of a column (the loop unrolled, so their loads are
// in flight together) and the grid holds at most EW_BLOCKS work-groups: with
// one element per thread and a work-group per column the launch spent its
// time dispatching work-groups (about a third of the HBM rate).
#define NLA_EW_BEGIN                                                                \
    for (int64_t c = blockIdx.y; c < cols; c += gridDim.y)                          \
        _Pragma("unroll 4") for (int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; r < rows; \
                                 r += (int64_t)gridDim.x * blockDim.x) {
#define NLA_EW_END }

constexpr int64_t EW_BLOCKS = 8192;

static inline dim3 ew_grid(int64_t rows, int64_t cols, int64_t max_blocks = EW_BLOCKS) {
    int64_t bx = (rows + 1023) / 1024;
    if (bx > 64) bx = 64;
    int64_t by = cols < 65535 ? cols : 65535;
    if (bx * by > max_blocks) by = max_blocks / bx > 0 ? max_blocks / bx : 1;
    return dim3((unsigned)bx, (unsigned)by);
}

// tab[i + j*nt] = max |A32| over tile (i, j) of the stored triangle: one
// work-group per tile, thread t on row t % 64 of 16 columns (the walk runs
// down the unit-stride rows).  Non-negative floats order like their bit
// patterns, so the maximum is an integer max.
__global__ __launch_bounds__(256) void tile_absmax_kernel(int64_t n, const float *__restrict__ A32, int64_t ld32,
                                                          int lower, unsigned *__restrict__ tab, int64_t nt) {
    __shared__ unsigned red[4];
    const int64_t ti = blockIdx.x, tj = blockIdx.y;
    const int t = threadIdx.x;
    unsigned m = 0;
    if (lower ? ti >= tj : ti <= tj) {
        const int64_t r = ti * 64 + (t & 63);
        const int64_t c0 = tj * 64 + (t >> 6) * 16;
        if (ti != tj && (ti + 1) * 64 <= n && (tj + 1) * 64 <= n) {
            // a whole off-diagonal tile: 16 loads in flight, no checks
            const float *p = A32 + r + c0 * ld32;
            float v[16];
#pragma unroll
            for (int u = 0; u < 16; ++u) v[u] = p[u * ld32];
#pragma unroll
            for (int u = 0; u < 16; ++u) {
                const unsigned b = __float_as_uint(fabsf(v[u]));
                m = b > m ? b : m;
            }
        } else if (r < n) {
#pragma unroll 4
            for (int u = 0; u < 16; ++u) {
                const int64_t c = c0 + u;
                if (c < n && (lower ? r >= c : r <= c)) {
                    const unsigned b = __float_as_uint(fabsf(A32[r + c * ld32]));
                    m = b > m ? b : m;
                }
            }
        }
    }
    for (int o = 32; o > 0; o >>= 1) {
        const unsigned v = __shfl_xor(m, o);
        m = v > m ? v : m;
    }
    if ((t & 63) == 0) red[t >> 6] = m;
    __syncthreads();
    if (t == 0) {
        unsigned r = red[0];
        for (int w = 1; w < 4; ++w) r = red[w] > r ? red[w] : r;
        tab[ti + tj * nt] = r;
    }
}

hipError_t launch_tile_absmax(int64_t n, const float *A32, int64_t ld32, bool lower, unsigned *tab, int64_t nt,
                              hipStream_t s) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(tile_absmax_kernel, dim3((unsigned)nt, (unsigned)nt), dim3(256), 0, s, n, A32, ld32,
                       lower ? 1 : 0, tab, nt);
    return hipGetLastError();
}

__global__ void fill_u32_kernel(unsigned *p, int64_t n, unsigned v) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        p[i] = v;
}

hipError_t launch_fill_u32(unsigned *p, int64_t n, unsigned v, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    int blocks = (int)((n + 255) / 256);
    if (blocks > 1024) blocks = 1024;
    hipLaunchKernelGGL(fill_u32_kernel, dim3(blocks), dim3(256), 0, s, p, n, v);
    return hipGetLastError();
}

// A32 = fp32(A) on the stored triangle (lower: r >= c, upper: r <= c).
__global__ void tri_to_f32_kernel(int64_t rows, int64_t cols, const double *__restrict__ A, int64_t lda, int lower,
                                  float *__restrict__ A32, int64_t ld32) {
    NLA_EW_BEGIN
    if (lower ? r >= c : r <= c) A32[r + c * ld32] = (float)A[r + c * lda];
    NLA_EW_END
}

// S = alpha * B (FP64).
__global__ void scale_copy_kernel(int64_t rows, int64_t cols, double alpha, const double *__restrict__ B, int64_t ldb,
                                  double *__restrict__ S, int64_t lds) {
    NLA_EW_BEGIN
    S[r + c * lds] = alpha * B[r + c * ldb];
    NLA_EW_END
}

// D = fp32(scale * R).
__global__ void down_kernel(int64_t rows, int64_t cols, double scale, const double *__restrict__ R, int64_t ldr,
                            float *__restrict__ D, int64_t ldd) {
    NLA_EW_BEGIN
    D[r + c * ldd] = (float)(scale * R[r + c * ldr]);
    NLA_EW_END
}

// X = beta * X + sigma * fp64(D)   (beta = 0: X = sigma * D)
__global__ void up_kernel(int64_t rows, int64_t cols, double sigma, const float *__restrict__ D, int64_t ldd, double beta,
                          double *__restrict__ X, int64_t ldx) {
    NLA_EW_BEGIN
    double d = sigma * (double)D[r + c * ldd];
    X[r + c * ldx] = beta != 0.0 ? beta * X[r + c * ldx] + d : d;
    NLA_EW_END
}

hipError_t launch_tri_to_f32(int64_t n, const double *A, int64_t lda, bool lower, float *A32, int64_t ld32,
                             hipStream_t s) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(tri_to_f32_kernel, ew_grid(n, n), dim3(256), 0, s, n, n, A, lda, lower ? 1 : 0, A32, ld32);
    return hipGetLastError();
}

hipError_t launch_scale_copy(int64_t rows, int64_t cols, double alpha, const double *B, int64_t ldb, double *S,
                             int64_t lds, hipStream_t s) {
    if (rows <= 0 || cols <= 0) return hipSuccess;
    hipLaunchKernelGGL(scale_copy_kernel, ew_grid(rows, cols), dim3(256), 0, s, rows, cols, alpha, B, ldb, S, lds);
    return hipGetLastError();
}

hipError_t launch_down(int64_t rows, int64_t cols, double scale, const double *R, int64_t ldr, float *D,
                       int64_t ldd, hipStream_t s) {
    if (rows <= 0 || cols <= 0) return hipSuccess;
    hipLaunchKernelGGL(down_kernel, ew_grid(rows, cols), dim3(256), 0, s, rows, cols, scale, R, ldr, D, ldd);
    return hipGetLastError();
}

hipError_t launch_up(int64_t rows, int64_t cols, double sigma, const float *D, int64_t ldd, double beta, double *X,
                     int64_t ldx, hipStream_t s) {
    if (rows <= 0 || cols <= 0) return hipSuccess;
    hipLaunchKernelGGL(up_kernel, ew_grid(rows, cols), dim3(256), 0, s, rows, cols, sigma, D, ldd, beta, X, ldx);
    return hipGetLastError();
}

// Sum of squares and absolute maximum of a column-major rows x cols matrix
// (restricted to a triangle when tri = 1 lower / 2 upper), in two passes
// with a fixed reduction order: NORM_PARTS partial pairs, then one
// work-group folds them into out[0] = sum of squares, out[1] = max |.|.
constexpr int NORM_PARTS = NORM_PARTS_MAX;

__global__ void norm_partial_kernel(int64_t rows, int64_t cols, const double *__restrict__ P, int64_t ld, int tri,
                                    double *__restrict__ part) {
    __shared__ double ss[256], mx[256];
    double sum = 0.0, m = 0.0;
    NLA_EW_BEGIN
    if ((tri == 1 && r < c) || (tri == 2 && r > c)) continue;
    double v = P[r + c * ld];
    sum += v * v;
    double a = fabs(v);
    m = a > m ? a : m;
    NLA_EW_END
    ss[threadIdx.x] = sum;
    mx[threadIdx.x] = m;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) {
            ss[threadIdx.x] += ss[threadIdx.x + w];
            mx[threadIdx.x] = mx[threadIdx.x] > mx[threadIdx.x + w] ? mx[threadIdx.x] : mx[threadIdx.x + w];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        const int b = blockIdx.y * gridDim.x + blockIdx.x;
        part[2 * b] = ss[0];
        part[2 * b + 1] = mx[0];
    }
}

__global__ void norm_final_kernel(int nparts, const double *__restrict__ part, double *__restrict__ out) {
    __shared__ double ss[256], mx[256];
    double s = 0.0, m = 0.0;
    for (int i = threadIdx.x; i < nparts; i += blockDim.x) {
        s += part[2 * i];
        m = part[2 * i + 1] > m ? part[2 * i + 1] : m;
    }
    ss[threadIdx.x] = s;
    mx[threadIdx.x] = m;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) {
            ss[threadIdx.x] += ss[threadIdx.x + w];
            mx[threadIdx.x] = mx[threadIdx.x] > mx[threadIdx.x + w] ? mx[threadIdx.x] : mx[threadIdx.x + w];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        out[0] = ss[0];
        out[1] = mx[0];
    }
}

hipError_t launch_norms(int64_t rows, int64_t cols, const double *P, int64_t ld, int tri, double *part,
                        double *out, hipStream_t s) {
    int nparts = 0;
    if (rows > 0 && cols > 0) {
        dim3 g = ew_grid(rows, cols, NORM_PARTS);
        nparts = (int)(g.x * g.y);
        hipLaunchKernelGGL(norm_partial_kernel, g, dim3(256), 0, s, rows, cols, P, ld, tri, part);
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    hipLaunchKernelGGL(norm_final_kernel, dim3(1), dim3(256), 0, s, nparts, part, out);
    return hipGetLastError();
}


// ---- per-system norms of the stopping test ---------------------------------
//
// A right-hand side and its solution form one system: a column of B for the
// left side, a row for the right side.  The stopping test is normwise per
// system, max_j ||r_j|| / (||A||_lb ||x_j|| + ||s_j||), with ||A||_lb the
// largest column 2-norm of A's stored triangle (a lower bound of ||A||_2, so
// the value bounds the error with the exact 2-norm from above).  Sums of
// squares and absolute maxima are taken per (system, chunk of the reduction
// dimension) and folded in a fixed order, so every result is deterministic.

// Systems are columns (unit stride along the reduction): one wave per column,
// lanes over the rows of chunk blockIdx.y.  tri: 1 lower / 2 upper triangle
// only (r >= c / r <= c), 0 the whole column.
__global__ __launch_bounds__(256) void sys_sumsq_col_kernel(int64_t rows, int64_t cols, const double *__restrict__ P,
                                                            int64_t ld, int tri, int64_t chunk,
                                                            double *__restrict__ part) {
    const int lane = threadIdx.x & 63;
    const int64_t c = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    if (c >= cols) return;
    int64_t r0 = (int64_t)blockIdx.y * chunk, r1 = r0 + chunk < rows ? r0 + chunk : rows;
    if (tri == 1 && r0 < c) r0 = c;
    if (tri == 2 && r1 > c + 1) r1 = c + 1;
    // four partial sums (rows r0 + lane + 64 (4 i + u) in sum u), so that
    // four loads are in flight and the adds do not wait on each other
    double su[4] = {0.0, 0.0, 0.0, 0.0}, mx = 0.0;
    const double *col = P + c * ld;
    int64_t r = r0 + lane;
    for (; r + 192 < r1; r += 256) {
        double v[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) v[u] = col[r + 64 * u];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            su[u] += v[u] * v[u];
            mx = fmax(mx, fabs(v[u]));
        }
    }
#pragma unroll
    for (int u = 0; u < 4; ++u, r += 64)
        if (r < r1) {
            const double v = col[r];
            su[u] += v * v;
            mx = fmax(mx, fabs(v));
        }
    double ss = (su[0] + su[1]) + (su[2] + su[3]);
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        ss += __shfl_xor(ss, o);
        mx = fmax(mx, __shfl_xor(mx, o));
    }
    if (lane == 0) {
        double *q = part + 2 * ((int64_t)blockIdx.y * cols + c);
        q[0] = ss;
        q[1] = mx;
    }
}

// Systems are rows (unit stride across the systems): one thread per row, over
// the columns of chunk blockIdx.y.
__global__ __launch_bounds__(256) void sys_sumsq_row_kernel(int64_t rows, int64_t cols, const double *__restrict__ P,
                                                            int64_t ld, int64_t chunk, double *__restrict__ part) {
    const int64_t r = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (r >= rows) return;
    const int64_t c0 = (int64_t)blockIdx.y * chunk, c1 = c0 + chunk < cols ? c0 + chunk : cols;
    double ss = 0.0, mx = 0.0;
    for (int64_t c = c0; c < c1; ++c) {
        const double v = P[r + c * ld];
        ss += v * v;
        mx = fmax(mx, fabs(v));
    }
    double *q = part + 2 * ((int64_t)blockIdx.y * rows + r);
    q[0] = ss;
    q[1] = mx;
}

int sys_chunks(int64_t nsys, int64_t len) {
    // enough work-groups to cover the chip, chunks of at least 64 elements
    const int64_t groups = nsys / 4 + 1;
    int64_t nch = (2048 + groups - 1) / groups;
    const int64_t cap = (len + 63) / 64;
    if (nch > cap) nch = cap;
    if (nch > SYS_CHUNKS_MAX) nch = SYS_CHUNKS_MAX;
    return nch < 1 ? 1 : (int)nch;
}

hipError_t launch_sys_sumsq(int64_t rows, int64_t cols, const double *P, int64_t ld, bool by_col, int tri, int nch,
                            double *part, hipStream_t s) {
    if (rows <= 0 || cols <= 0) return hipSuccess;
    if (by_col) {
        const int64_t chunk = (rows + nch - 1) / nch;
        hipLaunchKernelGGL(sys_sumsq_col_kernel, dim3((unsigned)((cols + 3) / 4), (unsigned)nch), dim3(256), 0, s,
                           rows, cols, P, ld, tri, chunk, part);
    } else {
        const int64_t chunk = (cols + nch - 1) / nch;
        hipLaunchKernelGGL(sys_sumsq_row_kernel, dim3((unsigned)((rows + 255) / 256), (unsigned)nch), dim3(256), 0,
                           s, rows, cols, P, ld, chunk, part);
    }
    return hipGetLastError();
}

__device__ inline void fold_block4(double *v, double *red, int nt) {
    // v[0], v[2] summed; v[1], v[3] max-reduced over the work-group (fixed order)
    const int t = threadIdx.x;
#pragma unroll
    for (int k = 0; k < 4; ++k) red[k * 256 + t] = v[k];
    __syncthreads();
    for (int w = nt / 2; w > 0; w >>= 1) {
        if (t < w) {
            red[t] += red[t + w];
            red[256 + t] = fmax(red[256 + t], red[256 + t + w]);
            red[512 + t] += red[512 + t + w];
            red[768 + t] = fmax(red[768 + t], red[768 + t + w]);
        }
        __syncthreads();
    }
}

// Folds the partials of one matrix: per-system sums of squares into sys2
// (when not null) and out[0] = total sum of squares, out[1] = max |.|,
// out[2] = max_j sqrt(system j's sum of squares).
__global__ __launch_bounds__(256) void sys_fold_kernel(int nch, int64_t nsys, const double *__restrict__ part,
                                                       double *__restrict__ sys2, double *__restrict__ out) {
    __shared__ double red[4 * 256];
    double v[4] = {0.0, 0.0, 0.0, 0.0};
    for (int64_t j = threadIdx.x; j < nsys; j += 256) {
        double ss = 0.0, mx = 0.0;
        for (int c = 0; c < nch; ++c) {
            ss += part[2 * ((int64_t)c * nsys + j)];
            mx = fmax(mx, part[2 * ((int64_t)c * nsys + j) + 1]);
        }
        if (sys2) sys2[j] = ss;
        v[0] += ss;
        v[1] = fmax(v[1], mx);
        v[3] = fmax(v[3], sqrt(ss));
    }
    fold_block4(v, red, 256);
    if (threadIdx.x == 0) {
        out[0] = red[0];
        out[1] = red[256];
        out[2] = red[768];
    }
}

// The stopping test from the partials of R and X: out[0] = ||R||_F^2,
// out[1] = max |R|, out[2] = ||X||_F^2, out[3] = max_j ||r_j|| /
// (anorm_lb ||x_j|| + ||s_j||) (0 for a system with r_j = 0; +inf if a
// norm is not finite).
__global__ __launch_bounds__(256) void sys_eta_kernel(int nch, int64_t nsys, const double *__restrict__ pr,
                                                      const double *__restrict__ px, const double *__restrict__ s2,
                                                      const double *__restrict__ anorm_lb, double *__restrict__ out) {
    __shared__ double red[4 * 256];
    const double alb = *anorm_lb;
    double v[4] = {0.0, 0.0, 0.0, 0.0};
    for (int64_t j = threadIdx.x; j < nsys; j += 256) {
        double r2 = 0.0, rmx = 0.0, x2 = 0.0;
        for (int c = 0; c < nch; ++c) {
            r2 += pr[2 * ((int64_t)c * nsys + j)];
            rmx = fmax(rmx, pr[2 * ((int64_t)c * nsys + j) + 1]);
            x2 += px[2 * ((int64_t)c * nsys + j)];
        }
        const double num = sqrt(r2), den = alb * sqrt(x2) + sqrt(s2[j]);
        double eta = num == 0.0 ? 0.0 : (den > 0.0 ? num / den : INFINITY);
        if (!(eta == eta) || !(r2 < INFINITY) || !(x2 < INFINITY)) eta = INFINITY;
        v[0] += r2;
        v[1] = fmax(v[1], rmx);
        v[2] += x2;
        v[3] = fmax(v[3], eta);
    }
    fold_block4(v, red, 256);
    if (threadIdx.x == 0) {
        out[0] = red[0];
        out[1] = red[256];
        out[2] = red[512];
        out[3] = red[768];
    }
}

hipError_t launch_sys_fold(int nch, int64_t nsys, const double *part, double *sys2, double *out, hipStream_t s) {
    hipLaunchKernelGGL(sys_fold_kernel, dim3(1), dim3(256), 0, s, nch, nsys, part, sys2, out);
    return hipGetLastError();
}

hipError_t launch_sys_eta(int nch, int64_t nsys, const double *pr, const double *px, const double *s2,
                          const double *anorm_lb, double *out, hipStream_t s) {
    hipLaunchKernelGGL(sys_eta_kernel, dim3(1), dim3(256), 0, s, nch, nsys, pr, px, s2, anorm_lb, out);
    return hipGetLastError();
}

}  // namespace nla
