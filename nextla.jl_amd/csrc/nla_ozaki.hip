// FP64-accurate residual R = S - A*X of the refined solve on the int8 MFMA
// (Ozaki's scheme II: products of integers modulo small moduli, combined by
// the Chinese remainder theorem).
//
// The residual the refinement needs (src/rectrxm.jl:62-68 is the solve it
// converges to) costs n^2 m flops, as much as the TRSM itself, and FP64 MFMA
// runs at the FP64 vector rate on MI355X; the int8 MFMA runs at 64 times
// that rate.  With A's rows and X's columns scaled to integers,
//     A_int(i,k) = round(A(i,k) 2^e_i),  |A_int| < 2^bA,
//     X_int(k,j) = round(X(k,j) 2^f_j),  |X_int| < 2^bX,
//     S_int(i,j) = round(S(i,j) 2^(e_i + f_j)),
// the integer R_int = S_int - sum_k A_int X_int is known exactly once it is
// known modulo T pairwise coprime moduli p_t <= 256 whose product M exceeds
// 2 |R_int|: each residue plane sum_k (A_int mod p_t)(X_int mod p_t) is an
// exact int32 sum of int8 products (the int8 MFMA), and the reconstruction
// recovers R_int in (-M/2, M/2] and returns R = R_int 2^-(e_i + f_j).
//
// Accuracy: the only rounding is that of A to bA bits per row and of X to
// bX bits per column.  X is the refinement's own iterate, so its rounded
// value is written back (the residual is then that of the returned X); the
// rounding of A is a backward perturbation of at most 2^-bA of each row's
// largest entry.  With bA, bX >= 55 both stay far below the 1e-14 target.
//
// Layout (all int8 / uint8, 256-aligned dimensions npad, mpad):
//   A residues: the triangle of op(A) (or its transpose for side R), in row
//     panels of 256 rows holding only the k range of the triangle, k
//     contiguous;
//   X residues: (system j, k) at j*npad + k;
//   product residues: column-major like R.
#include "nla_device.h"
#include "nla_internal.h"

namespace nla {

#define NLA_LDS __attribute__((address_space(3)))
typedef int v4i __attribute__((ext_vector_type(4)));

// v mod p in [0, p) for an integer-valued double |v| < 2^53; invp = 1/p.
// floor(v/p) from the product is off by at most one (its relative error is
// 2^-52), and v - p*q is an integer below 2p in magnitude, so the fma is
// exact.
__device__ __forceinline__ int oz_mod53(double v, double p, double invp) {
    const double q = floor(v * invp);
    int r = (int)fma(-p, q, v);
    const int pi = (int)p;
    r += (r < 0) ? pi : 0;
    r -= (r >= pi) ? pi : 0;
    return r;
}

// An integer-valued double |v| <= 2^62 as v = l3 2^48 + l2 2^32 + l1 2^16 + l0
// with l0..l2 in [0, 2^16) and l3 in [-2^14, 2^14] (each remainder has at
// most 53 significant bits, so every fma is exact).
__device__ __forceinline__ void oz_limbs16(double v, int &l0, int &l1, int &l2, int &l3) {
    const double h3 = floor(v * 0x1p-48);
    double r = fma(-h3, 0x1p48, v);
    const double h2 = floor(r * 0x1p-32);
    r = fma(-h2, 0x1p32, r);
    const double h1 = floor(r * 0x1p-16);
    l0 = (int)fma(-h1, 0x1p16, r);
    l1 = (int)h1;
    l2 = (int)h2;
    l3 = (int)h3;
}

// The low byte of the centred residue of v (given by its limbs) modulo p_t
// (the residue lies in [rhi_t - p_t + 1, rhi_t], rhi_t = floor((p_t - 1) / 2),
// so its byte is its int8 value).  With the limb weights' centred residues
// c_i (|c_i| <= 128), x = l0 + sum l_i c_i = v (mod p_t) and |x| < 2^25.  For
// odd p_t, q = round(x / p_t) in double is exact (the product's error is below
// 2^-34 and x / p_t lies at least 1 / (2 p_t) from a half-integer), so
// x - q p_t is the centred residue itself; for p_t = 256 the byte of x - q p_t
// is x's low byte whatever q is.  No correction steps.
__device__ __forceinline__ int oz_residue16(int l0, int l1, int l2, int l3, const OzConst &C, int t) {
    int x = __mul24(l1, C.rc[t][0]) + l0;
    x = __mul24(l2, C.rc[t][1]) + x;
    x = __mul24(l3, C.rc[t][2]) + x;
    const int q = __double2int_rn((double)x * C.invpd[t]);
    return __mul24(q, -C.p[t]) + x;
}

// ---- scales -----------------------------------------------------------------

// Source matrix of a conversion: element (o, k) at P[o*so + k*sk], o < no,
// k < nk; tri = 1 keeps k <= o, 2 keeps k >= o (the triangle of op(A)).
struct OzSrc {
    double *P;
    int64_t so, sk, no, nk;
    int tri;
};

__device__ __forceinline__ bool oz_keep(const OzSrc &s, int64_t o, int64_t k) {
    if (o >= s.no || k >= s.nk) return false;
    if (s.tri == 1) return k <= o;
    if (s.tri == 2) return k >= o;
    return true;
}

// Stages the 64 x 64 tile (o0, k0) of the source into LDS (zero outside the
// kept region), loading along whichever stride is 1.
__device__ __forceinline__ void oz_load_tile(const OzSrc &s, int64_t o0, int64_t k0, double (*tile)[65]) {
    // all 16 loads of a thread are issued before the first LDS store, so
    // they are in flight together
    const bool kfast = (s.sk == 1);
    double v[16];
#pragma unroll
    for (int u = 0; u < 16; ++u) {
        const int idx = threadIdx.x + 256 * u, a = idx & 63, b = idx >> 6;
        const int oo = kfast ? b : a, ko = kfast ? a : b;
        const int64_t o = o0 + oo, k = k0 + ko;
        v[u] = oz_keep(s, o, k) ? s.P[o * s.so + k * s.sk] : 0.0;
    }
#pragma unroll
    for (int u = 0; u < 16; ++u) {
        const int idx = threadIdx.x + 256 * u, a = idx & 63, b = idx >> 6;
        tile[kfast ? b : a][kfast ? a : b] = v[u];
    }
}

// mx[o] = max(mx[o], max_k |src(o, k)|) as the bit pattern of a
// non-negative double (which orders like the value; NaN orders above
// infinity).  Grid: (ceil(nk/64), ceil(no/64)).
__global__ __launch_bounds__(256) void oz_absmax_kernel(OzSrc s, unsigned long long *mx) {
    __shared__ double tile[64][65];
    const int64_t o0 = (int64_t)blockIdx.y * 64, k0 = (int64_t)blockIdx.x * 64;
    if (s.tri == 1 && k0 > o0 + 63) return;
    if (s.tri == 2 && k0 + 63 < o0) return;
    __shared__ unsigned long long part[3][64];
    oz_load_tile(s, o0, k0, tile);
    __syncthreads();
    // wave w reduces columns 16w .. 16w + 15 of every row; wave 0 combines
    const int oo = threadIdx.x & 63, w = threadIdx.x >> 6;
    unsigned long long best = 0;
#pragma unroll
    for (int k = 0; k < 16; ++k) {
        const unsigned long long b = (unsigned long long)__double_as_longlong(fabs(tile[oo][16 * w + k]));
        best = b > best ? b : best;
    }
    if (w > 0) part[w - 1][oo] = best;
    __syncthreads();
    if (w == 0) {
#pragma unroll
        for (int q = 0; q < 3; ++q) best = part[q][oo] > best ? part[q][oo] : best;
        if (o0 + oo < s.no && best != 0) atomicMax(mx + o0 + oo, best);
    }
}

// The same maxima for a source whose rows are contiguous (sk = 1) and whole
// (no triangle), as X's systems are: one wave per row and chunk of k, four
// coalesced loads in flight per lane, one atomicMax per wave (no staging of
// tiles through LDS).
__global__ __launch_bounds__(256) void oz_absmax_rows_kernel(OzSrc s, int64_t chunk, unsigned long long *mx) {
    const int lane = threadIdx.x & 63;
    const int64_t o = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    if (o >= s.no) return;
    const int64_t k0 = (int64_t)blockIdx.y * chunk, k1 = k0 + chunk < s.nk ? k0 + chunk : s.nk;
    const double *row = s.P + o * s.so;
    unsigned long long b[4] = {0, 0, 0, 0};
    int64_t k = k0 + lane;
    for (; k + 192 < k1; k += 256) {
        double v[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) v[u] = row[k + 64 * u];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const unsigned long long x = (unsigned long long)__double_as_longlong(fabs(v[u]));
            b[u] = x > b[u] ? x : b[u];
        }
    }
#pragma unroll
    for (int u = 0; u < 4; ++u, k += 64)
        if (k < k1) {
            const unsigned long long x = (unsigned long long)__double_as_longlong(fabs(row[k]));
            b[u] = x > b[u] ? x : b[u];
        }
    unsigned long long best = b[0] > b[1] ? b[0] : b[1];
    best = b[2] > best ? b[2] : best;
    best = b[3] > best ? b[3] : best;
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        const unsigned long long x = __shfl_xor(best, off);
        best = x > best ? x : best;
    }
    if (lane == 0 && best != 0) atomicMax(mx + o, best);
}

// e[o] = bits - 1 - floor(log2 mx[o]) (so that max |src(o,.)| 2^e[o] lies in
// [2^(bits-1), 2^bits)), 0 for an all-zero row; a non-finite maximum sets
// *flag and e[o] = OZ_BAD_EXP (the row is then neither converted nor
// rewritten).
constexpr int OZ_BAD_EXP = -(1 << 20);
__global__ void oz_exponent_kernel(int64_t no, const unsigned long long *mx, int bits, int *e, int *flag) {
    const int64_t o = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (o >= no) return;
    const double v = __longlong_as_double((long long)mx[o]);
    int r = 0;
    if (!(v <= 1.7976931348623157e308)) {
        *flag = 1;
        r = OZ_BAD_EXP;
    } else if (v > 0.0) {
        r = bits - 1 - ilogb(v);
    }
    e[o] = r;
}

// ---- residues ---------------------------------------------------------------

// Destination of the residue planes: plane t at base + t*plane.  Both
// operands of the modular products are stored as the product kernel stages
// them: in groups of 256 rows (a row panel of the triangle of op(A), or 256
// systems of X), each group a run of 16 KB blocks, one per 64-wide k step
// (the triangle's k range for A: k < (tm+1)*256 lower, k >= tm*256 upper;
// all of npad for X).  Row r of a block is 64 bytes at r*64, its 16-byte
// chunk c at chunk c ^ 2*bit2(r): the block is copied to LDS as it is, and
// the fragment reads of a wave (lane l: row l % 16, chunk l / 16) then meet
// 16 distinct 16-byte bank groups per phase.  panel: 0 = X (groups
// 256 * npad bytes apart), 1 = lower triangle, 2 = upper triangle.
struct OzDst {
    int8_t *base;
    int64_t plane, npad;
    int panel;
};

// Byte offset of the 16-byte chunk of row o holding k .. k + 15 (k a
// multiple of 16) in a plane of layout `panel`.
__device__ __forceinline__ int64_t oz_tiled_off(int panel, int64_t npad, int64_t o, int64_t k) {
    const int64_t g = o >> 8;
    const int r = (int)(o & 255);
    int64_t base, kk;
    if (panel == 0) {
        base = g * 256 * npad;
        kk = k;
    } else if (panel == 1) {
        base = (int64_t)65536 * (g * (g + 1) / 2);
        kk = k;
    } else {
        base = (int64_t)256 * (g * npad - 256 * (g * (g - 1) / 2));
        kk = k - g * 256;
    }
    const int c = (int)((kk >> 4) & 3);
    return base + (kk >> 6) * 16384 + r * 64 + ((c ^ ((r >> 1) & 2)) << 4);
}

// Whether the 64 x 64 tile (o0, k0) lies in the stored region of a panel
// layout (row panel tm = o0/256 holds k < (tm+1)*256 (lower) or k >= tm*256).
__device__ __forceinline__ bool oz_tile_stored(int panel, int64_t o0, int64_t k0) {
    const int64_t tm = o0 >> 8;
    if (panel == 1) return k0 < (tm + 1) * 256;
    if (panel == 2) return k0 >= tm * 256;
    return true;
}

// Residues of round(src(o,k) 2^e[o]) modulo every p_t, as int8 in the
// symmetric range, written to the T planes of dst; with writeback the
// rounded value round(src 2^e) 2^-e replaces src (the refinement's iterate).
// Grid: (npad_k/64, npad_o/64) over the whole padded destination, so the
// padding holds zeros.
__global__ __launch_bounds__(256) void oz_residue_kernel(OzSrc s, const int *e, OzDst d, int writeback, OzConst C) {
    __shared__ double tile[64][65];
    const int64_t o0 = (int64_t)blockIdx.y * 64, k0 = (int64_t)blockIdx.x * 64;
    if (!oz_tile_stored(d.panel, o0, k0)) return;
    oz_load_tile(s, o0, k0, tile);
    __syncthreads();
    const int oo = threadIdx.x >> 2, kq = (threadIdx.x & 3) * 16;
    const int64_t o = o0 + oo;
    const int eraw = (o < s.no) ? e[o] : 0;
    const bool bad = (eraw == OZ_BAD_EXP);  // a non-finite row: residues 0, values kept
    const int eo = bad ? 0 : eraw;
    // |v| <= 2^62: the scales keep every rounded entry within 2^bA / 2^bX
    int l0[16], l1[16], l2[16], l3[16];
#pragma unroll
    for (int u = 0; u < 16; ++u) {
        const double v = bad ? 0.0 : rint(ldexp(tile[oo][kq + u], eo));
        oz_limbs16(v, l0[u], l1[u], l2[u], l3[u]);
        if (writeback && !bad) tile[oo][kq + u] = ldexp(v, -eo);
    }
    const int64_t off = oz_tiled_off(d.panel, d.npad, o, k0 + kq);
    for (int t = 0; t < C.T; ++t) {
        v4i packed;
#pragma unroll
        for (int w = 0; w < 4; ++w) {
            unsigned word = 0;
#pragma unroll
            for (int b = 0; b < 4; ++b) {
                const int u = 4 * w + b;
                const int r = oz_residue16(l0[u], l1[u], l2[u], l3[u], C, t);
                word |= (unsigned)(r & 255) << (8 * b);
            }
            packed[w] = (int)word;
        }
        *reinterpret_cast<v4i *>(d.base + t * d.plane + off) = packed;
    }
    if (writeback) {
        __syncthreads();
        const bool kfast = (s.sk == 1);
        for (int idx = threadIdx.x; idx < 64 * 64; idx += 256) {
            const int a = idx & 63, b = idx >> 6;
            const int oq = kfast ? b : a, kk = kfast ? a : b;
            const int64_t og = o0 + oq, kg = k0 + kk;
            if (og < s.no && kg < s.nk) s.P[og * s.so + kg * s.sk] = tile[oq][kk];
        }
    }
}

// ---- the modular products ---------------------------------------------------

// Per modulus t and 256 x 256 output tile (tm, tn):
//   c(i, j) = sum_k Ares_t(i, k) Xres_t(j, k) (int32, exact: |c| < K 2^14),
// stored as c mod p_t in [0, p_t).  A persistent grid takes (modulus, tile)
// items from the queue counters (zero at launch).  Eight waves, each a
// 64 x 128 block of the tile as 4 x 8 v_mfma_i32_16x16x64_i8 blocks.
//
// Operands reach LDS without passing through registers: every 64-wide k
// step of the A panel and of the X panel is a 16 KB block of the residue
// planes (oz_tiled_off), and each wave copies its 2 KB share of both with
// four global_load_lds_dwordx4 (64 lanes x 16 bytes into 1 KB of LDS, lane
// l at byte 16 l).  S stages of 32 KB form a ring: the copies of step
// kt + S - 1 are issued when step kt starts, so S - 1 steps (96 KB with
// S = 4) are in flight per CU while one is multiplied.  At the top of step
// kt each wave waits until its own copies of step kt have landed (the later
// steps' 4 (S - 2) copies may still be outstanding: vmcnt counts in issue
// order) and then meets the others at a barrier, after which every copy of
// step kt is in LDS and every wave has finished reading the stage that the
// new copies overwrite.  The barrier and the fragment reads are written
// with their waits in inline assembly: the compiler cannot tell which
// stage an LDS-bound copy writes, and would otherwise wait for all of them
// before every read.  The copies past the last step repeat it into stages
// nobody reads; they land before the next tile (the queue's barrier waits
// for every outstanding memory operation).
//
// The planes' layout also keeps each copy instruction on 1 KB of
// contiguous memory.  Against the earlier kernel (rows of the plane L bytes
// apart, staged through registers into 96-byte padded LDS rows, two steps
// in flight) the products of the n = 16384, m = 4096 residual went from
// 9.0 to 8.2 ms.  What bounds them now, measured with variants of this
// kernel: the fragment reads (96 KB per CU per step; LDS delivers about
// 128 bytes per clock) and the MFMAs serialize at the per-step barrier
// (6.4 ms with no global traffic at all), and the operand copies alone
// take 5.2 ms.
//
// OUT_T: the output plane is (j, i) with j contiguous (side R) instead of
// (i, j) with i contiguous.
//
// Tile order.  sr == 0: one counter; tiles modulus by modulus, the longest
// rows of the triangle first, all column tiles of a row together.  sr > 0:
// one counter per XCD (work-group g runs on XCD g % 8) and tiles in
// super-tiles of sr row tiles x sc column tiles of one modulus; an XCD's 32
// CUs take the tiles of one super-tile together, so its L2 fetches sr A
// panels and sc X panels for sr * sc tiles.  Super-tiles, longest first, go
// to the XCDs in rounds of 8, in reverse order on odd rounds so that every
// XCD gets rows of every length.
struct OzGemmArgs {
    const int8_t *A;
    int64_t aplane;
    const int8_t *X;
    int64_t xplane;
    uint8_t *Cr;
    int64_t cplane, ldc;
    int64_t npad, mpad;
    int tiles_m, tiles_n, T, lower;
    int sr, sc;  // super-tile shape (sr == 0: one counter, row order)
    int *queue;  // 8 counters
};

template <int S>
__device__ __forceinline__ void oz_stage_barrier() {
    // this wave's copies of the current step have landed (4 (S - 2) later
    // ones may be outstanding), then every wave's have
    static_assert(S == 3 || S == 4, "ring depth");
    if constexpr (S == 4) asm volatile("s_waitcnt vmcnt(8)\n\ts_barrier" ::: "memory");
    else asm volatile("s_waitcnt vmcnt(4)\n\ts_barrier" ::: "memory");
}

template <bool OUT_T, int S>
__global__ __launch_bounds__(512) void oz_gemm_kernel(OzGemmArgs a, OzConst C) {
    constexpr int BM = 256, BN = 256, BK = 64, TILE = BM * BK;
    __shared__ __attribute__((aligned(1024))) int8_t sbuf[S * 2 * TILE];  // stage: A block, X block
    __shared__ int s_item;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm = wave & 3, wn = wave >> 2;
    const int per_mod = a.tiles_m * a.tiles_n, total = per_mod * a.T;
    const int fr = lane & 15, fc = lane >> 4;
    const int xcd = blockIdx.x & 7;
    const int nbands = a.sr > 0 ? (a.tiles_m + a.sr - 1) / a.sr : 0;
    const int ngroups = a.sr > 0 ? (a.tiles_n + a.sc - 1) / a.sc : 0;
    const int sper = a.sr * a.sc, ssup = nbands * ngroups * a.T;
    // fragment of lane l: row 16 b + l % 16, chunk l / 16 (stored at chunk
    // (l / 16) ^ 2*bit2(row); bit 2 of the row is bit 2 of l % 16)
    const int frag_off = fr * BK + ((fc ^ ((fr >> 1) & 2)) << 4);
    NLA_LDS int8_t *lds = (NLA_LDS int8_t *)sbuf;
    for (;;) {
        if (tid == 0) s_item = atomicAdd(a.queue + (a.sr > 0 ? xcd : 0), 1);
        __syncthreads();
        const int w = s_item;
        __syncthreads();
        int t, tr, tn;  // modulus, row tile counted from the longest, column tile
        if (a.sr == 0) {
            if (w >= total) break;
            t = w / per_mod;
            const int rest = w - t * per_mod;
            tr = rest / a.tiles_n;
            tn = rest % a.tiles_n;
        } else {
            const int k = w / sper, within = w - k * sper;
            const int sup = 8 * k + ((k & 1) ? 7 - xcd : xcd);
            if (sup >= ssup) break;
            t = sup / (nbands * ngroups);
            const int rest = sup - t * nbands * ngroups;
            tr = (rest / ngroups) * a.sr + within / a.sc;
            tn = (rest % ngroups) * a.sc + within % a.sc;
            if (tr >= a.tiles_m || tn >= a.tiles_n) continue;  // past the edge of a partial super-tile
        }
        const int tm = a.lower ? a.tiles_m - 1 - tr : tr;
        int64_t L, off, kbeg;
        if (a.lower) {
            L = (int64_t)(tm + 1) * BM;
            off = (int64_t)BM * BM * ((int64_t)tm * (tm + 1) / 2);
            kbeg = 0;
        } else {
            L = a.npad - (int64_t)tm * BM;
            off = (int64_t)BM * ((int64_t)tm * a.npad - (int64_t)BM * ((int64_t)tm * (tm - 1) / 2));
            kbeg = (int64_t)tm * BM;
        }
        // this wave's 2 KB of every 16 KB step block (lane l: 16 bytes at 16 l)
        const int8_t *Ag = a.A + t * a.aplane + off + wave * 2048 + lane * 16;
        const int8_t *Xg = a.X + t * a.xplane + (int64_t)tn * BN * a.npad + kbeg * 256 + wave * 2048 + lane * 16;
        const int nkt = (int)(L / BK);  // >= 4: L is a positive multiple of 256
        auto copy = [&](int step, int stage) {
            const int64_t k0 = (int64_t)step * TILE;
            NLA_LDS int8_t *dA = lds + stage * 2 * TILE + wave * 2048;
            NLA_LDS int8_t *dX = dA + TILE;
            __builtin_amdgcn_global_load_lds((const void *)(Ag + k0), (NLA_LDS void *)dA, 16, 0, 0);
            __builtin_amdgcn_global_load_lds((const void *)(Ag + k0 + 1024), (NLA_LDS void *)(dA + 1024), 16, 0, 0);
            __builtin_amdgcn_global_load_lds((const void *)(Xg + k0), (NLA_LDS void *)dX, 16, 0, 0);
            __builtin_amdgcn_global_load_lds((const void *)(Xg + k0 + 1024), (NLA_LDS void *)(dX + 1024), 16, 0, 0);
        };

        v4i acc[4][8];
#pragma unroll
        for (int rb = 0; rb < 4; ++rb)
#pragma unroll
            for (int cb = 0; cb < 8; ++cb) acc[rb][cb] = v4i{0, 0, 0, 0};

#pragma unroll
        for (int p = 0; p < S - 1; ++p) copy(p < nkt ? p : nkt - 1, p);
        for (int kt = 0; kt < nkt; ++kt) {
            oz_stage_barrier<S>();
            const int nxt = kt + S - 1;
            copy(nxt < nkt ? nxt : nkt - 1, nxt % S);
            const NLA_LDS int8_t *la = lds + (kt % S) * 2 * TILE + wm * 64 * BK + frag_off;
            const NLA_LDS int8_t *lx = lds + (kt % S) * 2 * TILE + TILE + wn * 128 * BK + frag_off;
            v4i af[4], xf[8];
            asm volatile(
                "ds_read_b128 %0, %12\n\t"
                "ds_read_b128 %1, %12 offset:1024\n\t"
                "ds_read_b128 %2, %12 offset:2048\n\t"
                "ds_read_b128 %3, %12 offset:3072\n\t"
                "ds_read_b128 %4, %13\n\t"
                "ds_read_b128 %5, %13 offset:1024\n\t"
                "ds_read_b128 %6, %13 offset:2048\n\t"
                "ds_read_b128 %7, %13 offset:3072\n\t"
                "ds_read_b128 %8, %13 offset:4096\n\t"
                "ds_read_b128 %9, %13 offset:5120\n\t"
                "ds_read_b128 %10, %13 offset:6144\n\t"
                "ds_read_b128 %11, %13 offset:7168\n\t"
                "s_waitcnt lgkmcnt(0)"
                : "=&v"(af[0]), "=&v"(af[1]), "=&v"(af[2]), "=&v"(af[3]), "=&v"(xf[0]), "=&v"(xf[1]),
                  "=&v"(xf[2]), "=&v"(xf[3]), "=&v"(xf[4]), "=&v"(xf[5]), "=&v"(xf[6]), "=&v"(xf[7])
                : "v"((unsigned)(uintptr_t)la), "v"((unsigned)(uintptr_t)lx));
            __builtin_amdgcn_s_setprio(1);
#pragma unroll
            for (int rb = 0; rb < 4; ++rb)
#pragma unroll
                for (int cb = 0; cb < 8; ++cb)
                    acc[rb][cb] = OUT_T ? __builtin_amdgcn_mfma_i32_16x16x64_i8(xf[cb], af[rb], acc[rb][cb], 0, 0, 0)
                                        : __builtin_amdgcn_mfma_i32_16x16x64_i8(af[rb], xf[cb], acc[rb][cb], 0, 0, 0);
            __builtin_amdgcn_s_setprio(0);
        }

        // residues of the tile: lane holds 4 consecutive rows (columns for
        // OUT_T) of one column (row) of each 16 x 16 block
        const double pd = (double)C.p[t], ipd = C.invpd[t];
        uint8_t *Cp = a.Cr + t * a.cplane;
#pragma unroll
        for (int rb = 0; rb < 4; ++rb)
#pragma unroll
            for (int cb = 0; cb < 8; ++cb) {
                unsigned word = 0;
#pragma unroll
                for (int r = 0; r < 4; ++r)
                    word |= (unsigned)oz_mod53((double)acc[rb][cb][r], pd, ipd) << (8 * r);
                const int64_t i0 = (int64_t)tm * BM + wm * 64 + rb * 16;
                const int64_t j0 = (int64_t)tn * BN + wn * 128 + cb * 16;
                if (!OUT_T)
                    *reinterpret_cast<unsigned *>(Cp + (j0 + fr) * a.ldc + i0 + fc * 4) = word;
                else
                    *reinterpret_cast<unsigned *>(Cp + (i0 + fr) * a.ldc + j0 + fc * 4) = word;
            }
    }
}

// ---- reconstruction ---------------------------------------------------------

// R(r, c) = S(r, c) - (A X) at (r, c) from the product residues c_t, in
// exact integer arithmetic on doubles.  With w_t = y_t (M / p_t) (1 modulo
// p_t, 0 modulo the other moduli), V = sum_t c_t w_t = P modulo M, where
// P = (A X)_int and |P|, |S_int| <= M/8.  Numbers are kept as four limbs of
// base 2^40 (weights 1, 2^40, 2^80, 2^120); every limb operation is an exact
// integer below 2^53:
//   V's limbs: sum_t c_t w_t,i with c_t < 256, w_t,i < 2^40, T <= 20;
//   D = V - S_int, S_int's limbs from |S_int| = m 2^sh (53-bit m) by exact
//       floor / fma steps;
//   k = round(D / M) from D's limbs in rounded double (|D - kM| = |R_int|
//       <= M/4 leaves a margin of 1/4 against an error near 2^-39);
//   E = D - kM = -R_int limb by limb, carries propagated (E0..E2 in
//       [0, 2^40)), and E's value formed by three fmas: exact below 2^93,
//       within 3 ulp above (no cancellation there).
// R_int is scaled back by 2^-(er + ec), the scale exponents of R's rows and
// columns.  |S_int| above C.smax, or a non-finite S, sets *flag (the modulus
// product would not cover R_int).
struct OzCrtArgs {
    int64_t rows, cols;
    const double *S;
    int64_t lds;
    double *R;
    int64_t ldr;
    const uint8_t *Cr;
    int64_t cplane, ldc;
    const int *er, *ec;
    int *flag;
};

// Four consecutive rows per thread: one 32-bit load gives their residues
// of one plane, and every plane's load is issued before the first is used.
// The planes are padded to multiples of 256 rows, so the 4-row word is
// always inside the plane; rows past `rows` are skipped.  TT: the number of
// moduli (C.T), or OZ_TMAX for any count (the basis padded with zeros); the
// basis limbs are read from LDS (80 doubles do not fit the scalar registers).
template <int TT>
__global__ __launch_bounds__(256) void oz_crt_kernel(OzCrtArgs a, OzConst C) {
    __shared__ double wl[TT][4];
    for (int i = threadIdx.x; i < TT * 4; i += 256) wl[i / 4][i % 4] = (i / 4 < C.T) ? C.w[i / 4][i % 4] : 0.0;
    __syncthreads();
    const int64_t r0 = ((int64_t)blockIdx.x * 256 + threadIdx.x) * 4;
    if (r0 >= a.rows) return;
    constexpr double B40 = 0x1p40, B80 = 0x1p80, B120 = 0x1p120;
    int er[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) er[j] = (r0 + j < a.rows) ? a.er[r0 + j] : 0;
    for (int64_t c = blockIdx.y; c < a.cols; c += gridDim.y) {
        const uint8_t *cp = a.Cr + c * a.ldc + r0;
        uint32_t cw[TT];
#pragma unroll
        for (int t = 0; t < TT; ++t)
            cw[t] = (TT != OZ_TMAX || t < C.T) ? *reinterpret_cast<const uint32_t *>(cp + t * a.cplane) : 0u;
        const int ecc = a.ec[c];
        double sv[4];
#pragma unroll
        for (int j = 0; j < 4; ++j) sv[j] = (r0 + j < a.rows) ? a.S[r0 + j + c * a.lds] : 0.0;
        // V = sum_t c_t w_t, limb by limb, for the four rows
        double v[4][4];
#pragma unroll
        for (int j = 0; j < 4; ++j)
#pragma unroll
            for (int i = 0; i < 4; ++i) v[j][i] = 0.0;
        // the limbs of w_{t+1} are read while w_t's products run (the
        // scheduler is held to that order, and the table's address is made
        // opaque so the reads are not hoisted out of the column loop:
        // holding every limb takes 136 registers)
        const NLA_LDS double(*wt)[4] = (const NLA_LDS double(*)[4])wl;
        asm volatile("" : "+v"(wt));
        double wn[4];
#pragma unroll
        for (int i = 0; i < 4; ++i) wn[i] = wt[0][i];
#pragma unroll
        for (int t = 0; t < TT; ++t) {
            double w[4];
#pragma unroll
            for (int i = 0; i < 4; ++i) w[i] = wn[i];
            if (t + 1 < TT) {
#pragma unroll
                for (int i = 0; i < 4; ++i) wn[i] = wt[t + 1][i];
            }
            __builtin_amdgcn_sched_barrier(0);
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const double ct = (double)((cw[t] >> (8 * j)) & 255u);
#pragma unroll
                for (int i = 0; i < 4; ++i) v[j][i] = fma(ct, w[i], v[j][i]);
            }
            __builtin_amdgcn_sched_barrier(0);
        }
        // the sums are formed here, not inside each row's store branch
#pragma unroll
        for (int j = 0; j < 4; ++j)
#pragma unroll
            for (int i = 0; i < 4; ++i) asm volatile("" : "+v"(v[j][i]));
        // every row is computed (no branch), rows past `rows` not stored
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int sc = er[j] + ecc;
            const double sint = rint(ldexp(sv[j], sc));
            const bool bad = !(fabs(sint) <= C.smax);
            // |S_int| as limbs (each remainder has at most 53 significant bits)
            const double sa = bad ? 0.0 : fabs(sint);
            const double s3 = floor(sa * 0x1p-120);
            double rem = fma(-s3, B120, sa);
            const double s2 = floor(rem * 0x1p-80);
            rem = fma(-s2, B80, rem);
            const double s1 = floor(rem * 0x1p-40);
            const double s0 = fma(-s1, B40, rem);
            const double sg = sint < 0.0 ? -1.0 : 1.0;
            // D = V - S_int
            double d0 = fma(-sg, s0, v[j][0]), d1 = fma(-sg, s1, v[j][1]);
            double d2 = fma(-sg, s2, v[j][2]), d3 = fma(-sg, s3, v[j][3]);
            const double k = rint((fma(d3, B120, d2 * B80) + fma(d1, B40, d0)) * C.invm);
            // E = D - k M = -R_int, then carries
            d0 = fma(-k, C.ml[0], d0);
            d1 = fma(-k, C.ml[1], d1);
            d2 = fma(-k, C.ml[2], d2);
            d3 = fma(-k, C.ml[3], d3);
            double cy = floor(d0 * 0x1p-40);
            d0 = fma(-cy, B40, d0);
            d1 += cy;
            cy = floor(d1 * 0x1p-40);
            d1 = fma(-cy, B40, d1);
            d2 += cy;
            cy = floor(d2 * 0x1p-40);
            d2 = fma(-cy, B40, d2);
            d3 += cy;
            const double e = fma(fma(fma(d3, B40, d2), B40, d1), B40, d0);
            if (r0 + j < a.rows) {
                a.R[r0 + j + c * a.ldr] = bad ? 0.0 : ldexp(0.0 - e, -sc);
                if (bad) *a.flag = 1;
            }
        }
    }
}

// ---- host side --------------------------------------------------------------

static int64_t oz_round256(int64_t x) { return (x + 255) / 256 * 256; }

static int64_t modinv(int64_t a, int64_t p) {
    int64_t t = 0, nt = 1, r = p, nr = ((a % p) + p) % p;
    while (nr != 0) {
        const int64_t q = r / nr;
        int64_t tmp = t - q * nt;
        t = nt;
        nt = tmp;
        tmp = r - q * nr;
        r = nr;
        nr = tmp;
    }
    return (t % p + p) % p;
}

OzPlan oz_plan(int64_t n, int64_t m) {
    OzPlan P{};
    if (n <= 0 || m <= 0 || n > OZ_N_MAX) return P;
    // bits of the modulus product needed: bA + bX + ceil(log2 n) + 3
    // (|sum A_int X_int| <= n 2^(bA+bX) <= M/8 and |S_int| <= M/8 keep |R_int|
    // <= M/4); take the fewest moduli that leave bA, bX >= OZ_BITS_MIN
    int lg = 0;
    while (((int64_t)1 << lg) < n) ++lg;
    double logm = 0.0;
    int T = 0;
    for (; T < OZ_TMAX; ++T) {
        logm += std::log2((double)OZ_MODULI[T]);
        if ((int)std::floor(logm) - 3 - lg >= 2 * OZ_BITS_MIN) {
            ++T;
            break;
        }
    }
    if (T > OZ_TMAX) T = OZ_TMAX;
    int budget = (int)std::floor(logm) - 3 - lg;
    if (budget > 2 * 62) budget = 2 * 62;
    P.T = T;
    P.bA = (budget + 1) / 2;
    P.bX = budget - P.bA;
    P.npad = oz_round256(n);
    P.mpad = oz_round256(m);
    const int64_t nt = P.npad / 256;
    P.aplane = (int64_t)65536 * (nt * (nt + 1) / 2);
    P.xplane = P.mpad * P.npad;
    P.cplane = P.npad * P.mpad;
    OzConst &C = P.C;
    C.T = T;
    double M = 1.0;
    for (int t = 0; t < T; ++t) M *= (double)OZ_MODULI[t];
    C.M = M;
    C.smax = std::ldexp(1.0, (int)std::floor(logm) - 3);
    for (int t = 0; t < T; ++t) {
        const int64_t p = OZ_MODULI[t];
        C.p[t] = (int)p;
        C.invpd[t] = 1.0 / (double)p;
        int64_t mt = 1;  // (M / p_t) mod p_t
        for (int s = 0; s < T; ++s)
            if (s != t) mt = (mt * (OZ_MODULI[s] % p)) % p;
        C.y[t] = (int)modinv(mt, p);
        // centred residues of 2^16, 2^32, 2^48 and the centred range's top
        C.rhi[t] = (int)((p - 1) / 2);
        int64_t w = 1;
        for (int i = 0; i < 3; ++i) {
            w = (w * 65536) % p;
            C.rc[t][i] = (int)(w > C.rhi[t] ? w - p : w);
        }
    }
    // w_t = y_t prod_{s != t} p_s and M as base-2^40 limbs (32-bit words
    // on the host: M < 256^20 = 2^160)
    constexpr int NWD = 6;
    auto limbs40 = [](const uint32_t *x, double *out) {
        for (int i = 0; i < 4; ++i) {
            uint64_t v = 0;
            for (int b = 0; b < 40; ++b) {
                const int bit = 40 * i + b;
                v |= (uint64_t)((x[bit / 32] >> (bit % 32)) & 1u) << b;
            }
            out[i] = (double)v;
        }
    };
    auto mul_small = [](uint32_t *x, uint32_t f) {
        uint64_t carry = 0;
        for (int i = 0; i < NWD; ++i) {
            const uint64_t v = (uint64_t)x[i] * f + carry;
            x[i] = (uint32_t)v;
            carry = v >> 32;
        }
    };
    uint32_t mw[NWD] = {1, 0, 0, 0, 0, 0};
    for (int t = 0; t < T; ++t) mul_small(mw, (uint32_t)OZ_MODULI[t]);
    limbs40(mw, C.ml);
    for (int t = 0; t < T; ++t) {
        uint32_t w[NWD] = {(uint32_t)C.y[t], 0, 0, 0, 0, 0};
        for (int s = 0; s < T; ++s)
            if (s != t) mul_small(w, (uint32_t)OZ_MODULI[s]);
        limbs40(w, C.w[t]);
    }
    C.invm = 1.0 / M;
    return P;
}

OzLayout oz_layout(const OzPlan &P, int64_t base) {
    OzLayout L{};
    auto al = [](int64_t x) { return (x + 255) & ~(int64_t)255; };
    int64_t o = al(base);
    L.a = o;     o = al(o + P.T * P.aplane);
    L.x = o;     o = al(o + P.T * P.xplane);
    L.c = o;     o = al(o + P.T * P.cplane);
    L.ea = o;    o = al(o + P.npad * 4);
    L.fx = o;    o = al(o + P.mpad * 4);
    L.mxa = o;   o = al(o + P.npad * 8);
    L.mxx = o;   o = al(o + P.mpad * 8);
    L.flag = o;  o = al(o + 4);
    L.queue = o; o = al(o + 8 * 4);
    L.total = o;
    return L;
}

static hipError_t oz_fill(void *p, int64_t words, hipStream_t s) {
    return launch_fill_u32((unsigned *)p, words, 0u, s);
}

static hipError_t oz_scales(const OzSrc &src, unsigned long long *mx, int bits, int *e, int *flag, hipStream_t s) {
    hipError_t err = oz_fill(mx, 2 * src.no, s);
    if (err != hipSuccess) return err;
    if (src.sk == 1 && src.tri == 0) {
        // chunks of at least 1024 k, about 16k work-groups in all
        const int64_t wg = (src.no + 3) / 4;
        int64_t nch = (16384 + wg - 1) / wg;
        const int64_t most = (src.nk + 1023) / 1024;
        nch = nch < 1 ? 1 : (nch > most ? most : nch);
        const int64_t chunk = (src.nk + nch - 1) / nch;
        hipLaunchKernelGGL(oz_absmax_rows_kernel, dim3((unsigned)wg, (unsigned)nch), dim3(256), 0, s, src, chunk,
                           mx);
    } else {
        dim3 g((unsigned)((src.nk + 63) / 64), (unsigned)((src.no + 63) / 64));
        hipLaunchKernelGGL(oz_absmax_kernel, g, dim3(256), 0, s, src, mx);
    }
    err = hipGetLastError();
    if (err != hipSuccess) return err;
    hipLaunchKernelGGL(oz_exponent_kernel, dim3((unsigned)((src.no + 255) / 256)), dim3(256), 0, s, src.no, mx, bits,
                       e, flag);
    return hipGetLastError();
}

hipError_t oz_prepare(const OzPlan &P, const OzLayout &L, char *ws, const double *A, int64_t ars, int64_t acs,
                      bool lower, int64_t n, hipStream_t s) {
    hipError_t err = oz_fill(ws + L.flag, 1, s);
    if (err != hipSuccess) return err;
    OzSrc src{const_cast<double *>(A), ars, acs, n, n, lower ? 1 : 2};
    int *ea = (int *)(ws + L.ea);
    err = oz_scales(src, (unsigned long long *)(ws + L.mxa), P.bA, ea, (int *)(ws + L.flag), s);
    if (err != hipSuccess) return err;
    OzDst d{(int8_t *)(ws + L.a), P.aplane, P.npad, lower ? 1 : 2};
    dim3 g((unsigned)(P.npad / 64), (unsigned)(P.npad / 64));
    hipLaunchKernelGGL(oz_residue_kernel, g, dim3(256), 0, s, src, (const int *)ea, d, 0, P.C);
    return hipGetLastError();
}

// Schedules of the modular-product kernel (nla_set_config(NLA_CFG_OZ_VARIANT)):
// 0 4 x 8 super-tiles per XCD, 4-stage ring; 1 row order, 4 stages; 2 row
// order, 3 stages; 3 8 x 4 super-tiles, 4 stages.
template <typename K>
static hipError_t oz_launch_persistent(K kern, const OzGemmArgs &g, const OzConst &C, int items, hipStream_t s,
                                       int threads = 512) {  // grid: CUs x occupancy, at most the item count
    int o = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&o, (const void *)kern, threads, 0) != hipSuccess || o < 1) o = 1;
    int dev = 0, cus = 0;
    hipError_t err = hipGetDevice(&dev);
    if (err == hipSuccess) err = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    if (err != hipSuccess) return err;
    const int grid = cus * o < items ? cus * o : items;
    OzGemmArgs h = g;
    if (grid < 8) h.sr = h.sc = 0;  // per-XCD counters need a work-group on each of the 8
    hipLaunchKernelGGL(kern, dim3(grid), dim3(threads), 0, s, h, C);
    return hipGetLastError();
}

template <bool OUT_T>
static hipError_t oz_gemm_variant(OzGemmArgs g, const OzConst &C, int items, hipStream_t s) {
    const int v = oz_variant();
    g.sr = 0;
    g.sc = 0;
    if (v == 0) {
        g.sr = 4;
        g.sc = 8;
    } else if (v == 3) {
        g.sr = 8;
        g.sc = 4;
    }
    if (v == 2) return oz_launch_persistent(oz_gemm_kernel<OUT_T, 3>, g, C, items, s);
    return oz_launch_persistent(oz_gemm_kernel<OUT_T, 4>, g, C, items, s);
}

hipError_t oz_convert_x(const OzPlan &P, const OzLayout &L, char *ws, double *X, int64_t xso, int64_t xsk,
                        int64_t n, int64_t m, hipStream_t s) {
    // column scales (a system's largest entry), residues, rounded iterate
    OzSrc xs{X, xso, xsk, m, n, 0};
    int *fx = (int *)(ws + L.fx);
    hipError_t err = oz_scales(xs, (unsigned long long *)(ws + L.mxx), P.bX, fx, (int *)(ws + L.flag), s);
    if (err != hipSuccess) return err;
    OzDst d{(int8_t *)(ws + L.x), P.xplane, P.npad, 0};
    hipLaunchKernelGGL(oz_residue_kernel, dim3((unsigned)(P.npad / 64), (unsigned)(P.mpad / 64)), dim3(256), 0, s,
                       xs, (const int *)fx, d, 1, P.C);
    return hipGetLastError();
}

hipError_t oz_products(const OzPlan &P, const OzLayout &L, char *ws, bool lower, bool out_t, hipStream_t s) {
    hipError_t err = oz_fill(ws + L.queue, 8, s);
    if (err != hipSuccess) return err;
    OzGemmArgs g{};
    g.A = (const int8_t *)(ws + L.a);
    g.aplane = P.aplane;
    g.X = (const int8_t *)(ws + L.x);
    g.xplane = P.xplane;
    g.Cr = (uint8_t *)(ws + L.c);
    g.cplane = P.cplane;
    g.ldc = out_t ? P.mpad : P.npad;
    g.npad = P.npad;
    g.mpad = P.mpad;
    g.tiles_m = (int)(P.npad / 256);
    g.tiles_n = (int)(P.mpad / 256);
    g.T = P.T;
    g.lower = lower ? 1 : 0;
    g.queue = (int *)(ws + L.queue);
    const int items = g.tiles_m * g.tiles_n * g.T;
    return out_t ? oz_gemm_variant<true>(g, P.C, items, s) : oz_gemm_variant<false>(g, P.C, items, s);
}

hipError_t oz_reconstruct(const OzPlan &P, const OzLayout &L, char *ws, bool out_t, int64_t n, int64_t m,
                          const double *S, int64_t lds, double *R, int64_t ldr, hipStream_t s) {
    // R is rows x cols column-major (side L: n x m with rows scaled by A's
    // exponents; side R: m x n with rows scaled by X's)
    const int *fx = (const int *)(ws + L.fx), *ea = (const int *)(ws + L.ea);
    OzCrtArgs c{};
    c.rows = out_t ? m : n;
    c.cols = out_t ? n : m;
    c.S = S;
    c.lds = lds;
    c.R = R;
    c.ldr = ldr;
    c.Cr = (const uint8_t *)(ws + L.c);
    c.cplane = P.cplane;
    c.ldc = out_t ? P.mpad : P.npad;
    c.er = out_t ? fx : ea;
    c.ec = out_t ? ea : fx;
    c.flag = (int *)(ws + L.flag);
    const unsigned gy = (unsigned)(c.cols < 65535 ? c.cols : 65535);
    const dim3 grid((unsigned)((c.rows + 1023) / 1024), gy);
    switch (P.T) {
        case 15: hipLaunchKernelGGL(oz_crt_kernel<15>, grid, dim3(256), 0, s, c, P.C); break;
        case 16: hipLaunchKernelGGL(oz_crt_kernel<16>, grid, dim3(256), 0, s, c, P.C); break;
        case 17: hipLaunchKernelGGL(oz_crt_kernel<17>, grid, dim3(256), 0, s, c, P.C); break;
        default: hipLaunchKernelGGL(oz_crt_kernel<OZ_TMAX>, grid, dim3(256), 0, s, c, P.C); break;
    }
    return hipGetLastError();
}

}  // namespace nla
