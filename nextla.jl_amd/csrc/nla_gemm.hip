// GEMM update of the recursive triangular solve / multiply on gfx950 MFMA.
//
//   C = beta*C + alpha*op(X)*op(Y)      op(X): M x K, op(Y): K x N, C: M x N
//
// Replaces matmul! (src/matmul.jl:5-66) as used by GEMM_ADD!/GEMM_SUB!
// (src/matmul.jl:69-80) from unified_rec (src/rectrxm.jl:164-192).  The
// reference kernel computes one output per work-item from 32x32 LDS tiles;
// here every wave owns a (BM/WM) x (BN/WN) block of C held in MFMA
// accumulators, operands are staged through double-buffered LDS, and the
// alpha pre/post scaling of unified_rectrxm! (src/rectrxm.jl:64,72) is folded
// into beta/alpha so B is never swept separately.
//
// Orientation: the MFMA computes D = C^T for each 16x16 (f64) / 32x32 (f32)
// block, so that the lane index of D's column walks down a column of C and the
// epilogue writes whole 128-byte segments of the column-major output.
//
// LDS layout per operand follows its contiguous dimension in HBM:
//   outer-contiguous (X 'N', Y 'T'):   tile[k][outer]   (row stride BO + PO)
//   k-contiguous     (X 'T', Y 'N'):   tile[outer][k]   (row stride BK + PK)
// so HBM reads are 16-byte vectors along the contiguous dimension and LDS
// writes are vectors too; the transpose happens in the per-lane MFMA operand
// read instead.
#include "nla_device.h"
#include "nla_internal.h"

namespace nla {

// Pointers into LDS keep their address space so that every access compiles
// to ds_read / ds_write (a generic pointer becomes a flat access, which also
// waits on outstanding global loads).
#define NLA_LDS __attribute__((address_space(3)))

template <typename T, int BO, int BK, int NT, bool KCONTIG, int VW>
struct TileLoader {
    // Each thread moves CH chunks of VW consecutive elements along the
    // contiguous dimension of the operand.
    static constexpr int CHUNKS = BO * BK / VW;
    static constexpr int CH = CHUNKS / NT;
    static_assert(CHUNKS % NT == 0, "tile does not divide among threads");
    static constexpr int PO = (sizeof(T) == 8) ? 2 : 4;  // pads keep 16-byte alignment
    static constexpr int PK = (sizeof(T) == 8) ? 2 : 4;
    static constexpr int LDS_ELEMS = KCONTIG ? BO * (BK + PK) : BK * (BO + PO);

    T r[CH][VW];
    uint32_t okbits;  // CHECK loads: bit (c*VW+e) set when element is in bounds

    // Coordinates (outer, k) inside the tile of chunk c of thread t.
    __device__ static inline void coords(int t, int c, int &o, int &k) {
        int q = t + c * NT;
        if (KCONTIG) {
            k = (q % (BK / VW)) * VW;
            o = q / (BK / VW);
        } else {
            o = (q % (BO / VW)) * VW;
            k = q / (BO / VW);
        }
    }

    // Element (outer, k) of the operand lives at P[outer*so + k*sk] with one
    // of so/sk equal to 1.  o0/k0: tile origin; no/nk: extents for bounds.
    // CHECK = false: the whole tile is in bounds, plain vector loads.
    // CHECK = true: element-wise loads without branches: an out-of-bounds
    // element is read from P itself (always valid) and replaced by zero, so
    // the loads of a k-step stay in flight together (a branch per element
    // would make the compiler wait for every earlier load at each join).
    template <bool CHECK>
    __device__ inline void load(const T *__restrict__ P, int64_t so, int64_t sk, int64_t o0, int64_t k0,
                                int64_t no, int64_t nk, int t) {
#pragma unroll
        for (int c = 0; c < CH; ++c) {
            int o, k;
            coords(t, c, o, k);
            int64_t go = o0 + o, gk = k0 + k;
            if (!CHECK) {
                const T *p = P + go * so + gk * sk;
                if (VW > 1) {
                    typedef T vec_t __attribute__((ext_vector_type(VW)));
                    vec_t v = *reinterpret_cast<const vec_t *>(p);
#pragma unroll
                    for (int e = 0; e < VW; ++e) r[c][e] = v[e];
                } else {
                    r[c][0] = *p;
                }
            } else {
                if (c == 0) okbits = 0;
                // the chunk's elements are consecutive in memory (the stride
                // along the chunk is 1): one address per chunk, not a 64-bit
                // multiply per element
                const T *pc = P + go * so + gk * sk;
#pragma unroll
                for (int e = 0; e < VW; ++e) {
                    const bool ok = (KCONTIG ? go : go + e) < no && (KCONTIG ? gk + e : gk) < nk;
                    r[c][e] = *(ok ? pc + e : P);
                    okbits |= (ok ? 1u : 0u) << (c * VW + e);
                }
            }
        }
    }

    // Zero the out-of-bounds elements of a CHECK load.  Done just before the
    // LDS store, so no instruction needs the loaded values any earlier.
    __device__ inline void apply_bounds() {
#pragma unroll
        for (int c = 0; c < CH; ++c)
#pragma unroll
            for (int e = 0; e < VW; ++e)
                if (!((okbits >> (c * VW + e)) & 1u)) r[c][e] = T(0);
    }

    // Zero the elements outside a triangle: keep (outer, k) when
    // (KEEP_LE ? k <= outer : k >= outer), in global coordinates.
    template <bool KEEP_LE>
    __device__ inline void mask(int64_t o0, int64_t k0, int t) {
#pragma unroll
        for (int c = 0; c < CH; ++c) {
            int o, k;
            coords(t, c, o, k);
#pragma unroll
            for (int e = 0; e < VW; ++e) {
                int64_t eo = o0 + (KCONTIG ? o : o + e);
                int64_t ek = k0 + (KCONTIG ? k + e : k);
                bool keep = KEEP_LE ? (ek <= eo) : (ek >= eo);
                if (!keep) r[c][e] = T(0);
            }
        }
    }

    __device__ inline void store(NLA_LDS T *lds, int t) const {
#pragma unroll
        for (int c = 0; c < CH; ++c) {
            int o, k;
            coords(t, c, o, k);
            NLA_LDS T *q = KCONTIG ? lds + o * (BK + PK) + k : lds + k * (BO + PO) + o;
            if (VW > 1) {
                typedef T vec_t __attribute__((ext_vector_type(VW)));
                vec_t v;
#pragma unroll
                for (int e = 0; e < VW; ++e) v[e] = r[c][e];
                *reinterpret_cast<NLA_LDS vec_t *>(q) = v;
            } else {
                q[0] = r[c][0];
            }
        }
    }

    // Operand element (outer, k) from the staged tile.
    __device__ static inline T at(const NLA_LDS T *lds, int o, int k) {
        return KCONTIG ? lds[o * (BK + PK) + k] : lds[k * (BO + PO) + o];
    }
};

// Hardware dispatches work-group g to XCD g % 8.  Renumber so each XCD owns
// a contiguous range of logical tiles, then walk tiles in groups of GM row
// tiles so the tiles resident on one XCD share X row panels and Y column
// panels in that XCD's L2.
__device__ inline void tile_coords(int G, int g, int tiles_m, int &tm, int &tn) {
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

// One BM x BN tile of C.  CHECK selects bounds-checked loads and stores
// (edge tiles, or K not a multiple of BK) or the unchecked vector path.
template <typename T, int BM, int BN, int BK, int WM, int WN, bool TX, bool TY, int VW, int TRI, bool CHECK>
__device__ __forceinline__ void gemm_tile(int64_t M, int64_t N, int64_t K, T alpha, const T *__restrict__ X,
                                          int64_t ldx, const T *__restrict__ Y, int64_t ldy, T beta, T *C,
                                          int64_t ldc, int64_t row0, int64_t col0, NLA_LDS T *lds_xb,
                                          NLA_LDS T *lds_yb) {
    typedef Mfma<T> MF;
    constexpr int NT = WM * WN * 64;
    constexpr int TWM = BM / WM, TWN = BN / WN;
    constexpr int RB = TWM / MF::BD, CB = TWN / MF::BD;
    static_assert(RB >= 1 && CB >= 1 && TWM % MF::BD == 0 && TWN % MF::BD == 0, "bad wave tile");
    static_assert(BK % MF::KD == 0, "bad BK");
    // X tile: outer = rows of C, contiguous along k when TX.
    typedef TileLoader<T, BM, BK, NT, TX, CHECK ? 1 : VW> LX;
    // Y tile: outer = columns of C, contiguous along k when !TY.
    typedef TileLoader<T, BN, BK, NT, !TY, CHECK ? 1 : VW> LY;
    NLA_LDS T *lds_x[2] = {lds_xb, lds_xb + LX::LDS_ELEMS};
    NLA_LDS T *lds_y[2] = {lds_yb, lds_yb + LY::LDS_ELEMS};

    const int t = threadIdx.x;
    const int lane = t & 63, wave = t >> 6;
    const int wm = wave % WM, wn = wave / WM;

    // op(X)(r, k): TX ? X[k + r*ldx] : X[r + k*ldx]
    const int64_t xso = TX ? ldx : 1, xsk = TX ? 1 : ldx;
    // op(Y)(k, c): TY ? Y[c + k*ldy] : Y[k + c*ldy]
    const int64_t yso = TY ? 1 : ldy, ysk = TY ? ldy : 1;

    typename MF::acc_t acc[CB][RB];
#pragma unroll
    for (int cb = 0; cb < CB; ++cb)
#pragma unroll
        for (int rb = 0; rb < RB; ++rb) acc[cb][rb] = MF::zero();

    LX lx;
    LY ly;
    // op(X)(r,k): outer r, inner k.  op(Y)(k,c): outer c, inner k; the Y
    // triangle conditions are c <= k (lower) / c >= k (upper).
    auto masks = [&](int64_t k0) {
        if (CHECK) {
            lx.apply_bounds();
            ly.apply_bounds();
        }
        if (TRI == 1) lx.template mask<true>(row0, k0, t);
        if (TRI == 2) lx.template mask<false>(row0, k0, t);
        if (TRI == 3) ly.template mask<false>(col0, k0, t);
        if (TRI == 4) ly.template mask<true>(col0, k0, t);
    };
    // k-steps that the triangle zeroes entirely are skipped: the operand's
    // rows (columns) of this tile only reach k up to (from) the tile edge.
    // Tile origins are multiples of BM/BN, which are multiples of BK.
    int64_t kbeg = 0, kend = K;
    if (TRI == 1) kend = row0 + BM < K ? row0 + BM : K;
    if (TRI == 2) kbeg = row0 < K ? row0 : K;
    if (TRI == 3) kbeg = col0 < K ? col0 : K;
    if (TRI == 4) kend = col0 + BN < K ? col0 + BN : K;
    const int nkt = (int)((kend - kbeg + BK - 1) / BK);
    if (nkt > 0) {
        lx.template load<CHECK>(X, xso, xsk, row0, kbeg, M, K, t);
        ly.template load<CHECK>(Y, yso, ysk, col0, kbeg, N, K, t);
        masks(kbeg);
        lx.store(lds_x[0], t);
        ly.store(lds_y[0], t);
    }
    __syncthreads();

    const int lo = lane % MF::BD, lk = lane / MF::BD;
    for (int kt = 0; kt < nkt; ++kt) {
        const int cur = kt & 1;
        const bool more = kt + 1 < nkt;
        const int64_t knext = kbeg + (int64_t)(kt + 1) * BK;
        if (more) {
            lx.template load<CHECK>(X, xso, xsk, row0, knext, M, K, t);
            ly.template load<CHECK>(Y, yso, ysk, col0, knext, N, K, t);
        }
        const NLA_LDS T *sx = lds_x[cur];
        const NLA_LDS T *sy = lds_y[cur];
        __builtin_amdgcn_s_setprio(1);  // issue priority to waves in their MFMA cluster
#pragma unroll
        for (int ks = 0; ks < BK; ks += MF::KD) {
            T a[CB], b[RB];
#pragma unroll
            for (int cb = 0; cb < CB; ++cb) a[cb] = LY::at(sy, wn * TWN + cb * MF::BD + lo, ks + lk);
#pragma unroll
            for (int rb = 0; rb < RB; ++rb) b[rb] = LX::at(sx, wm * TWM + rb * MF::BD + lo, ks + lk);
#pragma unroll
            for (int cb = 0; cb < CB; ++cb)
#pragma unroll
                for (int rb = 0; rb < RB; ++rb) acc[cb][rb] = MF::mma(a[cb], b[rb], acc[cb][rb]);
        }
        __builtin_amdgcn_s_setprio(0);
        if (more) {
            masks(knext);
            lx.store(lds_x[cur ^ 1], t);
            ly.store(lds_y[cur ^ 1], t);
        }
        __syncthreads();
    }

    // Epilogue: D[i][j] of block (cb, rb) is C(row j, column i).  beta is
    // uniform, so the choice is made once; with beta != 0 the C values of one
    // column block are all requested before any is used.
    if (beta == T(0)) {
#pragma unroll
        for (int cb = 0; cb < CB; ++cb)
#pragma unroll
            for (int rb = 0; rb < RB; ++rb)
#pragma unroll
                for (int r = 0; r < MF::NR; ++r) {
                    int64_t row = row0 + wm * TWM + rb * MF::BD + MF::dj(lane, r);
                    int64_t col = col0 + wn * TWN + cb * MF::BD + MF::di(lane, r);
                    if (!CHECK || (row < M && col < N)) C[row + col * ldc] = alpha * acc[cb][rb][r];
                }
    } else {
#pragma unroll
        for (int cb = 0; cb < CB; ++cb) {
            T cv[RB][MF::NR];
#pragma unroll
            for (int rb = 0; rb < RB; ++rb)
#pragma unroll
                for (int r = 0; r < MF::NR; ++r) {
                    int64_t row = row0 + wm * TWM + rb * MF::BD + MF::dj(lane, r);
                    int64_t col = col0 + wn * TWN + cb * MF::BD + MF::di(lane, r);
                    bool ok = !CHECK || (row < M && col < N);
                    cv[rb][r] = *(ok ? C + row + col * ldc : C);
                }
#pragma unroll
            for (int rb = 0; rb < RB; ++rb)
#pragma unroll
                for (int r = 0; r < MF::NR; ++r) {
                    int64_t row = row0 + wm * TWM + rb * MF::BD + MF::dj(lane, r);
                    int64_t col = col0 + wn * TWN + cb * MF::BD + MF::di(lane, r);
                    if (!CHECK || (row < M && col < N))
                        C[row + col * ldc] = alpha * acc[cb][rb][r] + beta * cv[rb][r];
                }
        }
    }
}

// TRI: triangle mask on one operand (see GemmArgs::tri).  C is not declared
// __restrict__: a leaf update writes the strip it reads.
template <typename T, int BM, int BN, int BK, int WM, int WN, bool TX, bool TY, int VW, int TRI>
__global__ __launch_bounds__(WM *WN * 64) void gemm_kernel(int64_t M, int64_t N, int64_t K, T alpha,
                                                           const T *__restrict__ X, int64_t ldx,
                                                           const T *__restrict__ Y, int64_t ldy, T beta,
                                                           T *C, int64_t ldc, int tiles_m) {
    constexpr int NT = WM * WN * 64;
    typedef TileLoader<T, BM, BK, NT, TX, VW> LX;
    typedef TileLoader<T, BN, BK, NT, !TY, VW> LY;
    __shared__ __attribute__((aligned(16))) T lds_x[2 * LX::LDS_ELEMS];
    __shared__ __attribute__((aligned(16))) T lds_y[2 * LY::LDS_ELEMS];
    int tm, tn;
    tile_coords(gridDim.x, blockIdx.x, tiles_m, tm, tn);
    const int64_t row0 = (int64_t)tm * BM, col0 = (int64_t)tn * BN;
    // interior tiles with whole k-steps take the unchecked vector path
    const bool interior = (row0 + BM <= M) && (col0 + BN <= N) && (K % BK == 0);
    if (interior)
        gemm_tile<T, BM, BN, BK, WM, WN, TX, TY, VW, TRI, false>(M, N, K, alpha, X, ldx, Y, ldy, beta, C, ldc,
                                                                 row0, col0, (NLA_LDS T *)lds_x,
                                                                 (NLA_LDS T *)lds_y);
    else
        gemm_tile<T, BM, BN, BK, WM, WN, TX, TY, VW, TRI, true>(M, N, K, alpha, X, ldx, Y, ldy, beta, C, ldc,
                                                                row0, col0, (NLA_LDS T *)lds_x,
                                                                (NLA_LDS T *)lds_y);
}

// A triangular operand gives the tiles k-ranges from one k-step to all of K.
// With one work-group per tile, the work-groups that draw the long tiles set
// the runtime (the hardware hands a CU its next work-group in a fixed
// rotation, so a CU keeps meeting tiles of the same rows).  Here a grid of
// resident work-groups takes tiles from a counter instead, longest first;
// all columns (rows) of a row (column) of tiles are taken together, so the
// work-groups running at one time share that operand panel in L2.  *queue
// must be zero at launch; every work-group leaves once the counter passes
// the tile count.
template <typename T, int BM, int BN, int BK, int WM, int WN, bool TX, bool TY, int VW, int TRI>
__global__ __launch_bounds__(WM *WN * 64) void gemm_tri_queue_kernel(int64_t M, int64_t N, int64_t K, T alpha,
                                                                     const T *__restrict__ X, int64_t ldx,
                                                                     const T *__restrict__ Y, int64_t ldy, T beta,
                                                                     T *C, int64_t ldc, int tiles_m, int tiles_n,
                                                                     int *queue) {
    static_assert(TRI >= 1 && TRI <= 4, "triangular operand expected");
    constexpr int NT = WM * WN * 64;
    typedef TileLoader<T, BM, BK, NT, TX, VW> LX;
    typedef TileLoader<T, BN, BK, NT, !TY, VW> LY;
    __shared__ __attribute__((aligned(16))) T lds_x[2 * LX::LDS_ELEMS];
    __shared__ __attribute__((aligned(16))) T lds_y[2 * LY::LDS_ELEMS];
    __shared__ int s_tile;
    const int ntiles = tiles_m * tiles_n;
    for (;;) {
        if (threadIdx.x == 0) s_tile = atomicAdd(queue, 1);
        __syncthreads();
        const int t = s_tile;
        __syncthreads();  // read by all before thread 0 takes the next one
        if (t >= ntiles) break;
        int tm, tn;
        if (TRI == 1) {  // k-range grows with the row: last rows first
            tm = tiles_m - 1 - t / tiles_n;
            tn = t % tiles_n;
        } else if (TRI == 2) {
            tm = t / tiles_n;
            tn = t % tiles_n;
        } else if (TRI == 3) {
            tn = t / tiles_m;
            tm = t % tiles_m;
        } else {
            tn = tiles_n - 1 - t / tiles_m;
            tm = t % tiles_m;
        }
        const int64_t row0 = (int64_t)tm * BM, col0 = (int64_t)tn * BN;
        const bool interior = (row0 + BM <= M) && (col0 + BN <= N) && (K % BK == 0);
        if (interior)
            gemm_tile<T, BM, BN, BK, WM, WN, TX, TY, VW, TRI, false>(M, N, K, alpha, X, ldx, Y, ldy, beta, C, ldc,
                                                                     row0, col0, (NLA_LDS T *)lds_x,
                                                                     (NLA_LDS T *)lds_y);
        else
            gemm_tile<T, BM, BN, BK, WM, WN, TX, TY, VW, TRI, true>(M, N, K, alpha, X, ldx, Y, ldy, beta, C, ldc,
                                                                    row0, col0, (NLA_LDS T *)lds_x,
                                                                    (NLA_LDS T *)lds_y);
    }
}

template <typename T, int BM, int BN, int BK, int WM, int WN, bool TX, bool TY, int VW, int TRI>
static hipError_t launch_cfg(const GemmArgs<T> &a, hipStream_t s) {
    int tiles_m = (int)((a.M + BM - 1) / BM);
    int tiles_n = (int)((a.N + BN - 1) / BN);
    int G = tiles_m * tiles_n;
    if constexpr (TRI != 0) {
        if (a.queue != nullptr && a.leaf == 0) {
            constexpr int NT = WM * WN * 64;
            static const int occ = [] {
                int o = 0;
                if (hipOccupancyMaxActiveBlocksPerMultiprocessor(
                        &o, (const void *)gemm_tri_queue_kernel<T, BM, BN, BK, WM, WN, TX, TY, VW, TRI>, NT, 0) !=
                    hipSuccess)
                    o = 1;
                return o > 0 ? o : 1;
            }();
            int dev = 0, cus = 0;
            hipError_t e = hipGetDevice(&dev);
            if (e == hipSuccess) e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
            if (e != hipSuccess) return e;
            int grid = cus * occ < G ? cus * occ : G;
            hipLaunchKernelGGL((gemm_tri_queue_kernel<T, BM, BN, BK, WM, WN, TX, TY, VW, TRI>), dim3(grid), dim3(NT),
                               0, s, a.M, a.N, a.K, a.alpha, a.X, a.ldx, a.Y, a.ldy, a.beta, a.C, a.ldc, tiles_m,
                               tiles_n, a.queue);
            return hipGetLastError();
        }
    }
    hipLaunchKernelGGL((gemm_kernel<T, BM, BN, BK, WM, WN, TX, TY, VW, TRI>), dim3(G), dim3(WM * WN * 64), 0, s, a.M,
                       a.N, a.K, a.alpha, a.X, a.ldx, a.Y, a.ldy, a.beta, a.C, a.ldc, tiles_m);
    return hipGetLastError();
}

template <typename T, int BM, int BN, int BK, int WM, int WN, int VW, int TRI>
static hipError_t launch_trans(const GemmArgs<T> &a, hipStream_t s) {
    if (!a.tx && !a.ty) return launch_cfg<T, BM, BN, BK, WM, WN, false, false, VW, TRI>(a, s);
    if (a.tx && !a.ty) return launch_cfg<T, BM, BN, BK, WM, WN, true, false, VW, TRI>(a, s);
    if (!a.tx && a.ty) return launch_cfg<T, BM, BN, BK, WM, WN, false, true, VW, TRI>(a, s);
    return launch_cfg<T, BM, BN, BK, WM, WN, true, true, VW, TRI>(a, s);
}

template <typename T, int BM, int BN, int BK, int WM, int WN, int VW>
static hipError_t launch_tri(const GemmArgs<T> &a, hipStream_t s) {
    switch (a.tri) {
        case 1: return launch_trans<T, BM, BN, BK, WM, WN, VW, 1>(a, s);
        case 2: return launch_trans<T, BM, BN, BK, WM, WN, VW, 2>(a, s);
        case 3: return launch_trans<T, BM, BN, BK, WM, WN, VW, 3>(a, s);
        case 4: return launch_trans<T, BM, BN, BK, WM, WN, VW, 4>(a, s);
        default: return launch_trans<T, BM, BN, BK, WM, WN, VW, 0>(a, s);
    }
}

// A dense product with one triangular operand: op(X) (TRI 1/2, op(Y) dense,
// untransposed) or op(Y) (TRI 3/4, op(X) dense, untransposed) -- the shapes
// of the residual alpha*B - op(A)*X and alpha*B - X*op(A).
template <typename T, int BM, int BN, int BK, int WM, int WN, int VW>
static hipError_t launch_tri_dense(const GemmArgs<T> &a, hipStream_t s) {
    switch (a.tri) {
        case 1:
            if (a.ty) return hipErrorInvalidValue;
            return a.tx ? launch_cfg<T, BM, BN, BK, WM, WN, true, false, VW, 1>(a, s)
                        : launch_cfg<T, BM, BN, BK, WM, WN, false, false, VW, 1>(a, s);
        case 2:
            if (a.ty) return hipErrorInvalidValue;
            return a.tx ? launch_cfg<T, BM, BN, BK, WM, WN, true, false, VW, 2>(a, s)
                        : launch_cfg<T, BM, BN, BK, WM, WN, false, false, VW, 2>(a, s);
        case 3:
            if (a.tx) return hipErrorInvalidValue;
            return a.ty ? launch_cfg<T, BM, BN, BK, WM, WN, false, true, VW, 3>(a, s)
                        : launch_cfg<T, BM, BN, BK, WM, WN, false, false, VW, 3>(a, s);
        case 4:
            if (a.tx) return hipErrorInvalidValue;
            return a.ty ? launch_cfg<T, BM, BN, BK, WM, WN, false, true, VW, 4>(a, s)
                        : launch_cfg<T, BM, BN, BK, WM, WN, false, false, VW, 4>(a, s);
        default:
            return launch_trans<T, BM, BN, BK, WM, WN, VW, 0>(a, s);
    }
}

static inline bool aligned16(const void *p) { return ((uintptr_t)p & 15) == 0; }

template <typename T>
hipError_t launch_gemm(const GemmArgs<T> &a, hipStream_t s) {
    if (a.M <= 0 || a.N <= 0) return hipSuccess;
    constexpr int VW = 16 / sizeof(T);
    // 16-byte vector loads need every column start of X, Y (and the tile
    // origins, which are multiples of the tile sizes) to be 16-byte aligned.
    bool vec = aligned16(a.X) && aligned16(a.Y) && (a.ldx % VW == 0) && (a.ldy % VW == 0);
    if (a.leaf) {
        // one leaf (<= 64) along M (left) or N (right) and along K: a single
        // tile covers the leaf dimension, K is consumed in two steps.
        if (a.leaf == 1) {  // left: C (nb x m) = W (nb x nb) * B (nb x m)
            if (a.M > 64 || a.K > 64) return hipErrorInvalidValue;
            if constexpr (sizeof(T) == 8) {
                if (vec) return launch_tri<T, 64, 32, 32, 2, 2, VW>(a, s);
                return launch_tri<T, 64, 32, 32, 2, 2, 1>(a, s);
            }
            if (vec) return launch_tri<T, 64, 64, 32, 2, 2, VW>(a, s);
            return launch_tri<T, 64, 64, 32, 2, 2, 1>(a, s);
        }
        // right: C (m x nb) = B (m x nb) * W (nb x nb)
        if (a.N > 64 || a.K > 64) return hipErrorInvalidValue;
        if constexpr (sizeof(T) == 8) {
            if (vec) return launch_tri<T, 32, 64, 32, 2, 2, VW>(a, s);
            return launch_tri<T, 32, 64, 32, 2, 2, 1>(a, s);
        }
        if (vec) return launch_tri<T, 64, 64, 32, 2, 2, VW>(a, s);
        return launch_tri<T, 64, 64, 32, 2, 2, 1>(a, s);
    }
    // Tile shape by how many tiles the problem yields: the largest shape that
    // still gives at least one tile per CU (256) wins; 8- and 16-wave
    // work-groups keep two or more waves on every SIMD.
    auto tiles = [&](int bm, int bn) { return ((a.M + bm - 1) / bm) * ((a.N + bn - 1) / bn); };
    if constexpr (sizeof(T) == 8) {
        if (tiles(256, 128) >= 256) {
            if (vec) return launch_tri_dense<T, 256, 128, 16, 4, 2, VW>(a, s);
            return launch_tri_dense<T, 256, 128, 16, 4, 2, 1>(a, s);
        }
        if (tiles(128, 128) >= 256) {
            if (vec) return launch_tri_dense<T, 128, 128, 16, 4, 2, VW>(a, s);
            return launch_tri_dense<T, 128, 128, 16, 4, 2, 1>(a, s);
        }
        if (vec) return launch_tri_dense<T, 64, 128, 16, 2, 2, VW>(a, s);
        return launch_tri_dense<T, 64, 128, 16, 2, 2, 1>(a, s);
    } else {
        if (a.tri) return hipErrorInvalidValue;  // triangular products are FP64 only
        if (tiles(256, 256) >= 256) {
            if (vec) return launch_trans<T, 256, 256, 32, 4, 4, VW, 0>(a, s);
            return launch_trans<T, 256, 256, 32, 4, 4, 1, 0>(a, s);
        }
        if (tiles(256, 128) >= 256) {
            if (vec) return launch_trans<T, 256, 128, 16, 4, 2, VW, 0>(a, s);
            return launch_trans<T, 256, 128, 16, 4, 2, 1, 0>(a, s);
        }
        if (tiles(128, 128) >= 256) {
            if (vec) return launch_trans<T, 128, 128, 32, 2, 2, VW, 0>(a, s);
            return launch_trans<T, 128, 128, 32, 2, 2, 1, 0>(a, s);
        }
        if (vec) return launch_trans<T, 64, 64, 32, 2, 2, VW, 0>(a, s);
        return launch_trans<T, 64, 64, 32, 2, 2, 1, 0>(a, s);
    }
}

// ---- grouped GEMM of the staged batched solver ----------------------------

// Work-groups take (gemm, tile) tasks from a counter until it passes ntasks;
// each task is one tile of one GEMM of the group, run by the same tile
// routine as the single GEMM.
template <typename T, int BM, int BN, int BK, int WM, int WN, bool TX, bool TY, int VW>
__global__ __launch_bounds__(WM *WN * 64) void grouped_gemm_kernel(const GroupGemm<T> *__restrict__ g,
                                                                   const int2 *__restrict__ tasks, int ntasks,
                                                                   int *counter) {
    constexpr int NT = WM * WN * 64;
    typedef TileLoader<T, BM, BK, NT, TX, VW> LX;
    typedef TileLoader<T, BN, BK, NT, !TY, VW> LY;
    __shared__ __attribute__((aligned(16))) T lds_x[2 * LX::LDS_ELEMS];
    __shared__ __attribute__((aligned(16))) T lds_y[2 * LY::LDS_ELEMS];
    __shared__ int s_task;
    for (;;) {
        if (threadIdx.x == 0) s_task = atomicAdd(counter, 1);
        __syncthreads();
        const int t = s_task;
        __syncthreads();
        if (t >= ntasks) break;
        const int2 tk = tasks[t];
        const GroupGemm<T> G = g[tk.x];
        const int64_t row0 = (int64_t)(tk.y % G.tiles_m) * BM, col0 = (int64_t)(tk.y / G.tiles_m) * BN;
        const bool interior = (row0 + BM <= G.M) && (col0 + BN <= G.N) && (G.K % BK == 0);
        if (interior)
            gemm_tile<T, BM, BN, BK, WM, WN, TX, TY, VW, 0, false>(G.M, G.N, G.K, G.alpha, G.X, G.ldx, G.Y, G.ldy,
                                                                   G.beta, G.C, G.ldc, row0, col0,
                                                                   (NLA_LDS T *)lds_x, (NLA_LDS T *)lds_y);
        else
            gemm_tile<T, BM, BN, BK, WM, WN, TX, TY, VW, 0, true>(G.M, G.N, G.K, G.alpha, G.X, G.ldx, G.Y, G.ldy,
                                                                  G.beta, G.C, G.ldc, row0, col0,
                                                                  (NLA_LDS T *)lds_x, (NLA_LDS T *)lds_y);
    }
}

__global__ void zero_counter_kernel(int *p) {
    if (threadIdx.x == 0 && blockIdx.x == 0) *p = 0;
}

// FP64: 64 x 32 tiles for side 'L', 32 x 64 for side 'R', four waves of
// 32 x 16 (16 x 32) in 16 k per step; narrow GEMMs 64 x 16 (16 x 64), four
// waves of one 16 x 16 block in 32 k per step.  FP32: 64 x 64, four waves
// of one 32 x 32 block in 32 k per step.
template <typename T, bool TX, bool TY, int VW>
static const void *gg_ptr(bool rows_m, bool narrow) {
    if constexpr (sizeof(T) == 8) {
        if (narrow) {
            if (rows_m) return (const void *)grouped_gemm_kernel<T, 64, 16, 32, 4, 1, TX, TY, VW>;
            return (const void *)grouped_gemm_kernel<T, 16, 64, 32, 1, 4, TX, TY, VW>;
        }
        if (rows_m) return (const void *)grouped_gemm_kernel<T, 64, 32, 16, 2, 2, TX, TY, VW>;
        return (const void *)grouped_gemm_kernel<T, 32, 64, 16, 2, 2, TX, TY, VW>;
    } else {
        (void)rows_m;
        (void)narrow;
        return (const void *)grouped_gemm_kernel<T, 64, 64, 32, 2, 2, TX, TY, VW>;
    }
}

template <typename T>
const void *grouped_gemm_ptr(bool tx, bool ty, bool vec, bool rows_m, bool narrow) {
    constexpr int V = 16 / sizeof(T);
    if (vec) {
        if (!tx && !ty) return gg_ptr<T, false, false, V>(rows_m, narrow);
        if (tx && !ty) return gg_ptr<T, true, false, V>(rows_m, narrow);
        if (!tx && ty) return gg_ptr<T, false, true, V>(rows_m, narrow);
        return gg_ptr<T, true, true, V>(rows_m, narrow);
    }
    if (!tx && !ty) return gg_ptr<T, false, false, 1>(rows_m, narrow);
    if (tx && !ty) return gg_ptr<T, true, false, 1>(rows_m, narrow);
    if (!tx && ty) return gg_ptr<T, false, true, 1>(rows_m, narrow);
    return gg_ptr<T, true, true, 1>(rows_m, narrow);
}

template <typename T>
hipError_t launch_grouped_gemm(const GroupGemm<T> *g, const int2 *tasks, int ntasks, int *counter, bool tx, bool ty,
                               bool vec, bool rows_m, bool narrow, int groups, hipStream_t s) {
    if (ntasks <= 0) return hipSuccess;
    hipLaunchKernelGGL(zero_counter_kernel, dim3(1), dim3(64), 0, s, counter);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    void *args[] = {(void *)&g, (void *)&tasks, (void *)&ntasks, (void *)&counter};
    return hipLaunchKernel(grouped_gemm_ptr<T>(tx, ty, vec, rows_m, narrow), dim3(groups), dim3(256), args, 0, s);
}

template const void *grouped_gemm_ptr<float>(bool, bool, bool, bool, bool);
template const void *grouped_gemm_ptr<double>(bool, bool, bool, bool, bool);
template hipError_t launch_grouped_gemm<float>(const GroupGemm<float> *, const int2 *, int, int *, bool, bool, bool,
                                               bool, bool, int, hipStream_t);
template hipError_t launch_grouped_gemm<double>(const GroupGemm<double> *, const int2 *, int, int *, bool, bool,
                                                bool, bool, bool, int, hipStream_t);

template hipError_t launch_gemm<float>(const GemmArgs<float> &, hipStream_t);
template hipError_t launch_gemm<double>(const GemmArgs<double> &, hipStream_t);

}  // namespace nla
