// Recursive LU factorization with partial pivoting (nla_getrf2 in
// include/nla.h): getrf2! of the reference (src/lu.jl:185-300, LAPACK
// dgetrf2's recursion) on the device, built on this library's unit-lower
// TRSM and GEMM update.
//
// The host walks the same recursion as the reference: split the columns at
// n1 = min(m, n) / 2, factor the left block, apply its row interchanges to
// the right block, solve A12 = L11^-1 A12 (unit lower), update
// A22 -= A21 A12, factor A22 and apply its interchanges to the left block.
// Blocks of up to 8 columns are one kernel (the same eliminations with
// rank-1 updates in place of the GEMMs of that width), which keeps the
// launch count near 5 per 8 columns instead of 5 per column.  Such a leaf
// is held in registers: by one work-group up to 4096 rows, by one
// work-group per 2048 rows with a grid barrier per column above that.  Pivots are
// kept as global rows throughout (the reference shifts them by n1 after
// each recursive call).  Everything is enqueued on the caller's stream;
// pivots and info stay on the device.
#include <hip/hip_runtime.h>

#include <limits>
#include <mutex>
#include <vector>

#include "../../include/nla.h"
#include "nla_internal.h"

namespace nla {

int32_t report_hip(hipError_t e, const char *where);
int32_t report_bad_arg(int k, const char *what);

namespace {

constexpr int PIV_THREADS = 1024;

// Leaf of up to LEAF_COLS columns, one work-group: for each column j the
// pivot search (src/lu.jl:223-253: first index of the largest |a|, as
// idamax), the row interchange across the leaf's columns, the scaling of the
// column below the pivot (by the reciprocal; a division below the smallest
// normal number) and the rank-1 update of the leaf's remaining columns (what
// the recursion's GEMM updates do at this width).  m == 1 is the reference's
// row case (src/lu.jl:216-221).  Pivots are written 1-based with the leaf's
// row offset r0 added (its rows start on the diagonal, r0 = its first
// column); info takes the first zero pivot's global column.
constexpr int LEAF_COLS = 8;
constexpr int LEAF_WAVES = PIV_THREADS / 64;

// (|a|, signed a, row) of the larger |a|, ties to the smaller row.
template <typename T>
__device__ inline void better(T &bv, T &bs, int64_t &bi, T v, T sv, int64_t i) {
    if (v > bv || (v == bv && i < bi)) {
        bv = v;
        bs = sv;
        bi = i;
    }
}

template <typename T>
__device__ inline void wave_argmax(T &bv, T &bs, int64_t &bi) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        const T v = __shfl_xor(bv, o, 64), sv = __shfl_xor(bs, o, 64);
        const int64_t i = __shfl_xor(bi, o, 64);
        better(bv, bs, bi, v, sv, i);
    }
}

template <typename T>
__device__ __forceinline__ void lu_leaf_body(int64_t m, int n, T *__restrict__ A, int64_t lda, int64_t *__restrict__ ipiv,
                             int64_t *__restrict__ info, int64_t r0) {
    __shared__ T s_val[2][LEAF_WAVES], s_sgn[2][LEAF_WAVES];
    __shared__ int64_t s_row[2][LEAF_WAVES];
    __shared__ T s_u[LEAF_COLS];  // row j of the leaf after the interchange
    const int t = threadIdx.x, lane = t & 63, wv = t >> 6;
    const int kmax = (int)(m < n ? m : n);
    // partial maxima of column 0 (each later column's are taken in the
    // update pass of the column before it); the two LDS buffers alternate
    // between consecutive columns
    T bv = T(-1), bs = T(0);
    int64_t bi = m;
    for (int64_t i = t; i < m; i += PIV_THREADS) {
        const T a = A[i];
        better(bv, bs, bi, (T)fabs(a), a, i);
    }
    for (int j = 0; j < kmax; ++j) {
        const int buf = j & 1;
        T *cj = A + (int64_t)j * lda;
        wave_argmax(bv, bs, bi);
        if (lane == 0) {
            s_val[buf][wv] = bv;
            s_sgn[buf][wv] = bs;
            s_row[buf][wv] = bi;
        }
        __syncthreads();
        T pv = T(-1), a = T(0);
        int64_t p = m;
#pragma unroll
        for (int w = 0; w < LEAF_WAVES; ++w) better(pv, a, p, s_val[buf][w], s_sgn[buf][w], s_row[buf][w]);
        if (p >= m) {  // no finite maximum (all NaN): keep row j
            p = j;
            a = cj[j];
        }
        if (t == 0) {
            ipiv[j] = r0 + p + 1;
            if (a == T(0) && *info == 0) *info = r0 + j + 1;
        }
        // interchange rows j and p of the leaf's columns; row j's new values
        // go to LDS for the update
        if (t < n) {
            T *ct = A + (int64_t)t * lda;
            const T xj = ct[j], xp = ct[p];
            if (a != T(0) && p != j) {
                ct[j] = xp;
                ct[p] = xj;
                s_u[t] = xp;
            } else {
                s_u[t] = xj;
            }
        }
        __syncthreads();
        bv = T(-1);
        bs = T(0);
        bi = m;
        if (a != T(0)) {
            // scale column j below the pivot, update the columns to its right
            // and take the partial maxima of column j + 1.  Rows go in chunks
            // of RU per thread; every load of a chunk (column j and all the
            // columns to its right) is issued before its first store, with
            // addresses clamped to row / column j and results masked.
            constexpr int RU = 4, NC = LEAF_COLS - 1;
            const bool recip = fabs(a) >= std::numeric_limits<T>::min();
            const T r = T(1) / a;
            for (int64_t i0 = j + 1 + t; i0 < m; i0 += RU * PIV_THREADS) {
                T l[RU], x[NC][RU];
#pragma unroll
                for (int u = 0; u < RU; ++u) {
                    const int64_t i = i0 + (int64_t)u * PIV_THREADS;
                    l[u] = cj[i < m ? i : j];
                }
#pragma unroll
                for (int q = 0; q < NC; ++q) {
                    const int c = j + 1 + q;
                    const T *cc = A + (int64_t)(c < n ? c : j) * lda;
#pragma unroll
                    for (int u = 0; u < RU; ++u) {
                        const int64_t i = i0 + (int64_t)u * PIV_THREADS;
                        x[q][u] = cc[i < m ? i : j];
                    }
                }
#pragma unroll
                for (int u = 0; u < RU; ++u) {
                    const int64_t i = i0 + (int64_t)u * PIV_THREADS;
                    l[u] = recip ? l[u] * r : l[u] / a;
                    if (i < m) cj[i] = l[u];
                }
#pragma unroll
                for (int q = 0; q < NC; ++q) {
                    const int c = j + 1 + q;
                    if (c >= n) break;
                    T *cc = A + (int64_t)c * lda;
                    const T uc = s_u[c];
#pragma unroll
                    for (int u = 0; u < RU; ++u) {
                        const int64_t i = i0 + (int64_t)u * PIV_THREADS;
                        if (i < m) {
                            const T y = x[q][u] - l[u] * uc;
                            cc[i] = y;
                            if (q == 0) better(bv, bs, bi, (T)fabs(y), y, i);
                        }
                    }
                }
            }
        } else if (j + 1 < kmax) {
            // a zero column (the max |a| is 0): nothing to eliminate
            T *cn = A + (int64_t)(j + 1) * lda;
            for (int64_t i = j + 1 + t; i < m; i += PIV_THREADS) {
                const T x = cn[i];
                better(bv, bs, bi, (T)fabs(x), x, i);
            }
        }
        // no barrier here: the next column's first barrier follows every
        // thread's updates, and s_u is rewritten only after it
    }
}

template <typename T>
__global__ __launch_bounds__(PIV_THREADS) void lu_leaf_kernel(int64_t m, int n, T *__restrict__ A, int64_t lda,
                                                              int64_t *__restrict__ ipiv, int64_t *__restrict__ info,
                                                              int64_t r0) {
    lu_leaf_body<T>(m, n, A, lda, ipiv, info, r0);
}

// The same leaf for tall blocks, spread over G work-groups of SPLIT_ROWS
// rows each, the block held in registers from the first load to the final
// store.  Per column: each work-group finds its candidate (first index of
// the largest |a| among its rows at or below the diagonal) and publishes it
// with its row, the group holding row j publishes that row, and after one
// grid barrier every work-group takes the same winner, makes the same
// interchange in its registers and eliminates its own rows.  The published
// values alternate between two buffers by column parity, so a work-group
// one column ahead never overwrites what a slower one still reads.  At most
// 64 groups of one CU each are launched, which an idle part of the chip
// holds at once; a cooperative launch would guarantee it, but costs about
// 150 us once other streams have been used.  Each wait is bounded, so a
// group that is not resident (another persistent grid holds the CUs) cannot
// hang the others.  Every group saves its rows before the first barrier (and
// group 0 the prior info), and a wait that runs out raises a flag; the
// one-work-group recovery launch that follows every split leaf returns at
// once unless the flag is up, and otherwise restores the saved block and
// factors it with the global-memory leaf, so the factorization stays exact.
// The last group to finish resets the counters for the next launch on the
// stream.
// With ONE (a single work-group) the group's own winner is the pivot and
// nothing is published: the whole leaf stays in registers with two
// work-group barriers per column and no memory round trip.
constexpr int SPLIT_THREADS = 256, SPLIT_RPT = 8, SPLIT_ROWS = SPLIT_THREADS * SPLIT_RPT;
constexpr int ONE_THREADS = 512, ONE_RPT = 8, ONE_ROWS = ONE_THREADS * ONE_RPT;
constexpr int SPLIT_MAX_WG = 64, SPLIT_SLOT = 16;  // 64-bit words per published candidate
constexpr int SYNC_SLOTS = 64, SYNC_DIAG = SYNC_SLOTS + 2 * SPLIT_MAX_WG * SPLIT_SLOT;
constexpr int SYNC_WORDS = SYNC_DIAG + 2 * SPLIT_SLOT;
// words 2, 3 of the buffer: the timeout flag and the info value saved before
// the leaf; after SYNC_WORDS: the saved block, column c at c * SAVE_LD
constexpr int SYNC_FLAG = 2, SYNC_INFO = 3;
constexpr int64_t SAVE_LD = (int64_t)SPLIT_ROWS * SPLIT_MAX_WG;
constexpr int64_t SYNC_BYTES = SYNC_WORDS * 8 + SAVE_LD * LEAF_COLS * 8;
constexpr int64_t SPLIT_SPIN_LIMIT = int64_t(1) << 20;  // about a second of polling

__device__ inline void st_agent(uint64_t *p, uint64_t v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ inline uint64_t ld_agent(uint64_t *p) { return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT); }
__device__ inline uint64_t dbits(double v) { return (uint64_t)__double_as_longlong(v); }
__device__ inline double bitsd(uint64_t v) { return __longlong_as_double((long long)v); }

template <typename T, int NT, int RPT, bool ONE>
__global__ __launch_bounds__(NT) void lu_leaf_reg_kernel(int64_t m, int n, T *__restrict__ A, int64_t lda,
                                                         int64_t *__restrict__ ipiv, int64_t *__restrict__ info,
                                                         int64_t r0, uint64_t *__restrict__ sync,
                                                         int64_t spin_limit) {
    constexpr int W = NT / 64, ROWS = NT * RPT;
    __shared__ T s_val[W], s_sgn[W];
    __shared__ int64_t s_row[W];
    __shared__ T s_u[LEAF_COLS], s_o[LEAF_COLS];  // candidate / pivot row, row j before the interchange
    __shared__ T s_a;
    __shared__ int64_t s_p;
    __shared__ int s_fail;
    const int t = threadIdx.x, lane = t & 63, wv = t >> 6;
    const int G = (int)gridDim.x, g = (int)blockIdx.x;
    const int64_t first = (int64_t)g * ROWS;  // this group's rows: [first, first + ROWS)
    const int64_t base = first + t;
    const int kmax = (int)(m < n ? m : n);
    uint64_t *arrive = sync, *leave = sync + 1;
    T a[RPT][LEAF_COLS];
#pragma unroll
    for (int r = 0; r < RPT; ++r) {
        const int64_t i = base + (int64_t)r * NT;
#pragma unroll
        for (int c = 0; c < LEAF_COLS; ++c) a[r][c] = (i < m && c < n) ? A[i + (int64_t)c * lda] : T(0);
    }
    if (t == 0) s_fail = 0;
    if constexpr (!ONE) {
        // the block as it was, for the recovery launch
        T *save = reinterpret_cast<T *>(sync + SYNC_WORDS);
#pragma unroll
        for (int r = 0; r < RPT; ++r) {
            const int64_t i = base + (int64_t)r * NT;
            if (i < m) {
#pragma unroll
                for (int c = 0; c < LEAF_COLS; ++c)
                    if (c < n) save[i + (int64_t)c * SAVE_LD] = a[r][c];
            }
        }
        // only group 0 writes info, and the same thread arrives first at
        // every barrier after this store
        if (g == 0 && t == 0) st_agent(sync + SYNC_INFO, (uint64_t)*info);
    }
    bool failed = false;  // lane 0 of wave 0: a wait ran out; later ones are skipped
#pragma unroll
    for (int j = 0; j < LEAF_COLS; ++j) {
        if (j >= kmax) break;
        const int par = j & 1;
        T bv = T(-1), bs = T(0);
        int64_t bi = m;
#pragma unroll
        for (int r = 0; r < RPT; ++r) {
            const int64_t i = base + (int64_t)r * NT;
            if (i >= j && i < m) better(bv, bs, bi, (T)fabs(a[r][j]), a[r][j], i);
        }
        wave_argmax(bv, bs, bi);
        if (lane == 0) {
            s_val[wv] = bv;
            s_sgn[wv] = bs;
            s_row[wv] = bi;
        }
        __syncthreads();
        T pv = T(-1), ps = T(0);
        int64_t pi = m;
#pragma unroll
        for (int w = 0; w < W; ++w) better(pv, ps, pi, s_val[w], s_sgn[w], s_row[w]);
#pragma unroll
        for (int r = 0; r < RPT; ++r) {
            const int64_t i = base + (int64_t)r * NT;
            if (i == pi) {
#pragma unroll
                for (int c = 0; c < LEAF_COLS; ++c) s_u[c] = a[r][c];
            }
            if (i == j) {
#pragma unroll
                for (int c = 0; c < LEAF_COLS; ++c) s_o[c] = a[r][c];
            }
        }
        __syncthreads();
        int64_t p = pi;
        T av = ps;
        if constexpr (!ONE) {
            if (wv == 0) {
                uint64_t *slot = sync + SYNC_SLOTS + (par * SPLIT_MAX_WG + g) * SPLIT_SLOT;
                if (lane == 0) {
                    st_agent(slot + 0, dbits((double)pv));
                    st_agent(slot + 1, dbits((double)ps));
                    st_agent(slot + 2, (uint64_t)pi);
                }
                if (lane < LEAF_COLS) {
                    st_agent(slot + 4 + lane, dbits(pi < m ? (double)s_u[lane] : 0.0));
                    if (j >= first && j < first + ROWS)
                        st_agent(sync + SYNC_DIAG + par * SPLIT_SLOT + lane, dbits((double)s_o[lane]));
                }
                if (lane == 0) {
                    // grid barrier: arrival count reaches G * (j + 1)
                    __hip_atomic_fetch_add(arrive, (uint64_t)1, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
                    const uint64_t target = (uint64_t)G * (uint64_t)(j + 1);
                    int64_t spins = 0;
                    while (!failed && ld_agent(arrive) < target) {
                        if (++spins > spin_limit) {
                            failed = true;
                            s_fail = 1;
                            break;
                        }
                        __builtin_amdgcn_s_sleep(1);
                    }
                    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
                }
                // every lane reads one group's candidate (and lanes < LEAF_COLS
                // row j), then the wave reduces to the winner; ties go to the
                // smaller row, as within a group
                double cv = -1.0, cs = 0.0, crow[LEAF_COLS];
                int64_t ci = m;
                if (lane < G) {
                    uint64_t *sl = sync + SYNC_SLOTS + (par * SPLIT_MAX_WG + lane) * SPLIT_SLOT;
                    cv = bitsd(ld_agent(sl + 0));
                    cs = bitsd(ld_agent(sl + 1));
                    ci = (int64_t)ld_agent(sl + 2);
#pragma unroll
                    for (int c = 0; c < LEAF_COLS; ++c) crow[c] = bitsd(ld_agent(sl + 4 + c));
                } else {
#pragma unroll
                    for (int c = 0; c < LEAF_COLS; ++c) crow[c] = 0.0;
                }
                const double rowj = lane < LEAF_COLS ? bitsd(ld_agent(sync + SYNC_DIAG + par * SPLIT_SLOT + lane)) : 0.0;
                wave_argmax(cv, cs, ci);
                // the winning candidate's lane is the group holding row ci
                const int wl = ci < m ? (int)(ci / ROWS) : 0;
                double u = 0.0;
#pragma unroll
                for (int c = 0; c < LEAF_COLS; ++c) {
                    const double x = __shfl(crow[c], wl, 64);
                    if (lane == c) u = x;
                }
                if (lane == 0) {
                    s_p = ci;
                    s_a = (T)cs;
                }
                if (lane < LEAF_COLS) {
                    s_u[lane] = (T)u;
                    s_o[lane] = (T)rowj;
                }
            }
            __syncthreads();
            p = s_p;
            av = s_a;
        }
        T u[LEAF_COLS], o[LEAF_COLS];
#pragma unroll
        for (int c = 0; c < LEAF_COLS; ++c) {
            u[c] = s_u[c];
            o[c] = s_o[c];
        }
        if (p >= m) {  // no finite maximum (all NaN): keep row j
            p = j;
            av = o[j];
#pragma unroll
            for (int c = 0; c < LEAF_COLS; ++c) u[c] = o[c];
        }
        if (g == 0 && t == 0) {
            ipiv[j] = r0 + p + 1;
            if (av == T(0) && *info == 0) *info = r0 + j + 1;
        }
        if (av != T(0)) {
            const bool swap = p != j;
            const bool recip = fabs(av) >= std::numeric_limits<T>::min();
            const T rc = T(1) / av;
#pragma unroll
            for (int r = 0; r < RPT; ++r) {
                const int64_t i = base + (int64_t)r * NT;
                if (swap && i == j) {
#pragma unroll
                    for (int c = 0; c < LEAF_COLS; ++c) a[r][c] = u[c];
                } else if (swap && i == p) {
#pragma unroll
                    for (int c = 0; c < LEAF_COLS; ++c) a[r][c] = o[c];
                }
                if (i > j && i < m) {
                    const T l = recip ? a[r][j] * rc : a[r][j] / av;
                    a[r][j] = l;
#pragma unroll
                    for (int c = j + 1; c < LEAF_COLS; ++c) a[r][c] = a[r][c] - l * u[c];
                }
            }
        }
    }
#pragma unroll
    for (int r = 0; r < RPT; ++r) {
        const int64_t i = base + (int64_t)r * NT;
        if (i < m) {
#pragma unroll
            for (int c = 0; c < LEAF_COLS; ++c)
                if (c < n) A[i + (int64_t)c * lda] = a[r][c];
        }
    }
    if constexpr (ONE) return;
    __syncthreads();
    if (t == 0) {
        if (s_fail) st_agent(sync + SYNC_FLAG, 1);
        // the last group out resets the counters (every group has passed
        // every barrier by then)
        const uint64_t prev = __hip_atomic_fetch_add(leave, (uint64_t)1, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
        if (prev == (uint64_t)G - 1) {
            st_agent(arrive, 0);
            st_agent(leave, 0);
        }
    }
}

// After every split leaf: nothing unless one of its waits ran out; then the
// block saved before the leaf is put back, info restored, and the block
// factored by the one-work-group global-memory leaf (pivots rewritten).
template <typename T>
__global__ __launch_bounds__(PIV_THREADS) void lu_leaf_recover_kernel(int64_t m, int n, T *__restrict__ A, int64_t lda,
                                                                      int64_t *__restrict__ ipiv,
                                                                      int64_t *__restrict__ info, int64_t r0,
                                                                      uint64_t *__restrict__ sync) {
    if (__hip_atomic_load(sync + SYNC_FLAG, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == 0) return;
    const T *save = reinterpret_cast<const T *>(sync + SYNC_WORDS);
    for (int64_t idx = threadIdx.x; idx < m * n; idx += PIV_THREADS) {
        const int64_t i = idx % m, c = idx / m;
        A[i + c * lda] = save[i + c * SAVE_LD];
    }
    if (threadIdx.x == 0) *info = (int64_t)sync[SYNC_INFO];
    __syncthreads();
    lu_leaf_body<T>(m, n, A, lda, ipiv, info, r0);
    __syncthreads();
    if (threadIdx.x == 0) st_agent(sync + SYNC_FLAG, 0);
}

__global__ void sync_zero_kernel(uint64_t *p, int words) {
    for (int i = threadIdx.x; i < words; i += blockDim.x) p[i] = 0;
}

// laswp (src/lu.jl:470): for k = k1..k2 (1-based) swap rows k and
// ipiv[k] - r0 of each of the ncols columns (ipiv holds global rows, the
// block's row 1 is global row r0 + 1); one thread per column walks the
// interchanges in order.
template <typename T>
__global__ void laswp_kernel(int64_t ncols, T *__restrict__ A, int64_t lda, int64_t k1, int64_t k2,
                             const int64_t *__restrict__ ipiv, int64_t r0) {
    const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= ncols) return;
    T *col = A + c * lda;
    for (int64_t k = k1; k <= k2; ++k) {
        const int64_t p = ipiv[k - 1] - 1 - r0;
        if (p != k - 1) {
            const T a = col[k - 1];
            col[k - 1] = col[p];
            col[p] = a;
        }
    }
}

// The same interchanges as one row permutation, for blocks whose rows fit
// the work-group's LDS.  Positions are taken relative to b = k1 - 1; swap k
// (0-based) exchanges positions k and p_k >= k.  Every work-group composes
// the K swaps into src (the final content of position x is the old content
// of position src[x]) with one thread walking them in LDS, lists the
// positions past K that the swaps moved, and then permutes its columns: each
// thread loads up to LASWP_R moved entries of a column, a barrier, and the
// stores.  The sequential part runs at LDS latency instead of one dependent
// round trip to memory per swap and column.
constexpr int LASWP_THREADS = 1024, LASWP_R = 16;
constexpr int LASWP_LDS = 40960 - 64;  // ints: src map (span) + pivots / moved list (K)

template <typename T>
__global__ __launch_bounds__(LASWP_THREADS) void laswp_perm_kernel(int64_t ncols, T *__restrict__ A, int64_t lda,
                                                                   int64_t b, int K, int span,
                                                                   const int64_t *__restrict__ ipiv, int64_t r0) {
    __shared__ int s_map[LASWP_LDS];
    __shared__ int s_moved;
    int *src = s_map, *lst = s_map + span;
    const int t = threadIdx.x;
    for (int x = t; x < span; x += LASWP_THREADS) src[x] = x;
    for (int k = t; k < K; k += LASWP_THREADS) {
        const int64_t p = ipiv[b + k] - 1 - r0 - b;
        lst[k] = (p >= k && p < span) ? (int)p : k;  // pivots lie in [k, span)
    }
    if (t == 0) s_moved = 0;
    __syncthreads();
    if (t == 0) {
        for (int k = 0; k < K; ++k) {
            const int p = lst[k];
            if (p != k) {
                const int u = src[k];
                src[k] = src[p];
                src[p] = u;
            }
        }
    }
    __syncthreads();
    // positions past K hold their own rows unless a swap moved them
    for (int x = K + t; x < span; x += LASWP_THREADS)
        if (src[x] != x) lst[atomicAdd(&s_moved, 1)] = x;
    __syncthreads();
    const int tot = K + s_moved;
    for (int64_t c = blockIdx.x; c < ncols; c += gridDim.x) {
        T *col = A + c * lda + b;
        T v[LASWP_R];
#pragma unroll
        for (int r = 0; r < LASWP_R; ++r) {
            const int i = t + r * LASWP_THREADS;
            const int x = i < K ? i : (i < tot ? lst[i - K] : 0);
            v[r] = (i < tot) ? col[src[x]] : T(0);
        }
        __syncthreads();
#pragma unroll
        for (int r = 0; r < LASWP_R; ++r) {
            const int i = t + r * LASWP_THREADS;
            if (i < tot) {
                const int x = i < K ? i : lst[i - K];
                if (src[x] != x) col[x] = v[r];
            }
        }
    }
}

// A12 = L11^-1 A12 for a unit lower L11 of n1 <= SMALL_TRSM rows (the
// reference's trsm!('L','L','N','U'), src/lu.jl:277): forward substitution,
// one thread per column of A12 holding it in registers; L11 is read through
// the cache (every thread reads the same entries).  Replaces the inverse and
// panel launches of the general TRSM at the bottom of the recursion.
constexpr int SMALL_TRSM = 32;

template <typename T>
__global__ __launch_bounds__(64) void lu_trsm_small_kernel(int n1, int64_t n2, const T *__restrict__ L, int64_t lda,
                                                           T *__restrict__ B) {
    const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= n2) return;
    T *b = B + c * lda;
    T x[SMALL_TRSM];
#pragma unroll
    for (int i = 0; i < SMALL_TRSM; ++i) x[i] = b[i < n1 ? i : 0];
#pragma unroll
    for (int k = 0; k < SMALL_TRSM - 1; ++k) {
        const T xk = x[k];
        const T *lk = L + (int64_t)(k < n1 ? k : 0) * lda;
#pragma unroll
        for (int i = k + 1; i < SMALL_TRSM; ++i)
            if (i < n1) x[i] -= lk[i] * xk;
    }
#pragma unroll
    for (int i = 0; i < SMALL_TRSM; ++i)
        if (i < n1) b[i] = x[i];
}

__global__ void info_zero_kernel(int64_t *info) {
    if (threadIdx.x == 0 && blockIdx.x == 0) *info = 0;
}

// Counters and published candidates of the split leaf, one buffer per
// (device, stream): launches on one stream run in order, so they share it.
struct SplitSync {
    int dev;
    hipStream_t stream;
    uint64_t *ptr;
};

uint64_t *split_sync(hipStream_t s, int32_t *st) {
    static std::mutex mu;
    static std::vector<SplitSync> cache;
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) {
        *st = report_hip(e, "lu split leaf device");
        return nullptr;
    }
    std::lock_guard<std::mutex> g(mu);
    for (auto &c : cache)
        if (c.dev == dev && c.stream == s) return c.ptr;
    uint64_t *p = nullptr;
    e = hipMalloc((void **)&p, SYNC_BYTES);
    if (e == hipSuccess) {
        hipLaunchKernelGGL(sync_zero_kernel, dim3(1), dim3(256), 0, s, p, SYNC_WORDS);
        e = hipGetLastError();
    }
    if (e != hipSuccess) {
        *st = report_hip(e, "lu split leaf sync buffer");
        return nullptr;
    }
    cache.push_back({dev, s, p});
    return p;
}

template <typename T>
struct LuRec {
    int32_t dtype;
    int64_t lda;
    int64_t *info;
    hipStream_t s;
    int32_t st = NLA_OK;
    int64_t split_rows = lu_split_rows();
    int64_t spin_limit = lu_spin_limit() > 0 ? lu_spin_limit() : SPLIT_SPIN_LIMIT;

    // an m x n leaf (n <= LEAF_COLS) over ceil(m / SPLIT_ROWS) work-groups
    void split_leaf(int64_t m, int n, T *A, int64_t *ipiv, int64_t r0) {
        uint64_t *sync = split_sync(s, &st);
        if (!sync) return;
        const unsigned g = (unsigned)((m + SPLIT_ROWS - 1) / SPLIT_ROWS);
        hipLaunchKernelGGL((lu_leaf_reg_kernel<T, SPLIT_THREADS, SPLIT_RPT, false>), dim3(g), dim3(SPLIT_THREADS), 0, s,
                           m, n, A, lda, ipiv, info, r0, sync, spin_limit);
        if (!ok(hipGetLastError(), "lu split leaf")) return;
        hipLaunchKernelGGL(lu_leaf_recover_kernel<T>, dim3(1), dim3(PIV_THREADS), 0, s, m, n, A, lda, ipiv, info, r0,
                           sync);
        ok(hipGetLastError(), "lu split leaf recovery");
    }

    bool ok(hipError_t e, const char *where) {
        if (e != hipSuccess && st == NLA_OK) st = report_hip(e, where);
        return st == NLA_OK;
    }

    // interchanges k1..k2 of the m-row block at A on ncols columns: the LDS
    // permutation when its map fits, else the sequential walk
    void laswp(T *A, int64_t m, int64_t ncols, int64_t k1, int64_t k2, const int64_t *ipiv, int64_t r0) {
        if (ncols <= 0 || k2 < k1 || st != NLA_OK) return;
        const int64_t K = k2 - k1 + 1, b = k1 - 1, span = m - b;
        if (2 * K <= (int64_t)LASWP_R * LASWP_THREADS && span + K <= LASWP_LDS) {
            const unsigned g = (unsigned)(ncols < 256 ? ncols : 256);
            hipLaunchKernelGGL(laswp_perm_kernel<T>, dim3(g), dim3(LASWP_THREADS), 0, s, ncols, A, lda, b, (int)K,
                               (int)span, ipiv, r0);
        } else {
            hipLaunchKernelGGL(laswp_kernel<T>, dim3((unsigned)((ncols + 63) / 64)), dim3(64), 0, s, ncols, A, lda, k1,
                               k2, ipiv, r0);
        }
        ok(hipGetLastError(), "laswp");
    }

    // Factor the m x n block at A, whose first row and column are the global
    // row / column r0 (blocks start on the diagonal); its pivots go to
    // ipiv[0 .. min(m,n)) as global 1-based rows, which is what the
    // reference's "ipiv[i] += n1" (src/lu.jl:289-291) makes of them.
    void rec(int64_t m, int64_t n, T *A, int64_t *ipiv, int64_t r0) {
        if (st != NLA_OK || m == 0 || n == 0) return;
        if (n <= LEAF_COLS && split_rows > 0 && m >= split_rows && m <= (int64_t)SPLIT_ROWS * SPLIT_MAX_WG) {
            if (m <= ONE_ROWS) {
                hipLaunchKernelGGL((lu_leaf_reg_kernel<T, ONE_THREADS, ONE_RPT, true>), dim3(1), dim3(ONE_THREADS), 0, s,
                                   m, (int)n, A, lda, ipiv, info, r0, (uint64_t *)nullptr, (int64_t)0);
                ok(hipGetLastError(), "lu register leaf");
            } else {
                split_leaf(m, (int)n, A, ipiv, r0);
            }
            return;
        }
        if (n <= LEAF_COLS || m == 1) {
            const int nl = (int)(n <= LEAF_COLS ? n : 1);  // m == 1: only the first column is examined
            hipLaunchKernelGGL(lu_leaf_kernel<T>, dim3(1), dim3(PIV_THREADS), 0, s, m, nl, A, lda, ipiv, info, r0);
            ok(hipGetLastError(), "lu leaf");
            return;
        }
        const int64_t mn = m < n ? m : n, n1 = mn / 2, n2 = n - n1;
        rec(m, n1, A, ipiv, r0);
        laswp(A + n1 * lda, m, n2, 1, n1, ipiv, r0);
        if (st != NLA_OK) return;
        // A12 = L11^-1 A12 (unit lower; the identity for one column)
        if (n1 > 1 && n1 <= SMALL_TRSM) {
            hipLaunchKernelGGL(lu_trsm_small_kernel<T>, dim3((unsigned)((n2 + 63) / 64)), dim3(64), 0, s, (int)n1, n2,
                               A, lda, A + n1 * lda);
            if (!ok(hipGetLastError(), "lu small trsm")) return;
        } else if (n1 > 1) {
            const int32_t r = nla_trsm('L', 'L', 'N', 'U', dtype, 1.0, n1, n2, A, lda, A + n1 * lda, lda, s);
            if (r != NLA_OK) {
                st = r;
                return;
            }
        }
        // A22 -= A21 A12
        if (m > n1) {
            const int32_t r = nla_gemm(dtype, 'N', 'N', m - n1, n2, n1, -1.0, A + n1, lda, A + n1 * lda, lda, 1.0,
                                       A + n1 + n1 * lda, lda, s);
            if (r != NLA_OK) {
                st = r;
                return;
            }
        }
        rec(m - n1, n2, A + n1 + n1 * lda, ipiv + n1, r0 + n1);
        laswp(A, m, n1, n1 + 1, mn, ipiv, r0);
    }
};

}  // namespace
}  // namespace nla

using namespace nla;

extern "C" int32_t nla_getrf2(int32_t dtype, int64_t m, int64_t n, void *A, int64_t lda, int64_t *ipiv,
                              int64_t *info, void *stream) {
    if (dtype != NLA_F32 && dtype != NLA_F64) return report_bad_arg(1, "dtype");
    if (m < 0) return report_bad_arg(2, "m");
    if (n < 0) return report_bad_arg(3, "n");
    if (lda < (m > 1 ? m : 1)) return report_bad_arg(5, "lda");
    if (m > 0 && n > 0 && A == nullptr) return report_bad_arg(4, "A");
    if (m > 0 && n > 0 && ipiv == nullptr) return report_bad_arg(6, "ipiv");
    if (info == nullptr) return report_bad_arg(7, "info");
    hipStream_t s = (hipStream_t)stream;
    hipLaunchKernelGGL(info_zero_kernel, dim3(1), dim3(64), 0, s, info);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return report_hip(e, "getrf2 info");
    if (m == 0 || n == 0) return NLA_OK;
    if (dtype == NLA_F64) {
        LuRec<double> r{dtype, lda, info, s};
        r.rec(m, n, (double *)A, ipiv, 0);
        return r.st;
    }
    LuRec<float> r{dtype, lda, info, s};
    r.rec(m, n, (float *)A, ipiv, 0);
    return r.st;
}
