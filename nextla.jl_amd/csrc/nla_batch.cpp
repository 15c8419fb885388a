// Batched triangular solves / multiplies: host-side planning for the
// persistent panel kernel and the grouped GEMM (nla_batch_* in
// include/nla.h).
//
// A plan puts every system into the panel kernel's canonical form (the same
// transformation nla_rectrxm applies to one system: op(A) as a strided view
// with flipped uplo for 'T'/'C'; right-side systems solved on the transpose).
// A system of up to BATCH_MAX rows (512 when the class mask has
// BATCH_STAGE_ABOVE_512 set) is one panel.  A larger one is cut into
// blocks of BATCH_STAGE rows and solved in stages, as the recursion would at
// that block size: per stage one diagonal block (a panel) and one GEMM that
// applies it to the rest of the system (TRSM: the rows still to solve; TRMM:
// the block itself from the rows not yet multiplied).  Stage s of every
// system runs in the same launches, so a plan of any mix of sizes takes
//   1 inversion launch (TRSM) + per stage: a panel launch per size class
//   and one grouped GEMM launch (two in FP64: systems of at most 16
//   right-hand sides take narrow tiles in the second),
// independent of the number of systems.  Panel tasks are one per (panel,
// 16-column strip), grouped by size class (the LDS strip of the kernel that
// solves a class fits its largest panel, so small panels run at higher
// occupancy) and longest first within a class; GEMM tasks are one per output
// tile (FP64: 64 rows by 32 right-hand sides, by 16 for systems of at most
// 16, as most have few; FP32: 64 x 64), longest k-range first.  Descriptors are uploaded in stream
// order into the caller's workspace; executing a plan launches kernels on
// the caller's stream and nothing else, so it can be captured in a hipGraph
// and replayed.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <new>
#include <vector>

#include "../../include/nla.h"
#include "nla_internal.h"

namespace nla {

int32_t report_hip(hipError_t e, const char *where);
int32_t report_bad_arg(int k, const char *what);

static inline char upc(char c) { return (c >= 'a' && c <= 'z') ? (char)(c - 'a' + 'A') : c; }
static inline int64_t align256(int64_t x) { return (x + 255) & ~(int64_t)255; }

// Stages of one system: one up to the single-panel limit, else 512-row blocks.
static inline int64_t panel_limit(int mask) { return (mask & BATCH_STAGE_ABOVE_512) ? BATCH_STAGE : BATCH_MAX; }
static inline int64_t stages_of(int64_t n, int mask) {
    return n <= panel_limit(mask) ? 1 : (n + BATCH_STAGE - 1) / BATCH_STAGE;
}

struct Layout {
    int64_t counter, sys, pan, ptasks, gem, gtasks, blk, inv, flg, total;
    int64_t npan, nptasks, ngem, ngtasks, nblk, nstages;
    int mask;  // size-class mask the plan was laid out for
};

// Counts and offsets; they depend on func, dtype and the items' n, m only.
static Layout layout_of(char func, int32_t dtype, const nla_batch_item *items, int64_t count, int mask) {
    const int64_t es = dtype == NLA_F64 ? 8 : 4;
    const int64_t sys_bytes = dtype == NLA_F64 ? sizeof(BatchSys<double>) : sizeof(BatchSys<float>);
    const int64_t gem_bytes = dtype == NLA_F64 ? sizeof(GroupGemm<double>) : sizeof(GroupGemm<float>);
    Layout L{};
    L.mask = mask;
    L.nstages = 0;
    for (int64_t i = 0; i < count; ++i) {
        const int64_t n = items[i].n, m = items[i].m;
        if (n <= 0 || m <= 0) continue;
        const int64_t st = stages_of(n, mask), strips = (m + 15) / 16;
        L.nstages = std::max(L.nstages, st);
        L.npan += st;
        L.nptasks += st * strips;
        if (func == 'S') L.nblk += (n + 63) / 64;
        if (st > 1) {
            // every stage but one has a GEMM; its output is at most n rows x m
            const int64_t tc = group_tile_c((int)es, group_narrow(m, (int)es));
            const int64_t tn = (m + tc - 1) / tc, tr = (n + GROUP_TILE_R - 1) / GROUP_TILE_R;
            L.ngem += st - 1;
            L.ngtasks += (st - 1) * tr * tn;
        }
    }
    L.counter = 0;
    L.sys = align256(std::max<int64_t>(256, L.nstages * 6 * 4));
    L.pan = align256(L.sys + count * sys_bytes);
    L.ptasks = align256(L.pan + L.npan * sys_bytes);
    L.gem = align256(L.ptasks + L.nptasks * (int64_t)sizeof(BatchTask));
    L.gtasks = align256(L.gem + L.ngem * gem_bytes);
    L.blk = align256(L.gtasks + L.ngtasks * (int64_t)sizeof(int2));
    L.inv = align256(L.blk + L.nblk * (int64_t)sizeof(int2));
    L.flg = align256(L.inv + L.nblk * 64 * 64 * es);
    L.total = align256(L.flg + L.nblk * (int64_t)sizeof(int));
    return L;
}

// One launch of the plan: a panel launch (cls 0..3) or a grouped GEMM (cls 4;
// cls 5: the FP64 GEMMs of few right-hand sides, in narrow tiles).
struct Launch {
    int stage, cls, start, count, groups;
    bool vec;
};

}  // namespace nla

struct nla_batch_plan_s {
    std::vector<char> host;  // staged descriptors, alive until the upload has run
    int32_t dtype;
    bool trmm;
    bool tx, ty;  // grouped GEMM operand orientation
    bool rows_m;  // the grouped GEMMs' M runs along the systems' rows (side 'L')
    char *ws;
    int nblk;
    std::vector<nla::Launch> launches;
    nla::Layout L;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
};

namespace nla {

static inline bool al16(const void *p) { return ((uintptr_t)p & 15) == 0; }

// The host side of a plan: descriptors, tasks and the launch partition
// (grid sizes come later, from the device's occupancy).  Besides what the
// device reads, it records for every panel and GEMM the item and rows it
// belongs to, in the canonical coordinates (rows: the order of op(A); for
// side 'R' the transposed system), for nla_batch_plan_describe.
template <typename T>
struct HostPlan {
    std::vector<BatchSys<T>> sys, pan;
    std::vector<int> pan_stage;
    std::vector<int64_t> pan_item, pan_r0;
    std::vector<BatchTask> ptasks;
    std::vector<GroupGemm<T>> gem;
    std::vector<int> gem_stage;
    std::vector<int64_t> gem_item, gem_r0, gem_rn, gem_k0, gem_kn;
    std::vector<char> gem_narrow;
    std::vector<int2> gtasks, blk;
    std::vector<Launch> launches;
    bool tx = false, ty = false;
};

// Fisher-Yates with a 64-bit LCG: the same permutation for the same seed.
template <typename V>
static void shuffle_range(std::vector<V> &v, size_t start, size_t count, uint64_t &state) {
    for (size_t k = count; k > 1; --k) {
        state = state * 6364136223846793005ULL + 1442695040888963407ULL;
        const size_t j = (size_t)((state >> 33) % k);
        std::swap(v[start + k - 1], v[start + j]);
    }
}

template <typename T>
static int32_t plan_host(char side, char uplo, char trans, char func, const nla_batch_item *items, int64_t count,
                         const Layout &L, uintptr_t ws, int order, HostPlan<T> &H) {
    const bool trmm = (func == 'M');
    std::vector<BatchSys<T>> &sys = H.sys, &pan = H.pan;
    std::vector<int> &pan_stage = H.pan_stage, &gem_stage = H.gem_stage;
    std::vector<BatchTask> &ptasks = H.ptasks;
    std::vector<GroupGemm<T>> &gem = H.gem;
    std::vector<int2> &blk = H.blk;
    sys.assign((size_t)count, BatchSys<T>{});
    pan.reserve((size_t)L.npan);
    ptasks.reserve((size_t)L.nptasks);
    gem.reserve((size_t)L.ngem);
    blk.reserve((size_t)L.nblk);
    T *inv = reinterpret_cast<T *>(ws + (uintptr_t)L.inv);
    int *flg = reinterpret_cast<int *>(ws + (uintptr_t)L.flg);
    int64_t inv_off = 0;
    const bool tr = (trans != 'N');
    const char ue = tr ? (uplo == 'L' ? 'U' : 'L') : uplo;  // uplo of op(A)
    // grouped GEMM orientation (see the descriptors below)
    H.tx = (side == 'L') ? tr : false;
    H.ty = (side == 'L') ? false : tr;
    for (int64_t i = 0; i < count; ++i) {
        const nla_batch_item &it = items[i];
        BatchSys<T> &S = sys[(size_t)i];
        std::memset(&S, 0, sizeof(S));
        S.n = (int)it.n;
        S.m = it.m;
        S.scale = (T)it.alpha;
        // op(A)(r, c) = A[r*ars + c*acs]
        const int64_t ars = tr ? it.lda : 1, acs = tr ? 1 : it.lda;
        S.M = (const T *)it.A;
        S.B = (T *)it.B;
        if (side == 'L') {
            S.lower = ue == 'L';
            S.mrs = ars;
            S.mcs = acs;
            S.brs = 1;
            S.bcs = it.ldb;
            S.wrs = 1;
            S.wcs = 64;
        } else {  // op(A)^T X^T = B^T
            S.lower = ue == 'U';
            S.mrs = acs;
            S.mcs = ars;
            S.brs = it.ldb;
            S.bcs = 1;
            S.wrs = 64;
            S.wcs = 1;
        }
        if (it.n <= 0 || it.m <= 0) continue;
        if (func == 'S') {
            S.W = inv + inv_off;
            S.F = flg + inv_off / (64 * 64);
            const int64_t nb = (it.n + 63) / 64;
            for (int64_t b = 0; b < nb; ++b) blk.push_back(make_int2((int)i, (int)b));
            inv_off += nb * 64 * 64;
        }
        const int64_t strips = (it.m + 15) / 16;
        const int J = (int)stages_of(it.n, L.mask);
        if (J == 1) {  // one panel, alpha at its first (TRSM) / last (TRMM) touch
            const int pi = (int)pan.size();
            pan.push_back(S);
            pan_stage.push_back(0);
            H.pan_item.push_back(i);
            H.pan_r0.push_back(0);
            for (int64_t st = 0; st < strips; ++st) ptasks.push_back(BatchTask{pi, (int)st});
            continue;
        }
        const T alpha = (T)it.alpha;
        const bool fwd = (S.lower != 0) != trmm;  // block order: first to last
        for (int s = 0; s < J; ++s) {
            const int j = fwd ? s : J - 1 - s;
            const int64_t r0 = (int64_t)j * BATCH_STAGE;
            const int64_t rows = std::min<int64_t>(BATCH_STAGE, it.n - r0);
            // the diagonal block as a panel
            BatchSys<T> D = S;
            D.n = (int)rows;
            D.M = S.M + r0 * S.mrs + r0 * S.mcs;
            D.B = S.B + r0 * S.brs;
            if (func == 'S') {
                D.W = S.W + (r0 / 64) * 64 * 64;
                D.F = S.F + r0 / 64;
            }
            D.scale = trmm ? alpha : (s == 0 ? alpha : T(1));
            const int pi = (int)pan.size();
            pan.push_back(D);
            pan_stage.push_back(s);
            H.pan_item.push_back(i);
            H.pan_r0.push_back(r0);
            for (int64_t st = 0; st < strips; ++st) ptasks.push_back(BatchTask{pi, (int)st});
            // the GEMM of the stage: rows R of the output, columns K of M
            int64_t R0, RN, K0, KN;
            T ga, gb;
            if (!trmm) {  // the solved block updates the rows still to solve
                if (s == J - 1) continue;
                K0 = r0;
                KN = rows;
                if (S.lower) {
                    R0 = r0 + rows;
                    RN = it.n - R0;
                } else {
                    R0 = 0;
                    RN = r0;
                }
                ga = T(-1);
                gb = (s == 0) ? alpha : T(1);
            } else {  // the block gathers the rows not yet multiplied
                R0 = r0;
                RN = rows;
                if (S.lower) {
                    K0 = 0;
                    KN = r0;
                } else {
                    K0 = r0 + rows;
                    KN = it.n - K0;
                }
                if (KN == 0) continue;
                ga = alpha;
                gb = T(1);
            }
            GroupGemm<T> g;
            const T *Mblk = S.M + R0 * S.mrs + K0 * S.mcs;  // op(M)(r, k), r in R, k in K
            if (side == 'L') {  // B(R, :) = gb*B(R, :) + ga * M(R, K) * B(K, :)
                g.M = RN; g.N = it.m; g.K = KN;
                g.X = Mblk; g.ldx = H.tx ? S.mrs : S.mcs;
                g.Y = S.B + K0; g.ldy = S.bcs;
                g.C = S.B + R0; g.ldc = S.bcs;
            } else {  // transposed: Bt(:, R) = gb*Bt(:, R) + ga * Bt(:, K) * M(R, K)^T
                g.M = it.m; g.N = RN; g.K = KN;
                g.X = S.B + K0 * S.brs; g.ldx = S.brs;
                g.Y = Mblk; g.ldy = H.ty ? S.mcs : S.mrs;
                g.C = S.B + R0 * S.brs; g.ldc = S.brs;
            }
            g.alpha = ga;
            g.beta = gb;
            const bool narrow = group_narrow(it.m, (int)sizeof(T));
            const int tm_edge = group_tile_m(side == 'L', (int)sizeof(T), narrow);
            g.tiles_m = (int)((g.M + tm_edge - 1) / tm_edge);
            gem.push_back(g);
            H.gem_narrow.push_back(narrow ? 1 : 0);
            gem_stage.push_back(s);
            H.gem_item.push_back(i);
            H.gem_r0.push_back(R0);
            H.gem_rn.push_back(RN);
            H.gem_k0.push_back(K0);
            H.gem_kn.push_back(KN);
        }
    }
    // Panel tasks: by stage, then size class (largest first), then cost.
    const int mask = L.mask;
    auto cls_of = [&](int n) {
        for (int c = 0; c < 4; ++c)
            if (((mask >> c) & 1 || c == 3) && n <= BATCH_CLASSES[c]) return c;
        return 3;
    };
    std::stable_sort(ptasks.begin(), ptasks.end(), [&](const BatchTask &a, const BatchTask &b) {
        const int sa = pan_stage[(size_t)a.sys], sb = pan_stage[(size_t)b.sys];
        if (sa != sb) return sa < sb;
        const int na = pan[(size_t)a.sys].n, nb2 = pan[(size_t)b.sys].n;
        const int ca = cls_of(na), cb = cls_of(nb2);
        if (ca != cb) return ca > cb;
        const int pa = (na + 63) / 64, pb = (nb2 + 63) / 64;
        return pa * (pa + 1) > pb * (pb + 1);
    });
    // GEMM tiles: by stage, wide tiles before narrow ones, longest k-range first.
    std::vector<int> gorder(gem.size());
    for (size_t k = 0; k < gorder.size(); ++k) gorder[k] = (int)k;
    std::stable_sort(gorder.begin(), gorder.end(), [&](int a, int b) {
        if (gem_stage[(size_t)a] != gem_stage[(size_t)b]) return gem_stage[(size_t)a] < gem_stage[(size_t)b];
        if (H.gem_narrow[(size_t)a] != H.gem_narrow[(size_t)b]) return H.gem_narrow[(size_t)a] < H.gem_narrow[(size_t)b];
        return gem[(size_t)a].K > gem[(size_t)b].K;
    });
    std::vector<int2> &gtasks = H.gtasks;
    gtasks.reserve((size_t)L.ngtasks);
    for (int gi : gorder) {
        const GroupGemm<T> &g = gem[(size_t)gi];
        const int tn_edge = group_tile_n(side == 'L', (int)sizeof(T), H.gem_narrow[(size_t)gi] != 0);
        const int tiles = g.tiles_m * (int)((g.N + tn_edge - 1) / tn_edge);
        for (int t = 0; t < tiles; ++t) gtasks.push_back(make_int2(gi, t));
    }
    if ((int64_t)pan.size() > L.npan || (int64_t)ptasks.size() > L.nptasks || (int64_t)gem.size() > L.ngem ||
        (int64_t)gtasks.size() > L.ngtasks || (int64_t)blk.size() > L.nblk)
        return report_hip(hipErrorInvalidValue, "batch plan sizes");

    // launches, in execution order (grids sized later)
    H.launches.clear();
    size_t pt = 0, gt = 0;
    for (int s = 0; s < (int)L.nstages; ++s) {
        for (int c = 3; c >= 0; --c) {
            const size_t start = pt;
            while (pt < ptasks.size() && pan_stage[(size_t)ptasks[pt].sys] == s &&
                   cls_of(pan[(size_t)ptasks[pt].sys].n) == c)
                ++pt;
            if (pt > start) H.launches.push_back(Launch{s, c, (int)start, (int)(pt - start), 0, false});
        }
        for (int nw = 0; nw < 2; ++nw) {  // wide tiles (cls 4), then narrow ones (cls 5)
            const size_t start = gt;
            bool vec = true;
            while (gt < gtasks.size() && gem_stage[(size_t)gtasks[gt].x] == s &&
                   H.gem_narrow[(size_t)gtasks[gt].x] == nw) {
                const GroupGemm<T> &g = gem[(size_t)gtasks[gt].x];
                constexpr int V = 16 / sizeof(T);
                vec = vec && al16(g.X) && al16(g.Y) && g.ldx % V == 0 && g.ldy % V == 0;
                ++gt;
            }
            if (gt > start) H.launches.push_back(Launch{s, 4 + nw, (int)start, (int)(gt - start), 0, vec});
        }
    }
    // a perturbed task order within every launch (testing: a task's result
    // must not depend on which work-group runs it or when)
    if (order > 0) {
        uint64_t state = 0x9E3779B97F4A7C15ULL * (uint64_t)order;
        for (const Launch &l : H.launches) {
            if (l.cls < 4) shuffle_range(ptasks, (size_t)l.start, (size_t)l.count, state);
            else shuffle_range(gtasks, (size_t)l.start, (size_t)l.count, state);
        }
        shuffle_range(blk, 0, blk.size(), state);
    }
    return NLA_OK;
}

template <typename T>
static int32_t build(char side, char uplo, char trans, char func, const nla_batch_item *items, int64_t count,
                     nla_batch_plan_s *P, hipStream_t stream) {
    const Layout &L = P->L;
    const bool trmm = (func == 'M');
    HostPlan<T> H;
    int32_t st = plan_host<T>(side, uplo, trans, func, items, count, L, (uintptr_t)P->ws, batch_task_order(), H);
    if (st != NLA_OK) return st;
    P->tx = H.tx;
    P->rows_m = (side == 'L');
    P->ty = H.ty;
    const std::vector<BatchSys<T>> &sys = H.sys, &pan = H.pan;
    const std::vector<BatchTask> &ptasks = H.ptasks;
    const std::vector<GroupGemm<T>> &gem = H.gem;
    const std::vector<int2> &gtasks = H.gtasks, &blk = H.blk;
    // persistent grids from occupancy
    int dev = 0, cus = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e == hipSuccess) e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    if (e != hipSuccess) return report_hip(e, "batch occupancy");
    auto grid = [&](const void *k, int count_) {
        int per_cu = 0;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k, 256, 0) != hipSuccess || per_cu < 1) per_cu = 1;
        return std::max(1, std::min(cus * per_cu, count_));
    };
    P->launches = H.launches;
    for (Launch &l : P->launches)
        l.groups = l.cls < 4 ? grid(batch_kernel_ptr<T>(trmm, BATCH_CLASSES[l.cls]), l.count)
                             : grid(grouped_gemm_ptr<T>(P->tx, P->ty, l.vec, P->rows_m, l.cls == 5), l.count);
    // one staged image of the descriptor region, copied in stream order
    P->host.assign((size_t)(L.inv - L.sys), 0);
    char *h = P->host.data() - L.sys;
    std::memcpy(h + L.sys, sys.data(), sys.size() * sizeof(BatchSys<T>));
    if (!pan.empty()) std::memcpy(h + L.pan, pan.data(), pan.size() * sizeof(BatchSys<T>));
    if (!ptasks.empty()) std::memcpy(h + L.ptasks, ptasks.data(), ptasks.size() * sizeof(BatchTask));
    if (!gem.empty()) std::memcpy(h + L.gem, gem.data(), gem.size() * sizeof(GroupGemm<T>));
    if (!gtasks.empty()) std::memcpy(h + L.gtasks, gtasks.data(), gtasks.size() * sizeof(int2));
    if (!blk.empty()) std::memcpy(h + L.blk, blk.data(), blk.size() * sizeof(int2));
    // stream-ordered: no wait on other streams; the image lives in the plan
    e = hipMemcpyAsync(P->ws + L.sys, P->host.data(), P->host.size(), hipMemcpyHostToDevice, stream);
    if (e != hipSuccess) return report_hip(e, "batch plan upload");
    P->nblk = (int)blk.size();
    return NLA_OK;
}

template <typename T>
static hipError_t enqueue(nla_batch_plan_s *P, hipStream_t s) {
    const Layout &L = P->L;
    hipError_t e = hipSuccess;
    if (!P->trmm) e = launch_batch_trtri<T>((const BatchSys<T> *)(P->ws + L.sys), (const int2 *)(P->ws + L.blk),
                                            P->nblk, s);
    const BatchSys<T> *pan = (const BatchSys<T> *)(P->ws + L.pan);
    const BatchTask *ptasks = (const BatchTask *)(P->ws + L.ptasks);
    const GroupGemm<T> *gem = (const GroupGemm<T> *)(P->ws + L.gem);
    const int2 *gtasks = (const int2 *)(P->ws + L.gtasks);
    int *counters = (int *)(P->ws + L.counter);
    for (const Launch &l : P->launches) {
        if (e != hipSuccess) break;
        int *ctr = counters + l.stage * 6 + l.cls;
        if (l.cls < 4)
            e = launch_batch<T>(pan, ptasks + l.start, l.count, ctr, P->trmm, BATCH_CLASSES[l.cls], l.groups, s);
        else
            e = launch_grouped_gemm<T>(gem, gtasks + l.start, l.count, ctr, P->tx, P->ty, l.vec, P->rows_m,
                                       l.cls == 5, l.groups, s);
    }
    return e;
}

}  // namespace nla

using namespace nla;

extern "C" {

int64_t nla_batch_workspace_bytes(char func, int32_t dtype, const nla_batch_item *items, int64_t count) {
    func = upc(func);
    if ((func != 'S' && func != 'M') || (dtype != NLA_F32 && dtype != NLA_F64) || count < 0 || (count && !items))
        return -1;
    return layout_of(func, dtype, items, count, batch_class_mask()).total;
}

int32_t nla_batch_plan_create(char side, char uplo, char trans, char func, int32_t dtype, const nla_batch_item *items,
                              int64_t count, void *workspace, int64_t workspace_bytes, void *stream,
                              nla_batch_plan *plan) {
    side = upc(side); uplo = upc(uplo); trans = upc(trans); func = upc(func);
    if (side != 'L' && side != 'R') return report_bad_arg(1, "side");
    if (uplo != 'L' && uplo != 'U') return report_bad_arg(2, "uplo");
    if (trans != 'N' && trans != 'T' && trans != 'C') return report_bad_arg(3, "trans");
    if (func != 'S' && func != 'M') return report_bad_arg(4, "func");
    if (dtype != NLA_F32 && dtype != NLA_F64) return report_bad_arg(5, "dtype");
    if (count < 0 || (count > 0 && items == nullptr)) return report_bad_arg(6, "items");
    if (count > (int64_t)1 << 30) return report_bad_arg(7, "count");
    for (int64_t i = 0; i < count; ++i) {
        const nla_batch_item &it = items[i];
        if (it.n < 0 || it.n > BATCH_N_MAX || it.m < 0) return report_bad_arg(6, "items (n must be in [0, 65536])");
        if (it.n == 0 || it.m == 0) continue;
        if (!it.A || !it.B || it.lda < it.n || it.ldb < (side == 'L' ? it.n : it.m))
            return report_bad_arg(6, "items (pointer or leading dimension)");
    }
    if (plan == nullptr) return report_bad_arg(11, "plan");
    Layout L = layout_of(func, dtype, items, count, batch_class_mask());
    if (workspace == nullptr || workspace_bytes < L.total) return report_bad_arg(9, "workspace_bytes");
    if (((uintptr_t)workspace & 255) != 0) return report_bad_arg(8, "workspace (needs 256-byte alignment)");
    nla_batch_plan_s *P = new (std::nothrow) nla_batch_plan_s();
    if (!P) return report_hip(hipErrorOutOfMemory, "batch plan");
    P->dtype = dtype;
    P->trmm = (func == 'M');
    P->ws = (char *)workspace;
    P->L = L;
    hipStream_t s = (hipStream_t)stream;
    int32_t st = dtype == NLA_F64 ? build<double>(side, uplo, trans, func, items, count, P, s)
                                  : build<float>(side, uplo, trans, func, items, count, P, s);
    if (st != NLA_OK) {
        delete P;
        return st;
    }
    *plan = P;
    return NLA_OK;
}

int64_t nla_batch_plan_describe(char side, char uplo, char trans, char func, int32_t dtype, const nla_batch_item *items,
                                int64_t count, nla_batch_task_desc *out, int64_t capacity) {
    side = upc(side); uplo = upc(uplo); trans = upc(trans); func = upc(func);
    if (side != 'L' && side != 'R') return report_bad_arg(1, "side");
    if (uplo != 'L' && uplo != 'U') return report_bad_arg(2, "uplo");
    if (trans != 'N' && trans != 'T' && trans != 'C') return report_bad_arg(3, "trans");
    if (func != 'S' && func != 'M') return report_bad_arg(4, "func");
    if (dtype != NLA_F32 && dtype != NLA_F64) return report_bad_arg(5, "dtype");
    if (count < 0 || (count > 0 && items == nullptr)) return report_bad_arg(6, "items");
    if (count > (int64_t)1 << 30) return report_bad_arg(7, "count");
    for (int64_t i = 0; i < count; ++i)
        if (items[i].n < 0 || items[i].n > BATCH_N_MAX || items[i].m < 0)
            return report_bad_arg(6, "items (n must be in [0, 65536])");
    if (capacity < 0 || (capacity > 0 && out == nullptr)) return report_bad_arg(8, "out");
    const Layout L = layout_of(func, dtype, items, count, batch_class_mask());
    auto describe = [&](auto tag) -> int64_t {
        typedef decltype(tag) T;
        HostPlan<T> H;
        const int32_t st = plan_host<T>(side, uplo, trans, func, items, count, L, 0, batch_task_order(), H);
        if (st != NLA_OK) return st > 0 ? -(int64_t)st : (int64_t)st;
        int64_t k = 0;
        auto put = [&](const nla_batch_task_desc &d) {
            if (k < capacity) out[k] = d;
            ++k;
        };
        const int first = func == 'S' ? 1 : 0;
        for (const int2 &b : H.blk) {
            const int64_t n = items[b.x].n, r0 = (int64_t)b.y * 64;
            put(nla_batch_task_desc{0, 0, b.x, r0, std::min<int64_t>(64, n - r0), 0, 0, 0, 0, 0.0, 0.0});
        }
        for (size_t li = 0; li < H.launches.size(); ++li) {
            const Launch &l = H.launches[li];
            for (int t = l.start; t < l.start + l.count; ++t) {
                nla_batch_task_desc d{};
                d.launch = (int32_t)li + first;
                if (l.cls < 4) {
                    const BatchTask &bt = H.ptasks[(size_t)t];
                    const BatchSys<T> &P = H.pan[(size_t)bt.sys];
                    d.kind = 1;
                    d.system = H.pan_item[(size_t)bt.sys];
                    d.r0 = H.pan_r0[(size_t)bt.sys];
                    d.rows = P.n;
                    d.c0 = (int64_t)bt.strip * 16;
                    d.cols = std::min<int64_t>(16, P.m - d.c0);
                    d.alpha = (double)P.scale;
                } else {
                    const int2 &gt = H.gtasks[(size_t)t];
                    const GroupGemm<T> &g = H.gem[(size_t)gt.x];
                    const int64_t tm = gt.y % g.tiles_m, tn = gt.y / g.tiles_m;
                    const bool nar = H.gem_narrow[(size_t)gt.x] != 0;
                    const int64_t em = group_tile_m(side == 'L', (int)sizeof(T), nar);
                    const int64_t en = group_tile_n(side == 'L', (int)sizeof(T), nar);
                    const int64_t a0 = tm * em, b0 = tn * en;
                    const int64_t an = std::min<int64_t>(em, g.M - a0), bn = std::min<int64_t>(en, g.N - b0);
                    d.kind = 2;
                    d.system = H.gem_item[(size_t)gt.x];
                    if (side == 'L') {  // tile rows along R, columns along the RHS
                        d.r0 = H.gem_r0[(size_t)gt.x] + a0;
                        d.rows = an;
                        d.c0 = b0;
                        d.cols = bn;
                    } else {  // transposed: tile rows along the RHS, columns along R
                        d.c0 = a0;
                        d.cols = an;
                        d.r0 = H.gem_r0[(size_t)gt.x] + b0;
                        d.rows = bn;
                    }
                    d.k0 = H.gem_k0[(size_t)gt.x];
                    d.kn = H.gem_kn[(size_t)gt.x];
                    d.alpha = (double)g.alpha;
                    d.beta = (double)g.beta;
                }
                put(d);
            }
        }
        return k;
    };
    return dtype == NLA_F64 ? describe(double{}) : describe(float{});
}

int32_t nla_batch_plan_execute(nla_batch_plan plan, void *stream, int32_t use_graph) {
    if (plan == nullptr) return report_bad_arg(1, "plan");
    hipStream_t s = (hipStream_t)stream;
    if (plan->launches.empty()) return NLA_OK;
    auto enq = [&](hipStream_t q) {
        return plan->dtype == NLA_F64 ? enqueue<double>(plan, q) : enqueue<float>(plan, q);
    };
    if (!use_graph) return report_hip(enq(s), "batch launch");
    if (plan->exec == nullptr) {
        hipStream_t cs = s;
        bool own = false;
        hipError_t e = hipSuccess;
        if (cs == nullptr) {  // the null stream cannot be captured
            e = hipStreamCreateWithFlags(&cs, hipStreamNonBlocking);
            own = true;
        }
        if (e == hipSuccess) e = hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal);
        if (e == hipSuccess) {
            hipError_t le = enq(cs);
            e = hipStreamEndCapture(cs, &plan->graph);
            if (le != hipSuccess) e = le;
        }
        if (e == hipSuccess) e = hipGraphInstantiate(&plan->exec, plan->graph, nullptr, nullptr, 0);
        if (own) (void)hipStreamDestroy(cs);
        if (e != hipSuccess) return report_hip(e, "batch graph capture");
    }
    return report_hip(hipGraphLaunch(plan->exec, s), "batch graph launch");
}

int32_t nla_batch_plan_destroy(nla_batch_plan plan) {
    if (plan == nullptr) return NLA_OK;
    if (plan->exec) (void)hipGraphExecDestroy(plan->exec);
    if (plan->graph) (void)hipGraphDestroy(plan->graph);
    delete plan;
    return NLA_OK;
}

}  // extern "C"
