// Batched triangular solves / multiplies: host-side planning for the
// persistent panel kernel (nla_batch_* in include/nla.h).
//
// A plan puts every system into the panel kernel's canonical form (the same
// transformation nla_rectrxm applies to one system: op(A) as a strided view
// with flipped uplo for 'T'/'C'; right-side systems solved on the transpose),
// lists one task per (system, 16-column strip), grouped by size class (the
// LDS strip of the kernel that solves a class fits its largest system, so
// small systems run at higher occupancy) and longest first within a class
// (the cost of a strip grows with the square of the number of 64-blocks), and
// uploads the descriptors into the caller's workspace.  Executing a plan
// launches the inversion kernel (TRSM) and, per non-empty class, a counter
// reset and the persistent solver, on the caller's stream and nothing else,
// so it can be captured in a hipGraph and replayed.
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

struct Layout {
    int64_t counter, sys, tasks, blk, inv, total;
    int64_t ntasks, nblk;
};

static Layout layout_of(char func, int32_t dtype, const nla_batch_item *items, int64_t count) {
    const int64_t es = dtype == NLA_F64 ? 8 : 4;
    const int64_t sys_bytes = dtype == NLA_F64 ? sizeof(BatchSys<double>) : sizeof(BatchSys<float>);
    Layout L{};
    for (int64_t i = 0; i < count; ++i) {
        int64_t n = items[i].n, m = items[i].m;
        if (n <= 0 || m <= 0) continue;
        L.ntasks += (m + 15) / 16;
        if (func == 'S') L.nblk += (n + 63) / 64;
    }
    L.counter = 0;
    L.sys = align256(256);
    L.tasks = align256(L.sys + count * sys_bytes);
    L.blk = align256(L.tasks + L.ntasks * (int64_t)sizeof(BatchTask));
    L.inv = align256(L.blk + L.nblk * (int64_t)sizeof(int2));
    L.total = align256(L.inv + L.nblk * 64 * 64 * es);
    return L;
}

}  // namespace nla

struct nla_batch_plan_s {
    std::vector<char> host;  // staged descriptors, alive until the upload has run
    int32_t dtype;
    bool trmm;
    char *ws;
    int ntasks, nblk;
    // per size class (nla::BATCH_CLASSES): first task, task count, grid
    int cls_nbig[4], cls_start[4], cls_count[4], cls_groups[4];
    nla::Layout L;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    hipStream_t graph_stream = nullptr;
};

namespace nla {

template <typename T>
static int32_t build(char side, char uplo, char trans, char func, const nla_batch_item *items, int64_t count,
                     nla_batch_plan_s *P, hipStream_t stream) {
    const Layout &L = P->L;
    std::vector<BatchSys<T>> sys((size_t)count);
    std::vector<BatchTask> tasks;
    std::vector<int2> blk;
    tasks.reserve((size_t)L.ntasks);
    blk.reserve((size_t)L.nblk);
    T *inv = (T *)(P->ws + L.inv);
    int64_t inv_off = 0;
    const bool tr = (trans != 'N');
    const char ue = tr ? (uplo == 'L' ? 'U' : 'L') : uplo;  // uplo of op(A)
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
            const int64_t nb = (it.n + 63) / 64;
            for (int64_t b = 0; b < nb; ++b) blk.push_back(make_int2((int)i, (int)b));
            inv_off += nb * 64 * 64;
        }
        for (int64_t s = 0; s < (it.m + 15) / 16; ++s) tasks.push_back(BatchTask{(int)i, (int)s});
    }
    // Size class of a system: the smallest enabled class that holds it (the
    // largest class is always enabled).  Tasks are grouped by class, largest
    // class first, and longest first within a class (the cost of a strip
    // grows with the number of 64-block phases).
    const int mask = batch_class_mask();
    auto cls_of = [&](int n) {
        for (int c = 0; c < 4; ++c)
            if (((mask >> c) & 1 || c == 3) && n <= BATCH_CLASSES[c]) return c;
        return 3;
    };
    std::stable_sort(tasks.begin(), tasks.end(), [&](const BatchTask &a, const BatchTask &b) {
        const int na = sys[(size_t)a.sys].n, nb2 = sys[(size_t)b.sys].n;
        const int ca = cls_of(na), cb = cls_of(nb2);
        if (ca != cb) return ca > cb;
        const int pa = (na + 63) / 64, pb = (nb2 + 63) / 64;
        return pa * (pa + 1) > pb * (pb + 1);
    });
    for (int c = 0; c < 4; ++c) {
        P->cls_nbig[c] = BATCH_CLASSES[c];
        P->cls_start[c] = 0;
        P->cls_count[c] = 0;
        P->cls_groups[c] = 0;
    }
    for (size_t k = 0; k < tasks.size(); ++k) {
        const int c = cls_of(sys[(size_t)tasks[k].sys].n);
        if (P->cls_count[c]++ == 0) P->cls_start[c] = (int)k;
    }
    // one staged image of the descriptor region, copied in stream order
    P->host.assign((size_t)(L.inv - L.sys), 0);
    std::memcpy(P->host.data(), sys.data(), sys.size() * sizeof(BatchSys<T>));
    if (!tasks.empty())
        std::memcpy(P->host.data() + (L.tasks - L.sys), tasks.data(), tasks.size() * sizeof(BatchTask));
    if (!blk.empty()) std::memcpy(P->host.data() + (L.blk - L.sys), blk.data(), blk.size() * sizeof(int2));
    // stream-ordered: no wait on other streams; the image lives in the plan
    hipError_t e = hipMemcpyAsync(P->ws + L.sys, P->host.data(), P->host.size(), hipMemcpyHostToDevice, stream);
    if (e != hipSuccess) return report_hip(e, "batch plan upload");
    P->ntasks = (int)tasks.size();
    P->nblk = (int)blk.size();
    // persistent grid per class: every CU filled to that kernel's occupancy
    int dev = 0, cus = 0;
    e = hipGetDevice(&dev);
    if (e == hipSuccess) e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    if (e != hipSuccess) return report_hip(e, "batch occupancy");
    for (int c = 0; c < 4; ++c) {
        if (P->cls_count[c] == 0) continue;
        int per_cu = 0;
        e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, batch_kernel_ptr<T>(func == 'M', BATCH_CLASSES[c]),
                                                         256, 0);
        if (e != hipSuccess) return report_hip(e, "batch occupancy");
        const int g = std::max(1, cus * std::max(1, per_cu));
        P->cls_groups[c] = std::min(g, P->cls_count[c]);
    }
    return NLA_OK;
}

template <typename T>
static hipError_t enqueue(nla_batch_plan_s *P, hipStream_t s) {
    const Layout &L = P->L;
    const BatchSys<T> *sys = (const BatchSys<T> *)(P->ws + L.sys);
    hipError_t e = hipSuccess;
    if (!P->trmm) e = launch_batch_trtri<T>(sys, (const int2 *)(P->ws + L.blk), P->nblk, s);
    if (e != hipSuccess) return e;
    const BatchTask *tasks = (const BatchTask *)(P->ws + L.tasks);
    int *counters = (int *)(P->ws + L.counter);
    for (int c = 3; c >= 0 && e == hipSuccess; --c)
        if (P->cls_count[c] > 0)
            e = launch_batch<T>(sys, tasks + P->cls_start[c], P->cls_count[c], counters + c, P->trmm, P->cls_nbig[c],
                                P->cls_groups[c], s);
    return e;
}

}  // namespace nla

using namespace nla;

extern "C" {

int64_t nla_batch_workspace_bytes(char func, int32_t dtype, const nla_batch_item *items, int64_t count) {
    func = upc(func);
    if ((func != 'S' && func != 'M') || (dtype != NLA_F32 && dtype != NLA_F64) || count < 0 || (count && !items))
        return -1;
    return layout_of(func, dtype, items, count).total;
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
        if (it.n < 0 || it.n > BATCH_MAX || it.m < 0) return report_bad_arg(6, "items (n must be in [0, 1024])");
        if (it.n == 0 || it.m == 0) continue;
        if (!it.A || !it.B || it.lda < it.n || it.ldb < (side == 'L' ? it.n : it.m))
            return report_bad_arg(6, "items (pointer or leading dimension)");
    }
    if (plan == nullptr) return report_bad_arg(11, "plan");
    Layout L = layout_of(func, dtype, items, count);
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

int32_t nla_batch_plan_execute(nla_batch_plan plan, void *stream, int32_t use_graph) {
    if (plan == nullptr) return report_bad_arg(1, "plan");
    hipStream_t s = (hipStream_t)stream;
    if (plan->ntasks == 0) return NLA_OK;
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
