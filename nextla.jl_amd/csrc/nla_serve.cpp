// Continuous-batching request server behind the C ABI (nla_serve_* in
// include/nla.h).  The reference serves one unified_rectrxm! call at a time
// (test/unified_rectrxm.jl:33); here requests whose A and B already live in
// HBM are
//   - admitted against an HBM budget (B, A unless it is a resident factor,
//     and the batched workspace the request needs),
//   - grouped by signature (dtype, side, uplo, trans, func) into one batched
//     plan (nla_batch_*: a fixed number of launches whatever the sizes),
//   - run on one of several in-flight streams, so requests that arrive while
//     others run overlap with them on the device,
//   - completed by polling each batch's event; the caller collects
//     (ticket, status) records from nla_serve_step.
// FP64 TRSM requests can instead be served by batched refinement
// (nla_batch_mixed_*: fp32 correction solves and FP64 residuals to a
// per-system backward error): those batches run on worker threads, each
// with its own stream, because the refinement reads its stopping test on
// the host once per step.
//
// Ordering with the submitter: nla_serve_submit records an event on the
// submitter's stream, and the stream that runs the request waits on it, so
// work already queued there (writing A or B) comes first.  The server never
// owns A or B: the caller keeps them allocated until the request completes.
// All nla_serve_* calls on one server must come from one host thread at a
// time; the worker threads are internal.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <condition_variable>
#include <cstdio>
#include <cstring>
#include <deque>
#include <list>
#include <map>
#include <memory>
#include <mutex>
#include <thread>
#include <vector>

#include "../../include/nla.h"
#include "nla_internal.h"

namespace nla {
namespace {

constexpr int64_t SERVE_N_BATCH_MAX = 65536;  // largest system a batched plan takes

struct SrvReq {
    int64_t ticket = 0;
    nla_serve_request r{};
    int64_t bytes = 0;        // charged to the budget while queued work runs
    hipEvent_t ready = nullptr;
    bool mixed = false;
};

int64_t elem_size(int32_t dtype) { return dtype == NLA_F64 ? 8 : 4; }

// A device buffer owned by one stream, grown in stream order.
struct StreamBuf {
    hipStream_t s = nullptr;
    void *p = nullptr;
    int64_t bytes = 0;
    hipError_t reserve(int64_t need) {
        if (need <= bytes) return hipSuccess;
        if (p) {
            hipError_t e = hipFreeAsync(p, s);
            if (e != hipSuccess) return e;
            p = nullptr;
            bytes = 0;
        }
        const int64_t b = std::max<int64_t>(need + need / 4, 1 << 20);
        hipError_t e = hipMallocAsync(&p, (size_t)b, s);
        if (e != hipSuccess) return e;
        bytes = b;
        return hipSuccess;
    }
};

struct MixedTask {
    std::vector<SrvReq> reqs;
    std::vector<double> bwd;
    std::vector<int32_t> iters;
    std::vector<int32_t> status;
    std::atomic<int> done{0};
};

struct Batch {
    std::vector<SrvReq> reqs;
    int slot = -1;
    nla_batch_plan plan = nullptr;
    hipEvent_t done = nullptr;
    int64_t bytes = 0;
    std::vector<int32_t> status;  // per request (plain batches)
    std::shared_ptr<MixedTask> task;
};

struct Resident {
    const void *A = nullptr;
    int64_t bytes = 0;
};

}  // namespace
}  // namespace nla

struct nla_server_s {
    int dev = 0;
    int64_t budget = 0, used = 0;
    int max_batch = 1, use_graph = 0, mode = NLA_SERVE_PLAIN, workers_n = 0, mixed_precision = NLA_LP_FP16;
    double tol = 1e-14;
    int max_iter = 10;
    int64_t next_ticket = 1;
    std::deque<nla::SrvReq> queue;
    std::list<nla::Batch> inflight;
    std::vector<nla::StreamBuf> slots;
    std::vector<int> free_slots;
    std::map<int64_t, nla::Resident> resident;
    std::list<int64_t> lru;  // resident keys, least recently used first
    std::deque<nla_serve_done> completed;
    std::deque<int64_t> evicted;
    // refinement workers
    std::vector<std::thread> workers;
    std::mutex mu;
    std::condition_variable cv;
    std::deque<std::shared_ptr<nla::MixedTask>> tasks;
    bool stop = false;
    int busy_tasks = 0;
};

namespace nla {
namespace {

// Runs `f` with the server's device current, restoring the caller's.
template <typename F>
int32_t on_device(nla_server_s *S, F f) {
    int old = 0;
    if (hipGetDevice(&old) != hipSuccess) old = S->dev;
    if (old != S->dev && hipSetDevice(S->dev) != hipSuccess) return report_status(NLA_ERR_HIP, "server device");
    const int32_t st = f();
    if (old != S->dev) (void)hipSetDevice(old);
    return st;
}

bool sig_equal(const nla_serve_request &a, const nla_serve_request &b) {
    return a.dtype == b.dtype && a.side == b.side && a.uplo == b.uplo && a.trans == b.trans && a.func == b.func;
}

// Bytes a request holds while it runs: B, A unless resident, the batched
// workspace's share (TRSM inverse blocks; refinement arrays).
int64_t request_bytes(const nla_serve_request &r, bool mixed, bool a_resident) {
    const int64_t es = elem_size(r.dtype);
    const int64_t rows = r.side == 'L' ? r.n : r.m, cols = r.side == 'L' ? r.m : r.n;
    int64_t b = rows * cols * es + 4096;
    if (!a_resident) b += r.n * r.n * es;
    if (mixed)
        b += 4 * r.n * r.n + 20 * r.n * r.m + 256 * r.n;
    else if (r.func == 'S')
        b += (r.n + 63) / 64 * 64 * 64 * es;
    return b;
}

void finish(nla_server_s *S, const SrvReq &q, int32_t status, int32_t iters, double bwd) {
    nla_serve_done d{};
    d.ticket = q.ticket;
    d.status = status;
    d.iterations = iters;
    d.backward_error = bwd;
    S->completed.push_back(d);
}

void release_ready(SrvReq &q) {
    if (q.ready) {
        (void)hipEventDestroy(q.ready);
        q.ready = nullptr;
    }
}

// Resident factors: touched on use; evicted (least recently used first) when
// no queued or running request refers to them and the budget needs room.
bool key_in_use(nla_server_s *S, int64_t key) {
    for (const auto &q : S->queue)
        if (q.r.a_key == key) return true;
    for (const auto &b : S->inflight)
        for (const auto &q : b.reqs)
            if (q.r.a_key == key) return true;
    return false;
}

void evict(nla_server_s *S, int64_t need) {
    for (auto it = S->lru.begin(); it != S->lru.end() && S->used + need > S->budget;) {
        const int64_t key = *it;
        if (key_in_use(S, key)) {
            ++it;
            continue;
        }
        S->used -= S->resident[key].bytes;
        S->resident.erase(key);
        S->evicted.push_back(key);
        it = S->lru.erase(it);
    }
}

void touch(nla_server_s *S, int64_t key) {
    auto it = std::find(S->lru.begin(), S->lru.end(), key);
    if (it != S->lru.end()) S->lru.splice(S->lru.end(), S->lru, it);
}

nla_batch_item item_of(const nla_serve_request &r) {
    nla_batch_item it{};
    it.n = r.n;
    it.m = r.m;
    it.A = r.A;
    it.lda = r.lda;
    it.B = r.B;
    it.ldb = r.ldb;
    it.alpha = r.alpha;
    return it;
}

// Takes the queued requests that share the first eligible request's
// signature and fit the budget (at most max_batch); 1 when some were taken,
// 0 when none could be, -1 when the first one can never fit (it is failed).
int take_batch(nla_server_s *S, bool mixed, std::vector<SrvReq> &take, int64_t &bytes) {
    auto head = std::find_if(S->queue.begin(), S->queue.end(), [&](const SrvReq &q) { return q.mixed == mixed; });
    if (head == S->queue.end()) return 0;
    const nla_serve_request sig = head->r;
    std::deque<SrvReq> keep;
    bytes = 0;
    while (!S->queue.empty() && (int)take.size() < S->max_batch) {
        SrvReq q = S->queue.front();
        S->queue.pop_front();
        if (q.mixed == mixed && sig_equal(q.r, sig) && S->used + bytes + q.bytes <= S->budget) {
            bytes += q.bytes;
            take.push_back(q);
        } else {
            keep.push_back(q);
        }
    }
    while (!keep.empty()) {
        S->queue.push_front(keep.back());
        keep.pop_back();
    }
    if (!take.empty()) return 1;
    head = std::find_if(S->queue.begin(), S->queue.end(), [&](const SrvReq &q) { return q.mixed == mixed; });
    if (S->inflight.empty()) {
        // nothing running will free memory: fail a request that cannot fit
        // even with every unused resident factor evicted (and evict none)
        int64_t evictable = 0;
        for (const auto &kv : S->resident)
            if (!key_in_use(S, kv.first)) evictable += kv.second.bytes;
        if (S->used - evictable + head->bytes > S->budget) {
            SrvReq q = *head;
            S->queue.erase(head);
            release_ready(q);
            finish(S, q, NLA_ERR_WORKSPACE, 0, 0.0);
            return -1;
        }
    }
    evict(S, head->bytes);
    return 0;
}

int32_t launch_plain(nla_server_s *S, bool &launched) {
    launched = false;
    if (S->free_slots.empty()) return NLA_OK;
    std::vector<SrvReq> take;
    int64_t bytes = 0;
    const int t = take_batch(S, false, take, bytes);
    if (t < 0) {
        launched = true;  // progress: a request was failed
        return NLA_OK;
    }
    if (t == 0) return NLA_OK;
    const int slot = S->free_slots.back();
    S->free_slots.pop_back();
    StreamBuf &sb = S->slots[slot];
    Batch b;
    b.slot = slot;
    b.bytes = bytes;
    b.status.assign(take.size(), NLA_OK);
    for (auto &q : take) {
        if (q.ready) (void)hipStreamWaitEvent(sb.s, q.ready, 0);
        release_ready(q);
    }
    const nla_serve_request &r0 = take[0].r;
    std::vector<nla_batch_item> items;
    std::vector<size_t> idx;
    for (size_t i = 0; i < take.size(); ++i) {
        if (take[i].r.n <= SERVE_N_BATCH_MAX) {
            items.push_back(item_of(take[i].r));
            idx.push_back(i);
        }
    }
    if (!items.empty()) {
        const int64_t need = nla_batch_workspace_bytes(r0.func, r0.dtype, items.data(), (int64_t)items.size());
        int32_t st = need < 0 ? NLA_ERR_UNSUPPORTED : NLA_OK;
        if (st == NLA_OK && sb.reserve(std::max<int64_t>(need, 256)) != hipSuccess)
            st = report_status(NLA_ERR_HIP, "server batch workspace");
        if (st == NLA_OK)
            st = nla_batch_plan_create(r0.side, r0.uplo, r0.trans, r0.func, r0.dtype, items.data(),
                                       (int64_t)items.size(), sb.p, sb.bytes, sb.s, &b.plan);
        if (st == NLA_OK) st = nla_batch_plan_execute(b.plan, sb.s, S->use_graph);
        if (st != NLA_OK)
            for (size_t i : idx) b.status[i] = st;
    }
    for (size_t i = 0; i < take.size(); ++i) {
        const nla_serve_request &r = take[i].r;
        if (r.n <= SERVE_N_BATCH_MAX) continue;
        b.status[i] = nla_rectrxm(r.side, r.uplo, r.trans, r.func, r.dtype, r.alpha, r.n, r.m, r.A, r.lda, r.B, r.ldb,
                                  sb.s);
    }
    if (hipEventCreateWithFlags(&b.done, hipEventDisableTiming) != hipSuccess ||
        hipEventRecord(b.done, sb.s) != hipSuccess) {
        (void)hipStreamSynchronize(sb.s);  // no event: complete it now
        if (b.done) (void)hipEventDestroy(b.done);
        b.done = nullptr;
    }
    b.reqs = std::move(take);
    S->used += bytes;
    for (const auto &q : b.reqs)
        if (q.r.a_key > 0) touch(S, q.r.a_key);
    S->inflight.push_back(std::move(b));
    launched = true;
    return NLA_OK;
}

int32_t launch_mixed(nla_server_s *S, bool &launched) {
    launched = false;
    if (S->workers.empty() || S->busy_tasks >= S->workers_n) return NLA_OK;
    std::vector<SrvReq> take;
    int64_t bytes = 0;
    const int t = take_batch(S, true, take, bytes);
    if (t < 0) {
        launched = true;
        return NLA_OK;
    }
    if (t == 0) return NLA_OK;
    auto task = std::make_shared<MixedTask>();
    task->reqs = take;
    task->bwd.assign(take.size(), 0.0);
    task->iters.assign(take.size(), 0);
    task->status.assign(take.size(), NLA_OK);
    Batch b;
    b.bytes = bytes;
    b.task = task;
    b.reqs = std::move(take);
    for (auto &q : b.reqs) q.ready = nullptr;  // the worker owns (and releases) the events
    S->used += bytes;
    for (const auto &q : b.reqs)
        if (q.r.a_key > 0) touch(S, q.r.a_key);
    S->inflight.push_back(std::move(b));
    {
        std::lock_guard<std::mutex> g(S->mu);
        S->tasks.push_back(task);
    }
    S->cv.notify_one();
    ++S->busy_tasks;
    launched = true;
    return NLA_OK;
}

// One refinement worker: its own stream and workspace.
void worker_main(nla_server_s *S) {
    (void)hipSetDevice(S->dev);
    StreamBuf sb;
    if (hipStreamCreateWithFlags(&sb.s, hipStreamNonBlocking) != hipSuccess) sb.s = nullptr;
    for (;;) {
        std::shared_ptr<MixedTask> task;
        {
            std::unique_lock<std::mutex> g(S->mu);
            S->cv.wait(g, [&] { return S->stop || !S->tasks.empty(); });
            if (S->tasks.empty()) break;
            task = S->tasks.front();
            S->tasks.pop_front();
        }
        const size_t nr = task->reqs.size();
        for (auto &q : task->reqs) {
            if (q.ready && sb.s) (void)hipStreamWaitEvent(sb.s, q.ready, 0);
            release_ready(q);
        }
        if (!sb.s) {
            std::fill(task->status.begin(), task->status.end(), (int32_t)NLA_ERR_HIP);
            task->done.store(1);
            continue;
        }
        const nla_serve_request &r0 = task->reqs[0].r;
        std::vector<nla_batch_item> items;
        std::vector<size_t> idx;
        for (size_t i = 0; i < nr; ++i)
            if (task->reqs[i].r.n <= SERVE_N_BATCH_MAX) {
                items.push_back(item_of(task->reqs[i].r));
                idx.push_back(i);
            }
        if (!items.empty()) {
            const int64_t need = nla_batch_mixed_workspace_bytes(r0.side, items.data(), (int64_t)items.size());
            int32_t st = need < 0 ? NLA_ERR_UNSUPPORTED : NLA_OK;
            if (st == NLA_OK && sb.reserve(std::max<int64_t>(need, 256)) != hipSuccess) st = NLA_ERR_HIP;
            nla_batch_mixed_plan plan = nullptr;
            if (st == NLA_OK)
                st = nla_batch_mixed_plan_create(r0.side, r0.uplo, r0.trans, items.data(), (int64_t)items.size(), sb.p,
                                                 sb.bytes, sb.s, &plan);
            std::vector<double> bwd(items.size(), 0.0);
            std::vector<int32_t> its(items.size(), 0);
            if (st == NLA_OK) st = nla_batch_mixed_plan_execute(plan, S->tol, S->max_iter, sb.s, 0, bwd.data(), its.data());
            if (plan) nla_batch_mixed_plan_destroy(plan);
            for (size_t k = 0; k < idx.size(); ++k) {
                task->status[idx[k]] = st;
                task->bwd[idx[k]] = bwd[k];
                task->iters[idx[k]] = its[k];
            }
        }
        for (size_t i = 0; i < nr; ++i) {  // beyond the batched size limit: one refined call each
            const nla_serve_request &r = task->reqs[i].r;
            if (r.n <= SERVE_N_BATCH_MAX) continue;
            const int64_t need = nla_mixed_workspace_bytes(r.side, r.n, r.m);
            int32_t st = sb.reserve(std::max<int64_t>(need, 256)) == hipSuccess ? NLA_OK : NLA_ERR_HIP;
            nla_mixed_info info{};
            if (st == NLA_OK)
                st = nla_rectrxm_mixed(r.side, r.uplo, r.trans, S->mixed_precision, r.alpha, r.n, r.m,
                                       (const double *)r.A, r.lda, (double *)r.B, r.ldb, S->tol, S->max_iter, sb.p,
                                       sb.bytes, sb.s, &info);
            task->status[i] = st;
            task->bwd[i] = info.backward_error;
            task->iters[i] = info.iterations;
        }
        if (hipStreamSynchronize(sb.s) != hipSuccess)
            for (auto &st : task->status)
                if (st == NLA_OK) st = NLA_ERR_HIP;
        task->done.store(1, std::memory_order_release);
    }
    if (sb.p) (void)hipFreeAsync(sb.p, sb.s);
    if (sb.s) {
        (void)hipStreamSynchronize(sb.s);
        (void)hipStreamDestroy(sb.s);
    }
}

int reap(nla_server_s *S) {
    int n = 0;
    for (auto it = S->inflight.begin(); it != S->inflight.end();) {
        Batch &b = *it;
        bool ready;
        if (b.task)
            ready = b.task->done.load(std::memory_order_acquire) != 0;
        else
            ready = b.done == nullptr || hipEventQuery(b.done) == hipSuccess;
        if (!ready) {
            ++it;
            continue;
        }
        for (size_t i = 0; i < b.reqs.size(); ++i) {
            if (b.task)
                finish(S, b.reqs[i], b.task->status[i], b.task->iters[i], b.task->bwd[i]);
            else
                finish(S, b.reqs[i], b.status[i], 0, 0.0);
        }
        if (b.task) {
            --S->busy_tasks;
        } else {
            if (b.plan) nla_batch_plan_destroy(b.plan);
            if (b.done) (void)hipEventDestroy(b.done);
            S->free_slots.push_back(b.slot);
        }
        S->used -= b.bytes;
        n += (int)b.reqs.size();
        it = S->inflight.erase(it);
    }
    return n;
}

int32_t step_once(nla_server_s *S, int *progress) {
    int n = reap(S);
    for (;;) {
        bool a = false, c = false;
        int32_t st = launch_mixed(S, a);
        if (st != NLA_OK) return st;
        st = launch_plain(S, c);
        if (st != NLA_OK) return st;
        if (!a && !c) break;
        ++n;
    }
    if (progress) *progress = n;
    return NLA_OK;
}

}  // namespace
}  // namespace nla

using namespace nla;

extern "C" int32_t nla_serve_create(int64_t budget_bytes, int32_t max_inflight, int32_t max_batch, int32_t fp64_trsm,
                                    int32_t workers, int32_t use_graph, nla_server *out) {
    if (budget_bytes <= 0) return report_bad_arg(1, "budget_bytes");
    if (max_inflight < 1 || max_inflight > 64) return report_bad_arg(2, "max_inflight (1..64)");
    if (max_batch < 1) return report_bad_arg(3, "max_batch");
    if (fp64_trsm != NLA_SERVE_PLAIN && fp64_trsm != NLA_SERVE_REFINED) return report_bad_arg(4, "fp64_trsm");
    if (workers < 0 || workers > 64 || (fp64_trsm == NLA_SERVE_REFINED && workers < 1))
        return report_bad_arg(5, "workers (1..64 for refined serving)");
    if (!out) return report_bad_arg(7, "out");
    *out = nullptr;
    auto S = std::make_unique<nla_server_s>();
    if (hipGetDevice(&S->dev) != hipSuccess) return report_status(NLA_ERR_NO_DEVICE, "nla_serve_create: no device");
    S->budget = budget_bytes;
    S->max_batch = max_batch;
    S->use_graph = use_graph ? 1 : 0;
    S->mode = fp64_trsm;
    S->slots.resize(max_inflight);
    for (int i = 0; i < max_inflight; ++i) {
        if (hipStreamCreateWithFlags(&S->slots[i].s, hipStreamNonBlocking) != hipSuccess) {
            for (int j = 0; j < i; ++j) (void)hipStreamDestroy(S->slots[j].s);
            return report_status(NLA_ERR_HIP, "nla_serve_create: stream");
        }
        S->free_slots.push_back(max_inflight - 1 - i);
    }
    if (fp64_trsm == NLA_SERVE_REFINED) {
        S->workers_n = workers;
        for (int i = 0; i < workers; ++i) S->workers.emplace_back(worker_main, S.get());
    }
    *out = S.release();
    return NLA_OK;
}

extern "C" int32_t nla_serve_set_refinement(nla_server S, double tol, int32_t max_iter, int32_t precision) {
    if (!S) return report_bad_arg(1, "server");
    if (!(tol > 0.0)) return report_bad_arg(2, "tol");
    if (max_iter < 0 || max_iter > NLA_MIXED_MAX_ITER) return report_bad_arg(3, "max_iter");
    if (precision < NLA_LP_BF16 || precision > NLA_LP_FP16X2) return report_bad_arg(4, "precision");
    S->tol = tol;
    S->max_iter = max_iter;
    S->mixed_precision = precision;
    return NLA_OK;
}

extern "C" int32_t nla_serve_submit(nla_server S, const nla_serve_request *req, void *stream, int64_t *ticket) {
    if (!S) return report_bad_arg(1, "server");
    if (!req) return report_bad_arg(2, "request");
    if (!ticket) return report_bad_arg(4, "ticket");
    nla_serve_request r = *req;
    char side = r.side, uplo = r.uplo, trans = r.trans, func = r.func;
    const bool resident_key = r.a_key > 0 && S->resident.count(r.a_key) != 0;
    if (resident_key && r.A == nullptr) r.A = S->resident[r.a_key].A;
    const int32_t chk = rectrxm_check(side, uplo, trans, func, r.dtype, r.n, r.m, r.A, r.lda, r.B, r.ldb);
    if (chk != NLA_OK) return chk;
    r.side = side;
    r.uplo = uplo;
    r.trans = trans;
    r.func = func;
    return on_device(S, [&]() -> int32_t {
        SrvReq q;
        q.r = r;
        q.mixed = S->mode == NLA_SERVE_REFINED && r.func == 'S' && r.dtype == NLA_F64 && r.n > 0 && r.m > 0;
        if (r.a_key > 0 && !resident_key) {  // first use of a resident factor: charged once from now on
            Resident res;
            res.A = r.A;
            res.bytes = r.n * r.lda * elem_size(r.dtype);
            S->resident[r.a_key] = res;
            S->lru.push_back(r.a_key);
            S->used += res.bytes;
        }
        q.bytes = request_bytes(r, q.mixed, r.a_key > 0);
        if (hipEventCreateWithFlags(&q.ready, hipEventDisableTiming) != hipSuccess ||
            hipEventRecord(q.ready, (hipStream_t)stream) != hipSuccess) {
            if (q.ready) (void)hipEventDestroy(q.ready);
            return report_status(NLA_ERR_HIP, "nla_serve_submit: ready event");
        }
        q.ticket = S->next_ticket++;
        *ticket = q.ticket;
        if (r.n == 0 || r.m == 0) {  // nothing to compute: complete at once
            release_ready(q);
            finish(S, q, NLA_OK, 0, 0.0);
            return NLA_OK;
        }
        S->queue.push_back(q);
        return NLA_OK;
    });
}

extern "C" int32_t nla_serve_step(nla_server S, nla_serve_done *done, int64_t capacity, int64_t *ndone) {
    if (!S) return report_bad_arg(1, "server");
    if (capacity < 0 || (capacity > 0 && !done)) return report_bad_arg(2, "done");
    if (!ndone) return report_bad_arg(4, "ndone");
    return on_device(S, [&]() -> int32_t {
        int progress = 0;
        const int32_t st = step_once(S, &progress);
        int64_t k = 0;
        while (k < capacity && !S->completed.empty()) {
            done[k++] = S->completed.front();
            S->completed.pop_front();
        }
        *ndone = k;
        if (k == 0 && progress == 0 && !S->inflight.empty() && S->busy_tasks > 0)
            std::this_thread::yield();  // only refinement workers can make progress
        return st;
    });
}

extern "C" int32_t nla_serve_drain(nla_server S) {
    if (!S) return report_bad_arg(1, "server");
    return on_device(S, [&]() -> int32_t {
        while (!S->queue.empty() || !S->inflight.empty()) {
            int progress = 0;
            const int32_t st = step_once(S, &progress);
            if (st != NLA_OK) return st;
            if (progress == 0) std::this_thread::yield();
        }
        return NLA_OK;
    });
}

extern "C" int32_t nla_serve_stats(nla_server S, nla_serve_stats_t *out) {
    if (!S) return report_bad_arg(1, "server");
    if (!out) return report_bad_arg(2, "out");
    out->budget_bytes = S->budget;
    out->used_bytes = S->used;
    out->queued = (int64_t)S->queue.size();
    int64_t running = 0;
    for (const auto &b : S->inflight) running += (int64_t)b.reqs.size();
    out->running = running;
    out->completed_pending = (int64_t)S->completed.size();
    out->resident = (int64_t)S->resident.size();
    return NLA_OK;
}

extern "C" int32_t nla_serve_evicted(nla_server S, int64_t *keys, int64_t capacity, int64_t *count) {
    if (!S) return report_bad_arg(1, "server");
    if (capacity < 0 || (capacity > 0 && !keys)) return report_bad_arg(2, "keys");
    if (!count) return report_bad_arg(4, "count");
    int64_t k = 0;
    while (k < capacity && !S->evicted.empty()) {
        keys[k++] = S->evicted.front();
        S->evicted.pop_front();
    }
    *count = k;
    return NLA_OK;
}

extern "C" int32_t nla_serve_destroy(nla_server S) {
    if (!S) return NLA_OK;
    int32_t st = nla_serve_drain(S);
    {
        std::lock_guard<std::mutex> g(S->mu);
        S->stop = true;
    }
    S->cv.notify_all();
    for (auto &t : S->workers) t.join();
    (void)on_device(S, [&]() -> int32_t {
        for (auto &sb : S->slots) {
            if (sb.p) (void)hipFreeAsync(sb.p, sb.s);
            (void)hipStreamSynchronize(sb.s);
            (void)hipStreamDestroy(sb.s);
        }
        for (auto &q : S->queue) release_ready(q);
        return NLA_OK;
    });
    delete S;
    return st;
}
