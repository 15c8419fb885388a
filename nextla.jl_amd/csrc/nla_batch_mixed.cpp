// Batched FP64-accurate TRSM by fp32 correction solves and FP64 refinement
// (nla_batch_mixed_* in include/nla.h).
//
// A plan is two batched plans over the same systems plus the elementwise
// steps between them:
//   p32: TRSM in fp32 on (A32_i, D_i)  -- the correction solves;
//   p64: TRMM in FP64 on (A_i, T_i)    -- op(A_i) X_i for the residuals.
// Every array a system needs (fp32 copy of its triangle, alpha*B, the
// correction, the product / residual) lives in the caller's workspace, each
// column-major with leading dimensions padded to a multiple of 4 so the
// grouped GEMM of large systems takes its vector path.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <new>
#include <vector>

#include "../../include/nla.h"
#include "nla_internal.h"

namespace nla {

int32_t report_hip(hipError_t e, const char *where);
int32_t report_bad_arg(int k, const char *what);
int32_t report_status(int32_t code, const char *msg);

namespace {

constexpr int64_t TASK_ELEMS = 8192;  // elements per elementwise task (at least one column)

inline char upc(char c) { return (c >= 'a' && c <= 'z') ? (char)(c - 'a' + 'A') : c; }
inline int64_t al256(int64_t x) { return (x + 255) & ~(int64_t)255; }
inline int64_t pad4(int64_t x) { return (x + 3) & ~(int64_t)3; }

struct Shape {
    int64_t rows, cols, ldw, lda32, bcb, bnt, acb, ant;
};

inline Shape shape_of(char side, const nla_batch_item &it) {
    Shape h{};
    if (it.n <= 0 || it.m <= 0) return h;
    h.rows = side == 'L' ? it.n : it.m;
    h.cols = side == 'L' ? it.m : it.n;
    h.ldw = pad4(h.rows);
    h.lda32 = pad4(it.n);
    h.bcb = (int64_t)std::max<int64_t>(1, TASK_ELEMS / h.rows);
    h.bnt = (h.cols + h.bcb - 1) / h.bcb;
    h.acb = (int64_t)std::max<int64_t>(1, TASK_ELEMS / it.n);
    h.ant = (it.n + h.acb - 1) / h.acb;
    return h;
}

struct MLayout {
    int64_t sys, state, btasks, atasks, bpart, apart, arrays, p32, p64, total;
    int64_t nbt, nat, arrays_bytes, p32_bytes, p64_bytes;
};

// Sub-plan items with the array addresses of the layout (base = workspace,
// or nullptr when only sizes are wanted).
void sub_items(char side, bool lower, const nla_batch_item *items, int64_t count, const MLayout &L, char *base,
               std::vector<nla_batch_item> &i32, std::vector<nla_batch_item> &i64, std::vector<MixSys> *sys) {
    i32.assign((size_t)count, nla_batch_item{});
    i64.assign((size_t)count, nla_batch_item{});
    if (sys) sys->assign((size_t)count, MixSys{});
    int64_t off = L.arrays, bt = 0, at = 0;
    for (int64_t i = 0; i < count; ++i) {
        const nla_batch_item &it = items[i];
        const Shape h = shape_of(side, it);
        nla_batch_item &a = i32[(size_t)i], &b = i64[(size_t)i];
        a.n = b.n = std::max<int64_t>(it.n, 0);
        a.m = b.m = std::max<int64_t>(it.m, 0);
        a.alpha = b.alpha = 1.0;
        if (h.rows == 0) continue;
        const int64_t o_a32 = off;
        off = al256(off + h.lda32 * it.n * 4);
        const int64_t o_d = off;
        off = al256(off + h.ldw * h.cols * 4);
        const int64_t o_s = off;
        off = al256(off + h.ldw * h.cols * 8);
        const int64_t o_t = off;
        off = al256(off + h.ldw * h.cols * 8);
        a.A = base ? base + o_a32 : nullptr;
        a.lda = h.lda32;
        a.B = base ? base + o_d : nullptr;
        a.ldb = h.ldw;
        b.A = it.A;
        b.lda = it.lda;
        b.B = base ? base + o_t : nullptr;
        b.ldb = h.ldw;
        if (sys) {
            MixSys &S = (*sys)[(size_t)i];
            S.A = (const double *)it.A;
            S.lda = it.lda;
            S.A32 = (float *)(base + o_a32);
            S.lda32 = h.lda32;
            S.B = (double *)it.B;
            S.ldb = it.ldb;
            S.S = (double *)(base + o_s);
            S.T = (double *)(base + o_t);
            S.D = (float *)(base + o_d);
            S.ldw = h.ldw;
            S.rows = h.rows;
            S.cols = h.cols;
            S.n = it.n;
            S.alpha = it.alpha;
            S.lower = lower ? 1 : 0;
            S.bcb = (int)h.bcb;
            S.acb = (int)h.acb;
            S.bt0 = (int)bt;
            S.bnt = (int)h.bnt;
            S.at0 = (int)at;
            S.ant = (int)h.ant;
        }
        bt += h.bnt;
        at += h.ant;
    }
}

MLayout layout_of(char side, const nla_batch_item *items, int64_t count) {
    MLayout L{};
    int64_t arrays = 0;
    for (int64_t i = 0; i < count; ++i) {
        const Shape h = shape_of(side, items[i]);
        if (h.rows == 0) continue;
        L.nbt += h.bnt;
        L.nat += h.ant;
        arrays = al256(arrays + h.lda32 * items[i].n * 4);
        arrays = al256(arrays + h.ldw * h.cols * 4);
        arrays = al256(arrays + h.ldw * h.cols * 8);
        arrays = al256(arrays + h.ldw * h.cols * 8);
    }
    L.sys = 0;
    L.state = al256(L.sys + count * (int64_t)sizeof(MixSys));
    L.btasks = al256(L.state + count * (int64_t)sizeof(MixState));
    L.atasks = al256(L.btasks + L.nbt * (int64_t)sizeof(int2));
    L.bpart = al256(L.atasks + L.nat * (int64_t)sizeof(int2));
    L.apart = al256(L.bpart + 3 * L.nbt * 8);
    L.arrays = al256(L.apart + L.nat * 8);
    L.arrays_bytes = arrays;
    std::vector<nla_batch_item> i32, i64;
    sub_items(side, true, items, count, L, nullptr, i32, i64, nullptr);
    L.p32_bytes = nla_batch_workspace_bytes('S', NLA_F32, i32.data(), count);
    L.p64_bytes = nla_batch_workspace_bytes('M', NLA_F64, i64.data(), count);
    L.p32 = al256(L.arrays + arrays);
    L.p64 = al256(L.p32 + std::max<int64_t>(L.p32_bytes, 0));
    L.total = al256(L.p64 + std::max<int64_t>(L.p64_bytes, 0));
    return L;
}

}  // namespace
}  // namespace nla

struct nla_batch_mixed_plan_s {
    nla_batch_plan p32 = nullptr, p64 = nullptr;
    char *ws = nullptr;
    int64_t count = 0;
    nla::MLayout L{};
    std::vector<char> host;  // staged descriptors, alive until the upload has run
    std::vector<nla::MixState> state;
};

using namespace nla;

extern "C" {

int64_t nla_batch_mixed_workspace_bytes(char side, const nla_batch_item *items, int64_t count) {
    side = upc(side);
    if ((side != 'L' && side != 'R') || count < 0 || (count && !items)) return -1;
    for (int64_t i = 0; i < count; ++i)
        if (items[i].n < 0 || items[i].n > BATCH_N_MAX || items[i].m < 0) return -1;
    const MLayout L = layout_of(side, items, count);
    if (L.p32_bytes < 0 || L.p64_bytes < 0) return -1;
    return L.total;
}

int32_t nla_batch_mixed_plan_create(char side, char uplo, char trans, const nla_batch_item *items, int64_t count,
                                    void *workspace, int64_t workspace_bytes, void *stream,
                                    nla_batch_mixed_plan *plan) {
    side = upc(side); uplo = upc(uplo); trans = upc(trans);
    if (side != 'L' && side != 'R') return report_bad_arg(1, "side");
    if (uplo != 'L' && uplo != 'U') return report_bad_arg(2, "uplo");
    if (trans != 'N' && trans != 'T' && trans != 'C') return report_bad_arg(3, "trans");
    if (count < 0 || (count > 0 && items == nullptr)) return report_bad_arg(4, "items");
    if (count > (int64_t)1 << 24) return report_bad_arg(5, "count");
    for (int64_t i = 0; i < count; ++i) {
        const nla_batch_item &it = items[i];
        if (it.n < 0 || it.n > BATCH_N_MAX || it.m < 0 || it.m > ((int64_t)1 << 30))
            return report_bad_arg(4, "items (n must be in [0, 65536], m in [0, 2^30])");
        if (it.n == 0 || it.m == 0) continue;
        if (!it.A || !it.B || it.lda < it.n || it.ldb < (side == 'L' ? it.n : it.m))
            return report_bad_arg(4, "items (pointer or leading dimension)");
    }
    if (plan == nullptr) return report_bad_arg(9, "plan");
    const MLayout L = layout_of(side, items, count);
    if (L.p32_bytes < 0 || L.p64_bytes < 0) return report_bad_arg(4, "items");
    if (L.nbt > INT32_MAX / 4 || L.nat > INT32_MAX / 4) return report_bad_arg(5, "count (too many tasks)");
    if (workspace == nullptr || workspace_bytes < L.total) return report_bad_arg(7, "workspace_bytes");
    if (((uintptr_t)workspace & 255) != 0) return report_bad_arg(6, "workspace (needs 256-byte alignment)");
    nla_batch_mixed_plan_s *P = new (std::nothrow) nla_batch_mixed_plan_s();
    if (!P) return report_hip(hipErrorOutOfMemory, "batch mixed plan");
    P->ws = (char *)workspace;
    P->count = count;
    P->L = L;
    std::vector<nla_batch_item> i32, i64;
    std::vector<MixSys> sys;
    sub_items(side, uplo == 'L', items, count, L, P->ws, i32, i64, &sys);
    std::vector<int2> bt, at;
    bt.reserve((size_t)L.nbt);
    at.reserve((size_t)L.nat);
    for (int64_t i = 0; i < count; ++i) {
        const MixSys &S = sys[(size_t)i];
        for (int t = 0; t < S.bnt; ++t) bt.push_back(make_int2((int)i, t * S.bcb));
        for (int t = 0; t < S.ant; ++t) at.push_back(make_int2((int)i, t * S.acb));
    }
    // descriptors, states (written by the first kernel of every execute) and tasks
    P->host.assign((size_t)L.bpart, 0);
    char *h = P->host.data();
    if (count) std::memcpy(h + L.sys, sys.data(), sys.size() * sizeof(MixSys));
    if (!bt.empty()) std::memcpy(h + L.btasks, bt.data(), bt.size() * sizeof(int2));
    if (!at.empty()) std::memcpy(h + L.atasks, at.data(), at.size() * sizeof(int2));
    hipStream_t s = (hipStream_t)stream;
    hipError_t e = P->host.empty() ? hipSuccess
                                   : hipMemcpyAsync(P->ws, P->host.data(), P->host.size(), hipMemcpyHostToDevice, s);
    if (e != hipSuccess) {
        delete P;
        return report_hip(e, "batch mixed plan upload");
    }
    int32_t st = nla_batch_plan_create(side, uplo, trans, 'S', NLA_F32, i32.data(), count, P->ws + L.p32,
                                       std::max<int64_t>(L.p32_bytes, 256), stream, &P->p32);
    if (st == NLA_OK)
        st = nla_batch_plan_create(side, uplo, trans, 'M', NLA_F64, i64.data(), count, P->ws + L.p64,
                                   std::max<int64_t>(L.p64_bytes, 256), stream, &P->p64);
    if (st != NLA_OK) {
        nla_batch_plan_destroy(P->p32);
        nla_batch_plan_destroy(P->p64);
        delete P;
        return st;
    }
    *plan = P;
    return NLA_OK;
}

int32_t nla_batch_mixed_plan_execute(nla_batch_mixed_plan plan, double tol, int32_t max_iter, void *stream,
                                     int32_t use_graph, double *backward_error, int32_t *iterations) {
    if (plan == nullptr) return report_bad_arg(1, "plan");
    if (!(tol > 0.0) || !std::isfinite(tol)) return report_bad_arg(2, "tol (must be positive and finite)");
    if (max_iter < 0 || max_iter > NLA_MIXED_MAX_ITER) return report_bad_arg(3, "max_iter (0..30)");
    const int64_t count = plan->count;
    if (count == 0) return NLA_OK;
    hipStream_t s = (hipStream_t)stream;
    const MLayout &L = plan->L;
    char *ws = plan->ws;
    const MixSys *sys = (const MixSys *)(ws + L.sys);
    MixState *st = (MixState *)(ws + L.state);
    const int2 *bt = (const int2 *)(ws + L.btasks), *at = (const int2 *)(ws + L.atasks);
    double *bpart = (double *)(ws + L.bpart), *apart = (double *)(ws + L.apart);
    const int nbt = (int)L.nbt, nat = (int)L.nat;
    const int g = use_graph ? 1 : 0;
    std::vector<MixState> &hs = plan->state;
    hs.assign((size_t)count, MixState{});
    std::vector<int32_t> its((size_t)count, -1);
    int64_t left = count;
    int step = 0;
#define NLA_CK(expr)                                                  \
    do {                                                              \
        hipError_t e_ = (expr);                                       \
        if (e_ != hipSuccess) return report_hip(e_, "batch refinement"); \
    } while (0)
#define NLA_CKS(expr)                      \
    do {                                   \
        int32_t s_ = (expr);               \
        if (s_ != NLA_OK) return s_;       \
    } while (0)
    NLA_CK(launch_mix_prep_a(sys, at, nat, apart, s));
    NLA_CK(launch_mix_prep_b(sys, bt, nbt, bpart, s));
    NLA_CK(launch_mix_final(sys, count, apart, bpart, st, true, tol, s));
    NLA_CK(launch_mix_down(sys, bt, nbt, st, false, s));
    NLA_CKS(nla_batch_plan_execute(plan->p32, stream, g));
    NLA_CK(launch_mix_up(sys, bt, nbt, st, true, s));
    for (;; ++step) {
        NLA_CKS(nla_batch_plan_execute(plan->p64, stream, g));
        NLA_CK(launch_mix_resid(sys, bt, nbt, bpart, s));
        NLA_CK(launch_mix_final(sys, count, apart, bpart, st, false, tol, s));
        NLA_CK(hipMemcpyAsync(hs.data(), st, (size_t)count * sizeof(MixState), hipMemcpyDeviceToHost, s));
        NLA_CK(hipStreamSynchronize(s));
        left = 0;
        for (int64_t i = 0; i < count; ++i) {
            if (hs[(size_t)i].done) {
                if (its[(size_t)i] < 0) its[(size_t)i] = step;
            } else {
                ++left;
            }
        }
        if (left == 0 || step >= max_iter) break;
        NLA_CK(launch_mix_down(sys, bt, nbt, st, true, s));
        NLA_CKS(nla_batch_plan_execute(plan->p32, stream, g));
        NLA_CK(launch_mix_up(sys, bt, nbt, st, false, s));
    }
#undef NLA_CK
#undef NLA_CKS
    double worst = 0.0;
    for (int64_t i = 0; i < count; ++i) {
        const double b = hs[(size_t)i].bwd;
        if (backward_error) backward_error[i] = b;
        if (iterations) iterations[i] = its[(size_t)i] < 0 ? step : its[(size_t)i];
        if (!hs[(size_t)i].done) worst = std::isnan(b) || b > worst ? b : worst;
    }
    if (left) {
        char msg[200];
        std::snprintf(msg, sizeof(msg),
                      "refinement left %lld of %lld systems above the target %.3e after %d steps (worst %.3e)",
                      (long long)left, (long long)count, tol, step, worst);
        return report_status(NLA_ERR_NOT_CONVERGED, msg);
    }
    return NLA_OK;
}

int32_t nla_batch_mixed_plan_destroy(nla_batch_mixed_plan plan) {
    if (plan == nullptr) return NLA_OK;
    nla_batch_plan_destroy(plan->p32);
    nla_batch_plan_destroy(plan->p64);
    delete plan;
    return NLA_OK;
}

}  // extern "C"
