// C ABI of libnlahip and the host-side recursion planner.
//
// nla_rectrxm restates unified_rectrxm! (src/rectrxm.jl:43-76) and its
// recursion unified_rec (src/rectrxm.jl:101-198) for column-major device
// matrices:
//   - 'T'/'C' become a strided view of A with swapped strides and a flipped
//     uplo (src/rectrxm.jl:56-59), so no transpose is ever materialised;
//   - the 2x2 split recurses until a diagonal block fits one leaf launch;
//     split points are multiples of the leaf size (64), which is also a
//     multiple of every MFMA tile edge, instead of the reference's
//     power-of-two rule (src/rectrxm.jl:129-134);
//   - forward (A11 -> update -> A22) or backward order per the reference's
//     table (src/rectrxm.jl:153-196), with the same GEMM_ADD!/GEMM_SUB!
//     operands (src/rectrxm.jl:164-192);
//   - alpha is folded into the kernels: for TRSM into the first kernel that
//     touches each block of B (the reference scales B up front,
//     src/rectrxm.jl:64), for TRMM into the last kernel that writes it (the
//     reference scales afterwards, src/rectrxm.jl:72).
// All launches go to the caller's stream in dependency order; nothing
// synchronizes and nothing is allocated.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <atomic>
#include <cmath>
#include <memory>
#include <mutex>
#include <type_traits>
#include <vector>

#include "../../include/nla.h"
#include "nla_internal.h"

namespace nla {

static thread_local char g_last_error[256] = "no error";

static int32_t hip_status(hipError_t e, const char *where) {
    if (e == hipSuccess) return NLA_OK;
    std::snprintf(g_last_error, sizeof(g_last_error), "%s: %s", where, hipGetErrorString(e));
    return NLA_ERR_HIP;
}

static int32_t bad_arg(int k, const char *what) {
    std::snprintf(g_last_error, sizeof(g_last_error), "argument %d (%s) is invalid", k, what);
    return -k;
}

// ---- kernel timing --------------------------------------------------------
struct ProfRec {
    int kind;
    double flops, bytes;
    hipEvent_t e0, e1;
};
static std::mutex g_prof_mu;
static bool g_prof_on = false;
static std::vector<ProfRec> g_prof;
static std::vector<hipEvent_t> g_event_pool;

static hipEvent_t take_event() {
    if (!g_event_pool.empty()) {
        hipEvent_t e = g_event_pool.back();
        g_event_pool.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    return e;
}

// Launch through `fn`, bracketed by events when timing is enabled.
template <typename F>
static hipError_t timed(int kind, double flops, double bytes, hipStream_t s, F fn) {
    if (!g_prof_on) return fn();
    std::lock_guard<std::mutex> g(g_prof_mu);
    ProfRec r{kind, flops, bytes, take_event(), take_event()};
    if (!r.e0 || !r.e1) return hipErrorOutOfMemory;
    hipError_t e = hipEventRecord(r.e0, s);
    if (e != hipSuccess) return e;
    e = fn();
    if (e != hipSuccess) return e;
    e = hipEventRecord(r.e1, s);
    if (e != hipSuccess) return e;
    g_prof.push_back(r);
    return hipSuccess;
}

int32_t report_hip(hipError_t e, const char *where) { return hip_status(e, where); }
int32_t report_bad_arg(int k, const char *what) { return bad_arg(k, what); }
int32_t report_status(int32_t code, const char *msg) {
    std::snprintf(g_last_error, sizeof(g_last_error), "%s", msg);
    return code;
}

static inline char up(char c) { return (c >= 'a' && c <= 'z') ? (char)(c - 'a' + 'A') : c; }

// A strided view: element (i, j) at p[i*rs + j*cs].
template <typename T>
struct View {
    T *p;
    int64_t rs, cs, nr, nc;
    View sub(int64_t r0, int64_t nr_, int64_t c0, int64_t nc_) const {
        return View{p + r0 * rs + c0 * cs, rs, cs, nr_, nc_};
    }
};

// Methods below the recursion (nla_set_config); -1 = not yet read from the
// environment.
static std::atomic<int> g_leaf_mode{-1};
static std::atomic<int> g_panel{-1};
// Panel shape per dtype (index NLA_F32 / NLA_F64): the largest diagonal
// block handed to one fused panel launch (a multiple of 64, <= PANEL_MAX,
// or 0: by layout) and its RHS columns per work-group.  By layout: where the
// panel's rows are contiguous in memory (left side with op(A) = A, right
// side with op(A) = A^T) its fragment loads are coalesced and 512 rows win;
// otherwise every fragment element is its own cache line and 256 rows
// win.  Measured on MI355X at n = 8192, 4 waves (rows 256 / 512): FP32 TRSM
// L-L-N 5.06 / 4.83 ms, R-L-T 5.48 / 5.17, L-L-T 5.49 / 6.06, R-L-N
// 5.84 / 6.33; FP64 TRSM L-L-N 9.39 / 9.10, L-L-T 9.89 / 10.10, R-L-N
// 10.08 / 10.19; FP64 TRMM L-L-T 9.96 / 10.22.  16 columns for both dtypes.
static std::atomic<int> g_panel_rows[2] = {{0}, {0}};
static int panel_rows_default(bool contiguous) { return contiguous ? 512 : 256; }
static std::atomic<int> g_panel_strip[2] = {{PANEL_STRIP}, {PANEL_STRIP}};
// Waves per panel work-group: 0 = the default, 4 (measured on MI355X at
// n = 8192 with 512-row panels: 4 waves against 8, which split each phase's
// k range, FP32 TRSM 4.83 / 4.95 ms, FP32 TRMM 4.80 / 4.87, FP64 TRMM
// 9.03 / 9.10, FP64 TRSM 9.36 / 9.62, the refined solve's fp32 panels
// 33.1 / 33.5 ms), else 4 or 8 for every panel.
static std::atomic<int> g_panel_waves{0};
// Batch size classes: all four, systems above 512 rows staged (measured on
// MI355X over 10,000 systems of 16..1024 rows: 5.5 ms vs 6.9 ms with
// 1024-row panels in FP64, 3.4 vs 3.6 ms in FP32).
static std::atomic<int> g_batch_classes{0x1F};

int batch_class_mask() { return g_batch_classes.load(); }
static std::atomic<int> g_batch_order{0};
int batch_task_order() { return g_batch_order.load(); }
// Residual of the refined solve: 1 = int8 modular products (Ozaki scheme II)
// where the shape allows, 0 = FP64 MFMA GEMM.
static std::atomic<int> g_residual{1};
static std::atomic<int> g_oz_variant{0};
int oz_variant() { return g_oz_variant.load(); }
// Smallest shape given the int8 residual: below it the FP64 product is a
// small part of the solve and the residue conversions cost as much.
constexpr int64_t OZ_MIN_N = 2048, OZ_MIN_M = 64;
// Rows from which an LU leaf is held in registers (one work-group up to 4096
// rows, several above); 0 = the global-memory leaf for every size.
static std::atomic<int> g_lu_split_rows{1};
int lu_split_rows() { return g_lu_split_rows.load(); }
static std::atomic<int> g_lu_spin_limit{0};
int lu_spin_limit() { return g_lu_spin_limit.load(); }

// Leaf method: 1 = inverted diagonal blocks applied on MFMA (default),
// 0 = register/LDS substitution kernel.
static int leaf_mode_default() {
    int v = g_leaf_mode.load();
    if (v < 0) {
        const char *e = std::getenv("NLA_LEAF_MODE");
        v = (e && std::strcmp(e, "subst") == 0) ? 0 : 1;
        int expect = -1;
        g_leaf_mode.compare_exchange_strong(expect, v);
        v = g_leaf_mode.load();
    }
    return v;
}

// Fused diagonal panels on (1) or off (0, recursion down to 64-leaves).
static bool panel_default() {
    int v = g_panel.load();
    if (v < 0) {
        const char *e = std::getenv("NLA_PANEL");
        v = (e && std::strcmp(e, "0") == 0) ? 0 : 1;
        int expect = -1;
        g_panel.compare_exchange_strong(expect, v);
        v = g_panel.load();
    }
    return v != 0;
}

static int leaf_nb_default() {
    static int nb = [] {
        const char *e = std::getenv("NLA_LEAF_NB");
        int v = e ? std::atoi(e) : LEAF_MAX_NB;
        if (v != 16 && v != 32 && v != 64) v = LEAF_MAX_NB;
        return v;
    }();
    return nb;
}

template <typename T>
struct Recursion {
    char func, side, uplo;  // uplo of the effective (possibly transposed) A
    int64_t nb;             // leaf size
    hipStream_t stream;
    const T *W = nullptr;   // inverted diagonal blocks (TRSM, leaf mode 1)
    const int *F = nullptr; // their ill-conditioning flags (refined in the panels)
    int mode = 1;
    bool use_panel = true;  // fuse diagonal blocks <= panel_rows into one launch
    int panel_rows = 512, panel_strip = PANEL_STRIP, panel_waves = 4;
    bool unit = false;  // unit diagonal: only the panels and the inverses read the diagonal
    hipError_t err = hipSuccess;
    // Low-precision GEMM updates (T = float only): lp = 0 bf16, 1 fp16,
    // 2 fp16 pairs (hi + lo, three fp16 products), -1 plain fp32.  fp16 and
    // fp16 pairs scale each work-group's operand panels by powers of two:
    // the blocks of A from a_tab, the 64 x 64 tile maxima of the stored fp32
    // triangle at a_base (leading dimension a_ld; the views are positive-
    // stride sub-views of it); the solved blocks of B from x_slots, where
    // every panel records the largest |X| it wrote at slot off / 64 (cleared
    // before each solve).  upd_koff: the first row (of the effective system)
    // of the solved block the next update reads.
    int lp = -1;
    const unsigned *a_tab = nullptr;
    int64_t a_tab_ld = 0;
    const float *a_base = nullptr;
    int64_t a_ld = 0;
    unsigned *x_slots = nullptr;
    int64_t upd_koff = 0;

    // Leaf as an MFMA product: TRSM applies inv(A_kk) from W, TRMM applies
    // A_kk itself with the other triangle masked to zero.
    void leaf_mfma(View<T> A, View<T> B, int64_t n, int64_t off, T scale) {
        GemmArgs<T> g;
        g.alpha = scale;
        g.beta = T(0);
        g.K = n;
        g.C = B.p;
        g.ldc = B.cs;
        const bool lower = (uplo == 'L');
        // the A_kk operand: inverse block (column-major, ld nb) or the view
        const T *ap;
        bool at;
        int64_t lda;
        if (func == 'S') {
            ap = W + (off / nb) * nb * nb;
            at = false;
            lda = nb;
        } else {
            ap = A.p;
            at = (A.rs != 1);
            lda = at ? A.rs : A.cs;
        }
        if (side == 'L') {  // B_k (n x m) <- op * B_k
            g.M = n; g.N = B.nc; g.leaf = 1;
            g.X = ap; g.tx = at; g.ldx = lda;
            g.Y = B.p; g.ty = false; g.ldy = B.cs;
            g.tri = (func == 'M') ? (lower ? 1 : 2) : 0;
        } else {  // B_k (m x n) <- B_k * op
            g.M = B.nr; g.N = n; g.leaf = 2;
            g.X = B.p; g.tx = false; g.ldx = B.cs;
            g.Y = ap; g.ty = at; g.ldy = lda;
            g.tri = (func == 'M') ? (lower ? 3 : 4) : 0;
        }
        const double fl = (double)n * (double)n * (double)(side == 'L' ? B.nc : B.nr);
        const double by = ((double)n * (n + 1) / 2 + 2.0 * n * (side == 'L' ? B.nc : B.nr)) * sizeof(T);
        err = timed(NLA_KIND_LEAF, fl, by, stream, [&] { return launch_gemm<T>(g, stream); });
    }

    // A whole diagonal block (n <= PANEL_MAX) in one fused launch.
    void panel(View<T> A, View<T> B, int64_t n, int64_t off, T scale) {
        PanelArgs<T> a;
        a.trmm = (func == 'M');
        a.nbig = (int)n;
        a.scale = scale;
        a.W = (func == 'S') ? W + (off / nb) * nb * nb : nullptr;
        a.F = (func == 'S' && F) ? F + off / nb : nullptr;
        if (side == 'L') {
            a.lower = (uplo == 'L');
            a.m = B.nc;
            a.M = A.p; a.mrs = A.rs; a.mcs = A.cs;
            a.B = B.p; a.brs = B.rs; a.bcs = B.cs;
            a.wrs = 1; a.wcs = nb;
        } else {  // op(A)^T X^T = B^T
            a.lower = (uplo == 'U');
            a.m = B.nr;
            a.M = A.p; a.mrs = A.cs; a.mcs = A.rs;
            a.B = B.p; a.brs = B.cs; a.bcs = B.rs;
            a.wrs = nb; a.wcs = 1;
        }
        a.strip = panel_strip;
        a.waves = n <= 512 ? panel_waves : 4;
        a.unit = unit;
        if constexpr (sizeof(T) == 4) {
            if (x_slots && func == 'S') a.xmax = x_slots + off / 64;
        }
        const double fl = (double)n * (double)n * (double)a.m;
        const double by = ((double)n * (n + 1) / 2 + 2.0 * n * a.m) * sizeof(T);
        err = timed(NLA_KIND_LEAF, fl, by, stream, [&] { return launch_panel<T>(a, stream); });
    }

    void leaf(View<T> A, View<T> B, int64_t n, int64_t off, T scale) {
        if (err != hipSuccess) return;
        if (x_slots) {  // only the panels record the maxima the scaled updates read
            err = hipErrorInvalidValue;
            return;
        }
        if (mode == 1) {
            leaf_mfma(A, B, n, off, scale);
            return;
        }
        LeafArgs<T> a;
        a.trmm = (func == 'M');
        a.nb = (int)n;
        a.scale = scale;
        const int64_t last = n - 1;
        if (side == 'L') {
            a.m = B.nc;
            if (uplo == 'L') {
                a.A = A.p; a.ars = A.rs; a.acs = A.cs;
                a.B = B.p; a.brs = B.rs; a.bcs = B.cs;
            } else {  // reversal J*U*J is lower
                a.A = A.p + last * (A.rs + A.cs); a.ars = -A.rs; a.acs = -A.cs;
                a.B = B.p + last * B.rs; a.brs = -B.rs; a.bcs = B.cs;
            }
        } else {  // X*op(A) = B solved as op(A)^T X^T = B^T
            a.m = B.nr;
            if (uplo == 'U') {  // op(A)^T is lower
                a.A = A.p; a.ars = A.cs; a.acs = A.rs;
                a.B = B.p; a.brs = B.cs; a.bcs = B.rs;
            } else {
                a.A = A.p + last * (A.rs + A.cs); a.ars = -A.cs; a.acs = -A.rs;
                a.B = B.p + last * B.cs; a.brs = -B.cs; a.bcs = B.rs;
            }
        }
        const double fl = (double)n * (double)n * (double)a.m;
        const double by = ((double)n * (n + 1) / 2 + 2.0 * n * a.m) * sizeof(T);
        err = timed(NLA_KIND_LEAF, fl, by, stream, [&] { return launch_leaf<T>(a, stream); });
    }

    // The update with operands rounded to bf16 / fp16 / fp16 pairs after
    // power-of-two scaling of each work-group's operand panels (fp16 kinds).
    void lp_gemm(View<float> C, View<float> X, View<float> Y, float alpha, float beta) {
        LpGemmArgs g;
        g.M = C.nr; g.N = C.nc; g.K = X.nc;
        g.tx = (X.rs != 1); g.ldx = g.tx ? X.rs : X.cs;
        g.ty = (Y.rs != 1); g.ldy = g.ty ? Y.rs : Y.cs;
        g.X = X.p; g.Y = Y.p; g.C = C.p; g.ldc = C.cs;
        g.alpha = alpha; g.beta = beta;
        g.lp = lp;
        if (lp == 1 || lp == 2) {
            if (!a_tab || !x_slots) {
                err = hipErrorInvalidValue;
                return;
            }
            // operand roles: left side X = block of A, Y = block of B; right side swapped
            const bool a_is_x = (side == 'L');
            const View<float> &Av = a_is_x ? X : Y;
            const int64_t e = (int64_t)(Av.p - a_base);
            LpScale sa, sb;
            sa.tab = a_tab;
            sa.ld = a_tab_ld;
            sa.r0 = e % a_ld;
            sa.c0 = e / a_ld;
            // operand element (o, k): X(o, k) = p[o*rs + k*cs] as X, Y(k, o) =
            // p[k*rs + o*cs] as Y; with rs == 1 the first index runs down the
            // stored rows
            sa.trans = a_is_x ? (Av.rs != 1) : (Av.rs == 1);
            sb.tab = x_slots;
            sb.ld = 0;
            sb.r0 = upd_koff / 64;
            sb.c0 = (g.K + 63) / 64;
            g.scx = a_is_x ? sa : sb;
            g.scy = a_is_x ? sb : sa;
        }
        const double fl = 2.0 * g.M * g.N * g.K;
        const double by = ((double)g.M * g.K + (double)g.K * g.N + 2.0 * g.M * g.N) * sizeof(float);
        err = timed(lp == 2 ? NLA_KIND_LP_PAIR_GEMM : NLA_KIND_LP_GEMM, fl, by, stream,
                    [&] { return launch_lp_gemm(g, stream); });
    }

    // C = beta*C + alpha*X*Y on views (GEMM_ADD!/GEMM_SUB!, src/matmul.jl:69-80).
    void gemm(View<T> C, View<T> X, View<T> Y, T alpha, T beta) {
        if (err != hipSuccess) return;
        if constexpr (std::is_same<T, float>::value) {
            if (lp >= 0) {
                lp_gemm(C, X, Y, alpha, beta);
                return;
            }
        }
        GemmArgs<T> g;
        g.M = C.nr; g.N = C.nc; g.K = X.nc;
        g.tx = (X.rs != 1); g.ldx = g.tx ? X.rs : X.cs;
        g.ty = (Y.rs != 1); g.ldy = g.ty ? Y.rs : Y.cs;
        g.X = X.p; g.Y = Y.p; g.C = C.p; g.ldc = C.cs;
        g.alpha = alpha; g.beta = beta;
        const double fl = 2.0 * g.M * g.N * g.K;
        const double by = ((double)g.M * g.K + (double)g.K * g.N + 2.0 * g.M * g.N) * sizeof(T);
        err = timed(NLA_KIND_GEMM, fl, by, stream, [&] { return launch_gemm<T>(g, stream); });
    }

    int64_t split(int64_t n) const {
        int64_t half = (n + 1) / 2;
        int64_t mid = ((half + nb - 1) / nb) * nb;
        if (mid >= n) mid = n - 1;
        if (mid < 1) mid = 1;
        return mid;
    }

    // pre: TRSM scale owed by B's untouched blocks; post: TRMM scale owed by
    // the last write of this B.
    // off: position of this diagonal block in the whole matrix.
    void rec(View<T> A, View<T> B, int64_t n, int64_t off, T pre, T post) {
        if (err != hipSuccess) return;
        if (mode == 1 && nb == 64 && n <= panel_rows && use_panel) {
            panel(A, B, n, off, func == 'S' ? pre : post);
            return;
        }
        if (n <= nb) {
            leaf(A, B, n, off, func == 'S' ? pre : post);
            return;
        }
        const int64_t mid = split(n), rem = n - mid;
        View<T> A11 = A.sub(0, mid, 0, mid), A22 = A.sub(mid, rem, mid, rem);
        View<T> A21 = A.sub(mid, rem, 0, mid), A12 = A.sub(0, mid, mid, rem);
        View<T> B1, B2;
        if (side == 'L') {
            B1 = B.sub(0, mid, 0, B.nc);
            B2 = B.sub(mid, rem, 0, B.nc);
        } else {
            B1 = B.sub(0, B.nr, 0, mid);
            B2 = B.sub(0, B.nr, mid, rem);
        }
        const bool forward = (side == 'L' && uplo == 'L' && func == 'S') || (side == 'R' && uplo == 'U' && func == 'S') ||
                             (side == 'L' && uplo == 'U' && func == 'M') || (side == 'R' && uplo == 'L' && func == 'M');
        const T one = T(1);
        if (func == 'S') {
            if (forward) {
                rec(A11, B1, mid, off, pre, one);
                upd_koff = off;
                if (side == 'L') gemm(B2, A21, B1, -one, pre);  // B2 -= A21*B1
                else gemm(B2, B1, A12, -one, pre);              // B2 -= B1*A12
                rec(A22, B2, rem, off + mid, one, one);
            } else {
                rec(A22, B2, rem, off + mid, pre, one);
                upd_koff = off + mid;
                if (side == 'L') gemm(B1, A12, B2, -one, pre);  // B1 -= A12*B2
                else gemm(B1, B2, A21, -one, pre);              // B1 -= B2*A21
                rec(A11, B1, mid, off, one, one);
            }
        } else {
            if (forward) {
                rec(A11, B1, mid, off, one, one);
                if (side == 'L') gemm(B1, A12, B2, post, post);  // B1 += A12*B2
                else gemm(B1, B2, A21, post, post);              // B1 += B2*A21
                rec(A22, B2, rem, off + mid, one, post);
            } else {
                rec(A22, B2, rem, off + mid, one, one);
                if (side == 'L') gemm(B2, A21, B1, post, post);  // B2 += A21*B1
                else gemm(B2, B1, A12, post, post);              // B2 += B1*A12
                rec(A11, B1, mid, off, one, post);
            }
        }
    }
};

// Workspace of the plain TRSM (the inverted diagonal blocks), cached per
// (device, stream) and grown on demand.  A stream-ordered malloc / free per
// call made the host wait once a stream had a backlog (the pool cannot hand
// back a block whose free has not executed yet), so a pipeline of solves ran
// no faster than one at a time.  Calls on one stream are ordered, so its
// buffer is reused without any wait; the entry's lock keeps two host threads
// that enqueue on the same stream from interleaving their launches.  A
// larger request frees the old buffer in stream order and takes a new one.
struct StreamWorkspace {
    int dev;
    hipStream_t stream;
    void *ptr = nullptr;
    size_t bytes = 0;
    std::mutex mu;
};

static StreamWorkspace *stream_workspace(hipStream_t s, hipError_t *err) {
    static std::mutex mu;
    static std::vector<std::unique_ptr<StreamWorkspace>> cache;
    int dev = 0;
    *err = hipGetDevice(&dev);
    if (*err != hipSuccess) return nullptr;
    std::lock_guard<std::mutex> g(mu);
    for (auto &w : cache)
        if (w->dev == dev && w->stream == s) return w.get();
    cache.emplace_back(new StreamWorkspace());
    cache.back()->dev = dev;
    cache.back()->stream = s;
    return cache.back().get();
}

// Called with w->mu held.
static hipError_t reserve(StreamWorkspace *w, size_t bytes) {
    if (bytes <= w->bytes) return hipSuccess;
    hipError_t e = hipSuccess;
    if (w->ptr) {
        e = hipFreeAsync(w->ptr, w->stream);
        w->ptr = nullptr;
        w->bytes = 0;
        if (e != hipSuccess) return e;
    }
    size_t want = bytes < ((size_t)1 << 20) ? ((size_t)1 << 20) : bytes;
    e = hipMallocAsync(&w->ptr, want, w->stream);
    if (e != hipSuccess) {
        w->ptr = nullptr;
        return e;
    }
    w->bytes = want;
    return hipSuccess;
}

// The diagonal-inverse workspace a TRSM call needs: the inverted leaf
// blocks (nb x nb each) followed by one int flag per block; none for TRMM
// or for the substitution leaf.
int64_t rectrxm_workspace_bytes(char func, int32_t dtype, int64_t n) {
    if (func != 'S' || leaf_mode_default() != 1 || n <= 0) return 0;
    const int64_t nb = leaf_nb_default(), blocks = (n + nb - 1) / nb;
    return blocks * nb * nb * (dtype == NLA_F64 ? 8 : 4) + blocks * (int64_t)sizeof(int);
}

// ws: a workspace of ws_bytes for the diagonal inverses (at least
// rectrxm_workspace_bytes for a TRSM with leaf mode 1), or nullptr to take
// this stream's cached buffer.
template <typename T>
static int32_t rectrxm_typed(char side, char uplo, char trans, char func, double alpha, int64_t n, int64_t m,
                             const void *Ap, int64_t lda, void *Bp, int64_t ldb, hipStream_t s, bool unit = false,
                             void *ws = nullptr, int64_t ws_bytes = 0) {
    View<T> A{(T *)Ap, 1, lda, n, n};
    View<T> B{(T *)Bp, 1, ldb, side == 'L' ? n : m, side == 'L' ? m : n};
    if (alpha == 0.0) return hip_status(launch_zero<T>(B.nr, B.nc, B.p, ldb, s), "zero fill");
    if (trans != 'N') {  // src/rectrxm.jl:56-59
        A.rs = lda;
        A.cs = 1;
        uplo = (uplo == 'L') ? 'U' : 'L';
    }
    Recursion<T> r;
    r.func = func;
    r.side = side;
    r.uplo = uplo;
    r.nb = leaf_nb_default();
    r.stream = s;
    r.mode = leaf_mode_default();
    r.use_panel = panel_default();
    if (unit) {  // every diagonal block goes through a panel (and, for TRSM, the inverses)
        r.unit = true;
        r.mode = 1;
        r.nb = 64;
        r.use_panel = true;
    }
    {
        const int di = std::is_same<T, double>::value ? NLA_F64 : NLA_F32;
        const int pr = g_panel_rows[di].load();
        // the panel's row stride: A's for the left side, its transpose's for the right
        const bool contiguous = (side == 'L' ? A.rs : A.cs) == 1;
        r.panel_rows = pr ? pr : panel_rows_default(contiguous);
        r.panel_strip = g_panel_strip[di].load();
        const int pw = g_panel_waves.load();
        r.panel_waves = pw ? pw : 4;
    }
    T *W = nullptr;
    std::unique_lock<std::mutex> hold;
    if (func == 'S' && r.mode == 1) {
        // inverses of all diagonal leaf blocks, from one batched launch, into
        // this stream's cached workspace
        const int64_t blocks = (n + r.nb - 1) / r.nb;
        hipError_t e = hipSuccess;
        const size_t wbytes = (size_t)(blocks * r.nb * r.nb) * sizeof(T);
        if (ws != nullptr && (int64_t)(wbytes + (size_t)blocks * sizeof(int)) > ws_bytes) {
            std::snprintf(g_last_error, sizeof(g_last_error),
                          "n = %lld needs %lld bytes of diagonal-inverse workspace, %lld given", (long long)n,
                          (long long)(wbytes + (size_t)blocks * sizeof(int)), (long long)ws_bytes);
            return NLA_ERR_WORKSPACE;
        }
        if (ws == nullptr) {
            StreamWorkspace *w = stream_workspace(s, &e);
            if (!w) return hip_status(e, "workspace");
            hold = std::unique_lock<std::mutex>(w->mu);
            e = reserve(w, wbytes + (size_t)blocks * sizeof(int));
            if (e != hipSuccess) return hip_status(e, "workspace");
            ws = w->ptr;
        }
        W = (T *)ws;
        int *F = (r.nb == 64) ? (int *)((char *)ws + wbytes) : nullptr;
        const double fl = (double)blocks * r.nb * r.nb * r.nb / 3.0;
        const double by = (double)blocks * r.nb * r.nb * 1.5 * sizeof(T);
        e = timed(NLA_KIND_LEAF, fl, by, s, [&] {
            return launch_trtri_blocks<T>(n, (int)r.nb, A.p, A.rs, A.cs, uplo == 'L', W, s, unit, F);
        });
        if (e != hipSuccess) return hip_status(e, "diagonal inverses");
        r.W = W;
        r.F = F;
    }
    r.rec(A, B, n, 0, (T)alpha, (T)alpha);
    return hip_status(r.err, "kernel launch");
}

// ---- mixed precision with FP64 refinement ---------------------------------

// Workspace of nla_rectrxm_mixed: an fp32 copy of A's stored triangle and
// the inverses of its diagonal 64-blocks (the correction solves), alpha*B and
// the residual in FP64, the fp32 correction, the absolute-maximum slots of
// the low-precision updates and the norm reductions.  256-byte aligned parts.
struct MixedLayout {
    int64_t rows, cols;  // shape of B
    int64_t ld32, ldr;   // leading dimension of the fp32 A copy / of S, R, D
    int64_t nt;  // 64-row blocks of A: the tile table is nt x nt, the panel slots nt
    int64_t nsys;  // systems: columns of B (side L) or rows (side R)
    int64_t a32, w, wf, s, r, d, tab, slots, part, norms, queue, spart, xpart, s2, t64, t64_bytes, total;
    // int8 residual: its plan and workspace (after the fields above)
    bool oz;
    OzPlan op;
    OzLayout ol;
};

static inline int64_t al256(int64_t x) { return (x + 255) & ~(int64_t)255; }

static MixedLayout mixed_layout(char side, int64_t n, int64_t m) {
    MixedLayout L{};
    L.rows = side == 'L' ? n : m;
    L.cols = side == 'L' ? m : n;
    L.ld32 = (n + 63) / 64 * 64;
    L.ldr = (L.rows + 63) / 64 * 64;
    const int64_t nblk = (n + 63) / 64;
    L.nt = nblk;
    int64_t o = 0;
    L.a32 = o;   o = al256(o + L.ld32 * n * 4);
    L.w = o;     o = al256(o + nblk * 64 * 64 * 4);
    L.wf = o;    o = al256(o + nblk * 4);
    L.s = o;     o = al256(o + L.ldr * L.cols * 8);
    L.r = o;     o = al256(o + L.ldr * L.cols * 8);
    L.d = o;     o = al256(o + L.ldr * L.cols * 4);
    L.tab = o;   o = al256(o + nblk * nblk * 4);
    L.slots = o; o = al256(o + nblk * 4);
    L.part = o;  o = al256(o + NORM_PARTS_MAX * 2 * 8);
    L.norms = o; o = al256(o + 16 * 8);
    L.queue = o; o = al256(o + 4);
    L.nsys = m;
    const int64_t np = (int64_t)SYS_CHUNKS_MAX * 2 * (n > m ? n : m);
    L.spart = o; o = al256(o + np * 8);
    L.xpart = o; o = al256(o + np * 8);
    L.s2 = o;    o = al256(o + (m > 0 ? m : 1) * 8);
    // the plain FP64 solve's scratch (NLA_LP_FP64)
    L.t64_bytes = rectrxm_workspace_bytes('S', NLA_F64, n);
    L.t64 = o;   o = al256(o + (L.t64_bytes > 0 ? L.t64_bytes : 1));
    L.total = o;
    L.oz = g_residual.load() == 1 && n >= OZ_MIN_N && n <= OZ_N_MAX && m >= OZ_MIN_M;
    if (L.oz) {
        L.op = oz_plan(n, m);
        L.ol = oz_layout(L.op, o);
        L.total = L.ol.total;
    }
    return L;
}

// A power of two at or just above |x| (1 for zero or non-finite x).
static double pow2_above(double x) {
    if (!(x > 0.0) || !std::isfinite(x)) return 1.0;
    int e;
    (void)std::frexp(x, &e);
    return std::ldexp(1.0, e);
}

// The int8 residual's phases, each recorded by the kernel timer: the
// conversions and the reconstruction as NLA_KIND_RESIDUAL, the modular
// products as NLA_KIND_RESIDUAL_GEMM (T n^2 m int8 operations).
static hipError_t oz_timed(const OzPlan &P, const OzLayout &L, char *ws, bool lower, bool out_t, double *X,
                           int64_t xso, int64_t xsk, int64_t n, int64_t m, const double *S, int64_t lds, double *R,
                           int64_t ldr, hipStream_t s) {
    const double nm = (double)n * m;
    hipError_t e = timed(NLA_KIND_RESIDUAL, 0.0, nm * (16.0 + P.T), s,
                         [&] { return oz_convert_x(P, L, ws, X, xso, xsk, n, m, s); });
    if (e != hipSuccess) return e;
    e = timed(NLA_KIND_RESIDUAL_GEMM, (double)P.T * n * n * m, (double)P.T * ((double)n * n / 2 + 2.0 * nm), s,
              [&] { return oz_products(P, L, ws, lower, out_t, s); });
    if (e != hipSuccess) return e;
    return timed(NLA_KIND_RESIDUAL, nm, nm * (16.0 + P.T), s,
                 [&] { return oz_reconstruct(P, L, ws, out_t, n, m, S, lds, R, ldr, s); });
}

// A second stream per (host thread, device) for the mixed solve's work that
// does not feed its first correction solve (A's norms and residue planes),
// so that it runs beside that solve; fork and join through two events.
// Default priority.  Kept until the thread ends (not destroyed).
struct SideStream {
    hipStream_t s = nullptr;
    hipEvent_t fork = nullptr, join = nullptr;
};

static hipError_t side_stream(SideStream **out) {
    constexpr int MAX_DEV = 64;
    thread_local SideStream cache[MAX_DEV];
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    if (dev < 0 || dev >= MAX_DEV) return hipErrorInvalidDevice;
    SideStream &x = cache[dev];
    if (x.s == nullptr) {
        e = hipStreamCreateWithFlags(&x.s, hipStreamNonBlocking);
        if (e == hipSuccess) e = hipEventCreateWithFlags(&x.fork, hipEventDisableTiming);
        if (e == hipSuccess) e = hipEventCreateWithFlags(&x.join, hipEventDisableTiming);
        if (e != hipSuccess) return e;
    }
    *out = &x;
    return hipSuccess;
}

static int32_t rectrxm_mixed_impl(char side, char uplo, char trans, int precision, double alpha, int64_t n,
                                  int64_t m, const double *A, int64_t lda, double *B, int64_t ldb, double tol,
                                  int max_iter, char *ws, const MixedLayout &L, hipStream_t s,
                                  nla_mixed_info *info) {
    float *A32 = (float *)(ws + L.a32);
    float *W = (float *)(ws + L.w);
    double *S = (double *)(ws + L.s), *R = (double *)(ws + L.r);
    float *D = (float *)(ws + L.d);
    unsigned *tab = (unsigned *)(ws + L.tab), *slots = (unsigned *)(ws + L.slots);
    double *part = (double *)(ws + L.part), *nrm = (double *)(ws + L.norms);
    const bool tr = (trans != 'N');
    const char ue = tr ? (uplo == 'L' ? 'U' : 'L') : uplo;  // uplo of op(A)
    hipError_t e = hipSuccess;
#define NLA_CK(x)                                          \
    do {                                                   \
        e = (x);                                           \
        if (e != hipSuccess) return hip_status(e, "mixed"); \
    } while (0)

    // fp32 copy of the stored triangle, viewed as op(A); inverses of its
    // diagonal 64-blocks for the correction solves
    NLA_CK(launch_tri_to_f32(n, A, lda, uplo == 'L', A32, L.ld32, s));
    View<float> A32v{A32, 1, L.ld32, n, n};
    if (tr) {
        A32v.rs = L.ld32;
        A32v.cs = 1;
    }
    int *WF = (int *)(ws + L.wf);
    NLA_CK(launch_trtri_blocks<float>(n, 64, A32v.p, A32v.rs, A32v.cs, ue == 'L', W, s, false, WF));
    // S = alpha*B; the norms of A and S enter every stopping test
    NLA_CK(launch_scale_copy(L.rows, L.cols, alpha, B, ldb, S, L.ldr, s));
    // ||A||_F and its largest column 2-norm (the lower bound of ||A||_2 in
    // the per-system test), ||S||_F, max |S| and every system's ||s_j||^2
    const bool by_col = (side == 'L');
    double *spart = (double *)(ws + L.spart), *xpart = (double *)(ws + L.xpart), *s2 = (double *)(ws + L.s2);
    const int ncha = sys_chunks(n, n), nchb = sys_chunks(L.nsys, n);
    NLA_CK(launch_sys_sumsq(L.rows, L.cols, S, L.ldr, by_col, 0, nchb, spart, s));
    NLA_CK(launch_sys_fold(nchb, L.nsys, spart, s2, nrm + 3, s));
    // On the side stream, beside X0's solve: A's norms (partials in xpart,
    // which the stopping tests use only after the join) and, for the int8
    // residual, A's row scales and residue planes (Ahat = op(A) for side L,
    // op(A)^T for side R, so that the residual is Ahat * Xhat or its
    // transpose), kept for every step.  A non-finite A raises the residual's
    // range flag, which the first stopping test reads (FP64 residuals then).
    bool use_oz = L.oz;
    const bool ahat_is_a = (side == 'L') != tr;
    const bool ahat_lower = ahat_is_a ? (uplo == 'L') : (uplo == 'U');
    int hflag = 0;
    SideStream *side_s = nullptr;
    NLA_CK(side_stream(&side_s));
    NLA_CK(hipEventRecord(side_s->fork, s));
    NLA_CK(hipStreamWaitEvent(side_s->s, side_s->fork, 0));
    NLA_CK(launch_sys_sumsq(n, n, A, lda, true, uplo == 'L' ? 1 : 2, ncha, xpart, side_s->s));
    NLA_CK(launch_sys_fold(ncha, n, xpart, nullptr, nrm + 0, side_s->s));
    if (use_oz)
        NLA_CK(oz_prepare(L.op, L.ol, ws, A, ahat_is_a ? 1 : lda, ahat_is_a ? lda : 1, ahat_lower, n, side_s->s));
    NLA_CK(hipEventRecord(side_s->join, side_s->s));
    // ||S||_F and max |S| (X0's scaling) from the main stream alone
    double h[6];
    NLA_CK(hipMemcpyAsync(h + 3, nrm + 3, 3 * sizeof(double), hipMemcpyDeviceToHost, s));
    NLA_CK(hipStreamSynchronize(s));
    double nA = 0.0, hA = 0.0;  // ||A||_F, read at the first stopping test
    const double nS = std::sqrt(h[3]);
    const double maxS = h[4];
    (void)part;

    // D <- op(A)^-1 D (or D op(A)^-1) on the fp32 copy, updates in `prec`
    bool tab_ready = false;  // tile maxima of the fp32 A (fp16 / fp16-pair updates)
    auto lp_solve = [&](int prec) -> hipError_t {
        Recursion<float> r;
        r.func = 'S';
        r.side = side;
        r.uplo = ue;
        r.nb = 64;
        r.stream = s;
        r.mode = 1;
        r.use_panel = true;
        r.W = W;
        r.F = WF;
        r.panel_waves = g_panel_waves.load() ? g_panel_waves.load() : 4;
        const bool scaled = (prec == NLA_LP_FP16 || prec == NLA_LP_FP16X2);
        if (prec != NLA_LP_FP32) r.lp = (prec == NLA_LP_FP16X2) ? 2 : prec;
        if (scaled) {
            hipError_t e2 = hipSuccess;
            if (!tab_ready) e2 = launch_tile_absmax(n, A32, L.ld32, uplo == 'L', tab, L.nt, s);
            if (e2 == hipSuccess) e2 = launch_fill_u32(slots, L.nt, 0u, s);
            if (e2 != hipSuccess) return e2;
            tab_ready = true;
            r.a_tab = tab;
            r.a_tab_ld = L.nt;
            r.a_base = A32;
            r.a_ld = L.ld32;
            r.x_slots = slots;
        }
        View<float> Dv{D, 1, L.ldr, L.rows, L.cols};
        r.rec(A32v, Dv, n, 0, 1.0f, 1.0f);
        return r.err;
    };

    // NLA_LP_AUTO: X0 with fp16-pair updates (as accurate as fp32 updates
    // on the measured workloads, in about 45 % of their time; a bf16 / fp16
    // X0 starts so far from the target at small cond that the saved solve
    // time is lost in an extra residual), then each correction in the
    // precision whose predicted cost to reach tol is least.  Costs are in
    // units of one int8 residual (the refinement's dominant phase, about
    // 10.5 ms at n = 16384, m = 4096): a correction solve costs about 0.41
    // (bf16, fp16), 0.55 (fp16 pairs) or 1.0 (fp32) of one on MI355X at the
    // sizes where refinement is used (n >= 4096).  A step in precision p
    // shrinks the backward error by a factor rho_p ~ 0.03 u_p for bf16 / fp16
    // (u_p the unit roundoff; fitted over cond 1e2..1e6) and ~1e-7 for fp16
    // pairs and fp32; once a step has been observed, its factor rescaled by
    // u_p / u_q replaces the prior (never below the 1e-7 floor).
    const bool autop = (precision == NLA_LP_AUTO);
    struct Cand {
        int prec;
        double u, cost, rho0;
    };
    const Cand cands[4] = {{NLA_LP_BF16, 0x1p-8, 0.41, 0.03 * 0x1p-8},
                           {NLA_LP_FP16, 0x1p-11, 0.41, 0.03 * 0x1p-11},
                           {NLA_LP_FP16X2, 0x1p-22, 0.55, 1e-7},
                           {NLA_LP_FP32, 0x1p-24, 1.0, 1e-7}};
    auto unit_of = [&](int p) {
        for (const Cand &c : cands)
            if (c.prec == p) return c.u;
        return 0x1p-24;
    };
    double rho_obs = -1.0;
    int rho_prec = NLA_LP_FP16X2;
    auto choose = [&](double err) {
        int best = NLA_LP_FP16X2;
        double best_cost = INFINITY;
        for (const Cand &c : cands) {
            double rho = c.rho0;
            if (rho_obs > 0.0) rho = std::max(rho_obs * c.u / unit_of(rho_prec), 1e-7);
            if (!(rho < 1.0)) continue;
            const double need = std::log(tol / err) / std::log(rho);
            const double steps = std::max(1.0, std::ceil(need - 1e-9));
            const double cost = steps * (c.cost + 1.0);
            if (cost < best_cost) {
                best_cost = cost;
                best = c.prec;
            }
        }
        return best;
    };

    // X0: the low-precision solve of S, scaled by a power of two so the fp32
    // buffer holds O(1) values; NLA_LP_FP64: the plain FP64 solve of alpha*B
    // in place, corrections (if any) in fp32
    int prec = autop ? NLA_LP_FP16X2 : precision;
    double sig = pow2_above(maxS);
    if (precision == NLA_LP_FP64) {
        const int32_t st = rectrxm_in_workspace(side, uplo, trans, 'S', NLA_F64, alpha, n, m, A, lda, B, ldb, s,
                                                ws + L.t64, L.t64_bytes);
        if (st != NLA_OK) return st;
        prec = NLA_LP_FP32;
    } else {
        NLA_CK(launch_down(L.rows, L.cols, 1.0 / sig, S, L.ldr, D, L.ldr, s));
        NLA_CK(lp_solve(prec));
        NLA_CK(launch_up(L.rows, L.cols, sig, D, L.ldr, 0.0, B, ldb, s));
    }

    // R = S - op(A) X (side L) or S - X op(A) (side R): the FP64 MFMA GEMM
    // with the triangle of op(A) masked and the k-steps beyond it skipped
    GemmArgs<double> g;
    g.alpha = -1.0;
    g.beta = 1.0;
    g.C = R;
    g.ldc = L.ldr;
    g.K = n;
    if (side == 'L') {
        g.M = n; g.N = m;
        g.X = A; g.tx = tr; g.ldx = lda;
        g.Y = B; g.ty = false; g.ldy = ldb;
        g.tri = (ue == 'L') ? 1 : 2;
    } else {
        g.M = m; g.N = n;
        g.X = B; g.tx = false; g.ldx = ldb;
        g.Y = A; g.ty = tr; g.ldy = lda;
        g.tri = (ue == 'L') ? 3 : 4;
    }
    g.queue = (int *)(ws + L.queue);
    const double res_fl = (double)n * n * m;
    const double res_by = ((double)n * (n + 1) / 2 + 3.0 * n * m) * 8;

    auto fp64_residual = [&]() -> hipError_t {
        hipError_t e2 = launch_scale_copy(L.rows, L.cols, 1.0, S, L.ldr, R, L.ldr, s);
        if (e2 == hipSuccess) e2 = launch_fill_u32((unsigned *)g.queue, 1, 0u, s);
        if (e2 == hipSuccess) e2 = launch_gemm<double>(g, s);
        return e2;
    };
    // Xhat(k, j) = X[j*xso + k*xsk]: a system is a column of X (side L) or a
    // row (side R); X is rounded in place to bX bits per system
    const int64_t xso = (side == 'L') ? ldb : 1, xsk = (side == 'L') ? 1 : ldb;
    auto int8_residual = [&]() -> hipError_t {
        return oz_timed(L.op, L.ol, ws, ahat_lower, side == 'R', B, xso, xsk, n, m, S, L.ldr, R, L.ldr, s);
    };
    // per-system test: max_j ||r_j|| / (||A||_lb ||x_j|| + ||s_j||), with the
    // int8 residual's range flag read in the same synchronization
    auto stop_norms = [&](bool with_flag) -> hipError_t {
        hipError_t e2 = launch_sys_sumsq(L.rows, L.cols, R, L.ldr, by_col, 0, nchb, spart, s);
        if (e2 == hipSuccess) e2 = launch_sys_sumsq(L.rows, L.cols, B, ldb, by_col, 0, nchb, xpart, s);
        if (e2 == hipSuccess) e2 = launch_sys_eta(nchb, L.nsys, spart, xpart, s2, nrm + 2, nrm + 6, s);
        if (e2 == hipSuccess) e2 = hipMemcpyAsync(h, nrm + 6, 4 * sizeof(double), hipMemcpyDeviceToHost, s);
        if (e2 == hipSuccess) e2 = hipMemcpyAsync(&hA, nrm, sizeof(double), hipMemcpyDeviceToHost, s);
        hflag = 0;
        if (e2 == hipSuccess && with_flag)
            e2 = hipMemcpyAsync(&hflag, ws + L.ol.flag, sizeof(int), hipMemcpyDeviceToHost, s);
        if (e2 == hipSuccess) e2 = hipStreamSynchronize(s);
        return e2;
    };

    // the side stream's work is complete before the first residual
    NLA_CK(hipStreamWaitEvent(s, side_s->join, 0));
    double prev = INFINITY, bwd = INFINITY, bwd_f = INFINITY;
    int it = 0;
    for (;; ++it) {
        if (use_oz)
            NLA_CK(int8_residual());
        else
            NLA_CK(timed(NLA_KIND_RESIDUAL, res_fl, res_by, s, fp64_residual));
        NLA_CK(stop_norms(use_oz));
        if (use_oz && hflag) {
            // S or X outside the moduli's range (or not finite): this and
            // every later residual in FP64
            use_oz = false;
            NLA_CK(timed(NLA_KIND_RESIDUAL, res_fl, res_by, s, fp64_residual));
            NLA_CK(stop_norms(false));
        }
        nA = std::sqrt(hA);
        const double nR = std::sqrt(h[0]), nX = std::sqrt(h[2]);
        const double den = nA * nX + nS;
        bwd_f = den > 0.0 ? nR / den : (nR > 0.0 ? INFINITY : 0.0);
        bwd = h[3];
        if (info && it <= NLA_MIXED_MAX_ITER) info->history[it] = bwd;
        if (bwd <= tol || it >= max_iter || !std::isfinite(bwd)) break;
        if (autop) {
            if (it > 0 && prev > 0.0 && bwd > 0.0) {
                rho_obs = bwd / prev;
                rho_prec = prec;
            }
            prec = choose(bwd);
        } else if (it > 0 && bwd > 0.5 * prev && prec != NLA_LP_FP32) {
            // too little progress: raise the update precision for the next steps
            prec = (prec == NLA_LP_BF16) ? NLA_LP_FP16 : (prec == NLA_LP_FP16) ? NLA_LP_FP16X2 : NLA_LP_FP32;
        }
        prev = bwd;
        sig = pow2_above(h[1]);
        NLA_CK(launch_down(L.rows, L.cols, 1.0 / sig, R, L.ldr, D, L.ldr, s));
        NLA_CK(lp_solve(prec));
        NLA_CK(launch_up(L.rows, L.cols, sig, D, L.ldr, 1.0, B, ldb, s));
    }
#undef NLA_CK
    if (info) {
        info->iterations = it;
        info->precision = (precision == NLA_LP_FP64 && it == 0) ? NLA_LP_FP64 : prec;
        info->backward_error = bwd;
        info->backward_error_frobenius = bwd_f;
        info->residual = use_oz ? NLA_RES_INT8 : NLA_RES_FP64;
    }
    if (!(bwd <= tol)) {
        std::snprintf(g_last_error, sizeof(g_last_error),
                      "refinement stopped at backward error %.3e after %d steps (target %.3e)", bwd, it, tol);
        return NLA_ERR_NOT_CONVERGED;
    }
    return NLA_OK;
}

// nla_rectrxm's work with the caller's workspace (rectrxm_workspace_bytes)
// instead of the per-stream buffer; arguments already validated.
int32_t rectrxm_in_workspace(char side, char uplo, char trans, char func, int32_t dtype, double alpha, int64_t n,
                             int64_t m, const void *A, int64_t lda, void *B, int64_t ldb, hipStream_t s, void *ws,
                             int64_t ws_bytes) {
    if (ws == nullptr) return report_status(NLA_ERR_WORKSPACE, "no workspace given");
    if (dtype == NLA_F64)
        return rectrxm_typed<double>(side, uplo, trans, func, alpha, n, m, A, lda, B, ldb, s, false, ws, ws_bytes);
    return rectrxm_typed<float>(side, uplo, trans, func, alpha, n, m, A, lda, B, ldb, s, false, ws, ws_bytes);
}

// Argument checks of nla_rectrxm (0 or -k); characters upper-cased in place.
int32_t rectrxm_check(char &side, char &uplo, char &trans, char &func, int32_t dtype, int64_t n, int64_t m,
                      const void *A, int64_t lda, const void *B, int64_t ldb) {
    side = up(side); uplo = up(uplo); trans = up(trans); func = up(func);
    if (side != 'L' && side != 'R') return bad_arg(1, "side");
    if (uplo != 'L' && uplo != 'U') return bad_arg(2, "uplo");
    if (trans != 'N' && trans != 'T' && trans != 'C') return bad_arg(3, "trans");
    if (func != 'S' && func != 'M') return bad_arg(4, "func");
    if (dtype != NLA_F32 && dtype != NLA_F64) return bad_arg(5, "dtype");
    if (n < 0) return bad_arg(7, "n");
    if (m < 0) return bad_arg(8, "m");
    if (n == 0 || m == 0) return NLA_OK;
    if (A == nullptr) return bad_arg(9, "A");
    if (lda < n) return bad_arg(10, "lda");
    if (B == nullptr) return bad_arg(11, "B");
    if (ldb < (side == 'L' ? n : m)) return bad_arg(12, "ldb");
    return NLA_OK;
}

}  // namespace nla

using namespace nla;

extern "C" {

int32_t nla_abi_version(void) { return 1; }

const char *nla_last_error(void) { return g_last_error; }

int32_t nla_rectrxm(char side, char uplo, char trans, char func, int32_t dtype, double alpha, int64_t n, int64_t m,
                    const void *A, int64_t lda, void *B, int64_t ldb, void *stream) {
    const int32_t chk = rectrxm_check(side, uplo, trans, func, dtype, n, m, A, lda, B, ldb);
    if (chk != NLA_OK || n == 0 || m == 0) return chk;
    hipStream_t s = (hipStream_t)stream;
    if (dtype == NLA_F64) return rectrxm_typed<double>(side, uplo, trans, func, alpha, n, m, A, lda, B, ldb, s);
    return rectrxm_typed<float>(side, uplo, trans, func, alpha, n, m, A, lda, B, ldb, s);
}

// BLAS-style xTRSM / xTRMM with the diagonal argument.
static int32_t trxm_diag(char func, char side, char uplo, char transa, char diag, int32_t dtype, double alpha,
                         int64_t n, int64_t m, const void *A, int64_t lda, void *B, int64_t ldb, void *stream) {
    side = up(side); uplo = up(uplo); transa = up(transa); diag = up(diag);
    if (side != 'L' && side != 'R') return bad_arg(1, "side");
    if (uplo != 'L' && uplo != 'U') return bad_arg(2, "uplo");
    if (transa != 'N' && transa != 'T' && transa != 'C') return bad_arg(3, "transa");
    if (diag != 'N' && diag != 'U') return bad_arg(4, "diag");
    if (dtype != NLA_F32 && dtype != NLA_F64) return bad_arg(5, "dtype");
    if (n < 0) return bad_arg(7, "n");
    if (m < 0) return bad_arg(8, "m");
    if (n == 0 || m == 0) return NLA_OK;
    if (A == nullptr) return bad_arg(9, "A");
    if (lda < n) return bad_arg(10, "lda");
    if (B == nullptr) return bad_arg(11, "B");
    if (ldb < (side == 'L' ? n : m)) return bad_arg(12, "ldb");
    hipStream_t s = (hipStream_t)stream;
    const bool unit = diag == 'U';
    if (dtype == NLA_F64)
        return rectrxm_typed<double>(side, uplo, transa, func, alpha, n, m, A, lda, B, ldb, s, unit);
    return rectrxm_typed<float>(side, uplo, transa, func, alpha, n, m, A, lda, B, ldb, s, unit);
}

int32_t nla_trsm(char side, char uplo, char transa, char diag, int32_t dtype, double alpha, int64_t n, int64_t m,
                 const void *A, int64_t lda, void *B, int64_t ldb, void *stream) {
    return trxm_diag('S', side, uplo, transa, diag, dtype, alpha, n, m, A, lda, B, ldb, stream);
}

int32_t nla_trmm(char side, char uplo, char transa, char diag, int32_t dtype, double alpha, int64_t n, int64_t m,
                 const void *A, int64_t lda, void *B, int64_t ldb, void *stream) {
    return trxm_diag('M', side, uplo, transa, diag, dtype, alpha, n, m, A, lda, B, ldb, stream);
}

int32_t nla_gemm(int32_t dtype, char transx, char transy, int64_t M, int64_t N, int64_t K, double alpha,
                 const void *X, int64_t ldx, const void *Y, int64_t ldy, double beta, void *C, int64_t ldc,
                 void *stream) {
    transx = up(transx); transy = up(transy);
    if (dtype != NLA_F32 && dtype != NLA_F64) return bad_arg(1, "dtype");
    if (transx != 'N' && transx != 'T' && transx != 'C') return bad_arg(2, "transx");
    if (transy != 'N' && transy != 'T' && transy != 'C') return bad_arg(3, "transy");
    if (M < 0) return bad_arg(4, "M");
    if (N < 0) return bad_arg(5, "N");
    if (K < 0) return bad_arg(6, "K");
    if (M == 0 || N == 0) return NLA_OK;
    bool tx = transx != 'N', ty = transy != 'N';
    if (K > 0 && X == nullptr) return bad_arg(8, "X");
    if (K > 0 && ldx < (tx ? K : M)) return bad_arg(9, "ldx");
    if (K > 0 && Y == nullptr) return bad_arg(10, "Y");
    if (K > 0 && ldy < (ty ? N : K)) return bad_arg(11, "ldy");
    if (C == nullptr) return bad_arg(13, "C");
    if (ldc < M) return bad_arg(14, "ldc");
    hipStream_t s = (hipStream_t)stream;
    if (dtype == NLA_F64) {
        GemmArgs<double> g{M, N, K, tx, ty, alpha, beta, (const double *)X, ldx, (const double *)Y, ldy, (double *)C, ldc};
        return hip_status(launch_gemm<double>(g, s), "gemm launch");
    }
    GemmArgs<float> g{M, N, K, tx, ty, (float)alpha, (float)beta, (const float *)X, ldx, (const float *)Y, ldy, (float *)C, ldc};
    return hip_status(launch_gemm<float>(g, s), "gemm launch");
}

int32_t nla_gemm_add(int32_t dtype, int64_t N, int64_t R, int64_t M, const void *X, int64_t ldx, const void *Y,
                     int64_t ldy, void *C, int64_t ldc, void *stream) {
    if (dtype != NLA_F32 && dtype != NLA_F64) return bad_arg(1, "dtype");
    if (N < 0) return bad_arg(2, "N");
    if (R < 0) return bad_arg(3, "R");
    if (M < 0) return bad_arg(4, "M");
    if (N == 0 || M == 0 || R == 0) return NLA_OK;
    if (X == nullptr) return bad_arg(5, "X");
    if (ldx < N) return bad_arg(6, "ldx");
    if (Y == nullptr) return bad_arg(7, "Y");
    if (ldy < R) return bad_arg(8, "ldy");
    if (C == nullptr) return bad_arg(9, "C");
    if (ldc < N) return bad_arg(10, "ldc");
    return nla_gemm(dtype, 'N', 'N', N, M, R, 1.0, X, ldx, Y, ldy, 1.0, C, ldc, stream);
}

int32_t nla_gemm_sub(int32_t dtype, int64_t N, int64_t R, int64_t M, void *C, int64_t ldc, const void *X,
                     int64_t ldx, const void *Y, int64_t ldy, void *stream) {
    if (dtype != NLA_F32 && dtype != NLA_F64) return bad_arg(1, "dtype");
    if (N < 0) return bad_arg(2, "N");
    if (R < 0) return bad_arg(3, "R");
    if (M < 0) return bad_arg(4, "M");
    if (N == 0 || M == 0 || R == 0) return NLA_OK;
    if (C == nullptr) return bad_arg(5, "C");
    if (ldc < N) return bad_arg(6, "ldc");
    if (X == nullptr) return bad_arg(7, "X");
    if (ldx < N) return bad_arg(8, "ldx");
    if (Y == nullptr) return bad_arg(9, "Y");
    if (ldy < R) return bad_arg(10, "ldy");
    return nla_gemm(dtype, 'N', 'N', N, M, R, -1.0, X, ldx, Y, ldy, 1.0, C, ldc, stream);
}

int64_t nla_mixed_workspace_bytes(char side, int64_t n, int64_t m) {
    side = up(side);
    if ((side != 'L' && side != 'R') || n < 0 || m < 0) return -1;
    return mixed_layout(side, n, m).total;
}

int32_t nla_rectrxm_mixed(char side, char uplo, char trans, int32_t precision, double alpha, int64_t n, int64_t m,
                          const double *A, int64_t lda, double *B, int64_t ldb, double tol, int32_t max_iter,
                          void *workspace, int64_t workspace_bytes, void *stream, nla_mixed_info *info) {
    side = up(side); uplo = up(uplo); trans = up(trans);
    if (side != 'L' && side != 'R') return bad_arg(1, "side");
    if (uplo != 'L' && uplo != 'U') return bad_arg(2, "uplo");
    if (trans != 'N' && trans != 'T' && trans != 'C') return bad_arg(3, "trans");
    if (precision != NLA_LP_BF16 && precision != NLA_LP_FP16 && precision != NLA_LP_FP32 && precision != NLA_LP_AUTO &&
        precision != NLA_LP_FP64 && precision != NLA_LP_FP16X2)
        return bad_arg(4, "precision");
    if (!std::isfinite(alpha)) return bad_arg(5, "alpha");
    if (n < 0) return bad_arg(6, "n");
    if (m < 0) return bad_arg(7, "m");
    if (!(tol > 0.0)) return bad_arg(12, "tol");
    if (max_iter < 0 || max_iter > NLA_MIXED_MAX_ITER) return bad_arg(13, "max_iter");
    if (info) {
        std::memset(info, 0, sizeof(*info));
        info->precision = precision;
    }
    if (n == 0 || m == 0) return NLA_OK;
    if (A == nullptr) return bad_arg(8, "A");
    if (lda < n) return bad_arg(9, "lda");
    if (B == nullptr) return bad_arg(10, "B");
    if (ldb < (side == 'L' ? n : m)) return bad_arg(11, "ldb");
    MixedLayout L = mixed_layout(side, n, m);
    if (workspace == nullptr || ((uintptr_t)workspace & 255) != 0)
        return bad_arg(14, "workspace (needs 256-byte alignment)");
    if (workspace_bytes < L.total) return bad_arg(15, "workspace_bytes");
    return rectrxm_mixed_impl(side, uplo, trans, precision, alpha, n, m, A, lda, B, ldb, tol, max_iter,
                              (char *)workspace, L, (hipStream_t)stream, info);
}

int64_t nla_residual_workspace_bytes(char side, int32_t method, int64_t n, int64_t m) {
    side = up(side);
    if ((side != 'L' && side != 'R') || n < 0 || m < 0) return -1;
    if (method == NLA_RES_FP64) return 256;
    if (method != NLA_RES_INT8 || n > OZ_N_MAX) return -1;
    if (n == 0 || m == 0) return 256;
    return oz_layout(oz_plan(n, m), 0).total;
}

int32_t nla_residual(char side, char uplo, char trans, int32_t method, double alpha, int64_t n, int64_t m,
                     const double *A, int64_t lda, double *X, int64_t ldx, const double *B, int64_t ldb, double *R,
                     int64_t ldr, void *workspace, int64_t workspace_bytes, void *stream) {
    side = up(side); uplo = up(uplo); trans = up(trans);
    if (side != 'L' && side != 'R') return bad_arg(1, "side");
    if (uplo != 'L' && uplo != 'U') return bad_arg(2, "uplo");
    if (trans != 'N' && trans != 'T' && trans != 'C') return bad_arg(3, "trans");
    if (method != NLA_RES_FP64 && method != NLA_RES_INT8) return bad_arg(4, "method");
    if (!std::isfinite(alpha)) return bad_arg(5, "alpha");
    if (n < 0) return bad_arg(6, "n");
    if (m < 0) return bad_arg(7, "m");
    if (n == 0 || m == 0) return NLA_OK;
    const int64_t rows = side == 'L' ? n : m;
    if (A == nullptr) return bad_arg(8, "A");
    if (lda < n) return bad_arg(9, "lda");
    if (X == nullptr) return bad_arg(10, "X");
    if (ldx < rows) return bad_arg(11, "ldx");
    if (B == nullptr) return bad_arg(12, "B");
    if (ldb < rows) return bad_arg(13, "ldb");
    if (R == nullptr) return bad_arg(14, "R");
    if (ldr < rows) return bad_arg(15, "ldr");
    if (method == NLA_RES_INT8 && n > OZ_N_MAX) {
        std::snprintf(g_last_error, sizeof(g_last_error), "int8 residual: n = %lld above %lld", (long long)n,
                      (long long)OZ_N_MAX);
        return NLA_ERR_UNSUPPORTED;
    }
    const int64_t need = nla_residual_workspace_bytes(side, method, n, m);
    if (workspace == nullptr || ((uintptr_t)workspace & 255) != 0)
        return bad_arg(16, "workspace (needs 256-byte alignment)");
    if (workspace_bytes < need) return bad_arg(17, "workspace_bytes");
    hipStream_t s = (hipStream_t)stream;
    char *ws = (char *)workspace;
    const bool tr = (trans != 'N');
    const char ue = tr ? (uplo == 'L' ? 'U' : 'L') : uplo;
    const int64_t cols = side == 'L' ? m : n;
    hipError_t e = launch_scale_copy(rows, cols, alpha, B, ldb, R, ldr, s);
    if (e != hipSuccess) return hip_status(e, "residual");
    GemmArgs<double> g;
    g.alpha = -1.0;
    g.beta = 1.0;
    g.C = R;
    g.ldc = ldr;
    g.K = n;
    if (side == 'L') {
        g.M = n; g.N = m;
        g.X = A; g.tx = tr; g.ldx = lda;
        g.Y = X; g.ty = false; g.ldy = ldx;
        g.tri = (ue == 'L') ? 1 : 2;
    } else {
        g.M = m; g.N = n;
        g.X = X; g.tx = false; g.ldx = ldx;
        g.Y = A; g.ty = tr; g.ldy = lda;
        g.tri = (ue == 'L') ? 3 : 4;
    }
    auto fp64 = [&]() -> hipError_t {
        g.queue = nullptr;
        return timed(NLA_KIND_RESIDUAL, (double)n * n * m, ((double)n * (n + 1) / 2 + 3.0 * n * m) * 8, s,
                     [&] { return launch_gemm<double>(g, s); });
    };
    if (method == NLA_RES_FP64) return hip_status(fp64(), "residual");
    const OzPlan P = oz_plan(n, m);
    const OzLayout L = oz_layout(P, 0);
    const bool ahat_is_a = (side == 'L') != tr;
    const bool ahat_lower = ahat_is_a ? (uplo == 'L') : (uplo == 'U');
    const int64_t xso = (side == 'L') ? ldx : 1, xsk = (side == 'L') ? 1 : ldx;
    e = oz_prepare(P, L, ws, A, ahat_is_a ? 1 : lda, ahat_is_a ? lda : 1, ahat_lower, n, s);
    if (e == hipSuccess) e = oz_timed(P, L, ws, ahat_lower, side == 'R', X, xso, xsk, n, m, R, ldr, R, ldr, s);
    if (e != hipSuccess) return hip_status(e, "residual");
    int flag = 0;
    e = hipMemcpyAsync(&flag, ws + L.flag, sizeof(int), hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) return hip_status(e, "residual");
    if (flag) {
        // an input outside the moduli's range or not finite: FP64 product
        e = launch_scale_copy(rows, cols, alpha, B, ldb, R, ldr, s);
        if (e == hipSuccess) e = fp64();
        if (e != hipSuccess) return hip_status(e, "residual");
    }
    return NLA_OK;
}

int32_t nla_set_config(int32_t key, int32_t value) {
    if (key == NLA_CFG_LEAF_MODE) {
        int prev = leaf_mode_default();
        g_leaf_mode.store(value ? 1 : 0);
        return prev;
    }
    if (key == NLA_CFG_PANEL) {
        int prev = panel_default() ? 1 : 0;
        g_panel.store(value ? 1 : 0);
        return prev;
    }
    if (key == NLA_CFG_PANEL_ROWS_F32 || key == NLA_CFG_PANEL_ROWS_F64) {
        if (value != 0 && (value < 64 || value > PANEL_MAX || value % 64 != 0)) return -1;
        return g_panel_rows[key == NLA_CFG_PANEL_ROWS_F64 ? NLA_F64 : NLA_F32].exchange(value);
    }
    if (key == NLA_CFG_BATCH_CLASSES) {
        if (value < 0 || value > 0x1F) return -1;
        return g_batch_classes.exchange(value | 0x8);
    }
    if (key == NLA_CFG_LU_SPLIT_ROWS) {
        if (value < 0) return -1;
        return g_lu_split_rows.exchange(value);
    }
    if (key == NLA_CFG_PANEL_WAVES) {
        if (value != 0 && value != 4 && value != 8) return -1;
        return g_panel_waves.exchange(value);
    }
    if (key == NLA_CFG_LU_SPIN_LIMIT) {
        if (value < 0) return -1;
        return g_lu_spin_limit.exchange(value);
    }
    if (key == NLA_CFG_BATCH_ORDER) {
        if (value < 0) return -1;
        return g_batch_order.exchange(value);
    }
    if (key == NLA_CFG_OZ_VARIANT) {
        if (value < 0 || value > 3) return -1;
        return g_oz_variant.exchange(value);
    }
    if (key == NLA_CFG_RESIDUAL) {
        if (value != NLA_RES_FP64 && value != NLA_RES_INT8) return -1;
        return g_residual.exchange(value);
    }
    if (key == NLA_CFG_PANEL_STRIP_F32 || key == NLA_CFG_PANEL_STRIP_F64) {
        if (value != 16 && value != 32) return -1;
        return g_panel_strip[key == NLA_CFG_PANEL_STRIP_F64 ? NLA_F64 : NLA_F32].exchange(value);
    }
    return -1;
}

int32_t nla_prof_enable(int32_t on) {
    std::lock_guard<std::mutex> g(g_prof_mu);
    for (auto &r : g_prof) {
        (void)hipEventSynchronize(r.e1);
        g_event_pool.push_back(r.e0);
        g_event_pool.push_back(r.e1);
    }
    g_prof.clear();
    g_prof_on = (on != 0);
    return NLA_OK;
}

int32_t nla_prof_summary(int32_t kind, int64_t *launches, double *ms, double *flops, double *bytes) {
    if (kind < NLA_KIND_GEMM || kind > NLA_KIND_LP_PAIR_GEMM) return bad_arg(1, "kind");
    std::lock_guard<std::mutex> g(g_prof_mu);
    int64_t n = 0;
    double t = 0, f = 0, b = 0;
    for (auto &r : g_prof) {
        if (r.kind != kind) continue;
        hipError_t e = hipEventSynchronize(r.e1);
        if (e != hipSuccess) return hip_status(e, "event sync");
        float dt = 0;
        e = hipEventElapsedTime(&dt, r.e0, r.e1);
        if (e != hipSuccess) return hip_status(e, "event time");
        ++n;
        t += dt;
        f += r.flops;
        b += r.bytes;
    }
    if (launches) *launches = n;
    if (ms) *ms = t;
    if (flops) *flops = f;
    if (bytes) *bytes = b;
    return NLA_OK;
}

int32_t nla_leaf(char func, char side, char uplo, int32_t dtype, int64_t n, int64_t m, const void *A, int64_t lda,
                 void *B, int64_t ldb, void *stream) {
    func = up(func); side = up(side); uplo = up(uplo);
    if (func != 'S' && func != 'M') return bad_arg(1, "func");
    if (side != 'L' && side != 'R') return bad_arg(2, "side");
    if (uplo != 'L' && uplo != 'U') return bad_arg(3, "uplo");
    if (dtype != NLA_F32 && dtype != NLA_F64) return bad_arg(4, "dtype");
    if (n < 0) return bad_arg(5, "n");
    if (m < 0) return bad_arg(6, "m");
    if (n == 0 || m == 0) return NLA_OK;
    if (A == nullptr) return bad_arg(7, "A");
    if (lda < n) return bad_arg(8, "lda");
    if (B == nullptr) return bad_arg(9, "B");
    if (ldb < (side == 'L' ? n : m)) return bad_arg(10, "ldb");
    return nla_rectrxm(side, uplo, 'N', func, dtype, 1.0, n, m, A, lda, B, ldb, stream);
}

}  // extern "C"
