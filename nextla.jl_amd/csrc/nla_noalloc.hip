// nla_rectrxm_noalloc: nla_rectrxm for callers that cannot allocate device
// memory (code under HIP graph capture, serving loops, harnesses that forbid
// hipMalloc).
//
// The only device memory the plain path needs beyond A and B is the TRSM's
// scratch: the inverses of A's diagonal 64-blocks and one flag per block,
// written by the inversion launch and read by the panels of the same call.
// nla_rectrxm keeps it in a per-stream buffer it allocates on first use; here
// it lives in `scratch`, a file-scope device array of NLA_NOALLOC_SCRATCH_BYTES
// that every device loading this library holds until the process ends.  Its
// device address is resolved once per device with hipGetSymbolAddress (in
// host code the array's name is only a stand-in, not a device address), and
// that address is what the kernels receive.  The call makes no hipMalloc,
// hipFree or other allocating call.
//
// Every call on one device shares that device's array: each call writes the
// blocks it reads before reading them (nothing needs clearing), and calls
// that overlap on one device (two streams, two host threads) race on it.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <mutex>

#include "../../include/nla.h"
#include "nla_internal.h"

namespace nla {
namespace {

static __device__ __align__(256) char scratch[NLA_NOALLOC_SCRATCH_BYTES];

constexpr int NOALLOC_MAX_DEVICES = 64;
std::mutex g_scratch_mu;
void *g_scratch_ptr[NOALLOC_MAX_DEVICES] = {};

// scratch's device address on the current device, resolved on the first
// call with that device current and kept for later calls.
int32_t scratch_address(void **out) {
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return report_hip(e, "nla_rectrxm_noalloc: current device");
    if (dev < 0 || dev >= NOALLOC_MAX_DEVICES)
        return report_status(NLA_ERR_UNSUPPORTED, "nla_rectrxm_noalloc: device index beyond the address table");
    std::lock_guard<std::mutex> g(g_scratch_mu);
    if (g_scratch_ptr[dev] == nullptr) {
        void *p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return report_hip(e, "nla_rectrxm_noalloc: scratch address");
        if (p == nullptr) return report_status(NLA_ERR_HIP, "nla_rectrxm_noalloc: scratch address is null");
        g_scratch_ptr[dev] = p;
    }
    *out = g_scratch_ptr[dev];
    return NLA_OK;
}

}  // namespace
}  // namespace nla

extern "C" int32_t nla_rectrxm_noalloc(char side, char uplo, char trans, char func, int32_t dtype, double alpha,
                                       int64_t n, int64_t m, const void *A, int64_t lda, void *B, int64_t ldb,
                                       void *stream) {
    using namespace nla;
    const int32_t chk = rectrxm_check(side, uplo, trans, func, dtype, n, m, A, lda, B, ldb);
    if (chk != NLA_OK || n == 0 || m == 0) return chk;
    const int64_t need = rectrxm_workspace_bytes(func, dtype, n);
    if (need > (int64_t)NLA_NOALLOC_SCRATCH_BYTES) {
        char msg[200];
        std::snprintf(msg, sizeof(msg),
                      "nla_rectrxm_noalloc: n = %lld needs %lld bytes of scratch, above the %lld bytes it holds "
                      "(use nla_rectrxm)",
                      (long long)n, (long long)need, (long long)NLA_NOALLOC_SCRATCH_BYTES);
        return report_status(NLA_ERR_WORKSPACE, msg);
    }
    void *ptr = nullptr;
    const int32_t st = scratch_address(&ptr);
    if (st != NLA_OK) return st;
    return rectrxm_in_workspace(side, uplo, trans, func, dtype, alpha, n, m, A, lda, B, ldb, (hipStream_t)stream,
                                ptr, (int64_t)NLA_NOALLOC_SCRATCH_BYTES);
}
