// lqg_static.hip -- the batch search entry point that allocates nothing, for
// callers that cannot call hipMalloc (code under HIP-graph capture, tight
// serving loops, harnesses that forbid allocation).  It is
// lqg_search_batch_device with the launch scratch (work-queue words, the
// FastScan overflow flag and GEMM-rotated queries) taken from one static
// device array instead of buffers the library allocates.
//
// The array lives in device memory of every GPU the library runs on until
// the process ends.  Host code never uses its name as an address: each
// device's address is asked once with hipGetSymbolAddress, with that device
// current, and kept per device.  That happens when an index is first put on
// a device (ensure_device), so the runtime has placed the array there before
// any allocation-free call; the entry point asks again only on a device no
// index has been uploaded to.
#include <hip/hip_runtime.h>

#include <mutex>

#include "lqg_index.h"

namespace {

using lqg::fail;

// Scratch regions, in 32-bit words: the work-queue block of one launch
// (2 x kQueueWords + 1 for the SQ8 path, 1 for FastScan), the FastScan
// overflow flag, and the queries rotated by the GEMM (nq x P floats).
constexpr size_t kCounterWords = 1024;
constexpr size_t kErrWords = 64;
constexpr size_t kRotFloats = size_t{1} << 18;  // 1 MiB: nq * P <= 262,144
constexpr size_t kScratchWords = kCounterWords + kErrWords + kRotFloats;
static_assert(2 * lqg::kQueueWords + 1 <= kCounterWords, "queue block does not fit");

}  // namespace

// 1,052,928 bytes on every GPU the library runs on.  Calls on one GPU share
// it: the launches zero the words that must start at zero (queue, flag)
// themselves, and two calls that overlap on one GPU race.
static __device__ __align__(256) uint32_t scratch[kScratchWords];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_mu;
uint32_t* g_scratch[kMaxDevices] = {};

// The array's address on the current device (asked once per device).
lqg_status scratch_on_device(uint32_t** out) {
    int dev = 0;
    LQG_HIP_TRY(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) return fail(LQG_ERR_DEVICE, "device ordinal out of range");
    std::lock_guard<std::mutex> g(g_mu);
    if (!g_scratch[dev]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return lqg::hip_fail(e, "hipGetSymbolAddress(scratch)");
        if (!p) return fail(LQG_ERR_DEVICE, "hipGetSymbolAddress(scratch) returned null");
        g_scratch[dev] = static_cast<uint32_t*>(p);
    }
    *out = g_scratch[dev];
    return LQG_OK;
}

}  // namespace

namespace lqg {

lqg_status static_scratch_setup() {
    uint32_t* p = nullptr;
    return scratch_on_device(&p);
}

}  // namespace lqg

extern "C" {

lqg_status lqg_search_batch_device_noalloc(lqg_index* h, const float* d_queries, uint64_t nq,
                                           const uint32_t* d_ef, const uint32_t* d_k,
                                           uint32_t ef_max, uint32_t kmax, lqg_order order,
                                           uint32_t* d_results, float* d_dists,
                                           uint32_t* d_stats, void* stream) {
    if (!h || !d_queries || !d_results) return fail(LQG_ERR_ARG, "null argument");
    if (nq == 0 || nq > 0x7FFFFFFFull) return fail(LQG_ERR_ARG, "nq must be in [1, 2^31)");
    if (kmax == 0) return fail(LQG_ERR_ARG, "kmax must be positive");
    if (lqg::abi_enum(order) != LQG_ORDER_REFERENCE && lqg::abi_enum(order) != LQG_ORDER_NEAREST)
        return fail(LQG_ERR_ARG, "unknown order");
    if (ef_max == 0) {
        if (d_ef) return fail(LQG_ERR_ARG, "ef_max must bound the per-query ef array");
        if (!h->ef_set) return fail(LQG_ERR_STATE, "set_ef() or ef_max required");
        ef_max = h->ef;
    }
    if (lqg::ef_class(ef_max) == 0)
        return fail(LQG_ERR_UNSUPPORTED, "ef above 1024 is not supported");
    // uploading the index allocates: it must already be resident (any search
    // call on the index, or a graph capture, has put it there)
    if (!h->on_device)
        return fail(LQG_ERR_STATE,
                    "the index is not on the device yet: run one search on it before calling the "
                    "allocation-free entry point");
    const bool fs = h->mode == LQG_MODE_FASTSCAN;
    if (fs && !h->h2_ready) return fail(LQG_ERR_STATE, "FastScan mode needs a sidecar");
    if (fs && h->h2_rotation == LQG_ROTATION_GEMM && nq * h->h2_p > kRotFloats)
        return fail(LQG_ERR_ARG,
                    "nq x P exceeds the static rotation buffer (262,144 floats): split the batch "
                    "or use the FHT rotation");
    uint32_t* base = nullptr;
    lqg_status st = scratch_on_device(&base);
    if (st != LQG_OK) return st;
    lqg::Scratch sc;
    sc.counter.p = base;
    sc.counter.cap = kCounterWords;
    sc.counter.fixed = true;
    sc.err.p = base + kCounterWords;
    sc.err.cap = kErrWords;
    sc.err.fixed = true;
    sc.rot.p = reinterpret_cast<float*>(base + kCounterWords + kErrWords);
    sc.rot.cap = kRotFloats;
    sc.rot.fixed = true;
    sc.gvis.fixed = true;  // cap 0: never used, since no_overflow is set
    sc.no_overflow = true;
    hipStream_t s = stream ? static_cast<hipStream_t>(stream) : h->stream;
    return lqg::run_batch(h, d_queries, nq, d_ef, d_k, ef_max, kmax, order, d_results, d_dists,
                          d_stats, s, false, &sc);
}

}  // extern "C"
