// lqg_index.h -- the index handle behind the C ABI, shared by the host files.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <algorithm>
#include <string>
#include <vector>

#include "../../include/liteqg.h"
#include "lqg_internal.h"

namespace lqg {

lqg_status fail(lqg_status s, const std::string& msg);
lqg_status hip_fail(hipError_t e, const char* what);
const char* last_error();

#define LQG_HIP_TRY(expr)                                     \
    do {                                                      \
        hipError_t e_ = (expr);                               \
        if (e_ != hipSuccess) return ::lqg::hip_fail(e_, #expr); \
    } while (0)

template <typename T>
struct DevBuf {
    T* p = nullptr;
    size_t cap = 0;  // elements
    // A view of memory the buffer does not own (a region of a static device
    // array): never allocated, grown or freed; a reserve beyond cap fails.
    bool fixed = false;
    hipError_t reserve(size_t n) {
        if (n <= cap) return hipSuccess;
        if (fixed) return hipErrorOutOfMemory;
        if (p) (void)hipFree(p);
        p = nullptr;
        cap = 0;
        hipError_t e = hipMalloc(reinterpret_cast<void**>(&p), std::max<size_t>(n, 1) * sizeof(T));
        if (e == hipSuccess) cap = n;
        return e;
    }
    void release() {
        if (fixed) return;
        if (p) (void)hipFree(p);
        p = nullptr;
        cap = 0;
    }
};

// Device scratch of one search launch.  Launches that may run at the same
// time (different streams, graph replays, serving lanes) each own one.
struct Scratch {
    DevBuf<uint32_t> counter;  // work-queue head
    DevBuf<uint32_t> err;      // FastScan visited-set overflow flag
    DevBuf<uint32_t> gvis;     // FastScan overflow kernel's per-wave visited tables
    DevBuf<uint32_t> ovf;      // FastScan overflow list (count, counter, query indices)
    DevBuf<float> rot;         // FastScan queries rotated by the GEMM (nq x P)
    // Per-query completion flags the launch sets (host-coherent memory owned by
    // the caller, e.g. a serving lane), or null.
    uint32_t* done = nullptr;
    // FastScan launches without the per-wave device overflow visited tables
    // (a query then fails once its LDS table is half full).
    bool no_overflow = false;
    void release() {
        counter.release();
        err.release();
        gvis.release();
        ovf.release();
        rot.release();
    }
};

// Pool-capacity classes: 64, 128, 192, 256, 512, 1024 (a kernel variant each;
// 192 = three pool slots per lane, so ef up to 192 does not pay for four).
constexpr int kClasses = 6;

// Smallest pool-capacity class that holds ef, 0 if none does.
int ef_class(uint32_t ef);

// Index 0..kClasses-1 of a capacity class, and the capacity of an index.
int class_slot(int ef_cap);
int class_cap(int slot);

// Launch settings used when none were set for a class, on an index of
// dimension dim.
LaunchCfg default_launch(int ef_cap, uint64_t dim);

// Smallest dimension whose SQ8 float-path launches split each query's
// distance evaluation over two waves by default.
constexpr uint64_t kSplitMinDim = 512;
// Largest launch the default split applies to (four waves per SIMD of the
// device with one wave per query).
constexpr uint64_t kSplitMaxQueries = 4096;

}  // namespace lqg

struct lqg_index {
    uint64_t n = 0, dim = 0, deg = 0;
    uint32_t ep = 0;
    uint32_t ef = 0;
    bool ef_set = false;
    bool have_data = false;  // loaded or built
    lqg_order order = LQG_ORDER_REFERENCE;

    // host copy: exactly the file's content, split by section
    std::vector<float> raw;      // n x dim
    std::vector<uint32_t> nbr;   // n x deg
    std::vector<uint8_t> qdata;  // n x dim
    std::vector<float> sq8;      // dim x {min, scale}

    // Bumped whenever device memory the kernels read may be freed, moved or
    // rewritten (upload, reload, rebuild, sidecar replaced): a captured graph
    // holds those addresses and refuses to replay across a change.
    uint64_t generation = 0;

    // device copy
    bool on_device = false;
    bool sq8_identity = false;  // all SQ8 params are min 0, scale 1
    bool sq8_integer = true;    // the exact integer kernel may be used (lqg_set_sq8_integer)
    hipStream_t stream = nullptr;
    lqg::DevBuf<uint8_t> d_qdata;
    lqg::DevBuf<uint32_t> d_nbr;
    lqg::DevBuf<float> d_min, d_scale;
    // per-call scratch (grow only)
    lqg::DevBuf<float> d_q;
    lqg::DevBuf<uint32_t> d_ids, d_ef, d_k, d_stats;
    lqg::Scratch scratch;  // for the entry points that run on the handle's stream
    lqg::DevBuf<float> d_dist;
    lqg::DevBuf<uint32_t> d_visit;  // exact visit counting bitmaps
    // plain search (one query per call): pinned, host-coherent buffers the
    // kernels read the query from and write the ids and the completion flag
    // to (word 0 of p_out, ids from word 16), and their device addresses
    float* p_q = nullptr;
    uint32_t* p_out = nullptr;
    uint32_t p_out_k = 0;  // ids p_out holds
    float* g_pq = nullptr;
    uint32_t* g_pout = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    // launch settings per pool-capacity class (64, 128, 192, 256, 512, 1024)
    lqg::LaunchCfg launch[lqg::kClasses];
    bool launch_set[lqg::kClasses] = {};

    // FastScan RaBitQ search (search_qg, qg.hpp:252-287): node records on the
    // device (raw | codes | factors | ids) and the rotation signs
    lqg_mode mode = LQG_MODE_SQ8;
    bool h2_ready = false;
    uint32_t h2_p = 0, h2_stride = 0, h2_off_codes = 0, h2_off_fac = 0, h2_off_ids = 0;
    std::vector<float> h2_signs;
    lqg::DevBuf<uint8_t> d_h2rec;
    lqg::DevBuf<float> d_h2signs;
    lqg::H2Cfg h2_launch[lqg::kClasses];
    bool h2_launch_set[lqg::kClasses] = {};
    lqg_rotation h2_rotation = LQG_ROTATION_FHT;
    bool h2_generic = false;     // lqg_h2_set_generic
    int build_candidates = 0;    // lqg_build_candidates of build_index
    uint32_t h2_gvis_bits = 14;  // per-wave overflow visited table, log2 entries (0: none)

    lqg::H2Dev h2dev() const {
        lqg::H2Dev ix;
        ix.rec = d_h2rec.p;
        ix.signs = d_h2signs.p;
        ix.n = static_cast<uint32_t>(n);
        ix.dim = static_cast<uint32_t>(dim);
        ix.deg = static_cast<uint32_t>(deg);
        ix.p = h2_p;
        ix.ep = ep;
        ix.stride = h2_stride;
        ix.off_codes = h2_off_codes;
        ix.off_fac = h2_off_fac;
        ix.off_ids = h2_off_ids;
        ix.generic = h2_generic ? 1u : 0u;
        return ix;
    }

    lqg::DevIndex dev() const {
        lqg::DevIndex ix;
        ix.qdata = d_qdata.p;
        ix.nbr = d_nbr.p;
        ix.sq8_min = d_min.p;
        ix.sq8_scale = d_scale.p;
        ix.n = static_cast<uint32_t>(n);
        ix.dim = static_cast<uint32_t>(dim);
        ix.deg = static_cast<uint32_t>(deg);
        ix.ep = ep;
        ix.sq8_identity = sq8_identity ? 1 : 0;
        return ix;
    }
};

namespace lqg {

// Checks the device, creates the stream and uploads the host copy (once).
lqg_status ensure_device(lqg_index* h);

// Takes the current device's address of the static scratch array of the
// allocation-free entry point (lqg_static.hip), once per device.
lqg_status static_scratch_setup();

// Enqueues one search launch over device arrays, in the handle's mode, with
// scratch `sc` (null: the handle's own).
lqg_status run_batch(lqg_index* h, const float* d_queries, uint64_t nq, const uint32_t* d_ef,
                     const uint32_t* d_k, uint32_t ef_max, uint32_t kmax, lqg_order order,
                     uint32_t* d_results, float* d_dists, uint32_t* d_stats, hipStream_t s,
                     bool count_visits = false, Scratch* sc = nullptr, int queries_int = -1);

// FastScan counterpart of run_batch (lqg_h2.cpp); d_stats gets (E, pops).
lqg_status run_h2_batch(lqg_index* h, const float* d_queries, uint64_t nq, const uint32_t* d_ef,
                        const uint32_t* d_k, uint32_t ef_max, uint32_t kmax, uint32_t* d_results,
                        float* d_dists, uint32_t* d_stats, hipStream_t s, Scratch* sc);

// Reserves the device scratch a FastScan batch of up to nq queries (ef up to
// ef_max, k up to kmax) needs, so that later launches allocate nothing (no-op
// without a sidecar).
lqg_status h2_prepare_scratch(lqg_index* h, uint64_t nq, uint32_t ef_max, uint32_t kmax,
                              Scratch* sc);

// Launch settings of the FastScan kernel used when none were set for a class.
H2Cfg h2_default_launch(int ef_cap);

// The settings a FastScan launch of class ef_cap uses on this index: the ones
// set, or the defaults without a register cap the index has no kernel for.
H2Cfg h2_effective_launch(const lqg_index* h, int ef_cap);

// Frees the FastScan records (index reloaded or rebuilt).
void h2_reset(lqg_index* h);

// After a synchronised FastScan launch with scratch sc (null: the handle's):
// fails if a visited set overflowed.
lqg_status h2_check_error(lqg_index* h, Scratch* sc = nullptr);

// Graph build with FastScan candidates (lqg_h2.cpp): records and rotated rows
// once, then per iteration the expanded nodes of every node's search on the
// current graph; h2_build_end drops the build's records.
lqg_status h2_build_begin(lqg_index* h, uint64_t seed, DevBuf<float>& rot);
lqg_status h2_run_expansions(lqg_index* h, const float* d_queries, uint64_t nq, const float* rot,
                             uint32_t ef, uint32_t* exp_ids, uint32_t exp_cap,
                             DevBuf<uint32_t>& out_k1, hipStream_t s);
lqg_status h2_build_candidates(lqg_index* h, const float* d_raw, const float* rot, uint32_t ef,
                               uint32_t* exp_ids, uint32_t exp_cap, DevBuf<uint32_t>& out_k1,
                               hipStream_t s);
void h2_build_end(lqg_index* h);

// Device graph build for lqg_build_index.
lqg_status build_index(lqg_index* h, const float* data, uint32_t ef_build, uint32_t num_iter);

}  // namespace lqg
