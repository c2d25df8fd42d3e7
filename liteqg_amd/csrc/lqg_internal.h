// lqg_internal.h -- declarations shared by the host library and the kernels.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <string.h>

namespace lqg {

// The bits of an enum argument that crossed the C ABI.  A C caller may pass
// any int there, and reading a value outside the enumerators' range as the
// enum type is undefined in C++, so arguments are validated on their bits
// and converted back only once they are known to be valid.
template <typename E>
inline int abi_enum(const E& e) {
    static_assert(sizeof(E) == sizeof(int), "C enums of this ABI are int-sized");
    int v;
    memcpy(&v, &e, sizeof v);
    return v;
}


constexpr uint32_t kInvalid = 0xFFFFFFFFu;
constexpr int kWave = 64;

// Completion flag values a launch writes per query (host-coherent memory,
// SearchJob::done / H2Job::done).
constexpr uint32_t kFlagDone = 1u;    // results written
constexpr uint32_t kFlagFailed = 2u;  // the query could not finish (ids 0xFFFFFFFF)

// Device view of an index for the SQ8 search (the reference's live path,
// qg.hpp:298-344).  qdata rows are `dim` bytes, neighbour rows `deg` ids.
struct DevIndex {
    const uint8_t* qdata;   // N x dim SQ8 codes (file section 3)
    const uint32_t* nbr;    // N x deg neighbour ids (file section 2, float offset D)
    const float* sq8_min;   // dim
    const float* sq8_scale; // dim
    uint32_t n, dim, deg, ep;
    int sq8_identity;       // every dim has min 0 and scale 1 (decode is exact)
};

// One batch of queries for the search kernel.
// The search kernel's work queues: one counter per XCD, kQueueStride words
// apart (a 128-byte line each).  A launch's queue block is kQueueWords words.
constexpr uint32_t kXcds = 8;
constexpr uint32_t kQueueStride = 32;
constexpr uint32_t kQueueWords = kXcds * kQueueStride;

struct SearchJob {
    const float* queries;   // nq x dim
    const uint32_t* ef;     // per query, or null -> ef_all
    const uint32_t* k;      // per query, or null -> k_all
    uint32_t nq, ef_all, k_all, kstride;
    int order;              // 0 reference, 1 nearest
    uint32_t* out_ids;      // nq x kstride
    float* out_d;           // nq x kstride or null
    uint32_t* out_stats;    // nq x 2 (E, V) or null
    uint32_t* counter;      // kQueueWords queue words (device), zeroed before launch
    // Exact visit counting (diagnostic launch only; null otherwise): one bitmap
    // of visit_words 32-bit words per wave of the grid.  out_stats then holds
    // (E, number of distinct nodes evaluated) instead of (E, evaluations).
    uint32_t* visit_bits;
    uint32_t visit_words;
    uint32_t visit_waves = 0;  // bitmaps allocated; a wave beyond them counts nothing
    // Float-path launch that follows an integer-path launch over the same
    // queries: skip the queries the integer path took.
    int skip_int = 0;
    // Queries the integer-path launch left to the float path (counted by the
    // integer launch, read by the float launch, which returns when it is 0).
    uint32_t* left = nullptr;
    // Per-query completion flags in host-coherent memory, or null: the launch
    // that writes a query's results then sets its flag to 1 (after a
    // system-scope fence), so a host can take each result as soon as it exists.
    uint32_t* done = nullptr;
};

// Launch geometry chosen by the host (autotuner or defaults).
struct LaunchCfg {
    int ef_cap;          // pool capacity class: 64, 128, 192, 256, 512 or 1024
    int waves_per_group; // waves per workgroup
    int cache_bits;      // log2 of the per-wave visited cache (4..14)
    int groups;          // workgroups in the persistent grid (0 = fill the device)
    int occupancy;       // register cap: 8 = 64 VGPRs, 6 = 80 VGPRs, 0 = uncapped
    int eval = 0;        // 0: SQ8 float distances; 1: exact integer distances (v_dot4);
                         // 3: the same for rows of at most 128 bytes
    int prefetch = 0;    // launches of at most sq8_latency_max(dim) queries: 0 = the
                         // batch kernel; 1..4 = latency mode with that prefetch depth
    int split = 1;       // SQ8 float path: waves per query sharing each pass's
                         // distance evaluation (h1_split_kernel); 1 = off, 2 or 4
};

// Largest launch that takes the latency-mode kernel (one query per
// workgroup, a prefetch wave beside each search wave) when its class's
// prefetch depth is set: FastScan, and the SQ8 path on rows of more than 256
// dims.  On 1,000-query launches latency mode was slower for 960-d rows in
// both modes (FastScan 4.56 -> 4.74 ms; SQ8 without split evaluation 11.3 ->
// 19.6 ms, 200k x 960 degree 64, ef U{64..1024}).
constexpr uint32_t kLatencyMaxQueries = 256;

// The same limit for the SQ8 path on rows of at most 256 dims: 1,000
// queries of 100k x 64 uniform, latency mode depth 2 against the batch
// kernel at ef 50 / 100 / 200: 0.191 / 0.319 / 0.614 -> 0.175 / 0.291 /
// 0.598 ms.
constexpr uint32_t kLatencyMaxQueriesNarrow = 1024;

__host__ __device__ inline uint32_t sq8_latency_max(uint32_t dim) {
    return dim <= 256u ? kLatencyMaxQueriesNarrow : kLatencyMaxQueries;
}

// Device view of an index for the FastScan RaBitQ search (search_qg,
// qg.hpp:252-287).  One record per node, `stride` bytes (a multiple of 128):
//   raw f32[dim] | codes[deg * p/8] | triple_x f32[deg] | fac_dq f32[deg] |
//   fac_vq f32[deg] | ids u32[deg]
// so that expanding a node is one contiguous read.  Codes are vector-major:
// neighbour j's p/8 bytes, byte b holding codebook 2b (dims 8b..8b+3, bit 3 =
// dim 8b) in its low nibble and codebook 2b+1 in its high nibble.
struct H2Dev {
    const uint8_t* rec;
    const float* signs;     // p rotation signs, +-1/sqrt(p)
    uint32_t n, dim, deg, p, ep;
    uint32_t stride, off_codes, off_fac, off_ids;  // bytes
    uint32_t generic = 0;  // 1: the generic search kernel where a specialised one exists
};

// One batch of queries for the FastScan search kernel.
struct H2Job {
    const float* queries;   // nq x dim
    const uint32_t* ef;     // per query, or null -> ef_all
    const uint32_t* k;      // per query, or null -> k_all
    uint32_t nq, ef_all, k_all, kstride;
    uint32_t* out_ids;      // nq x kstride
    float* out_d;           // nq x kstride or null
    uint32_t* out_stats;    // nq x 2 (expansions, pops) or null
    uint32_t* counter;      // work-queue head, zeroed before launch
    uint32_t* gvis;         // per-wave overflow visited tables (gvis_bits each)
    uint32_t gvis_bits;
    uint32_t* err;          // set to 1 if a visited set overflows both tables
    const float* rot = nullptr;  // nq x p queries rotated beforehand (GEMM), or null (FHT)
    uint32_t* done = nullptr;    // per-query completion flags, as in SearchJob
    // Overflow list (2 + nq words: count, overflow-kernel work counter, query
    // indices), zeroed before launch; null: a query whose LDS visited table
    // would pass half full fails instead.
    uint32_t* ovf = nullptr;
    // Graph build (find_candidates, qg.hpp:439-473): the ids of the first
    // exp_cap nodes each query expands, in expansion order, to exp_ids[q *
    // exp_cap ...] (slots past its expansions keep what the caller put
    // there).  General kernel only; null in searches.
    uint32_t* exp_ids = nullptr;
    uint32_t exp_cap = 0;
    // Latency mode (launches of at most kLatencyMaxQueries queries): records
    // of the best unexpanded pool entries a query's prefetch wave loads ahead.
    uint32_t prefetch_targets = 0;
};

struct H2Cfg {
    int ef_cap;          // pool capacity class, as for LaunchCfg
    int waves_per_group; // 1..4
    int groups;          // persistent grid size (0 = fill the device)
    int hash_bits;       // log2 entries of the per-wave LDS visited table
    int occupancy = 0;   // 0 = uncapped registers, 6 or 8 = capped for that many waves per SIMD
    int prefetch = 0;    // launches of at most kLatencyMaxQueries queries: 0 = the batch
                         // kernel; d = 1..3: latency mode, a prefetch wave per query loading
                         // the records of the 2^d best unexpanded pool entries
};

// Kernel-table key of the FastScan latency-mode kernels (in place of a
// register cap).
constexpr int kOccLatency = -2;

// Returns LDS bytes one workgroup needs for this configuration.
size_t search_lds_bytes(const DevIndex& ix, const LaunchCfg& cfg);

// FastScan search launches (h2_search.hip).
size_t h2_lds_bytes(const H2Dev& ix, const H2Cfg& cfg, uint32_t kmax);
// Rotated queries (nq x p): method 0 = the FHT of the search kernel (one
// wave per query, bit-identical to it), 1 = one FP32 GEMM on the matrix cores.
hipError_t h2_launch_rotate_queries(const H2Dev& ix, const float* queries, uint32_t nq,
                                    int method, float* out, hipStream_t stream);
// Whether the register-capped variants (occupancy 6, 8) exist for this index
// shape (at most 4 code words per lane).
bool h2_occupancy_ok(const H2Dev& ix);
// Whether the FastScan kernels take this kmax (k <= 256, in registers).
bool h2_kmax_ok(uint32_t kmax);
// Whether a kernel variant exists for this index shape, register cap and kmax.
bool h2_variant_ok(const H2Dev& ix, int occupancy, uint32_t kmax);
int h2_groups(const H2Dev& ix, const H2Cfg& cfg, uint32_t kmax, uint32_t nq);
hipError_t h2_launch_search(const H2Dev& ix, const H2Job& job, const H2Cfg& cfg,
                            hipStream_t stream);
// The overflow kernel after a search launch with job.ovf set: `groups`
// one-wave workgroups, wave g with the device visited table at
// job.gvis + (g << job.gvis_bits).
constexpr uint32_t kOverflowWaves = 256;
hipError_t h2_launch_overflow(const H2Dev& ix, const H2Job& job, const H2Cfg& cfg,
                              uint32_t groups, hipStream_t stream);
// Query preparation (and optionally one node's FastScan estimates) for each
// query: lut (nq x 4p), params (nq x 3: lo, width, sumq bits), sqr_y (nq), est
// (nq x deg); nodes may be null (no scan).
hipError_t h2_launch_probe(const H2Dev& ix, const float* queries, uint32_t nq,
                           const uint32_t* nodes, uint8_t* lut, float* params, float* sqr_y,
                           float* est, hipStream_t stream);
// Sidecar build: rotates every raw vector, then fills codes and factors of
// every record (raw and ids must already be in the records).
hipError_t h2_launch_sidecar(const H2Dev& ix, float* rot_scratch, hipStream_t stream);
// The two halves of h2_launch_sidecar: the records' raw rows rotated into rot
// (n x p), then codes and factors of every record from rot and its ids.
hipError_t h2_launch_rotate_rows(const H2Dev& ix, float* rot, hipStream_t stream);
hipError_t h2_launch_codes(const H2Dev& ix, const float* rot, hipStream_t stream);
// Copies w bytes per node between a contiguous device array (n x w) and
// byte offset `off` of every record (to_rec: into the records).
hipError_t h2_launch_rows(const H2Dev& ix, uint32_t off, void* buf, uint32_t w, bool to_rec,
                          hipStream_t stream);

// Enqueues the search kernel (queries pulled from job.counter).
hipError_t launch_search(const DevIndex& ix, const SearchJob& job, const LaunchCfg& cfg,
                         hipStream_t stream);

// Zeroes the work-queue counter with a kernel (capturable in a HIP graph).
hipError_t launch_reset_counter(uint32_t* counter, hipStream_t stream, uint32_t n = 1);

// Number of workgroups that fill the device for this configuration.
int search_full_grid(const DevIndex& ix, const LaunchCfg& cfg);

// Workgroups launch_search will use for this job (grid after clamping).
int search_groups(const DevIndex& ix, const SearchJob& job, const LaunchCfg& cfg);

}  // namespace lqg
