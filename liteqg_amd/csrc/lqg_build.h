// lqg_build.h -- device buffers and entry points of the graph builder.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace lqg {

struct BuildParams {
    uint32_t n, dim, deg;
    uint32_t ncand;   // beam-search results kept per node (min(EF, 750))
    uint32_t pr_cap;  // pruned candidates kept per node for the refine pass (300)
    uint64_t seed;
    int refine;       // last iteration: graph_refine fill
};

struct BuildBuffers {
    const float* raw;    // N x dim
    const uint8_t* qdata;     // N x dim SQ8 codes
    const float* sq8_min;     // dim
    const float* sq8_scale;   // dim
    uint32_t* graph;     // N x deg, the device graph the search reads
    uint32_t* cand_ids;  // N x ncand
    uint32_t* kept_ids;  // N x deg
    float* kept_d;       // N x deg
    uint32_t* kept_n;    // N
    uint32_t* pr_ids;    // N x pr_cap
    float* pr_d;         // N x pr_cap
    uint32_t* pr_n;      // N
    uint32_t* in_cnt;    // N + 1
    uint32_t* in_off;    // N + 1
    uint32_t* in_ids;    // N x deg: reverse-edge sources, CSR rows by target
    uint32_t* rev_keys;      // N x deg: edge targets (sort input)
    uint32_t* rev_keys_out;  // N x deg: sorted targets
    uint32_t* rev_vals;      // N x deg: edge sources (sort input)
    uint32_t* counter;   // 1
    void* scan_tmp;      // build_temp_bytes(N, deg) bytes: scan and sort scratch
    size_t scan_tmp_bytes;
};

// Bytes of temporary storage exclusive_scan_u32 needs for n elements.
size_t scan_temp_bytes(uint32_t n);

// Bytes of temporary storage one builder iteration needs (the larger of the
// reverse-edge scan and sort), 0 on error.
size_t build_temp_bytes(uint32_t n, uint32_t deg);

// out[i] = sum(in[0..i)), i < n.
hipError_t exclusive_scan_u32(const uint32_t* in, uint32_t* out, uint32_t n, void* tmp,
                              size_t tmp_bytes, hipStream_t s);

// One builder iteration after the beam search has filled cand_ids:
// heuristic prune, reverse edges, re-prune (and refine fill), new graph rows.
hipError_t build_run_iteration(const BuildBuffers& b, const BuildParams& p, hipStream_t s);

}  // namespace lqg
