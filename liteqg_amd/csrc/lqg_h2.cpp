// lqg_h2.cpp -- host side of the FastScan RaBitQ path (search_qg,
// qg.hpp:252-287): the node-record layout, sidecar construction on the
// device, installation from host arrays (either code layout), persistence,
// the parity probe and the batch launcher.
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <cmath>
#include <string>
#include <vector>

#include "lqg_index.h"

namespace lqg {

namespace {

uint32_t padded_dim(uint64_t d) {
    uint32_t p = 1;
    while (p < d) p <<= 1;
    return p;
}

uint32_t align_up(uint32_t x, uint32_t a) { return (x + a - 1) / a * a; }

// Record layout for the loaded index; fails for shapes the kernels do not
// take.  P = 1 << ceil(log2 D) (utils/tools.hpp:25-34); FHTRotator handles
// P = 2^6 .. 2^11 (utils/rotator.hpp:54-77).
lqg_status set_layout(lqg_index* h) {
    const uint32_t p = padded_dim(h->dim);
    if (p < 64 || p > 2048)
        return fail(LQG_ERR_UNSUPPORTED,
                    "FastScan needs a padded dimension in [64, 2048] (33 <= D <= 2048)");
    const uint32_t lpv = h->deg >= 64 ? 1u : 2u;
    if (p / 8 / lpv / 4 > 32)
        return fail(LQG_ERR_UNSUPPORTED, "FastScan: P = 2048 with degree >= 64 is not supported");
    const uint32_t deg = static_cast<uint32_t>(h->deg), dim = static_cast<uint32_t>(h->dim);
    h->h2_p = p;
    h->h2_off_codes = align_up(4 * dim, 16);
    h->h2_off_fac = h->h2_off_codes + deg * (p / 8);
    h->h2_off_ids = h->h2_off_fac + 12 * deg;
    h->h2_stride = align_up(h->h2_off_ids + 4 * deg, 128);
    return LQG_OK;
}

// Moves w bytes per node between host memory (n x w, contiguous) and offset
// `off` of every record: one contiguous copy to or from a device staging
// array, then the row kernel (a pitched host copy of millions of short rows
// is far slower).
lqg_status host_rows(lqg_index* h, uint32_t off, void* host, size_t w, bool to_rec) {
    DevBuf<unsigned char> tmp;
    LQG_HIP_TRY(tmp.reserve(h->n * w));
    hipError_t e = hipSuccess;
    if (to_rec) {
        e = hipMemcpyAsync(tmp.p, host, h->n * w, hipMemcpyHostToDevice, h->stream);
        if (e == hipSuccess)
            e = h2_launch_rows(h->h2dev(), off, tmp.p, static_cast<uint32_t>(w), true, h->stream);
    } else {
        e = h2_launch_rows(h->h2dev(), off, tmp.p, static_cast<uint32_t>(w), false, h->stream);
        if (e == hipSuccess)
            e = hipMemcpyAsync(host, tmp.p, h->n * w, hipMemcpyDeviceToHost, h->stream);
    }
    if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
    tmp.release();
    if (e != hipSuccess) return hip_fail(e, "record row copy");
    return LQG_OK;
}

// Records with the raw rows and neighbour ids of the host copy in place
// (codes and factors still to be written).
lqg_status alloc_records(lqg_index* h) {
    lqg_status st = set_layout(h);
    if (st != LQG_OK) return st;
    h->h2_ready = false;
    ++h->generation;
    // one spare record after the last: the last pass of a degree that is
    // not a multiple of the pass width reads past the neighbour lists
    LQG_HIP_TRY(h->d_h2rec.reserve((h->n + 1) * static_cast<size_t>(h->h2_stride)));
    LQG_HIP_TRY(hipMemset(h->d_h2rec.p + h->n * static_cast<size_t>(h->h2_stride), 0,
                          h->h2_stride));
    LQG_HIP_TRY(h->d_h2signs.reserve(h->h2_p));
    st = host_rows(h, 0, h->raw.data(), 4 * h->dim, true);
    if (st != LQG_OK) return st;
    st = host_rows(h, h->h2_off_ids, h->nbr.data(), 4 * h->deg, true);
    if (st != LQG_OK) return st;
    LQG_HIP_TRY(hipMemcpy(h->d_h2signs.p, h->h2_signs.data(), 4ull * h->h2_p,
                          hipMemcpyHostToDevice));
    return LQG_OK;
}

uint64_t splitmix64(uint64_t& s) {
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

// FHTRotator draws (2 * bernoulli - 1) / sqrt(P) per sign (utils/rotator.hpp:37-43);
// here the bits come from a seeded splitmix64 stream.
std::vector<float> draw_signs(uint32_t p, uint64_t seed) {
    std::vector<float> sg(p);
    uint64_t s = seed;
    uint64_t bits = 0;
    for (uint32_t i = 0; i < p; ++i) {
        if (i % 64 == 0) bits = splitmix64(s);
        const int b = static_cast<int>((bits >> (i % 64)) & 1u);
        sg[i] = static_cast<float>(2 * b - 1) / sqrtf(static_cast<float>(p));
    }
    return sg;
}

// Nibble of neighbour v for codebook c in the reference's FastScan block
// layout (pack_codes / pack_codes_helper, fastscan_impl.hpp:67-124): blocks
// of 32 neighbours, 4P bytes each; inside a block, 32 bytes per pair of
// codebooks, the first 16 for the even codebook and the next 16 for the odd
// one; byte j of those 16 holds neighbour perm[j] in its low nibble and
// perm[j] + 16 in its high nibble, perm = 0, 8, 1, 9, ..., 7, 15.
uint32_t block_nibble(const uint8_t* codes, uint32_t p, uint32_t v, uint32_t c) {
    const uint32_t w = v & 31u, u = w & 15u;
    const uint32_t j = u < 8 ? 2 * u : 2 * (u - 8) + 1;
    const uint8_t b = codes[static_cast<size_t>(v >> 5) * 4 * p + (c >> 1) * 32 + (c & 1u) * 16 + j];
    return w < 16 ? (b & 15u) : (b >> 4);
}

// FNV-1a over 32-bit words of the entry point, neighbour ids and vectors:
// ties a sidecar file to the index it was built from.
uint64_t fingerprint(const lqg_index* h) {
    uint64_t x = 0xCBF29CE484222325ull;
    auto mix = [&x](uint32_t w) {
        x ^= w;
        x *= 0x100000001B3ull;
    };
    mix(h->ep);
    for (uint32_t v : h->nbr) mix(v);
    for (float f : h->raw) {
        uint32_t w;
        memcpy(&w, &f, 4);
        mix(w);
    }
    return x;
}

struct SideHeader {
    char magic[8];
    uint32_t version, n, dim, deg, p, reserved;
    uint64_t fingerprint;
};
constexpr char kMagic[8] = {'L', 'Q', 'G', 'S', 'I', 'D', 'E', '1'};

lqg_status need_data(lqg_index* h) {
    if (!h->have_data)
        return fail(LQG_ERR_STATE, "index has no data: call load() or build_index() first");
    return ensure_device(h);
}

}  // namespace

H2Cfg h2_default_launch(int ef_cap) {
    H2Cfg c;
    c.ef_cap = ef_cap;
    c.waves_per_group = 4;
    c.groups = 0;
    // LDS visited table: at least 4x the largest ef of the class, so that
    // expansions (typically 1x-3x ef) stay under half of it (the limit is
    // three quarters, vis_limit)
    c.hash_bits = ef_cap <= 64 ? 9 : ef_cap <= 256 ? 10 : ef_cap <= 512 ? 11 : 12;
    // 6 waves per SIMD up to the 192 class: with these defaults on the 1M x 128
    // workload, 4-12% faster than uncapped at ef 50-150 (at ef 200 a 5-wave cap
    // was 3% slower, so larger classes stay uncapped); a launch drops the cap
    // where the index has no capped kernel (more than 4 code words per lane)
    c.occupancy = ef_cap <= 192 ? 6 : 0;
    // launches of up to kLatencyMaxQueries queries: latency mode, the records
    // of the two best unexpanded entries loaded ahead (lone searches, one
    // session: 1M x 128 at ef 128 0.196-0.202 -> 0.191 ms; 768-d and 960-d
    // rows 5-7% faster; four or eight records were no faster)
    c.prefetch = 1;
    return c;
}

H2Cfg h2_effective_launch(const lqg_index* h, int ef_cap) {
    const int slot = class_slot(ef_cap);
    H2Cfg cfg = h->h2_launch_set[slot] ? h->h2_launch[slot] : h2_default_launch(ef_cap);
    cfg.ef_cap = ef_cap;
    if (!h->h2_launch_set[slot] && cfg.occupancy != 0 &&
        !(h->h2_ready && h2_occupancy_ok(h->h2dev())))
        cfg.occupancy = 0;
    return cfg;
}

void h2_reset(lqg_index* h) {
    ++h->generation;
    h->d_h2rec.release();
    h->h2_ready = false;
    h->mode = LQG_MODE_SQ8;
}

namespace {

// Launch geometry of a FastScan batch of nq queries (ef up to ef_max, k up to
// kmax): the class's settings with waves per group and groups fitted.
lqg_status h2_batch_cfg(const lqg_index* h, uint64_t nq, uint32_t ef_max, uint32_t kmax,
                        H2Cfg* out) {
    const int cls = ef_class(ef_max);
    if (cls == 0) return fail(LQG_ERR_UNSUPPORTED, "ef above 1024 is not supported on the device");
    if (!h2_kmax_ok(kmax))
        return fail(LQG_ERR_UNSUPPORTED, "FastScan: k above 128 is not supported on the device");
    H2Cfg cfg = h2_effective_launch(h, cls);
    const H2Dev ix = h->h2dev();
    if (!h2_variant_ok(ix, cfg.occupancy, kmax)) {
        if (h->h2_launch_set[class_slot(cls)])
            return fail(LQG_ERR_UNSUPPORTED,
                        "FastScan: no kernel with this register cap for this index shape and k "
                        "(the caps need at most 4 code words per lane; 8 waves needs k <= 64)");
        cfg.occupancy = 0;
    }
    if (nq < static_cast<uint64_t>(cfg.waves_per_group)) cfg.waves_per_group = 1;
    while (cfg.waves_per_group > 1 && h2_lds_bytes(ix, cfg, kmax) > 160 * 1024)
        cfg.waves_per_group /= 2;
    if (h2_lds_bytes(ix, cfg, kmax) > 160 * 1024)
        return fail(LQG_ERR_UNSUPPORTED, "FastScan: kmax too large for the LDS of one workgroup");
    const int groups = h2_groups(ix, cfg, kmax, static_cast<uint32_t>(nq));
    if (groups <= 0) return fail(LQG_ERR_DEVICE, "FastScan: no launch geometry");
    cfg.groups = groups;
    *out = cfg;
    return LQG_OK;
}

// Waves of the overflow kernel after a batch of nq queries.
uint32_t overflow_waves(uint64_t nq) {
    return static_cast<uint32_t>(std::min<uint64_t>(nq, kOverflowWaves));
}

// Device scratch of such a batch: the overflow list (nq + 2 words) and the
// overflow kernel's per-wave visited tables in device memory (2^h2_gvis_bits
// entries each), for queries whose set outgrows three quarters of their LDS
// table (none
// when h2_gvis_bits is 0 or the scratch has no_overflow set: such queries
// fail); the failure flag and the work-queue head; the GEMM-rotated queries.
lqg_status h2_reserve(const lqg_index* h, const H2Cfg& cfg, uint64_t nq, Scratch* sc) {
    (void)cfg;
    const uint32_t gbits = sc->no_overflow ? 0u : h->h2_gvis_bits;
    if (gbits) {
        LQG_HIP_TRY(sc->gvis.reserve(static_cast<size_t>(overflow_waves(nq)) << gbits));
        LQG_HIP_TRY(sc->ovf.reserve(nq + 2));
    }
    LQG_HIP_TRY(sc->err.reserve(1));
    LQG_HIP_TRY(sc->counter.reserve(1));
    if (h->h2_rotation == LQG_ROTATION_GEMM) LQG_HIP_TRY(sc->rot.reserve(nq * h->h2_p));
    return LQG_OK;
}

}  // namespace

lqg_status h2_prepare_scratch(lqg_index* h, uint64_t nq, uint32_t ef_max, uint32_t kmax,
                              Scratch* sc) {
    if (!h->h2_ready) return LQG_OK;
    H2Cfg cfg;
    lqg_status st = h2_batch_cfg(h, nq, ef_max, kmax, &cfg);
    if (st != LQG_OK) return st;
    return h2_reserve(h, cfg, nq, sc);
}

lqg_status run_h2_batch(lqg_index* h, const float* d_queries, uint64_t nq, const uint32_t* d_ef,
                        const uint32_t* d_k, uint32_t ef_max, uint32_t kmax, uint32_t* d_results,
                        float* d_dists, uint32_t* d_stats, hipStream_t s, Scratch* sc) {
    if (!h->h2_ready) return fail(LQG_ERR_STATE, "FastScan mode needs a sidecar");
    H2Cfg cfg;
    lqg_status st = h2_batch_cfg(h, nq, ef_max, kmax, &cfg);
    if (st != LQG_OK) return st;
    st = h2_reserve(h, cfg, nq, sc);
    if (st != LQG_OK) return st;
    const H2Dev ix = h->h2dev();
    const uint32_t gbits = sc->no_overflow ? 0u : h->h2_gvis_bits;
    if (h->h2_rotation == LQG_ROTATION_GEMM) {
        LQG_HIP_TRY(h2_launch_rotate_queries(ix, d_queries, static_cast<uint32_t>(nq), 1,
                                             sc->rot.p, s));
    }
    H2Job job;
    job.queries = d_queries;
    job.ef = d_ef;
    job.k = d_k;
    job.nq = static_cast<uint32_t>(nq);
    job.ef_all = ef_max;
    job.k_all = kmax;
    job.kstride = kmax;
    job.out_ids = d_results;
    job.out_d = d_dists;
    job.out_stats = d_stats;
    job.counter = sc->counter.p;
    job.gvis = gbits ? sc->gvis.p : nullptr;
    job.gvis_bits = gbits;
    job.ovf = gbits ? sc->ovf.p : nullptr;
    job.err = sc->err.p;
    job.done = sc->done;
    job.rot = h->h2_rotation == LQG_ROTATION_GEMM ? sc->rot.p : nullptr;
    job.prefetch_targets = cfg.prefetch > 0 ? 1u << cfg.prefetch : 0u;
    LQG_HIP_TRY(launch_reset_counter(sc->counter.p, s));
    LQG_HIP_TRY(launch_reset_counter(sc->err.p, s));
    if (job.ovf) LQG_HIP_TRY(launch_reset_counter(job.ovf, s, 2));
    LQG_HIP_TRY(h2_launch_search(ix, job, cfg, s));
    if (job.ovf) LQG_HIP_TRY(h2_launch_overflow(ix, job, cfg, overflow_waves(nq), s));
    return LQG_OK;
}

// Graph build with FastScan candidates.  The reference's builder finds each
// node's candidates with the FastScan search from the entry point, the node's
// own vector as the query, and takes every node that search expands
// (find_candidates, qg.hpp:439-473; qg_builder.hpp:202-231).  Here the
// records of the graph being built carry codes and factors computed for the
// current graph (SymphonyQG computes them as neighbour lists change; LiteQG's
// rows carry none), and the general FastScan kernel writes each query's
// expanded nodes in expansion order.
//
// h2_build_begin: records with the raw rows and the initial graph's ids,
// rotation signs from `seed`, and every raw row rotated once into rot (n x P;
// the same rows the search's query preparation would compute).
lqg_status h2_build_begin(lqg_index* h, uint64_t seed, DevBuf<float>& rot) {
    lqg_status st = set_layout(h);
    if (st != LQG_OK) return st;
    h->h2_signs = draw_signs(h->h2_p, seed);
    st = alloc_records(h);
    if (st != LQG_OK) return st;
    LQG_HIP_TRY(rot.reserve(h->n * static_cast<size_t>(h->h2_p)));
    LQG_HIP_TRY(h2_launch_rotate_rows(h->h2dev(), rot.p, h->stream));
    return LQG_OK;
}

// Every query searched at ef with the general kernel, the first exp_cap
// nodes each expands written to exp_ids (nq x exp_cap, 0xFFFFFFFF past the
// last).  rot: the queries rotated beforehand (nq x P), or null.
lqg_status h2_run_expansions(lqg_index* h, const float* d_queries, uint64_t nq, const float* rot,
                             uint32_t ef, uint32_t* exp_ids, uint32_t exp_cap,
                             DevBuf<uint32_t>& out_k1, hipStream_t s) {
    const bool saved = h->h2_generic;
    h->h2_generic = true;
    H2Cfg cfg;
    lqg_status st = h2_batch_cfg(h, nq, ef, 1, &cfg);
    const H2Dev ix = h->h2dev();
    h->h2_generic = saved;
    if (st != LQG_OK) return st;
    Scratch* sc = &h->scratch;
    st = h2_reserve(h, cfg, nq, sc);
    if (st != LQG_OK) return st;
    LQG_HIP_TRY(out_k1.reserve(nq));
    LQG_HIP_TRY(hipMemsetAsync(exp_ids, 0xFF, nq * exp_cap * 4ull, s));
    const uint32_t gbits = sc->no_overflow ? 0u : h->h2_gvis_bits;
    H2Job job;
    job.queries = d_queries;
    job.ef = nullptr;
    job.k = nullptr;
    job.nq = static_cast<uint32_t>(nq);
    job.ef_all = ef;
    job.k_all = 1;
    job.kstride = 1;
    job.out_ids = out_k1.p;
    job.out_d = nullptr;
    job.out_stats = nullptr;
    job.counter = sc->counter.p;
    job.gvis = gbits ? sc->gvis.p : nullptr;
    job.gvis_bits = gbits;
    job.ovf = gbits ? sc->ovf.p : nullptr;
    job.err = sc->err.p;
    job.rot = rot;
    job.exp_ids = exp_ids;
    job.exp_cap = exp_cap;
    LQG_HIP_TRY(launch_reset_counter(sc->counter.p, s));
    LQG_HIP_TRY(launch_reset_counter(sc->err.p, s));
    if (job.ovf) LQG_HIP_TRY(launch_reset_counter(job.ovf, s, 2));
    LQG_HIP_TRY(h2_launch_search(ix, job, cfg, s));
    if (job.ovf) LQG_HIP_TRY(h2_launch_overflow(ix, job, cfg, overflow_waves(nq), s));
    return LQG_OK;
}

// One iteration's candidates: the records take the current device graph's
// rows, their codes and factors are recomputed, then every node is searched
// at ef from the entry point, its own rotated row as the query, and its
// first exp_cap expanded nodes written to exp_ids.  d_raw: the n x dim raw
// rows on the device (the queries).
lqg_status h2_build_candidates(lqg_index* h, const float* d_raw, const float* rot, uint32_t ef,
                               uint32_t* exp_ids, uint32_t exp_cap, DevBuf<uint32_t>& out_k1,
                               hipStream_t s) {
    {
        const H2Dev ix = h->h2dev();
        LQG_HIP_TRY(h2_launch_rows(ix, h->h2_off_ids, h->d_nbr.p, 4u * ix.deg, true, s));
        LQG_HIP_TRY(h2_launch_codes(ix, rot, s));
    }
    h->h2_ready = true;
    return h2_run_expansions(h, d_raw, h->n, rot, ef, exp_ids, exp_cap, out_k1, s);
}

void h2_build_end(lqg_index* h) { h2_reset(h); }

lqg_status h2_check_error(lqg_index* h, Scratch* sc) {
    if (!sc) sc = &h->scratch;
    if (h->mode != LQG_MODE_FASTSCAN || !sc->err.p) return LQG_OK;
    uint32_t e = 0;
    LQG_HIP_TRY(hipMemcpy(&e, sc->err.p, 4, hipMemcpyDeviceToHost));
    if (e)
        return fail(LQG_ERR_UNSUPPORTED,
                    "FastScan: a query's visited set overflowed (its ids are 0xFFFFFFFF; the "
                    "other queries' results are complete)");
    return LQG_OK;
}

}  // namespace lqg

using lqg::fail;

extern "C" {

lqg_status lqg_set_mode(lqg_index* h, lqg_mode mode) {
    if (!h) return fail(LQG_ERR_ARG, "null index");
    if (lqg::abi_enum(mode) == LQG_MODE_SQ8) {
        h->mode = LQG_MODE_SQ8;
        return LQG_OK;
    }
    if (lqg::abi_enum(mode) != LQG_MODE_FASTSCAN) return fail(LQG_ERR_ARG, "unknown mode");
    if (!h->h2_ready)
        return fail(LQG_ERR_STATE, "FastScan mode needs a sidecar: build, set or load one first");
    h->mode = mode;
    return LQG_OK;
}

lqg_status lqg_get_mode(const lqg_index* h, lqg_mode* mode) {
    if (!h || !mode) return fail(LQG_ERR_ARG, "null argument");
    *mode = h->mode;
    return LQG_OK;
}

lqg_status lqg_h2_build_sidecar(lqg_index* h, const float* signs, uint64_t seed) {
    if (!h) return fail(LQG_ERR_ARG, "null index");
    lqg_status st = lqg::need_data(h);
    if (st != LQG_OK) return st;
    st = lqg::set_layout(h);
    if (st != LQG_OK) return st;
    const uint32_t p = h->h2_p;
    std::vector<float> sg(p);
    if (signs) {
        for (uint32_t i = 0; i < p; ++i) {
            if (!(fabsf(signs[i]) > 0.0f) || !std::isfinite(signs[i]))
                return fail(LQG_ERR_ARG, "signs must be finite and non-zero");
            sg[i] = signs[i];
        }
    } else {
        sg = lqg::draw_signs(p, seed);
    }
    h->h2_signs = sg;
    st = lqg::alloc_records(h);
    if (st != LQG_OK) return st;
    lqg::DevBuf<float> rot;
    LQG_HIP_TRY(rot.reserve(h->n * static_cast<size_t>(p)));
    hipError_t e = lqg::h2_launch_sidecar(h->h2dev(), rot.p, h->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
    rot.release();
    if (e != hipSuccess) return lqg::hip_fail(e, "sidecar build");
    h->h2_ready = true;
    return LQG_OK;
}

lqg_status lqg_h2_set_sidecar(lqg_index* h, const float* signs, const uint8_t* codes, int layout,
                              const float* factors) {
    if (!h || !signs || !codes || !factors) return fail(LQG_ERR_ARG, "null argument");
    if (layout != 0 && layout != 1) return fail(LQG_ERR_ARG, "layout must be 0 or 1");
    lqg_status st = lqg::need_data(h);
    if (st != LQG_OK) return st;
    st = lqg::set_layout(h);
    if (st != LQG_OK) return st;
    const uint32_t p = h->h2_p, deg = static_cast<uint32_t>(h->deg);
    const size_t cb = static_cast<size_t>(deg) * (p / 8);
    h->h2_signs.assign(signs, signs + p);
    st = lqg::alloc_records(h);
    if (st != LQG_OK) return st;
    if (layout == 1) {
        st = lqg::host_rows(h, h->h2_off_codes, const_cast<uint8_t*>(codes), cb, true);
    } else {
        // block layout -> vector-major
        std::vector<uint8_t> vm;
        try {
            vm.assign(h->n * cb, 0);
        } catch (...) {
            return fail(LQG_ERR_NOMEM, "host allocation failed");
        }
        for (uint64_t u = 0; u < h->n; ++u) {
            const uint8_t* src = codes + u * cb;
            uint8_t* dst = vm.data() + u * cb;
            for (uint32_t v = 0; v < deg; ++v)
                for (uint32_t c = 0; c < p / 4; ++c) {
                    const uint32_t nib = lqg::block_nibble(src, p, v, c);
                    dst[static_cast<size_t>(v) * (p / 8) + c / 2] |=
                        static_cast<uint8_t>((c & 1u) ? nib << 4 : nib);
                }
        }
        st = lqg::host_rows(h, h->h2_off_codes, vm.data(), cb, true);
    }
    if (st != LQG_OK) return st;
    st = lqg::host_rows(h, h->h2_off_fac, const_cast<float*>(factors), 12ull * deg, true);
    if (st != LQG_OK) return st;
    h->h2_ready = true;
    return LQG_OK;
}

lqg_status lqg_h2_get_sidecar(const lqg_index* h, float* signs, uint8_t* codes, float* factors) {
    if (!h) return fail(LQG_ERR_ARG, "null index");
    if (!h->h2_ready) return fail(LQG_ERR_STATE, "no sidecar");
    const uint32_t p = h->h2_p, deg = static_cast<uint32_t>(h->deg);
    const size_t cb = static_cast<size_t>(deg) * (p / 8);
    if (signs) memcpy(signs, h->h2_signs.data(), 4ull * p);
    lqg_index* hm = const_cast<lqg_index*>(h);  // the copies use the handle's stream
    if (codes) {
        lqg_status st = lqg::host_rows(hm, h->h2_off_codes, codes, cb, false);
        if (st != LQG_OK) return st;
    }
    if (factors) {
        lqg_status st = lqg::host_rows(hm, h->h2_off_fac, factors, 12ull * deg, false);
        if (st != LQG_OK) return st;
    }
    return LQG_OK;
}

lqg_status lqg_h2_save_sidecar(const lqg_index* h, const char* path) {
    if (!h || !path) return fail(LQG_ERR_ARG, "null argument");
    if (!h->h2_ready) return fail(LQG_ERR_STATE, "no sidecar");
    const uint32_t p = h->h2_p, deg = static_cast<uint32_t>(h->deg);
    const size_t cb = static_cast<size_t>(deg) * (p / 8), fb = 12ull * deg;
    FILE* f = fopen(path, "wb");
    if (!f) return fail(LQG_ERR_IO, std::string("cannot create ") + path);
    lqg::SideHeader hd;
    memcpy(hd.magic, lqg::kMagic, 8);
    hd.version = 1;
    hd.n = static_cast<uint32_t>(h->n);
    hd.dim = static_cast<uint32_t>(h->dim);
    hd.deg = deg;
    hd.p = p;
    hd.reserved = 0;
    hd.fingerprint = lqg::fingerprint(h);
    bool ok = fwrite(&hd, sizeof hd, 1, f) == 1;
    ok = ok && fwrite(h->h2_signs.data(), 4, p, f) == p;
    // codes of every node, then factors of every node
    for (int part = 0; ok && part < 2; ++part) {
        const size_t w = part == 0 ? cb : fb;
        std::vector<uint8_t> buf;
        try {
            buf.resize(h->n * w);
        } catch (...) {
            fclose(f);
            return fail(LQG_ERR_NOMEM, "host allocation failed");
        }
        lqg_status st = lqg::host_rows(const_cast<lqg_index*>(h),
                                       part == 0 ? h->h2_off_codes : h->h2_off_fac, buf.data(), w,
                                       false);
        if (st != LQG_OK) {
            fclose(f);
            return st;
        }
        ok = fwrite(buf.data(), 1, buf.size(), f) == buf.size();
    }
    ok = (fclose(f) == 0) && ok;
    return ok ? LQG_OK : fail(LQG_ERR_IO, std::string("write failed: ") + path);
}

lqg_status lqg_h2_load_sidecar(lqg_index* h, const char* path) {
    if (!h || !path) return fail(LQG_ERR_ARG, "null argument");
    lqg_status st = lqg::need_data(h);
    if (st != LQG_OK) return st;
    st = lqg::set_layout(h);
    if (st != LQG_OK) return st;
    const uint32_t p = h->h2_p, deg = static_cast<uint32_t>(h->deg);
    const size_t cb = static_cast<size_t>(deg) * (p / 8), fb = 12ull * deg;
    FILE* f = fopen(path, "rb");
    if (!f) return fail(LQG_ERR_IO, std::string("cannot open ") + path);
    lqg::SideHeader hd;
    bool ok = fread(&hd, sizeof hd, 1, f) == 1;
    if (!ok || memcmp(hd.magic, lqg::kMagic, 8) != 0 || hd.version != 1) {
        fclose(f);
        return fail(LQG_ERR_IO, std::string(path) + " is not a sidecar file");
    }
    if (hd.n != h->n || hd.dim != h->dim || hd.deg != deg || hd.p != p ||
        hd.fingerprint != lqg::fingerprint(h)) {
        fclose(f);
        return fail(LQG_ERR_IO, std::string(path) + " was built for another index");
    }
    std::vector<float> sg(p);
    ok = fread(sg.data(), 4, p, f) == p;
    h->h2_signs = sg;
    if (ok) {
        st = lqg::alloc_records(h);
        if (st != LQG_OK) {
            fclose(f);
            return st;
        }
    }
    for (int part = 0; ok && part < 2; ++part) {
        const size_t w = part == 0 ? cb : fb;
        std::vector<uint8_t> buf;
        try {
            buf.resize(h->n * w);
        } catch (...) {
            fclose(f);
            return fail(LQG_ERR_NOMEM, "host allocation failed");
        }
        ok = fread(buf.data(), 1, buf.size(), f) == buf.size();
        if (ok) {
            st = lqg::host_rows(h, part == 0 ? h->h2_off_codes : h->h2_off_fac, buf.data(), w,
                                true);
            if (st != LQG_OK) {
                fclose(f);
                return st;
            }
        }
    }
    fclose(f);
    if (!ok) return fail(LQG_ERR_IO, std::string("short read from ") + path);
    h->h2_ready = true;
    return LQG_OK;
}

lqg_status lqg_h2_expansions(lqg_index* h, const float* queries, uint64_t nq, uint32_t ef,
                             uint32_t cap, uint32_t* out) {
    if (!h || !queries || !out) return fail(LQG_ERR_ARG, "null argument");
    if (nq == 0 || nq > 0x7FFFFFFFull) return fail(LQG_ERR_ARG, "nq must be in [1, 2^31)");
    if (cap == 0 || cap > 65536) return fail(LQG_ERR_ARG, "cap must be in [1, 65536]");
    if (ef == 0 || lqg::ef_class(ef) == 0) return fail(LQG_ERR_UNSUPPORTED, "ef above 1024");
    if (!h->h2_ready) return fail(LQG_ERR_STATE, "no sidecar: build, set or load one first");
    for (uint64_t i = 0; i < nq * h->dim; ++i)
        if (!std::isfinite(queries[i])) return fail(LQG_ERR_ARG, "query contains a non-finite value");
    lqg::DevBuf<float> dq;
    lqg::DevBuf<uint32_t> dout, k1;
    LQG_HIP_TRY(dq.reserve(nq * h->dim));
    LQG_HIP_TRY(dout.reserve(nq * cap));
    LQG_HIP_TRY(hipMemcpy(dq.p, queries, nq * h->dim * 4, hipMemcpyHostToDevice));
    lqg_status st = lqg::h2_run_expansions(h, dq.p, nq, nullptr, ef, dout.p, cap, k1, h->stream);
    if (st != LQG_OK) return st;
    LQG_HIP_TRY(hipStreamSynchronize(h->stream));
    st = lqg::h2_check_error(h);
    if (st != LQG_OK) return st;
    LQG_HIP_TRY(hipMemcpy(out, dout.p, nq * cap * 4ull, hipMemcpyDeviceToHost));
    return LQG_OK;
}

lqg_status lqg_h2_set_overflow_bits(lqg_index* h, uint32_t bits) {
    if (!h) return fail(LQG_ERR_ARG, "null index");
    if (bits != 0 && (bits < 8 || bits > 16))
        return fail(LQG_ERR_ARG, "overflow table bits must be 0 or in [8, 16]");
    h->h2_gvis_bits = bits;
    return LQG_OK;
}

lqg_status lqg_h2_set_generic(lqg_index* h, int generic) {
    if (!h) return fail(LQG_ERR_ARG, "null index");
    if (generic != 0 && generic != 1) return fail(LQG_ERR_ARG, "generic must be 0 or 1");
    h->h2_generic = generic != 0;
    return LQG_OK;
}

lqg_status lqg_h2_set_rotation(lqg_index* h, lqg_rotation method) {
    if (!h) return fail(LQG_ERR_ARG, "null index");
    if (lqg::abi_enum(method) != LQG_ROTATION_FHT && lqg::abi_enum(method) != LQG_ROTATION_GEMM)
        return fail(LQG_ERR_ARG, "rotation must be LQG_ROTATION_FHT or LQG_ROTATION_GEMM");
    h->h2_rotation = method;
    return LQG_OK;
}

lqg_status lqg_h2_get_rotation(const lqg_index* h, lqg_rotation* method) {
    if (!h || !method) return fail(LQG_ERR_ARG, "null argument");
    *method = h->h2_rotation;
    return LQG_OK;
}

lqg_status lqg_h2_rotate(lqg_index* h, const float* queries, uint64_t nq, lqg_rotation method,
                         float* out) {
    if (!h || !queries || !out) return fail(LQG_ERR_ARG, "null argument");
    if (lqg::abi_enum(method) != LQG_ROTATION_FHT && lqg::abi_enum(method) != LQG_ROTATION_GEMM)
        return fail(LQG_ERR_ARG, "rotation must be LQG_ROTATION_FHT or LQG_ROTATION_GEMM");
    if (nq == 0 || nq > 0xFFFFFFFFull) return fail(LQG_ERR_ARG, "nq must be in [1, 2^32)");
    if (!h->h2_ready) return fail(LQG_ERR_STATE, "no sidecar (rotation signs)");
    for (uint64_t i = 0; i < nq * h->dim; ++i)
        if (!std::isfinite(queries[i])) return fail(LQG_ERR_ARG, "query contains a non-finite value");
    lqg::DevBuf<float> dq, dout;
    LQG_HIP_TRY(dq.reserve(nq * h->dim));
    LQG_HIP_TRY(dout.reserve(nq * h->h2_p));
    LQG_HIP_TRY(hipMemcpy(dq.p, queries, nq * h->dim * 4, hipMemcpyHostToDevice));
    LQG_HIP_TRY(lqg::h2_launch_rotate_queries(h->h2dev(), dq.p, static_cast<uint32_t>(nq),
                                              method == LQG_ROTATION_GEMM ? 1 : 0, dout.p,
                                              h->stream));
    LQG_HIP_TRY(hipStreamSynchronize(h->stream));
    LQG_HIP_TRY(hipMemcpy(out, dout.p, nq * h->h2_p * 4, hipMemcpyDeviceToHost));
    return LQG_OK;
}

lqg_status lqg_h2_probe(lqg_index* h, const float* queries, uint64_t nq, const uint32_t* nodes,
                        uint8_t* lut, float* lo_width, int32_t* sumq, float* sqr_y, float* est) {
    if (!h || !queries || !lut || !lo_width || !sumq) return fail(LQG_ERR_ARG, "null argument");
    if (nodes && (!sqr_y || !est)) return fail(LQG_ERR_ARG, "nodes given without sqr_y / est");
    if (nq == 0 || nq > 65535) return fail(LQG_ERR_ARG, "nq must be in [1, 65535]");
    if (!h->h2_ready) return fail(LQG_ERR_STATE, "no sidecar");
    if (nodes)
        for (uint64_t i = 0; i < nq; ++i)
            if (nodes[i] >= h->n) return fail(LQG_ERR_ARG, "node id out of range");
    const uint32_t p = h->h2_p, deg = static_cast<uint32_t>(h->deg);
    lqg::DevBuf<float> dq, dpar, dsq, dest;
    lqg::DevBuf<uint32_t> dn;
    lqg::DevBuf<uint8_t> dlut;
    LQG_HIP_TRY(dq.reserve(nq * h->dim));
    LQG_HIP_TRY(dpar.reserve(3 * nq));
    LQG_HIP_TRY(dlut.reserve(nq * 4ull * p));
    LQG_HIP_TRY(hipMemcpy(dq.p, queries, nq * h->dim * 4, hipMemcpyHostToDevice));
    if (nodes) {
        LQG_HIP_TRY(dn.reserve(nq));
        LQG_HIP_TRY(dsq.reserve(nq));
        LQG_HIP_TRY(dest.reserve(nq * deg));
        LQG_HIP_TRY(hipMemcpy(dn.p, nodes, nq * 4, hipMemcpyHostToDevice));
    }
    LQG_HIP_TRY(lqg::h2_launch_probe(h->h2dev(), dq.p, static_cast<uint32_t>(nq),
                                     nodes ? dn.p : nullptr, dlut.p, dpar.p, dsq.p, dest.p,
                                     h->stream));
    LQG_HIP_TRY(hipStreamSynchronize(h->stream));
    std::vector<float> par(3 * nq);
    LQG_HIP_TRY(hipMemcpy(par.data(), dpar.p, 12 * nq, hipMemcpyDeviceToHost));
    LQG_HIP_TRY(hipMemcpy(lut, dlut.p, nq * 4ull * p, hipMemcpyDeviceToHost));
    for (uint64_t i = 0; i < nq; ++i) {
        lo_width[2 * i] = par[3 * i];
        lo_width[2 * i + 1] = par[3 * i + 1];
        memcpy(&sumq[i], &par[3 * i + 2], 4);
    }
    if (nodes) {
        LQG_HIP_TRY(hipMemcpy(sqr_y, dsq.p, 4 * nq, hipMemcpyDeviceToHost));
        LQG_HIP_TRY(hipMemcpy(est, dest.p, 4 * nq * deg, hipMemcpyDeviceToHost));
    }
    return LQG_OK;
}

}  // extern "C"
