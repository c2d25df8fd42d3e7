// builder.cpp -- host orchestration of Index.build_index on the device.
//
// QGBuilder (qg/qg_builder.hpp:48-91) steps, with where each runs:
//   centroid + exact nearest node as entry point   host  (space.hpp:16-78)
//   SQ8 codes of the data                          host  (quantization/sq8.hpp:11-26)
//   random initial graph                           host  (qg_builder.hpp:263-291)
//   per iteration: beam search of every node       device, h1_search.hip
//                  prune / reverse / refine        device, build_graph.hip
#include <math.h>
#include <string.h>

#include <thread>
#include <vector>

#include "lqg_build.h"
#include "lqg_index.h"

namespace lqg {
namespace {

constexpr uint32_t kMaxCandidatePool = 750;  // qg_builder.hpp:30
constexpr uint32_t kCandBuffer = 1024;       // candidates the prune kernels hold per node
constexpr uint32_t kMaxPruned = 300;         // qg_builder.hpp:31
constexpr uint64_t kSeed = 0x51A7E5EEDull;

uint64_t mix64(uint64_t x) {
    x += 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
}

template <typename F>
void parallel_rows(uint64_t n, F&& fn) {
    unsigned t = std::thread::hardware_concurrency();
    t = t == 0 ? 1 : (t > 32 ? 32 : t);
    std::vector<std::thread> pool;
    const uint64_t step = (n + t - 1) / t;
    for (unsigned w = 0; w < t; ++w) {
        const uint64_t s = w * step, e = std::min(n, s + step);
        if (s >= e) break;
        pool.emplace_back([=, &fn] { fn(s, e); });
    }
    for (auto& th : pool) th.join();
}

// l2_sqr in the reference's order (space/l2.hpp:16-34)
float l2_ref(const float* a, const float* b, uint64_t dim) {
    float s[16] = {0};
    const uint64_t mul16 = dim & ~15ull;
    uint64_t i = 0;
    for (; i < mul16; i += 16)
        for (int j = 0; j < 16; ++j) {
            const float t = a[i + j] - b[i + j];
            s[j] = fmaf(t, t, s[j]);
        }
    for (int j = 0; j < 8; ++j) s[j] += s[j + 8];
    for (int j = 0; j < 4; ++j) s[j] += s[j + 4];
    float r = (s[0] + s[2]) + (s[1] + s[3]);
    for (; i < dim; ++i) {
        const float t = a[i] - b[i];
        r = fmaf(t, t, r);
    }
    return r;
}

// sq8_quantize (quantization/sq8.hpp:11-26) as the reference's -Ofast build
// evaluates it: scale = (max - min) * fl(1/255), code = round((v - min) / scale).
void sq8_quantize(const float* x, uint64_t n, uint64_t d, uint8_t* codes, float* params) {
    std::vector<float> mn(d), mx(d);
    for (uint64_t j = 0; j < d; ++j) mn[j] = mx[j] = x[j];
    for (uint64_t i = 1; i < n; ++i)
        for (uint64_t j = 0; j < d; ++j) {
            const float v = x[i * d + j];
            mn[j] = std::min(mn[j], v);
            mx[j] = std::max(mx[j], v);
        }
    const float inv255 = 1.0f / 255.0f;
    std::vector<float> sc(d);
    for (uint64_t j = 0; j < d; ++j) {
        sc[j] = (mx[j] - mn[j]) * inv255;
        params[2 * j] = mn[j];
        params[2 * j + 1] = sc[j];
    }
    parallel_rows(n, [&](uint64_t s, uint64_t e) {
        for (uint64_t i = s; i < e; ++i)
            for (uint64_t j = 0; j < d; ++j) {
                const float r = (x[i * d + j] - mn[j]) / sc[j];
                codes[i * d + j] = std::isfinite(r) ? static_cast<uint8_t>(std::round(r)) : 0;
            }
    });
}

// compute_centroid (space.hpp:50-78, FP64 sums) then exact_nn (space.hpp:16-48)
uint32_t entry_point(const float* x, uint64_t n, uint64_t d) {
    std::vector<double> c(d, 0.0);
    for (uint64_t i = 0; i < n; ++i)
        for (uint64_t j = 0; j < d; ++j) c[j] += x[i * d + j];
    std::vector<float> cf(d);
    const double inv = 1.0 / static_cast<double>(n);
    for (uint64_t j = 0; j < d; ++j) cf[j] = static_cast<float>(c[j] * inv);
    unsigned t = std::thread::hardware_concurrency();
    t = t == 0 ? 1 : (t > 32 ? 32 : t);
    std::vector<float> best_d(t, INFINITY);
    std::vector<uint32_t> best_i(t, 0);
    std::vector<std::thread> pool;
    const uint64_t step = (n + t - 1) / t;
    for (unsigned w = 0; w < t; ++w) {
        pool.emplace_back([&, w] {
            const uint64_t s = w * step, e = std::min(n, s + step);
            for (uint64_t i = s; i < e; ++i) {
                const float dd = l2_ref(x + i * d, cf.data(), d);
                if (dd < best_d[w]) {
                    best_d[w] = dd;
                    best_i[w] = static_cast<uint32_t>(i);
                }
            }
        });
    }
    for (auto& th : pool) th.join();
    uint32_t bi = 0;
    float bd = INFINITY;
    for (unsigned w = 0; w < t; ++w)
        if (best_d[w] < bd) {
            bd = best_d[w];
            bi = best_i[w];
        }
    return bi;
}

// random_init (qg_builder.hpp:263-291): deg distinct ids != i per node
void random_graph(uint64_t n, uint64_t deg, uint32_t* nbr) {
    parallel_rows(n, [&](uint64_t s, uint64_t e) {
        for (uint64_t i = s; i < e; ++i) {
            uint32_t* row = nbr + i * deg;
            uint64_t ctr = 0, k = 0;
            while (k < deg) {
                const uint32_t r = static_cast<uint32_t>(mix64(kSeed ^ (i << 24) ^ ctr++) % n);
                bool dup = r == i;
                for (uint64_t j = 0; j < k && !dup; ++j) dup = row[j] == r;
                if (!dup) row[k++] = r;
            }
        }
    });
}

}  // namespace

lqg_status build_index(lqg_index* h, const float* data, uint32_t ef_build, uint32_t num_iter) {
    const uint64_t n = h->n, d = h->dim, g = h->deg;
    try {
        h->raw.assign(data, data + n * d);
        h->qdata.resize(n * d);
        h->sq8.resize(2 * d);
        h->nbr.resize(n * g);
    } catch (...) {
        return fail(LQG_ERR_NOMEM, "host allocation failed");
    }
    h->ep = entry_point(h->raw.data(), n, d);
    sq8_quantize(h->raw.data(), n, d, h->qdata.data(), h->sq8.data());
    random_graph(n, g, h->nbr.data());
    h->have_data = true;
    h->on_device = false;
    ++h->generation;
    lqg_status st = ensure_device(h);
    if (st != LQG_OK) return st;

    // Candidates of a node: the SQ8 beam's nearest min(EF, 750), or, with
    // FastScan candidates, every node the FastScan search expands (up to the
    // prune buffer's room beside the current neighbours), as find_candidates
    // collects them (qg.hpp:439-473).
    const bool fs_cand = h->build_candidates == LQG_BUILD_FASTSCAN;
    BuildParams p;
    p.n = static_cast<uint32_t>(n);
    p.dim = static_cast<uint32_t>(d);
    p.deg = static_cast<uint32_t>(g);
    p.ncand = fs_cand ? kCandBuffer - static_cast<uint32_t>(g) : std::min(ef_build, kMaxCandidatePool);
    p.pr_cap = kMaxPruned;
    p.seed = kSeed;

    DevBuf<float> raw, kept_d, pr_d, rot;
    DevBuf<uint32_t> out_k1;
    DevBuf<uint32_t> cand, kept_ids, kept_n, pr_ids, pr_n, in_cnt, in_off, in_ids, rev_k, rev_ko,
        rev_v, ctr;
    DevBuf<unsigned char> tmp;
    if (n * g > static_cast<size_t>(INT32_MAX))
        return fail(LQG_ERR_UNSUPPORTED, "build: N x degree above 2^31 edge slots");
    const size_t tmp_bytes = build_temp_bytes(p.n, p.deg);
    if (tmp_bytes == 0) return fail(LQG_ERR_DEVICE, "builder scratch size query failed");
    LQG_HIP_TRY(raw.reserve(n * d));
    LQG_HIP_TRY(cand.reserve(n * p.ncand));
    LQG_HIP_TRY(kept_ids.reserve(n * g));
    LQG_HIP_TRY(kept_d.reserve(n * g));
    LQG_HIP_TRY(kept_n.reserve(n));
    LQG_HIP_TRY(pr_ids.reserve(n * p.pr_cap));
    LQG_HIP_TRY(pr_d.reserve(n * p.pr_cap));
    LQG_HIP_TRY(pr_n.reserve(n));
    LQG_HIP_TRY(in_cnt.reserve(n + 1));
    LQG_HIP_TRY(in_off.reserve(n + 1));
    LQG_HIP_TRY(in_ids.reserve(n * g));
    LQG_HIP_TRY(rev_k.reserve(n * g));
    LQG_HIP_TRY(rev_ko.reserve(n * g));
    LQG_HIP_TRY(rev_v.reserve(n * g));
    LQG_HIP_TRY(ctr.reserve(1));
    LQG_HIP_TRY(tmp.reserve(tmp_bytes));
    LQG_HIP_TRY(hipMemcpy(raw.p, h->raw.data(), n * d * 4, hipMemcpyHostToDevice));

    BuildBuffers b;
    b.raw = raw.p;
    b.qdata = h->d_qdata.p;
    b.sq8_min = h->d_min.p;
    b.sq8_scale = h->d_scale.p;
    b.graph = h->d_nbr.p;
    b.cand_ids = cand.p;
    b.kept_ids = kept_ids.p;
    b.kept_d = kept_d.p;
    b.kept_n = kept_n.p;
    b.pr_ids = pr_ids.p;
    b.pr_d = pr_d.p;
    b.pr_n = pr_n.p;
    b.in_cnt = in_cnt.p;
    b.in_off = in_off.p;
    b.in_ids = in_ids.p;
    b.rev_keys = rev_k.p;
    b.rev_keys_out = rev_ko.p;
    b.rev_vals = rev_v.p;
    b.counter = ctr.p;
    b.scan_tmp = tmp.p;
    b.scan_tmp_bytes = tmp_bytes;

    lqg_status rs = LQG_OK;
    if (fs_cand) rs = h2_build_begin(h, kSeed, rot);
    for (uint32_t it = 0; it < num_iter && rs == LQG_OK; ++it) {
        p.refine = it + 1 == num_iter;
        p.seed = kSeed + it;
        if (fs_cand)
            rs = h2_build_candidates(h, raw.p, rot.p, ef_build, cand.p, p.ncand, out_k1,
                                     h->stream);
        else
            rs = run_batch(h, raw.p, n, nullptr, nullptr, ef_build, p.ncand, LQG_ORDER_NEAREST,
                           cand.p, nullptr, nullptr, h->stream);
        if (rs == LQG_OK) {
            hipError_t e = build_run_iteration(b, p, h->stream);
            if (e != hipSuccess) rs = hip_fail(e, "build_run_iteration");
        }
    }
    hipError_t e = hipStreamSynchronize(h->stream);
    if (rs == LQG_OK && e != hipSuccess) rs = hip_fail(e, "build synchronize");
    if (fs_cand) h2_build_end(h);
    rot.release();
    out_k1.release();
    if (rs == LQG_OK) {
        e = hipMemcpy(h->nbr.data(), h->d_nbr.p, n * g * 4, hipMemcpyDeviceToHost);
        if (e != hipSuccess) rs = hip_fail(e, "graph download");
    }
    raw.release();
    cand.release();
    kept_ids.release();
    kept_d.release();
    kept_n.release();
    pr_ids.release();
    pr_d.release();
    pr_n.release();
    in_cnt.release();
    in_off.release();
    in_ids.release();
    rev_k.release();
    rev_ko.release();
    rev_v.release();
    ctr.release();
    tmp.release();
    if (rs != LQG_OK) {
        h->have_data = false;
        h->on_device = false;
        ++h->generation;
    }
    return rs;
}

}  // namespace lqg
