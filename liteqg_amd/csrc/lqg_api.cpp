// lqg_api.cpp -- host side of the C ABI declared in include/liteqg.h.
//
// Owns the index (host copy for save(), device copy for search), validates
// every call the way the reference's binding should have (python/bindings.cpp),
// and launches the gfx950 kernels through lqg_internal.h.
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <string>
#include <thread>
#include <vector>

#include "lqg_index.h"

namespace lqg {

namespace {
thread_local std::string g_err;
}

lqg_status fail(lqg_status s, const std::string& msg) {
    g_err = msg;
    return s;
}

lqg_status hip_fail(hipError_t e, const char* what) {
    return fail(LQG_ERR_DEVICE, std::string(what) + ": " + hipGetErrorString(e));
}

const char* last_error() { return g_err.c_str(); }

int ef_class(uint32_t ef) {
    for (int slot = 0; slot < kClasses; ++slot)
        if (ef <= static_cast<uint32_t>(class_cap(slot))) return class_cap(slot);
    return 0;
}

int class_cap(int slot) {
    static const int caps[kClasses] = {64, 128, 192, 256, 512, 1024};
    return caps[slot];
}

lqg_status ensure_device(lqg_index* h) {
    if (h->on_device) return LQG_OK;
    if (!h->have_data)
        return fail(LQG_ERR_STATE, "index has no data: call load() or build_index() first");
    int ndev = 0;
    hipError_t e = hipGetDeviceCount(&ndev);
    if (e != hipSuccess || ndev == 0) return fail(LQG_ERR_DEVICE, "no HIP device available");
    int dev = 0;
    LQG_HIP_TRY(hipGetDevice(&dev));
    hipDeviceProp_t prop;
    LQG_HIP_TRY(hipGetDeviceProperties(&prop, dev));
    if (strncmp(prop.gcnArchName, "gfx950", 6) != 0)
        return fail(LQG_ERR_DEVICE, std::string("device is ") + prop.gcnArchName +
                                        ", this library is built for gfx950 only");
    if (!h->stream) LQG_HIP_TRY(hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking));
    lqg_status sst = static_scratch_setup();
    if (sst != LQG_OK) return sst;
    if (!h->ev0) LQG_HIP_TRY(hipEventCreate(&h->ev0));
    if (!h->ev1) LQG_HIP_TRY(hipEventCreate(&h->ev1));
    const size_t nd = h->n * h->dim, ng = h->n * h->deg;
    LQG_HIP_TRY(h->d_qdata.reserve(nd));
    LQG_HIP_TRY(h->d_nbr.reserve(ng));
    LQG_HIP_TRY(h->d_min.reserve(h->dim));
    LQG_HIP_TRY(h->d_scale.reserve(h->dim));
    std::vector<float> mn(h->dim), sc(h->dim);
    h->sq8_identity = true;
    for (uint64_t i = 0; i < h->dim; ++i) {
        mn[i] = h->sq8[2 * i];
        sc[i] = h->sq8[2 * i + 1];
        h->sq8_identity = h->sq8_identity && mn[i] == 0.0f && sc[i] == 1.0f;
    }
    LQG_HIP_TRY(hipMemcpy(h->d_qdata.p, h->qdata.data(), nd, hipMemcpyHostToDevice));
    // Device copy of the neighbour table: a repeated id later in a row, or the
    // row's own node, is replaced by kInvalid.  The search skips both exactly as
    // the reference does (the node is already visited when its row is scanned),
    // so the kernel needs no per-expansion duplicate test.
    std::vector<uint32_t> dev_nbr(h->nbr);
    {
        const uint64_t g = h->deg;
        const unsigned nt = std::max(1u, std::min(32u, std::thread::hardware_concurrency()));
        std::vector<std::thread> pool;
        for (unsigned t = 0; t < nt; ++t)
            pool.emplace_back([&, t] {
                for (uint64_t i = t; i < h->n; i += nt) {
                    uint32_t* row = dev_nbr.data() + i * g;
                    for (uint64_t a = 0; a < g; ++a) {
                        if (row[a] == i) {
                            row[a] = kInvalid;
                            continue;
                        }
                        for (uint64_t b = 0; b < a; ++b)
                            if (row[b] == row[a]) {
                                row[a] = kInvalid;
                                break;
                            }
                    }
                }
            });
        for (auto& th : pool) th.join();
    }
    LQG_HIP_TRY(hipMemcpy(h->d_nbr.p, dev_nbr.data(), ng * 4, hipMemcpyHostToDevice));
    LQG_HIP_TRY(hipMemcpy(h->d_min.p, mn.data(), h->dim * 4, hipMemcpyHostToDevice));
    LQG_HIP_TRY(hipMemcpy(h->d_scale.p, sc.data(), h->dim * 4, hipMemcpyHostToDevice));
    LQG_HIP_TRY(h->scratch.counter.reserve(1));
    h->on_device = true;
    ++h->generation;
    return LQG_OK;
}

int class_slot(int ef_cap) {
    for (int slot = 0; slot < kClasses - 1; ++slot)
        if (ef_cap <= class_cap(slot)) return slot;
    return kClasses - 1;
}

LaunchCfg default_launch(int ef_cap, uint64_t dim) {
    LaunchCfg c;
    c.ef_cap = ef_cap;
    c.waves_per_group = 4;
    c.occupancy = 0;
    // the visited cache must hold more of a query's touched nodes as ef
    // grows: on 1,000 960-d queries at ef U{64..1024} (pool class 1024), 2^11
    // entries evaluated 24.3 rows per expansion and 2^12 16.1 (18.8 -> 14.9
    // ms per launch)
    c.cache_bits = ef_cap >= 512 ? 12 : 10;
    c.groups = 0;
    // launches of up to sq8_latency_max(dim) queries (a lone search(); 1,024
    // queries on rows of at most 256 dims, else 256) run in latency mode, loading the next expansions' id rows and SQ8 rows ahead:
    // on the 1M x 128 workload at ef 50 a lone search() went from p50 0.115-
    // 0.120 ms to 0.100 ms, a 256-query batch from 0.216 to 0.186 ms
    c.prefetch = 2;
    // SQ8 float path at D >= 512: two waves per query share each pass's
    // distance evaluation (h1_split_kernel).  200k-node graphs, 1,000 queries,
    // ef U{64..1024} in one launch: 960-d 26.2 -> 21.3 ms, 768-d 12.1 -> 9.6 ms
    // (four waves: 23.1 and 10.6 ms); a lone 960-d query at ef 64 1.19 ms
    // (latency mode) -> 0.92 ms.  At 96-d the batch was slower (5.7 -> 6.1 ms).
    c.split = dim >= kSplitMinDim ? 2 : 1;
    return c;
}

lqg_status run_batch(lqg_index* h, const float* d_queries, uint64_t nq, const uint32_t* d_ef,
                     const uint32_t* d_k, uint32_t ef_max, uint32_t kmax, lqg_order order,
                     uint32_t* d_results, float* d_dists, uint32_t* d_stats, hipStream_t s,
                     bool count_visits, Scratch* sc, int queries_int) {
    if (!sc) sc = &h->scratch;
    if (h->mode == LQG_MODE_FASTSCAN) {
        if (count_visits) return fail(LQG_ERR_UNSUPPORTED, "visit counting is for the SQ8 path");
        return run_h2_batch(h, d_queries, nq, d_ef, d_k, ef_max, kmax, d_results, d_dists, d_stats,
                            s, sc);
    }
    if (nq > 0x7FFFFFFFull) return fail(LQG_ERR_ARG, "nq must be below 2^31");
    // queue words: [0, kQueueWords) integer launch, [kQueueWords, 2 kQueueWords)
    // float launch, then the count of queries the integer launch left
    const size_t nwords = 2 * kQueueWords + 1;
    LQG_HIP_TRY(sc->counter.reserve(nwords));
    if (ef_class(ef_max) == 0)
        return fail(LQG_ERR_UNSUPPORTED, "ef above 1024 is not supported on the device");
    SearchJob job;
    job.queries = d_queries;
    job.ef = d_ef;
    job.k = d_k;
    job.nq = static_cast<uint32_t>(nq);
    job.ef_all = ef_max;
    job.k_all = kmax;
    job.kstride = kmax;
    job.order = order == LQG_ORDER_NEAREST ? 1 : 0;
    job.out_ids = d_results;
    job.out_d = d_dists;
    job.out_stats = d_stats;
    job.counter = sc->counter.p;
    job.visit_bits = nullptr;
    job.visit_words = 0;
    job.visit_waves = 0;
    job.done = sc->done;
    const int cls = ef_class(ef_max);
    LaunchCfg cfg = h->launch_set[class_slot(cls)] ? h->launch[class_slot(cls)]
                                                   : default_launch(cls, h->dim);
    cfg.ef_cap = cls;
    if (nq < static_cast<uint64_t>(cfg.waves_per_group)) cfg.waves_per_group = 1;
    // the default split applies to launches too small to fill the device
    // with one wave per query; larger ones (a graph build's beam search of
    // every node: 300k x 960, 8.3 s unsplit against 8.6 s split) run the
    // batch kernel
    if (!h->launch_set[class_slot(cls)] && nq > kSplitMaxQueries) {
        cfg.split = 1;
        // and the default visited cache of the high-ef classes is for launches
        // of a few queries: a graph build's beam search of 1M nodes (EF 400,
        // class 512) took 1.00 s per iteration with 2^12 entries, 0.32 s
        // with 2^10 (LDS per wave sets its occupancy)
        cfg.cache_bits = std::min(cfg.cache_bits, 10);
    }
    // An index whose SQ8 decode is exact (min 0, scale 1; D % 16 == 0, D <=
    // 256) runs the integer-path kernel first; a float-path launch then takes
    // only the queries that were not integer valued.  Each kernel variant
    // holds only its own distance code's registers.
    const bool int_index =
        h->sq8_integer && h->sq8_identity && h->dim % 16 == 0 && h->dim <= 256;
    // rows of at most 128 bytes take the integer variant with half the row buffer
    const int int_eval = h->dim <= 128 ? 3 : 1;
    if (count_visits) {
        // one N-bit bitmap per wave, for the larger grid of the two kernel
        // variants (their occupancy, hence their grid, differs)
        LaunchCfg ci = cfg, cf = cfg;
        ci.eval = int_eval;
        cf.eval = 0;
        int groups = search_groups(h->dev(), job, cf);
        if (int_index) groups = std::max(groups, search_groups(h->dev(), job, ci));
        job.visit_words = static_cast<uint32_t>((h->n + 31) / 32);
        job.visit_waves = static_cast<uint32_t>(groups) * static_cast<uint32_t>(cfg.waves_per_group);
        const size_t words = static_cast<size_t>(job.visit_waves) * job.visit_words;
        LQG_HIP_TRY(h->d_visit.reserve(words));
        job.visit_bits = h->d_visit.p;
    }
    // a single latency-mode or split launch (one query per workgroup) uses no queue
    const bool one_launch = !count_visits && (!int_index || queries_int == 0 || queries_int == 1);
    const bool float_only = !int_index || queries_int == 0;
    if (!(one_launch && ((cfg.prefetch > 0 && nq <= sq8_latency_max(h->dim)) ||
                         (float_only && cfg.split > 1))))
        LQG_HIP_TRY(launch_reset_counter(sc->counter.p, s, static_cast<uint32_t>(nwords)));
    uint32_t* left = sc->counter.p + 2 * kQueueWords;
    // queries_int: the caller has checked on the host that every query is
    // (1) or that none is (0) integer valued in [0, 255]; then only the
    // launch that takes them runs.  -1: unknown, both run.
    if (int_index && queries_int == 1 && !count_visits) {
        cfg.eval = int_eval;
        job.counter = sc->counter.p;
        job.skip_int = 0;
        job.left = nullptr;
        LQG_HIP_TRY(launch_search(h->dev(), job, cfg, s));
        return LQG_OK;
    }
    if (int_index && queries_int == 0 && !count_visits) {
        cfg.eval = 0;
        job.skip_int = 0;
        job.left = nullptr;
        job.counter = sc->counter.p + kQueueWords;
        LQG_HIP_TRY(launch_search(h->dev(), job, cfg, s));
        return LQG_OK;
    }
    if (int_index) {
        cfg.eval = int_eval;
        job.counter = sc->counter.p;
        job.skip_int = 0;
        job.left = left;
        LQG_HIP_TRY(launch_search(h->dev(), job, cfg, s));
    }
    cfg.eval = 0;
    job.skip_int = int_index ? 1 : 0;
    job.left = int_index ? left : nullptr;
    job.counter = sc->counter.p + kQueueWords;
    LQG_HIP_TRY(launch_search(h->dev(), job, cfg, s));
    return LQG_OK;
}

}  // namespace lqg

namespace {

using lqg::ef_class;
using lqg::fail;

lqg_status check_queries(const float* q, uint64_t count) {
    for (uint64_t i = 0; i < count; ++i)
        if (!std::isfinite(q[i])) return fail(LQG_ERR_ARG, "query contains a non-finite value");
    return LQG_OK;
}

bool valid_order(const lqg_order& o) {
    const int v = lqg::abi_enum(o);
    return v == LQG_ORDER_REFERENCE || v == LQG_ORDER_NEAREST;
}

}  // namespace

#define HIP_TRY LQG_HIP_TRY

extern "C" {

const char* lqg_last_error(void) { return lqg::last_error(); }

lqg_status lqg_create(const char* index_type, const char* metric, uint64_t num_elements,
                      uint64_t dimension, uint64_t degree_bound, lqg_index** out) {
    if (!out) return fail(LQG_ERR_ARG, "out is null");
    *out = nullptr;
    if (!metric || std::string(metric) != "L2")
        return fail(LQG_ERR_ARG, "Only L2 distance supported currently");
    if (degree_bound < 32 || degree_bound % 32 != 0)
        return fail(LQG_ERR_ARG, "The degree bound must be a multiple of 32");
    if (!index_type || std::string(index_type) != "QG")
        return fail(LQG_ERR_ARG, std::string("Index type [") + (index_type ? index_type : "") +
                                     "] not supported");
    if (num_elements == 0 || num_elements >= 0x7FFFFFFFull)
        return fail(LQG_ERR_ARG, "num_elements must be in [1, 2^31-1)");
    if (dimension == 0 || dimension > 4096)
        return fail(LQG_ERR_ARG, "dimension must be in [1, 4096]");
    lqg_index* h = new (std::nothrow) lqg_index();
    if (!h) return fail(LQG_ERR_NOMEM, "allocation failed");
    h->n = num_elements;
    h->dim = dimension;
    h->deg = degree_bound;
    *out = h;
    return LQG_OK;
}

lqg_status lqg_destroy(lqg_index* h) {
    if (!h) return LQG_OK;
    if (h->stream) (void)hipStreamSynchronize(h->stream);
    if (h->p_q) (void)hipHostFree(h->p_q);
    if (h->p_out) (void)hipHostFree(h->p_out);
    h->d_qdata.release();
    h->d_nbr.release();
    h->d_min.release();
    h->d_scale.release();
    h->d_q.release();
    h->d_ids.release();
    h->d_ef.release();
    h->d_k.release();
    h->d_stats.release();
    h->scratch.release();
    h->d_dist.release();
    h->d_visit.release();
    h->d_h2rec.release();
    h->d_h2signs.release();
    if (h->ev0) (void)hipEventDestroy(h->ev0);
    if (h->ev1) (void)hipEventDestroy(h->ev1);
    if (h->stream) (void)hipStreamDestroy(h->stream);
    delete h;
    return LQG_OK;
}

lqg_status lqg_load(lqg_index* h, const char* filename) {
    if (!h || !filename) return fail(LQG_ERR_ARG, "null argument");
    FILE* f = fopen(filename, "rb");
    if (!f) return fail(LQG_ERR_IO, std::string("cannot open ") + filename);
    const uint64_t n = h->n, d = h->dim, g = h->deg;
    const uint64_t expect = 4 + n * (d + g) * 4 + n * d + 8 * d;
    fseek(f, 0, SEEK_END);
    const long long got = ftell(f);
    fseek(f, 0, SEEK_SET);
    if (got < 0 || static_cast<uint64_t>(got) != expect) {
        fclose(f);
        char msg[256];
        snprintf(msg, sizeof msg, "%s is %lld bytes, expected %llu for N=%llu D=%llu deg=%llu",
                 filename, got, static_cast<unsigned long long>(expect),
                 static_cast<unsigned long long>(n), static_cast<unsigned long long>(d),
                 static_cast<unsigned long long>(g));
        return fail(LQG_ERR_IO, msg);
    }
    uint32_t ep = 0;
    std::vector<float> raw, sq8;
    std::vector<uint32_t> nbr;
    std::vector<uint8_t> qdata;
    try {
        raw.resize(n * d);
        nbr.resize(n * g);
        qdata.resize(n * d);
        sq8.resize(2 * d);
    } catch (...) {
        fclose(f);
        return fail(LQG_ERR_NOMEM, "host allocation failed");
    }
    bool ok = fread(&ep, 4, 1, f) == 1;
    // rows: D floats then deg ids (qg.hpp:420-424), read in blocks
    const uint64_t row = d + g;
    const uint64_t block = std::max<uint64_t>(1, (64ull << 20) / (row * 4));
    std::vector<uint32_t> buf;
    try {
        buf.resize(std::min(n, block) * row);
    } catch (...) {
        fclose(f);
        return fail(LQG_ERR_NOMEM, "host allocation failed");
    }
    for (uint64_t s = 0; ok && s < n; s += block) {
        const uint64_t e = std::min(n, s + block);
        ok = fread(buf.data(), 4, (e - s) * row, f) == (e - s) * row;
        for (uint64_t i = s; ok && i < e; ++i) {
            memcpy(&raw[i * d], &buf[(i - s) * row], d * 4);
            memcpy(&nbr[i * g], &buf[(i - s) * row + d], g * 4);
        }
    }
    ok = ok && fread(qdata.data(), 1, n * d, f) == n * d;
    ok = ok && fread(sq8.data(), 4, 2 * d, f) == 2 * d;
    fclose(f);
    if (!ok) return fail(LQG_ERR_IO, std::string("short read from ") + filename);
    if (ep >= n) return fail(LQG_ERR_IO, "entry point is out of range: file does not match N");
    for (uint64_t i = 0; i < n * g; ++i)
        if (nbr[i] >= n) return fail(LQG_ERR_IO, "neighbour id out of range: file does not match N/D/deg");
    h->ep = ep;
    h->raw.swap(raw);
    h->nbr.swap(nbr);
    h->qdata.swap(qdata);
    h->sq8.swap(sq8);
    h->have_data = true;
    h->on_device = false;
    ++h->generation;
    lqg::h2_reset(h);
    return LQG_OK;
}

lqg_status lqg_save(lqg_index* h, const char* filename) {
    if (!h || !filename) return fail(LQG_ERR_ARG, "null argument");
    if (!h->have_data) return fail(LQG_ERR_STATE, "index has no data to save");
    FILE* f = fopen(filename, "wb");
    if (!f) return fail(LQG_ERR_IO, std::string("cannot create ") + filename);
    const uint64_t n = h->n, d = h->dim, g = h->deg, row = d + g;
    bool ok = fwrite(&h->ep, 4, 1, f) == 1;
    std::vector<uint32_t> r(row);
    for (uint64_t i = 0; ok && i < n; ++i) {
        memcpy(r.data(), &h->raw[i * d], d * 4);
        memcpy(r.data() + d, &h->nbr[i * g], g * 4);
        ok = fwrite(r.data(), 4, row, f) == row;
    }
    ok = ok && fwrite(h->qdata.data(), 1, n * d, f) == n * d;
    ok = ok && fwrite(h->sq8.data(), 4, 2 * d, f) == 2 * d;
    ok = (fclose(f) == 0) && ok;
    return ok ? LQG_OK : fail(LQG_ERR_IO, std::string("write failed: ") + filename);
}

lqg_status lqg_set_ef(lqg_index* h, uint32_t ef) {
    if (!h) return fail(LQG_ERR_ARG, "null index");
    if (ef == 0) return fail(LQG_ERR_ARG, "EF must be positive");
    if (ef_class(ef) == 0) return fail(LQG_ERR_UNSUPPORTED, "EF above 1024 is not supported");
    h->ef = ef;
    h->ef_set = true;
    return LQG_OK;
}

lqg_status lqg_set_order(lqg_index* h, lqg_order order) {
    if (!h) return fail(LQG_ERR_ARG, "null index");
    if (!valid_order(order)) return fail(LQG_ERR_ARG, "unknown order");
    h->order = order;
    return LQG_OK;
}

lqg_status lqg_set_sq8_integer(lqg_index* h, int enable) {
    if (!h) return fail(LQG_ERR_ARG, "null index");
    h->sq8_integer = enable != 0;
    return LQG_OK;
}

lqg_status lqg_search(lqg_index* h, const float* query, uint32_t k, uint32_t* results) {
    if (!h || !query || !results) return fail(LQG_ERR_ARG, "null argument");
    if (!h->ef_set) return fail(LQG_ERR_STATE, "set_ef() must be called before search()");
    if (k == 0) return fail(LQG_ERR_ARG, "k must be positive");
    lqg_status st = check_queries(query, h->dim);
    if (st != LQG_OK) return st;
    st = lqg::ensure_device(h);
    if (st != LQG_OK) return st;
    // The kernel reads the query from pinned host memory and writes the ids and
    // a completion flag back there: no copies through pageable memory.  The
    // stream is still synchronised before returning (what is left of it then
    // is at most the float-path launch finding nothing to do), so no later
    // call can free memory a launch of this one still reads.
    if (!h->p_q) {
        HIP_TRY(hipHostMalloc(reinterpret_cast<void**>(&h->p_q), 4 * h->dim, hipHostMallocCoherent));
        HIP_TRY(hipHostGetDevicePointer(reinterpret_cast<void**>(&h->g_pq), h->p_q, 0));
    }
    if (h->p_out_k < k) {
        if (h->p_out) {
            HIP_TRY(hipStreamSynchronize(h->stream));
            HIP_TRY(hipHostFree(h->p_out));
            h->p_out = nullptr;
            h->p_out_k = 0;
        }
        HIP_TRY(hipHostMalloc(reinterpret_cast<void**>(&h->p_out), 4ull * (16 + k),
                              hipHostMallocCoherent));
        HIP_TRY(hipHostGetDevicePointer(reinterpret_cast<void**>(&h->g_pout), h->p_out, 0));
        h->p_out_k = k;
    }
    memcpy(h->p_q, query, 4 * h->dim);
    __atomic_store_n(&h->p_out[0], 0u, __ATOMIC_RELEASE);
    // the integer-path test of the kernels, on the host: one launch instead of two
    int qint = 1;
    for (uint64_t i = 0; i < h->dim; ++i) {
        const float v = query[i];
        if (!(v >= 0.f && v <= 255.f && v == std::rint(v))) {
            qint = 0;
            break;
        }
    }
    h->scratch.done = h->g_pout;
    st = lqg::run_batch(h, h->g_pq, 1, nullptr, nullptr, h->ef, k, h->order, h->g_pout + 16,
                        nullptr, nullptr, h->stream, false, nullptr, qint);
    h->scratch.done = nullptr;
    if (st != LQG_OK) {
        (void)hipStreamSynchronize(h->stream);
        return st;
    }
    // wait for the flag; a stream that finishes without it had a query fail
    // (a FastScan visited set that overflowed) or a launch error
    for (uint32_t spin = 1;; ++spin) {
        const uint32_t flag = __atomic_load_n(&h->p_out[0], __ATOMIC_ACQUIRE);
        if (flag != 0u) {
            HIP_TRY(hipStreamSynchronize(h->stream));
            if (flag != lqg::kFlagDone)
                return fail(LQG_ERR_UNSUPPORTED, "FastScan: the query's visited set overflowed");
            memcpy(results, h->p_out + 16, 4ull * k);
            return LQG_OK;
        }
        if ((spin & 255u) == 0u) {
            const hipError_t q = hipStreamQuery(h->stream);
            if (q == hipErrorNotReady) continue;
            if (__atomic_load_n(&h->p_out[0], __ATOMIC_ACQUIRE) != 0u) continue;
            HIP_TRY(q);
            HIP_TRY(hipStreamSynchronize(h->stream));
            st = lqg::h2_check_error(h);
            return st != LQG_OK ? st : fail(LQG_ERR_DEVICE, "search finished without a result");
        }
    }
}

lqg_status lqg_search_batch(lqg_index* h, const float* queries, uint64_t nq, const uint32_t* ef,
                            const uint32_t* k, uint32_t kmax, lqg_order order, uint32_t* results,
                            float* dists, lqg_stats* stats) {
    if (!h || !queries || !results) return fail(LQG_ERR_ARG, "null argument");
    if (nq == 0 || nq > 0x7FFFFFFFull) return fail(LQG_ERR_ARG, "nq must be in [1, 2^31)");
    if (kmax == 0) return fail(LQG_ERR_ARG, "kmax must be positive");
    if (!valid_order(order)) return fail(LQG_ERR_ARG, "unknown order");
    if (!ef && !h->ef_set) return fail(LQG_ERR_STATE, "set_ef() or per-query ef required");
    uint32_t ef_max = ef ? 0 : h->ef;
    if (ef) {
        for (uint64_t i = 0; i < nq; ++i) {
            if (ef[i] == 0) return fail(LQG_ERR_ARG, "per-query ef must be positive");
            ef_max = std::max(ef_max, ef[i]);
        }
    }
    if (k) {
        for (uint64_t i = 0; i < nq; ++i)
            if (k[i] == 0 || k[i] > kmax) return fail(LQG_ERR_ARG, "per-query k must be in [1, kmax]");
    }
    if (ef_class(ef_max) == 0) return fail(LQG_ERR_UNSUPPORTED, "ef above 1024 is not supported");
    lqg_status st = check_queries(queries, nq * h->dim);
    if (st != LQG_OK) return st;
    st = lqg::ensure_device(h);
    if (st != LQG_OK) return st;
    HIP_TRY(h->d_q.reserve(nq * h->dim));
    HIP_TRY(h->d_ids.reserve(nq * kmax));
    if (dists) HIP_TRY(h->d_dist.reserve(nq * kmax));
    if (ef) HIP_TRY(h->d_ef.reserve(nq));
    if (k) HIP_TRY(h->d_k.reserve(nq));
    if (stats) HIP_TRY(h->d_stats.reserve(2 * nq));
    HIP_TRY(hipMemcpyAsync(h->d_q.p, queries, nq * h->dim * 4, hipMemcpyHostToDevice, h->stream));
    if (ef) HIP_TRY(hipMemcpyAsync(h->d_ef.p, ef, nq * 4, hipMemcpyHostToDevice, h->stream));
    if (k) HIP_TRY(hipMemcpyAsync(h->d_k.p, k, nq * 4, hipMemcpyHostToDevice, h->stream));
    HIP_TRY(hipEventRecord(h->ev0, h->stream));
    st = lqg::run_batch(h, h->d_q.p, nq, ef ? h->d_ef.p : nullptr, k ? h->d_k.p : nullptr, ef_max, kmax,
                   order, h->d_ids.p, dists ? h->d_dist.p : nullptr,
                   stats ? h->d_stats.p : nullptr, h->stream);
    if (st != LQG_OK) return st;
    HIP_TRY(hipEventRecord(h->ev1, h->stream));
    HIP_TRY(hipMemcpyAsync(results, h->d_ids.p, nq * kmax * 4, hipMemcpyDeviceToHost, h->stream));
    if (dists)
        HIP_TRY(hipMemcpyAsync(dists, h->d_dist.p, nq * kmax * 4, hipMemcpyDeviceToHost, h->stream));
    std::vector<uint32_t> st_host;
    if (stats) {
        st_host.resize(2 * nq);
        HIP_TRY(hipMemcpyAsync(st_host.data(), h->d_stats.p, 2 * nq * 4, hipMemcpyDeviceToHost,
                               h->stream));
    }
    HIP_TRY(hipStreamSynchronize(h->stream));
    st = lqg::h2_check_error(h);
    if (st != LQG_OK) return st;
    if (stats) {
        stats->queries = nq;
        stats->expansions = 0;
        stats->evaluations = 0;
        for (uint64_t i = 0; i < nq; ++i) {
            stats->expansions += st_host[2 * i];
            stats->evaluations += st_host[2 * i + 1];
        }
        float ms = 0.f;
        HIP_TRY(hipEventElapsedTime(&ms, h->ev0, h->ev1));
        stats->device_ms = ms;
    }
    return LQG_OK;
}

lqg_status lqg_search_batch_device(lqg_index* h, const float* d_queries, uint64_t nq,
                                   const uint32_t* d_ef, const uint32_t* d_k, uint32_t ef_max,
                                   uint32_t kmax, lqg_order order, uint32_t* d_results,
                                   float* d_dists, uint32_t* d_stats, void* stream) {
    if (!h || !d_queries || !d_results) return fail(LQG_ERR_ARG, "null argument");
    if (nq == 0 || nq > 0x7FFFFFFFull) return fail(LQG_ERR_ARG, "nq must be in [1, 2^31)");
    if (kmax == 0) return fail(LQG_ERR_ARG, "kmax must be positive");
    if (!valid_order(order)) return fail(LQG_ERR_ARG, "unknown order");
    if (ef_max == 0) {
        if (d_ef) return fail(LQG_ERR_ARG, "ef_max must bound the per-query ef array");
        if (!h->ef_set) return fail(LQG_ERR_STATE, "set_ef() or ef_max required");
        ef_max = h->ef;
    }
    if (ef_class(ef_max) == 0) return fail(LQG_ERR_UNSUPPORTED, "ef above 1024 is not supported");
    lqg_status st = lqg::ensure_device(h);
    if (st != LQG_OK) return st;
    hipStream_t s = stream ? static_cast<hipStream_t>(stream) : h->stream;
    return lqg::run_batch(h, d_queries, nq, d_ef, d_k, ef_max, kmax, order, d_results, d_dists, d_stats,
                     s);
}

lqg_status lqg_set_launch(lqg_index* h, uint32_t ef, const lqg_launch* cfg) {
    if (!h) return fail(LQG_ERR_ARG, "null index");
    if (ef != 0 && ef_class(ef) == 0) return fail(LQG_ERR_UNSUPPORTED, "ef above 1024");
    if (cfg && (cfg->waves_per_group < 1 || cfg->waves_per_group > 4 || cfg->cache_bits < 4 ||
                cfg->cache_bits > 14 || cfg->groups < 0 || cfg->prefetch < 0 || cfg->prefetch > 4 ||
                cfg->split < 0 || cfg->split == 3 || cfg->split > 4))
        return fail(LQG_ERR_ARG, "launch settings out of range");
    const bool fs = h->mode == LQG_MODE_FASTSCAN;
    if (fs && cfg && cfg->prefetch > 3)
        return fail(LQG_ERR_ARG, "FastScan prefetch depth must be in [0, 3]");
    if (fs && cfg && cfg->cache_bits < 6)
        return fail(LQG_ERR_ARG, "FastScan visited-table bits must be in [6, 14]");
    if (cfg && !fs && cfg->occupancy != -1 && cfg->occupancy != 0 && cfg->occupancy != 6 &&
        cfg->occupancy != 8)
        return fail(LQG_ERR_ARG, "SQ8 occupancy must be -1, 0, 6 or 8");
    if (cfg && fs && cfg->occupancy != 0 && cfg->occupancy != 6 && cfg->occupancy != 8)
        return fail(LQG_ERR_ARG, "FastScan occupancy must be 0, 6 or 8");
    if (cfg && fs && cfg->occupancy != 0 && h->h2_ready && !lqg::h2_occupancy_ok(h->h2dev()))
        return fail(LQG_ERR_UNSUPPORTED,
                    "FastScan occupancy caps need at most 4 code words per lane (P / lanes per "
                    "neighbour <= 128 bits)");
    for (int slot = 0; slot < lqg::kClasses; ++slot) {
        const int cap = lqg::class_cap(slot);
        if (ef != 0 && ef_class(ef) != cap) continue;
        if (fs) {
            if (!cfg) {
                h->h2_launch_set[slot] = false;
                continue;
            }
            lqg::H2Cfg c = lqg::h2_default_launch(cap);
            c.waves_per_group = cfg->waves_per_group;
            c.hash_bits = cfg->cache_bits;
            c.occupancy = cfg->occupancy;
            c.groups = cfg->groups;
            c.prefetch = cfg->prefetch;
            h->h2_launch[slot] = c;
            h->h2_launch_set[slot] = true;
            continue;
        }
        if (!cfg) {
            h->launch_set[slot] = false;
            continue;
        }
        lqg::LaunchCfg c;
        c.ef_cap = cap;
        c.waves_per_group = cfg->waves_per_group;
        c.cache_bits = cfg->cache_bits;
        c.occupancy = cfg->occupancy;
        c.groups = cfg->groups;
        c.prefetch = cfg->prefetch;
        c.split = cfg->split > 1 ? cfg->split : 1;
        h->launch[slot] = c;
        h->launch_set[slot] = true;
    }
    return LQG_OK;
}

lqg_status lqg_get_launch(const lqg_index* h, uint32_t ef, lqg_launch* cfg) {
    if (!h || !cfg) return fail(LQG_ERR_ARG, "null argument");
    const int cls = ef_class(ef == 0 ? 1 : ef);
    if (cls == 0) return fail(LQG_ERR_UNSUPPORTED, "ef above 1024");
    const int slot = lqg::class_slot(cls);
    if (h->mode == LQG_MODE_FASTSCAN) {
        const lqg::H2Cfg c = lqg::h2_effective_launch(h, cls);
        cfg->waves_per_group = c.waves_per_group;
        cfg->cache_bits = c.hash_bits;
        cfg->occupancy = c.occupancy;
        cfg->groups = c.groups;
        cfg->prefetch = c.prefetch;
        cfg->split = 1;
        return LQG_OK;
    }
    const lqg::LaunchCfg c =
        h->launch_set[slot] ? h->launch[slot] : lqg::default_launch(cls, h->dim);
    cfg->waves_per_group = c.waves_per_group;
    cfg->cache_bits = c.cache_bits;
    cfg->occupancy = c.occupancy;
    cfg->groups = c.groups;
    cfg->prefetch = c.prefetch;
    cfg->split = c.split;
    return LQG_OK;
}

lqg_status lqg_autotune(lqg_index* h, const float* queries, uint64_t nq, uint32_t ef, uint32_t k,
                        lqg_launch* best) {
    if (!h || !queries) return fail(LQG_ERR_ARG, "null argument");
    if (nq == 0 || nq > 0xFFFFFFFFull || k == 0 || ef == 0) return fail(LQG_ERR_ARG, "empty tuning set");
    const int cls = ef_class(ef);
    if (cls == 0) return fail(LQG_ERR_UNSUPPORTED, "ef above 1024");
    lqg_status st = check_queries(queries, nq * h->dim);
    if (st != LQG_OK) return st;
    st = lqg::ensure_device(h);
    if (st != LQG_OK) return st;
    lqg::DevBuf<float> dq;
    lqg::DevBuf<uint32_t> dids;
    HIP_TRY(dq.reserve(nq * h->dim));
    HIP_TRY(dids.reserve(nq * k));
    HIP_TRY(hipMemcpy(dq.p, queries, nq * h->dim * 4, hipMemcpyHostToDevice));
    const int slot = lqg::class_slot(cls);
    if (h->mode == LQG_MODE_FASTSCAN) {
        // register cap x waves per workgroup x LDS visited-table size
        const bool had = h->h2_launch_set[slot];
        const lqg::H2Cfg saved = h->h2_launch[slot];
        const lqg::H2Cfg base = lqg::h2_default_launch(cls);
        float best_ms = 1e30f;
        lqg::H2Cfg best_cfg = base;
        // a tuning set as small as a latency-mode launch also tries that mode
        // at each prefetch depth; a larger one keeps the class's depth
        const bool small = nq <= lqg::kLatencyMaxQueries;
        const int keep_pf = had ? saved.prefetch : base.prefetch;
        std::vector<lqg::H2Cfg> cands;
        for (int occ : {0, 6, 8}) {
            if (!h->h2_ready || !lqg::h2_variant_ok(h->h2dev(), occ, k)) continue;
            for (int extra = -1; extra < 2; ++extra)
                for (int wg : {1, 2, 4}) {
                    lqg::H2Cfg c = base;
                    c.occupancy = occ;
                    c.waves_per_group = wg;
                    c.hash_bits = base.hash_bits + extra;
                    c.prefetch = small ? 0 : keep_pf;
                    if (c.hash_bits >= 6) cands.push_back(c);
                }
        }
        if (small && h->h2_ready)
            for (int pf : {1, 2, 3})
                for (int extra = -1; extra < 2; ++extra) {
                    lqg::H2Cfg c = base;
                    c.hash_bits = base.hash_bits + extra;
                    c.prefetch = pf;
                    if (c.hash_bits >= 6) cands.push_back(c);
                }
        for (const lqg::H2Cfg& c : cands) {
            h->h2_launch[slot] = c;
            h->h2_launch_set[slot] = true;
            float ms_sum = 0.f;
            bool ok = true;
            for (int rep = 0; rep < 3 && ok; ++rep) {
                if (hipEventRecord(h->ev0, h->stream) != hipSuccess) ok = false;
                if (ok && lqg::run_batch(h, dq.p, nq, nullptr, nullptr, ef, k, LQG_ORDER_NEAREST,
                                         dids.p, nullptr, nullptr, h->stream) != LQG_OK)
                    ok = false;
                if (ok && hipEventRecord(h->ev1, h->stream) != hipSuccess) ok = false;
                if (ok && hipEventSynchronize(h->ev1) != hipSuccess) ok = false;
                float ms = 0.f;
                if (ok && hipEventElapsedTime(&ms, h->ev0, h->ev1) != hipSuccess) ok = false;
                if (rep > 0) ms_sum += ms;
            }
            (void)hipGetLastError();
            if (ok && ms_sum / 2 < best_ms) {
                best_ms = ms_sum / 2;
                best_cfg = c;
            }
        }
        dq.release();
        dids.release();
        if (best_ms >= 1e30f) {
            h->h2_launch[slot] = saved;
            h->h2_launch_set[slot] = had;
            return fail(LQG_ERR_DEVICE, "no launch setting ran");
        }
        h->h2_launch[slot] = best_cfg;
        h->h2_launch_set[slot] = true;
        if (best) {
            best->waves_per_group = best_cfg.waves_per_group;
            best->cache_bits = best_cfg.hash_bits;
            best->occupancy = best_cfg.occupancy;
            best->groups = best_cfg.groups;
            best->prefetch = best_cfg.prefetch;
            best->split = 1;
        }
        return LQG_OK;
    }
    const bool had = h->launch_set[slot];
    const lqg::LaunchCfg saved = h->launch[slot];
    float best_ms = 1e30f;
    lqg::LaunchCfg best_cfg = lqg::default_launch(cls, h->dim);
    // -1 (uncapped) differs from the default only up to the 256 class.  Pools
    // of 512 entries and more take the uncapped kernel only: capped at 64 or
    // 80 VGPRs they spill hundreds of registers (never the fastest), and one
    // such variant with 16-chunk row blocks faulted in a tuning run
    const int occs[] = {0, 6, 8, -1};
    const int n_occ = cls <= 256 ? 4 : 1;
    // visited-cache sizes: 2^8 .. 2^11, and up to 2^13 for pools of 512 or more
    const std::vector<int> bits =
        cls >= 512 ? std::vector<int>{9, 10, 11, 12, 13} : std::vector<int>{8, 9, 10, 11};
    const int wpg[] = {1, 2, 4};
    // A tuning set larger than a latency-mode launch times the batch kernel
    // only, and keeps the class's prefetch depth for small launches (a lone
    // search() after tuning on a large batch still runs in latency mode).
    const bool small = nq <= lqg::sq8_latency_max(h->dim);
    const int keep_pf = had ? saved.prefetch : lqg::default_launch(cls, h->dim).prefetch;
    std::vector<lqg::LaunchCfg> cands;
    for (int oi = 0; oi < n_occ; ++oi)
        for (int cb : bits)
            for (int wg : wpg) {
                lqg::LaunchCfg c = lqg::default_launch(cls, h->dim);
                c.occupancy = occs[oi];
                c.cache_bits = cb;
                c.waves_per_group = wg;
                c.prefetch = small ? 0 : keep_pf;
                c.split = 1;
                cands.push_back(c);
            }
    // split evaluation (the float path only: an index without an exact
    // integer decode), 2 or 4 waves per query
    const bool int_index = h->sq8_integer && h->sq8_identity && h->dim % 16 == 0 && h->dim <= 256;
    if (!int_index)
        for (int sw : {2, 4})
            for (int cb : bits) {
                lqg::LaunchCfg c = lqg::default_launch(cls, h->dim);
                c.cache_bits = cb;
                c.split = sw;
                cands.push_back(c);
            }
    // a tuning set as small as a latency-mode launch also tries that mode at
    // each prefetch depth (one query per workgroup: waves per group and the
    // register cap do not apply)
    if (small)
        for (int pf : {1, 2, 3, 4})
            for (int cb : bits) {
                lqg::LaunchCfg c = lqg::default_launch(cls, h->dim);
                c.cache_bits = cb;
                c.prefetch = pf;
                c.split = 1;
                cands.push_back(c);
            }
    for (const lqg::LaunchCfg& c : cands) {
        h->launch[slot] = c;
        h->launch_set[slot] = true;
        float ms_sum = 0.f;
        bool ok = true;
        for (int rep = 0; rep < 3 && ok; ++rep) {
            if (hipEventRecord(h->ev0, h->stream) != hipSuccess) ok = false;
            if (ok && lqg::run_batch(h, dq.p, nq, nullptr, nullptr, ef, k, LQG_ORDER_NEAREST,
                                     dids.p, nullptr, nullptr, h->stream) != LQG_OK)
                ok = false;
            if (ok && hipEventRecord(h->ev1, h->stream) != hipSuccess) ok = false;
            if (ok && hipEventSynchronize(h->ev1) != hipSuccess) ok = false;
            float ms = 0.f;
            if (ok && hipEventElapsedTime(&ms, h->ev0, h->ev1) != hipSuccess) ok = false;
            if (rep > 0) ms_sum += ms;  // the first run warms up
        }
        (void)hipGetLastError();
        if (ok && ms_sum / 2 < best_ms) {
            best_ms = ms_sum / 2;
            best_cfg = c;
        }
    }
    dq.release();
    dids.release();
    if (best_ms >= 1e30f) {
        h->launch[slot] = saved;
        h->launch_set[slot] = had;
        return fail(LQG_ERR_DEVICE, "no launch setting ran");
    }
    h->launch[slot] = best_cfg;
    h->launch_set[slot] = true;
    if (best) {
        best->waves_per_group = best_cfg.waves_per_group;
        best->cache_bits = best_cfg.cache_bits;
        best->occupancy = best_cfg.occupancy;
        best->groups = best_cfg.groups;
        best->prefetch = best_cfg.prefetch;
        best->split = best_cfg.split;
    }
    return LQG_OK;
}

lqg_status lqg_count_visits(lqg_index* h, const float* queries, uint64_t nq, uint32_t ef,
                            uint32_t k, uint32_t* stats) {
    if (!h || !queries || !stats) return fail(LQG_ERR_ARG, "null argument");
    if (nq == 0 || nq > 0xFFFFFFFFull || k == 0 || ef == 0) return fail(LQG_ERR_ARG, "empty query set");
    if (ef_class(ef) == 0) return fail(LQG_ERR_UNSUPPORTED, "ef above 1024");
    lqg_status st = check_queries(queries, nq * h->dim);
    if (st != LQG_OK) return st;
    st = lqg::ensure_device(h);
    if (st != LQG_OK) return st;
    lqg::DevBuf<float> dq;
    lqg::DevBuf<uint32_t> dids, dst;
    HIP_TRY(dq.reserve(nq * h->dim));
    HIP_TRY(dids.reserve(nq * k));
    HIP_TRY(dst.reserve(2 * nq));
    HIP_TRY(hipMemcpy(dq.p, queries, nq * h->dim * 4, hipMemcpyHostToDevice));
    st = lqg::run_batch(h, dq.p, nq, nullptr, nullptr, ef, k, LQG_ORDER_NEAREST, dids.p, nullptr,
                        dst.p, h->stream, true);
    if (st != LQG_OK) return st;
    HIP_TRY(hipMemcpyAsync(stats, dst.p, 2 * nq * 4, hipMemcpyDeviceToHost, h->stream));
    HIP_TRY(hipStreamSynchronize(h->stream));
    h->d_visit.release();
    return LQG_OK;
}

lqg_status lqg_info(const lqg_index* h, uint64_t* num_elements, uint64_t* dimension,
                    uint64_t* degree_bound, uint32_t* entry_point, uint32_t* ef) {
    if (!h) return fail(LQG_ERR_ARG, "null index");
    if (num_elements) *num_elements = h->n;
    if (dimension) *dimension = h->dim;
    if (degree_bound) *degree_bound = h->deg;
    if (entry_point) *entry_point = h->ep;
    if (ef) *ef = h->ef_set ? h->ef : 0;
    return LQG_OK;
}

lqg_status lqg_get_graph(const lqg_index* h, uint32_t* neighbors) {
    if (!h || !neighbors) return fail(LQG_ERR_ARG, "null argument");
    if (!h->have_data) return fail(LQG_ERR_STATE, "index has no data");
    memcpy(neighbors, h->nbr.data(), h->nbr.size() * 4);
    return LQG_OK;
}

lqg_status lqg_set_build_candidates(lqg_index* h, lqg_build_candidates source) {
    if (!h) return fail(LQG_ERR_ARG, "null index");
    const int v = lqg::abi_enum(source);
    if (v != LQG_BUILD_SQ8 && v != LQG_BUILD_FASTSCAN)
        return fail(LQG_ERR_ARG, "candidate source must be LQG_BUILD_SQ8 or LQG_BUILD_FASTSCAN");
    h->build_candidates = v;
    return LQG_OK;
}

lqg_status lqg_build_index(lqg_index* h, const float* data, uint64_t num_points,
                           uint64_t dimension, uint32_t ef_build, uint32_t num_iter,
                           uint32_t num_thread) {
    (void)num_thread;  // the build runs on the device
    if (!h || !data) return fail(LQG_ERR_ARG, "null argument");
    if (num_points != h->n || dimension != h->dim)
        return fail(LQG_ERR_ARG, "The shape of data is different with initialization!");
    if (num_iter <= 1) return fail(LQG_ERR_ARG, "The number of iter for building qg should >= 3");
    if (ef_build == 0) return fail(LQG_ERR_ARG, "EF must be positive");
    if (ef_class(ef_build) == 0)
        return fail(LQG_ERR_UNSUPPORTED, "build EF above 1024 is not supported");
    if (h->deg > 64) return fail(LQG_ERR_UNSUPPORTED, "build_index supports degree_bound <= 64");
    if (h->n <= h->deg) return fail(LQG_ERR_ARG, "num_elements must exceed degree_bound");
    if (check_queries(data, num_points * dimension) != LQG_OK)
        return fail(LQG_ERR_ARG, "data contains a non-finite value");
    lqg_status st = lqg::build_index(h, data, ef_build, num_iter);
    if (st == LQG_OK) lqg::h2_reset(h);
    return st;
}

}  // extern "C"
