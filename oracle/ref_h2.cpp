// ref_h2.cpp -- harness over the reference's FastScan/RaBitQ search (search_qg,
// qg.hpp:252-287), which LiteQG ships as dead code: its rows hold no codes or
// factors (code_offset_ = factor_offset_ = 0, qg.hpp:79-82, 420-436).  This
// driver restores the row layout the class documents (qg.hpp:73-78:
// raw | codes | factors | ids), fills codes and factors with the reference's
// own rabitq_codes (quantization/rabitq.hpp:23-59) on vectors rotated by the
// reference's FHTRotator (utils/rotator.hpp:88-114) under a sign vector read
// from a file, and runs the reference's search_qg.  It is a checker built
// into oracle/_ref; it never ships.
//
// qg.hpp:37 declares `friend class QGBuilder;`.  This translation unit does
// not include qg_builder.hpp and defines its own symqg::QGBuilder as the
// accessor, so the private members are reached without editing the
// reference.
//
// Commands (raw little-endian arrays):
//   search <index> N D deg <signs.f32 (P)> <queries.f32> NQ ef k <out_ids.u32>
//          <out_sidecar.bin> <out_prep.bin> NPREP
//     out_sidecar: per node, the rebuilt row's codes (deg*P/8 bytes, the
//     reference's FastScan block layout) then factors (3*deg floats:
//     triple_x[deg], factor_dq[deg], factor_vq[deg]).
//     out_prep: for the first NPREP queries: rotated query (P f32), lower,
//     width (f32), sumq (i32), LUT (4P u8), then for the entry point's row the
//     exact distance sqr_y (f32) and the deg estimates (f32).
//   time <index> N D deg <signs.f32 (P)> <queries.f32> NQ ef k threads rounds
//        <out_ids.u32 | -> <out_sidecar.bin | ->
//     the same rows, then the reference's search_qg timed as the FastScan CPU
//     baseline: single-thread per-query latency p50/p99 over the NQ queries,
//     and all-thread QPS with one QuantizedGraph copy per thread (search
//     scratch is per object, qg.hpp:66-71), the median of `rounds` passes
//     after a warm-up pass.  Prints one JSON line.  out_ids (NQ x k) are the
//     ids of the single-thread pass, out_sidecar the rebuilt rows' codes and
//     factors as `search` writes them ("-" skips either).
//   buffers <ops.u32> NOPS cap <out.bin>
//     replays (kind, id, dist-bits) ops on a SearchBuffer(cap) and a
//     ResultBuffer(cap) (utils/buffer.hpp) and dumps the sequences of popped
//     ids and the final ResultBuffer ids.
#include <omp.h>

#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <memory>
#include <string>
#include <vector>

#include "qg/qg.hpp"
#include "quantization/rabitq.hpp"

namespace symqg {

class QGBuilder {
   public:
    // Rebuild every row as raw | codes | factors | ids and fill codes and
    // factors with rabitq_codes over FHT-rotated vectors.
    static void add_sidecar(QuantizedGraph& g) {
        const size_t n = g.num_points_, d = g.dimension_, deg = g.degree_bound_,
                     p = g.padded_dim_;
        const size_t code_floats = deg * p / 32;  // deg * P / 8 bytes
        const size_t code_off = d, fac_off = d + code_floats, nb_off = fac_off + 3 * deg;
        const size_t row = nb_off + deg;
        decltype(g.data_) nd(std::vector<size_t>{n, row});
        for (size_t i = 0; i < n; ++i) {
            const float* src = &g.data_.at(g.row_offset_ * i);
            float* dst = &nd.at(row * i);
            std::memset(dst, 0, row * sizeof(float));
            std::memcpy(dst, src, d * sizeof(float));
            std::memcpy(dst + nb_off, src + g.neighbor_offset_, deg * sizeof(float));
        }
#pragma omp parallel for schedule(dynamic, 256)
        for (size_t i = 0; i < n; ++i) {
            float* dst = &nd.at(row * i);
            const PID* nb = reinterpret_cast<const PID*>(dst + nb_off);
            RowMatrix<float> x(deg, p), c(1, p);
            g.rotator_.rotate(dst, c.data());
            for (size_t j = 0; j < deg; ++j) g.rotator_.rotate(&nd.at(row * nb[j]), &x(j, 0));
            float* fac = dst + fac_off;
            rabitq_codes(x, c, reinterpret_cast<uint8_t*>(dst + code_off), fac, fac + deg,
                         fac + 2 * deg);
        }
        g.data_ = std::move(nd);
        g.code_offset_ = code_off;
        g.factor_offset_ = fac_off;
        g.neighbor_offset_ = nb_off;
        g.row_offset_ = row;
    }

    // A second graph with the same rows (a copy of the data array and the
    // offsets), for one search scratch per thread.
    static void clone_rows(const QuantizedGraph& g, QuantizedGraph& o) {
        decltype(g.data_) nd(std::vector<size_t>{g.num_points_, g.row_offset_});
        std::memcpy(&nd.at(0), &g.data_.at(0), g.num_points_ * g.row_offset_ * sizeof(float));
        o.data_ = std::move(nd);
        o.entry_point_ = g.entry_point_;
        o.code_offset_ = g.code_offset_;
        o.factor_offset_ = g.factor_offset_;
        o.neighbor_offset_ = g.neighbor_offset_;
        o.row_offset_ = g.row_offset_;
    }

    static void load_signs(QuantizedGraph& g, const char* path) {
        std::ifstream sf(path, std::ios::binary);
        g.rotator_.load(sf);  // FHTRotator::load, utils/rotator.hpp:116
    }

    static void search_h2(QuantizedGraph& g, const float* q, uint32_t k, uint32_t* out) {
        g.visited_.clear();
        g.search_pool_.clear();
        g.search_qg(q, k, out);
    }

    static void dump_sidecar(const QuantizedGraph& g, FILE* f) {
        const size_t deg = g.degree_bound_, p = g.padded_dim_;
        for (size_t i = 0; i < g.num_points_; ++i) {
            const float* r = &g.data_.at(g.row_offset_ * i);
            std::fwrite(r + g.code_offset_, 1, deg * p / 8, f);
            std::fwrite(r + g.factor_offset_, sizeof(float), 3 * deg, f);
        }
    }

    static void dump_prep(QuantizedGraph& g, const float* q, FILE* f) {
        const size_t p = g.padded_dim_, deg = g.degree_bound_;
        std::vector<float> rq(p);
        g.rotator_.rotate(q, rq.data());
        std::fwrite(rq.data(), sizeof(float), p, f);
        QGQuery qo(q, p);
        qo.query_prepare(g.rotator_, g.scanner_);
        std::fwrite(&qo.lower_val(), sizeof(float), 1, f);
        std::fwrite(&qo.width(), sizeof(float), 1, f);
        std::fwrite(&qo.sumq(), sizeof(int32_t), 1, f);
        std::fwrite(qo.lut().data(), 1, 4 * p, f);
        const float* row = &g.data_.at(g.row_offset_ * g.entry_point_);
        std::vector<float> est(deg);
        const float sqr_y = space::l2_sqr(q, row, g.dimension_);
        g.scanner_.scan_neighbors(est.data(), qo.lut().data(), sqr_y, qo.lower_val(),
                                  qo.width(), qo.sumq(),
                                  reinterpret_cast<const uint8_t*>(row + g.code_offset_),
                                  row + g.factor_offset_);
        std::fwrite(&sqr_y, sizeof(float), 1, f);
        std::fwrite(est.data(), sizeof(float), deg, f);
    }
};

}  // namespace symqg

namespace {

template <typename T>
std::vector<T> read_raw(const char* path, size_t count) {
    std::vector<T> v(count);
    FILE* f = std::fopen(path, "rb");
    if (!f || std::fread(v.data(), sizeof(T), count, f) != count) {
        std::fprintf(stderr, "cannot read %s\n", path);
        std::exit(2);
    }
    std::fclose(f);
    return v;
}

size_t arg(char** a, int i) { return static_cast<size_t>(std::strtoull(a[i], nullptr, 10)); }

int cmd_search(int argc, char** argv) {
    if (argc != 15) return 1;
    const size_t n = arg(argv, 3), d = arg(argv, 4), deg = arg(argv, 5);
    const size_t nq = arg(argv, 8), ef = arg(argv, 9), k = arg(argv, 10), nprep = arg(argv, 14);
    symqg::QuantizedGraph g(n, deg, d);
    g.load_index(argv[2]);
    symqg::QGBuilder::load_signs(g, argv[6]);
    symqg::QGBuilder::add_sidecar(g);
    g.set_ef(ef);
    auto q = read_raw<float>(argv[7], nq * d);
    std::vector<uint32_t> out(nq * k, 0xFFFFFFFFu);
    for (size_t i = 0; i < nq; ++i)
        symqg::QGBuilder::search_h2(g, &q[i * d], static_cast<uint32_t>(k), &out[i * k]);
    FILE* f = std::fopen(argv[11], "wb");
    std::fwrite(out.data(), 4, out.size(), f);
    std::fclose(f);
    f = std::fopen(argv[12], "wb");
    symqg::QGBuilder::dump_sidecar(g, f);
    std::fclose(f);
    f = std::fopen(argv[13], "wb");
    for (size_t i = 0; i < nprep && i < nq; ++i) symqg::QGBuilder::dump_prep(g, &q[i * d], f);
    std::fclose(f);
    return 0;
}

int cmd_time(int argc, char** argv) {
    if (argc != 15) return 1;
    const size_t n = arg(argv, 3), d = arg(argv, 4), deg = arg(argv, 5);
    const size_t nq = arg(argv, 8), ef = arg(argv, 9), k = arg(argv, 10);
    const int threads = static_cast<int>(arg(argv, 11));
    const size_t rounds = arg(argv, 12);
    if (threads < 1 || nq == 0 || rounds == 0) return 1;
    using clk = std::chrono::steady_clock;
    const auto t_load = clk::now();
    symqg::QuantizedGraph g(n, deg, d);
    g.load_index(argv[2]);
    symqg::QGBuilder::load_signs(g, argv[6]);
    symqg::QGBuilder::add_sidecar(g);
    g.set_ef(ef);
    const double side_s = std::chrono::duration<double>(clk::now() - t_load).count();
    auto q = read_raw<float>(argv[7], nq * d);

    // single thread: a warm-up pass (its ids are the output), then latency
    std::vector<uint32_t> out(nq * k, 0xFFFFFFFFu);
    for (size_t i = 0; i < nq; ++i)
        symqg::QGBuilder::search_h2(g, &q[i * d], static_cast<uint32_t>(k), &out[i * k]);
    std::vector<double> lat(nq);
    std::vector<uint32_t> o1(k);
    for (size_t i = 0; i < nq; ++i) {
        const auto t0 = clk::now();
        symqg::QGBuilder::search_h2(g, &q[i * d], static_cast<uint32_t>(k), o1.data());
        lat[i] = std::chrono::duration<double, std::milli>(clk::now() - t0).count();
    }
    std::sort(lat.begin(), lat.end());
    const double p50 = lat[nq / 2], p99 = lat[std::min(nq - 1, (nq * 99) / 100)];

    // all threads, one graph per thread
    std::vector<std::unique_ptr<symqg::QuantizedGraph>> gs(threads);
    for (int t = 0; t < threads; ++t) {
        gs[t] = std::make_unique<symqg::QuantizedGraph>(n, deg, d);
        symqg::QGBuilder::clone_rows(g, *gs[t]);
        symqg::QGBuilder::load_signs(*gs[t], argv[6]);
        gs[t]->set_ef(ef);
    }
    std::vector<double> qps;
    for (size_t r = 0; r < rounds + 1; ++r) {
        const auto t0 = clk::now();
#pragma omp parallel num_threads(threads)
        {
            const int t = omp_get_thread_num();
            std::vector<uint32_t> o(k);
#pragma omp for schedule(dynamic, 16)
            for (long i = 0; i < static_cast<long>(nq); ++i)
                symqg::QGBuilder::search_h2(*gs[t], &q[i * d], static_cast<uint32_t>(k), o.data());
        }
        const double s = std::chrono::duration<double>(clk::now() - t0).count();
        if (r > 0) qps.push_back(static_cast<double>(nq) / s);  // pass 0 is warm-up
    }
    std::sort(qps.begin(), qps.end());
    if (std::strcmp(argv[13], "-") != 0) {
        FILE* f = std::fopen(argv[13], "wb");
        if (!f || std::fwrite(out.data(), 4, out.size(), f) != out.size()) return 2;
        std::fclose(f);
    }
    if (std::strcmp(argv[14], "-") != 0) {
        FILE* f = std::fopen(argv[14], "wb");
        if (!f) return 2;
        symqg::QGBuilder::dump_sidecar(g, f);
        std::fclose(f);
    }
    std::printf(
        "{\"p50_ms\": %.6f, \"p99_ms\": %.6f, \"qps_all\": %.3f, \"threads\": %d, "
        "\"nq\": %zu, \"ef\": %zu, \"k\": %zu, \"sidecar_s\": %.3f}\n",
        p50, p99, qps[qps.size() / 2], threads, nq, ef, k, side_s);
    return 0;
}

// ops: triples (kind, id, dist bits); kind 0 = SearchBuffer.insert unless
// is_full(dist), 1 = SearchBuffer.pop (if has_next), 2 = ResultBuffer.insert
int cmd_buffers(int argc, char** argv) {
    if (argc != 6) return 1;
    const size_t nops = arg(argv, 3), cap = arg(argv, 4);
    auto ops = read_raw<uint32_t>(argv[2], 3 * nops);
    symqg::buffer::SearchBuffer sb(cap);
    symqg::buffer::ResultBuffer rb(cap);
    std::vector<uint32_t> popped;
    for (size_t i = 0; i < nops; ++i) {
        const uint32_t kind = ops[3 * i], id = ops[3 * i + 1];
        float dist;
        std::memcpy(&dist, &ops[3 * i + 2], 4);
        if (kind == 0) {
            if (!sb.is_full(dist)) sb.insert(id, dist);
        } else if (kind == 1) {
            popped.push_back(sb.has_next() ? sb.pop() : 0xFFFFFFFFu);
        } else {
            rb.insert(id, dist);
        }
    }
    std::vector<uint32_t> res(cap);
    rb.copy_results(res.data());
    FILE* f = std::fopen(argv[5], "wb");
    const uint32_t np = static_cast<uint32_t>(popped.size());
    std::fwrite(&np, 4, 1, f);
    std::fwrite(popped.data(), 4, popped.size(), f);
    std::fwrite(res.data(), 4, res.size(), f);
    std::fclose(f);
    return 0;
}

}  // namespace

int main(int argc, char** argv) {
    if (argc < 2) {
        std::fprintf(stderr, "usage: ref_h2 search|time|buffers ...\n");
        return 1;
    }
    const std::string c = argv[1];
    int rc = 1;
    if (c == "search") rc = cmd_search(argc, argv);
    else if (c == "time") rc = cmd_time(argc, argv);
    else if (c == "buffers") rc = cmd_buffers(argc, argv);
    if (rc == 1) std::fprintf(stderr, "bad arguments for %s\n", c.c_str());
    return rc;
}
