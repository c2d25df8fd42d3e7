// ref_driver.cpp -- a small command-line driver over the reference's own
// header-only C++ (symqglib/), compiled from /root/reference by
// oracle/Makefile into oracle/_ref/.  It is a checker: it writes golden
// vectors for tests/golden and times the reference's search as the CPU
// baseline.  It never ships with the library.
//
// Commands (all arrays are raw little-endian files, as numpy.tofile writes):
//   build  <data.f32> N D deg EF num_iter threads <out.index>
//            QGBuilder(index, EF, data, threads).build(num_iter)
//            (python/bindings.cpp:69-90), then save_index (qg.hpp:206-215)
//   search <index> N D deg <queries.f32> NQ ef k <out.u32>
//            set_ef(ef) (qg.hpp:226-230), then search(q, k, out) per query
//            (qg.hpp:235-243), exactly what Index.search does
//            (python/bindings.cpp:92-98)
//   funcs  <codes.u8> <params.f32> <queries.f32> D M <out_dec.f32> <out_l2.f32>
//            sq8_dequantize (quantization/sq8.hpp:28-32) of M code rows and
//            space::l2_sqr (space/l2.hpp:16-59) of query row i vs decoded row i
//   time   <index> N D deg <queries.f32> NQ ef k threads rounds
//            prints one JSON line: single-thread per-query latency p50/p99 and
//            all-thread QPS, each thread with its own QuantizedGraph (search
//            scratch is per object, qg.hpp:66-71)
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
#include "qg/qg_builder.hpp"

namespace {

template <typename T>
std::vector<T> read_raw(const char* path, size_t count) {
    std::vector<T> v(count);
    FILE* f = std::fopen(path, "rb");
    if (!f) {
        std::fprintf(stderr, "cannot open %s\n", path);
        std::exit(2);
    }
    size_t got = std::fread(v.data(), sizeof(T), count, f);
    std::fclose(f);
    if (got != count) {
        std::fprintf(stderr, "short read %s: %zu of %zu\n", path, got, count);
        std::exit(2);
    }
    return v;
}

template <typename T>
void write_raw(const char* path, const T* p, size_t count) {
    FILE* f = std::fopen(path, "wb");
    if (!f || std::fwrite(p, sizeof(T), count, f) != count) {
        std::fprintf(stderr, "cannot write %s\n", path);
        std::exit(2);
    }
    std::fclose(f);
}

size_t arg(char** a, int i) { return static_cast<size_t>(std::strtoull(a[i], nullptr, 10)); }

int cmd_build(int argc, char** argv) {
    if (argc != 10) return 1;
    size_t n = arg(argv, 3), d = arg(argv, 4), deg = arg(argv, 5), ef = arg(argv, 6);
    size_t iters = arg(argv, 7), threads = arg(argv, 8);
    auto data = read_raw<float>(argv[2], n * d);
    symqg::QuantizedGraph qg(n, deg, d);
    symqg::QGBuilder builder(qg, static_cast<uint32_t>(ef), data.data(), threads);
    builder.build(iters);
    qg.save_index(argv[9]);
    return 0;
}

int cmd_search(int argc, char** argv) {
    if (argc != 11) return 1;
    size_t n = arg(argv, 3), d = arg(argv, 4), deg = arg(argv, 5);
    size_t nq = arg(argv, 7), ef = arg(argv, 8), k = arg(argv, 9);
    auto q = read_raw<float>(argv[6], nq * d);
    symqg::QuantizedGraph qg(n, deg, d);
    qg.load_index(argv[2]);
    qg.set_ef(ef);
    std::vector<uint32_t> out(nq * k, 0xFFFFFFFFu);
    for (size_t i = 0; i < nq; ++i) qg.search(&q[i * d], static_cast<uint32_t>(k), &out[i * k]);
    write_raw(argv[10], out.data(), out.size());
    return 0;
}

int cmd_funcs(int argc, char** argv) {
    if (argc != 9) return 1;
    size_t d = arg(argv, 5), m = arg(argv, 6);
    auto codes = read_raw<uint8_t>(argv[2], m * d);
    auto params = read_raw<float>(argv[3], 2 * d);
    auto q = read_raw<float>(argv[4], m * d);
    std::vector<symqg::SQ8Params> p(d);
    for (size_t i = 0; i < d; ++i) p[i] = symqg::SQ8Params{params[2 * i], params[2 * i + 1]};
    std::vector<float> dec(m * d), l2(m);
    for (size_t i = 0; i < m; ++i) {
        symqg::sq8_dequantize(&codes[i * d], d, &dec[i * d], p.data());
        l2[i] = symqg::space::l2_sqr(&q[i * d], &dec[i * d], d);
    }
    write_raw(argv[7], dec.data(), dec.size());
    write_raw(argv[8], l2.data(), l2.size());
    return 0;
}

int cmd_time(int argc, char** argv) {
    if (argc != 12) return 1;
    size_t n = arg(argv, 3), d = arg(argv, 4), deg = arg(argv, 5);
    size_t nq = arg(argv, 7), ef = arg(argv, 8), k = arg(argv, 9);
    int threads = static_cast<int>(arg(argv, 10));
    size_t rounds = arg(argv, 11);
    auto q = read_raw<float>(argv[6], nq * d);
    using clk = std::chrono::steady_clock;

    // single thread latency
    symqg::QuantizedGraph qg(n, deg, d);
    qg.load_index(argv[2]);
    qg.set_ef(ef);
    std::vector<uint32_t> out(k);
    for (size_t i = 0; i < nq; ++i) qg.search(&q[i * d], static_cast<uint32_t>(k), out.data());
    std::vector<double> lat(nq);
    for (size_t i = 0; i < nq; ++i) {
        auto t0 = clk::now();
        qg.search(&q[i * d], static_cast<uint32_t>(k), out.data());
        lat[i] = std::chrono::duration<double, std::milli>(clk::now() - t0).count();
    }
    std::sort(lat.begin(), lat.end());
    double p50 = lat[nq / 2], p99 = lat[std::min(nq - 1, (nq * 99) / 100)];

    // all-thread throughput, one QuantizedGraph per thread
    std::vector<std::unique_ptr<symqg::QuantizedGraph>> g(threads);
    for (int t = 0; t < threads; ++t) {
        g[t] = std::make_unique<symqg::QuantizedGraph>(n, deg, d);
        g[t]->load_index(argv[2]);
        g[t]->set_ef(ef);
    }
    std::vector<double> qps_r;
    for (size_t r = 0; r < rounds + 1; ++r) {
        auto t0 = clk::now();
#pragma omp parallel num_threads(threads)
        {
            int t = omp_get_thread_num();
            std::vector<uint32_t> o(k);
#pragma omp for schedule(dynamic, 16)
            for (long i = 0; i < static_cast<long>(nq); ++i)
                g[t]->search(&q[i * d], static_cast<uint32_t>(k), o.data());
        }
        double s = std::chrono::duration<double>(clk::now() - t0).count();
        if (r > 0) qps_r.push_back(static_cast<double>(nq) / s);  // pass 0 is warm-up
    }
    std::sort(qps_r.begin(), qps_r.end());
    std::printf(
        "{\"p50_ms\": %.6f, \"p99_ms\": %.6f, \"qps_all\": %.3f, \"threads\": %d, "
        "\"nq\": %zu, \"ef\": %zu, \"k\": %zu}\n",
        p50, p99, qps_r[qps_r.size() / 2], threads, nq, ef, k);
    return 0;
}

}  // namespace

int main(int argc, char** argv) {
    if (argc < 2) {
        std::fprintf(stderr, "usage: ref_driver build|search|funcs|time ...\n");
        return 1;
    }
    std::string c = argv[1];
    int rc = 1;
    if (c == "build") rc = cmd_build(argc, argv);
    else if (c == "search") rc = cmd_search(argc, argv);
    else if (c == "funcs") rc = cmd_funcs(argc, argv);
    else if (c == "time") rc = cmd_time(argc, argv);
    if (rc == 1) std::fprintf(stderr, "bad arguments for %s\n", c.c_str());
    return rc;
}
