/*
 * liteqg.h -- C ABI of the MI355X (gfx950) LiteQG query path.
 *
 * Drop-in boundary for the reference's pybind11 class symphonyqg.Index
 * (python/bindings.cpp:35-125): every Index method maps to one entry point
 * here, with the same argument meaning.  Plain C types only; device memory
 * and streams are owned by the handle unless an entry point says otherwise.
 *
 * Errors: every function returns an lqg_status; on failure lqg_last_error()
 * returns a thread-local message.  Where the reference prints to std::cerr
 * and carries on (bindings.cpp:45-60, 80-86) these entry points fail instead.
 */
#ifndef LITEQG_H
#define LITEQG_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
    LQG_OK = 0,
    LQG_ERR_ARG = 1,     /* invalid argument (shape, range, null pointer)      */
    LQG_ERR_STATE = 2,   /* call order (e.g. search before set_ef / load)      */
    LQG_ERR_IO = 3,      /* file missing, short, or of the wrong size          */
    LQG_ERR_DEVICE = 4,  /* HIP runtime error or no gfx950 device              */
    LQG_ERR_NOMEM = 5,   /* host or device allocation failed                   */
    LQG_ERR_UNSUPPORTED = 6,
    LQG_ERR_BUSY = 7     /* serving: no free request slot, or result not ready  */
} lqg_status;

/* Output order of the k ids (see DESIGN: the reference returns the k farthest
 * entries of its final pool, qg.hpp:289-296,340-343). */
typedef enum {
    LQG_ORDER_REFERENCE = 0, /* k farthest of the ef pool, ascending (as shipped) */
    LQG_ORDER_NEAREST = 1    /* k nearest of the same pool, ascending            */
} lqg_order;

/* Search path used by lqg_search / lqg_search_batch / lqg_search_batch_device. */
typedef enum {
    /* QuantizedGraph::search_qg_with_heap (qg.hpp:298-344), the path the
     * reference's Index.search runs (qg.hpp:235-243): SQ8 distances, dual heap */
    LQG_MODE_SQ8 = 0,
    /* QuantizedGraph::search_qg (qg.hpp:252-287): FastScan RaBitQ estimates
     * steer a SearchBuffer(ef), exact L2 of expanded nodes fills a top-k
     * ResultBuffer.  Needs a sidecar (lqg_h2_build_sidecar / _set_sidecar /
     * _load_sidecar).  Output order is nearest first; `order` is ignored. */
    LQG_MODE_FASTSCAN = 1
} lqg_mode;

/* How the FastScan path rotates each query (QGQuery::query_prepare,
 * qg_query.hpp:28-43, via FHTRotator::rotate, utils/rotator.hpp:88-114). */
typedef enum {
    /* In the search kernel: sign multiply, then the unnormalised
     * Walsh-Hadamard butterflies in the reference's order.  Bit-identical
     * rotated queries, hence LUTs, to the reference's (default). */
    LQG_ROTATION_FHT = 0,
    /* One batched FP32 GEMM on the matrix cores before the search:
     * Q[nq x D] times H_P diag(signs).  Same products, another summation
     * order, so rotated values differ from the FHT's in the last bits and a
     * LUT entry can differ by one where a value sits on a quantisation
     * boundary. */
    LQG_ROTATION_GEMM = 1
} lqg_rotation;

/* Where lqg_build_index takes each node's candidate neighbours from. */
typedef enum {
    /* A beam search of width EF on SQ8 distances from the entry point, its
     * nearest min(EF, 750) entries (default). */
    LQG_BUILD_SQ8 = 0,
    /* The reference's find_candidates (qg.hpp:439-473, qg_builder.hpp:
     * 202-231): a FastScan RaBitQ search of width EF from the entry point
     * over the graph being built, whose codes and factors are recomputed
     * for each iteration's graph, and every node it expands. */
    LQG_BUILD_FASTSCAN = 1
} lqg_build_candidates;

typedef struct lqg_index lqg_index;

/* Per-call counters (optional out-parameter of the batch entry points). */
typedef struct {
    uint64_t queries;
    uint64_t expansions;  /* E: nodes popped and expanded                       */
    uint64_t evaluations; /* SQ8 mode: distance evaluations done on the device;
                             FastScan mode: SearchBuffer pops (expanded or
                             skipped as visited)                                */
    double device_ms;     /* kernel time measured with HIP events               */
} lqg_stats;

/* Thread-local description of the last failure on this thread. */
const char* lqg_last_error(void);

/* symphonyqg.Index(index_type, metric, num_elements, dimension, degree_bound)
 * -- python/bindings.cpp:35-60, qg.hpp:183-193.  index_type must be "QG",
 * metric "L2", degree_bound a positive multiple of 32. */
lqg_status lqg_create(const char* index_type, const char* metric, uint64_t num_elements,
                      uint64_t dimension, uint64_t degree_bound, lqg_index** out);

/* Releases host and device memory of the handle. */
lqg_status lqg_destroy(lqg_index* index);

/* Index.load(filename) -- bindings.cpp:63, qg.hpp:217-224.  Reads the
 * reference's file byte for byte (no header) and uploads it to the device.
 * The file size must equal 4 + N*(D+deg)*4 + N*D + 8*D. */
lqg_status lqg_load(lqg_index* index, const char* filename);

/* Index.save(filename) -- bindings.cpp:65, qg.hpp:206-215.  Byte-identical
 * to the reference's writer. */
lqg_status lqg_save(lqg_index* index, const char* filename);

/* Index.set_ef(EF) -- bindings.cpp:67, qg.hpp:226-230. */
lqg_status lqg_set_ef(lqg_index* index, uint32_t ef);

/* Output order for lqg_search (default LQG_ORDER_REFERENCE, as shipped). */
lqg_status lqg_set_order(lqg_index* index, lqg_order order);

/* Index.build_index(data, EF, num_iter, num_thread) -- bindings.cpp:69-90,
 * qg_builder.hpp:48-91.  data is N x D float32, host memory.  The graph is
 * built on the device; num_thread is accepted for signature parity. */
lqg_status lqg_build_index(lqg_index* index, const float* data, uint64_t num_points,
                           uint64_t dimension, uint32_t ef_build, uint32_t num_iter,
                           uint32_t num_thread);

/* Candidate source of later lqg_build_index calls (default LQG_BUILD_SQ8).
 * The reference's builder always runs its FastScan search
 * (qg_builder.hpp:202-231); this setter has no counterpart there. */
lqg_status lqg_set_build_candidates(lqg_index* index, lqg_build_candidates source);

/* Index.search(query, k) -- bindings.cpp:92-98, qg.hpp:235-244,298-344.
 * One query (D floats, host memory), one kernel launch; writes k ids to
 * results (host memory).  Slots the pool cannot fill are 0xFFFFFFFF. */
lqg_status lqg_search(lqg_index* index, const float* query, uint32_t k, uint32_t* results);

/* Many queries in one persistent launch (no counterpart in the reference,
 * which searches one query per call).  Host pointers.
 *   queries nq x D;  ef/k: per-query arrays or NULL (then the index's ef and
 *   kmax are used);  results nq x kmax ids;  dists (optional) nq x kmax
 *   squared distances;  order as lqg_set_order. */
lqg_status lqg_search_batch(lqg_index* index, const float* queries, uint64_t nq,
                            const uint32_t* ef, const uint32_t* k, uint32_t kmax,
                            lqg_order order, uint32_t* results, float* dists,
                            lqg_stats* stats);

/* Same as lqg_search_batch with every array in device memory, enqueued on
 * `stream` (a hipStream_t, NULL = the handle's stream) and not synchronised.
 * d_stats (optional) receives 2 x nq uint32 counters per query: (E, V) in SQ8
 * mode, (E, pops) in FastScan mode (E = 0xFFFFFFFF marks a query whose
 * visited set overflowed; the host entry points fail instead).
 * Every call on one index uses the index's single work-queue counter: two
 * calls must not run at the same time (for concurrent launches use graphs,
 * lqg_graph_create, or the serving engine, lqg_serve_start, which own their
 * scratch). */
lqg_status lqg_search_batch_device(lqg_index* index, const float* d_queries, uint64_t nq,
                                   const uint32_t* d_ef, const uint32_t* d_k, uint32_t ef_max,
                                   uint32_t kmax, lqg_order order, uint32_t* d_results,
                                   float* d_dists, uint32_t* d_stats, void* stream);

/* lqg_search_batch_device that allocates nothing.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code under
 * HIP-graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * Same arguments and results.  The launch scratch (work-queue words, the
 * FastScan overflow flag, GEMM-rotated queries) comes from one static device
 * array of 1,052,928 bytes per GPU, held until the process ends, instead of
 * buffers the library allocates.  Limits:
 *   - the index must already be on the device (any earlier search on it);
 *   - FastScan mode has no device overflow visited tables: a query whose
 *     visited set passes three quarters of its LDS table (384 to 3,072
 *     expansions by pool class) fails as in lqg_h2_set_overflow_bits(index, 0);
 *   - with LQG_ROTATION_GEMM, nq x P must be at most 262,144 (LQG_ERR_ARG).
 * Calls share the GPU's array: two calls that overlap on one GPU race. */
lqg_status lqg_search_batch_device_noalloc(lqg_index* index, const float* d_queries,
                                           uint64_t nq, const uint32_t* d_ef, const uint32_t* d_k,
                                           uint32_t ef_max, uint32_t kmax, lqg_order order,
                                           uint32_t* d_results, float* d_dists,
                                           uint32_t* d_stats, void* stream);

/* A batch launch captured into a HIP graph: lqg_search_batch_device with
 * fixed arguments (device arrays whose contents may change between replays),
 * recorded once and replayed with one hipGraphLaunch.  The graph holds the
 * work-queue reset and the persistent search kernel (kernel nodes only), in
 * the mode and with the launch settings in force at creation.  It refers to
 * the index's device memory: once the index is reloaded, rebuilt or given a
 * new sidecar, lqg_graph_launch refuses to replay (LQG_ERR_STATE); destroy
 * the graph before the index.  No counterpart in the reference. */
typedef struct lqg_batch_graph lqg_batch_graph;

lqg_status lqg_graph_create(lqg_index* index, const float* d_queries, uint64_t nq,
                            const uint32_t* d_ef, const uint32_t* d_k, uint32_t ef_max,
                            uint32_t kmax, lqg_order order, uint32_t* d_results, float* d_dists,
                            uint32_t* d_stats, lqg_batch_graph** out);
/* Enqueues one replay on `stream` (a hipStream_t, NULL = the index's stream). */
lqg_status lqg_graph_launch(lqg_batch_graph* graph, void* stream);
lqg_status lqg_graph_destroy(lqg_batch_graph* graph);

/* Serving engine: requests (one query each, its own ef and k) are submitted
 * as they arrive and complete asynchronously.  A dispatcher thread keeps up to
 * `lanes` batch launches in flight on separate streams, each stream with a
 * hardware queue of its own (a CU-masked stream over every CU, so launches of
 * different lanes do not wait for each other in a shared queue); requests that arrive
 * while launches run are admitted to the next free lane (continuous
 * batching), so their work overlaps the work already on the device.
 * Admission: at most `slots` requests outstanding (host pinned request
 * table); ef <= ef_max and k <= kmax, which fix the kernel's LDS layout for
 * the life of the server.  The search path is the index's mode at dispatch
 * time.  Results of a request do not depend on what else is in flight.  No
 * counterpart in the reference (one query per call, qg.hpp:235-243). */
typedef struct lqg_server lqg_server;

typedef struct {
    uint32_t lanes;     /* launches in flight, 1..16                              */
    uint32_t max_batch; /* requests per launch, 1..65536                          */
    uint32_t slots;     /* outstanding requests, >= max_batch                     */
    uint32_t ef_max;    /* largest ef a request may ask, <= 1024                  */
    uint32_t kmax;      /* largest k a request may ask                            */
    lqg_order order;    /* SQ8 mode output order                                  */
} lqg_serve_config;

typedef struct {
    uint64_t submitted, completed, batches;
    uint32_t max_lanes_busy; /* most launches seen in flight at once            */
    uint32_t max_batch_seen; /* largest batch dispatched                       */
} lqg_serve_stats;

lqg_status lqg_serve_start(lqg_index* index, const lqg_serve_config* cfg, lqg_server** out);
/* Copies the query (D floats); LQG_ERR_BUSY when every slot is taken. */
lqg_status lqg_serve_submit(lqg_server* server, const float* query, uint32_t ef, uint32_t k,
                            uint64_t* ticket);
/* LQG_OK and the request's k ids (dists optional) once done, LQG_ERR_BUSY
 * before.  capacity = entries results and dists hold: below the request's k
 * the call fails with LQG_ERR_ARG and the ticket stays valid.  A request that
 * could not finish (its launch failed, or its own FastScan visited set
 * overflowed; the other requests of its batch are unaffected) returns
 * LQG_ERR_DEVICE.  The ticket is released by a poll that returns LQG_OK or
 * LQG_ERR_DEVICE. */
lqg_status lqg_serve_poll(lqg_server* server, uint64_t ticket, uint32_t* results, float* dists,
                          uint32_t capacity);
/* Blocks until done, then as lqg_serve_poll. */
lqg_status lqg_serve_wait(lqg_server* server, uint64_t ticket, uint32_t* results, float* dists,
                          uint32_t capacity);
lqg_status lqg_serve_get_stats(lqg_server* server, lqg_serve_stats* stats);
/* Drains in-flight work and frees the server. */
lqg_status lqg_serve_stop(lqg_server* server);

/* Load replay for latency measurement: submits query i at arrival_ns[i]
 * (nanoseconds after the call starts; non-decreasing) with ef[i], k[i],
 * retrying while the server is full, and waits for all.  Writes latency_ns[i]
 * (completion minus scheduled arrival), results (nq x kmax, 0xFFFFFFFF past
 * k[i]) if given. */
lqg_status lqg_serve_replay(lqg_server* server, const float* queries, uint64_t nq,
                            const uint32_t* ef, const uint32_t* k, const uint64_t* arrival_ns,
                            uint64_t* latency_ns, uint32_t* results);

/* Launch settings of the search kernel for one pool-capacity class
 * (ef <= 64, 128, 192, 256, 512, 1024).  No counterpart in the reference (CPU
 * code has no launch geometry). */
typedef struct {
    int32_t waves_per_group; /* wavefronts (queries in flight) per workgroup, 1..4      */
    int32_t cache_bits;      /* log2 entries of the per-wave visited cache, 4..14 (SQ8
                                mode); of the LDS visited table, 6..14 (FastScan)    */
    int32_t occupancy;       /* kernel register cap.  SQ8: 0 = default (at least 4
                                waves/SIMD, 128 VGPRs, for pool classes up to 256;
                                uncapped above), -1 = uncapped, 8 = 64 VGPRs
                                (8 waves/SIMD), 6 = 80 VGPRs (6 waves/SIMD).
                                FastScan: 0 = uncapped, 6 or 8 waves/SIMD (80 or
                                64 VGPRs), for indexes with P / lanes per
                                neighbour <= 128 code bits (8: k <= 64)          */
    int32_t groups;          /* workgroups of the persistent grid; 0 = fill the device  */
    int32_t prefetch;        /* launches of at most 256 queries (a lone search()),
                                and SQ8 launches of at most 1,024 on indexes of
                                at most 256 dims:
                                0 = the batch kernel; 1, 2 or 3 = latency mode, one
                                query per workgroup with a second wave that loads
                                ahead what the next expansions read.  SQ8 -- 1: the
                                id rows of the two best unexpanded pool entries,
                                2: also their neighbours' SQ8 rows, 3: both for
                                the four best, 4: the second wave computes the
                                keys of all neighbours of the best unexpanded
                                entry ahead of its expansion (speculative
                                evaluation).  FastScan (0..3): the node records
                                of the 2^prefetch best.  Results are the same;
                                only time changes.  Default 2                     */
    int32_t split;           /* SQ8 float path (an index whose SQ8 decode is not the
                                identity, or D > 256): waves per query sharing each
                                expansion's distance evaluation, one query per
                                workgroup; 0 or 1 = off, 2 or 4.  Takes precedence
                                over prefetch.  Default: 2 for D >= 512 in launches
                                of at most 4,096 queries, else off.
                                FastScan: unused                                  */
} lqg_launch;

/* Sets the launch settings used for every ef in the class of `ef` (ef = 0:
 * all classes); cfg = NULL restores the defaults. */
lqg_status lqg_set_launch(lqg_index* index, uint32_t ef, const lqg_launch* cfg);

/* Reads the launch settings in force for the class of `ef`. */
lqg_status lqg_get_launch(const lqg_index* index, uint32_t ef, lqg_launch* cfg);

/* Autotuner: times each candidate launch setting on nq queries (host
 * memory) at (ef, k), keeps the fastest for ef's class and returns it in
 * best (optional).  Results are unaffected by the setting; only speed is. */
lqg_status lqg_autotune(lqg_index* index, const float* queries, uint64_t nq, uint32_t ef,
                        uint32_t k, lqg_launch* best);

/* Diagnostic: runs the batch search of nq queries (host memory) at (ef, k)
 * with an exact N-bit visited bitmap per wavefront beside the kernel's lossy
 * cache, and writes per query (E, V) to stats (2 x nq uint32): expansions, and
 * distinct nodes evaluated -- the counts of the reference's visited set
 * (qg.hpp:320-328).  Used for algorithmic-byte accounting; not for speed. */
lqg_status lqg_count_visits(lqg_index* index, const float* queries, uint64_t nq, uint32_t ef,
                            uint32_t k, uint32_t* stats);

/* Diagnostic: enable (default) or disable the exact integer SQ8 kernel.  An
 * index whose SQ8 parameters are all min 0 / scale 1, searched with queries
 * integer valued in [0, 255] (D % 16 == 0, D <= 256), has every FP32 step of
 * the reference's distance exact, and a kernel computing the distances with
 * integer dot products gives bit-identical results.  Disabled, every query
 * runs the general float kernel (dequantise, FMA partials).  Results do not
 * change; only speed does.  No counterpart in the reference. */
lqg_status lqg_set_sq8_integer(lqg_index* index, int enable);

/* Selects the search path.  LQG_MODE_FASTSCAN requires a sidecar.  No
 * counterpart in the reference, whose search() is hard-wired to the SQ8 path
 * (qg.hpp:241-242). */
lqg_status lqg_set_mode(lqg_index* index, lqg_mode mode);
lqg_status lqg_get_mode(const lqg_index* index, lqg_mode* mode);

/* FastScan sidecar: per node, RaBitQ codes (deg x P/8 bytes, P = padded
 * dimension, 1 << ceil(log2 D), 64 <= P <= 2048) and factors (triple_x[deg],
 * fac_dq[deg], fac_vq[deg]) of the node's neighbours relative to the node,
 * under the rotation x -> H_P diag(signs) [x, 0]; the row layout the reference
 * documents (qg.hpp:73-78) and fills with rabitq_codes (rabitq.hpp:23-59),
 * which LiteQG itself never calls (qg.hpp:489-501).
 *
 * Builds it on the device from the loaded index: signs = +-1/sqrt(P) (as
 * FHTRotator draws them, utils/rotator.hpp:37-43) from `signs` (P floats) if
 * given, else from a generator seeded with `seed`. */
lqg_status lqg_h2_build_sidecar(lqg_index* index, const float* signs, uint64_t seed);

/* Installs a sidecar from host arrays: signs (P), codes (N x deg*P/8 bytes) in
 * the reference's FastScan block layout (layout 0, fastscan_impl.hpp:67-124)
 * or vector-major (layout 1: neighbour j's P/8 bytes, byte b = codebook 2b in
 * the low nibble, 2b+1 in the high nibble, bit 3 = first dim), factors
 * (N x 3*deg floats). */
lqg_status lqg_h2_set_sidecar(lqg_index* index, const float* signs, const uint8_t* codes,
                              int layout, const float* factors);

/* Copies the sidecar out (codes vector-major); any pointer may be NULL. */
lqg_status lqg_h2_get_sidecar(const lqg_index* index, float* signs, uint8_t* codes,
                              float* factors);

/* Sidecar file: header (magic, N, D, deg, P, fingerprint of the index's
 * vectors and neighbour ids), signs, codes (vector-major), factors.  Loading
 * checks the header against the loaded index. */
lqg_status lqg_h2_save_sidecar(const lqg_index* index, const char* path);
lqg_status lqg_h2_load_sidecar(lqg_index* index, const char* path);

/* Rotation used by the FastScan searches of this index (default FHT).  No
 * counterpart in the reference, which has only the FHT. */
lqg_status lqg_h2_set_rotation(lqg_index* index, lqg_rotation method);
lqg_status lqg_h2_get_rotation(const lqg_index* index, lqg_rotation* method);

/* The rotated queries (nq x P floats, host memory in and out) by either
 * method, for checks against an FP64 rotation.  Needs a sidecar (the signs).
 * Replaces FHTRotator::rotate (utils/rotator.hpp:88-114) for a batch. */
lqg_status lqg_h2_rotate(lqg_index* index, const float* queries, uint64_t nq,
                         lqg_rotation method, float* out);

/* Parity probe of the FastScan path: for each of nq queries (host memory) the
 * device's query preparation (QGQuery::query_prepare, qg_query.hpp:28-43):
 * lut (nq x 4P bytes, the reference's layout), lo_width (nq x 2), sumq (nq);
 * and if nodes (nq ids) is given, the exact distance to nodes[i] (sqr_y, nq)
 * and the FastScan estimates of its deg neighbours (est, nq x deg). */
lqg_status lqg_h2_probe(lqg_index* index, const float* queries, uint64_t nq,
                        const uint32_t* nodes, uint8_t* lut, float* lo_width, int32_t* sumq,
                        float* sqr_y, float* est);

/* Diagnostic: the first cap nodes the FastScan search of each query (host
 * memory, nq x D) expands at ef, in expansion order, to out (nq x cap,
 * 0xFFFFFFFF past a query's last expansion) -- the candidate list the
 * reference's builder takes from find_candidates (qg.hpp:439-473), and
 * what lqg_build_index with LQG_BUILD_FASTSCAN takes per node. */
lqg_status lqg_h2_expansions(lqg_index* index, const float* queries, uint64_t nq, uint32_t ef,
                             uint32_t cap, uint32_t* out);

/* FastScan visited set: each query's exact set lives in an LDS table sized by
 * its pool class and moves, past three quarters full, to a per-wave table of
 * 2^bits entries in device memory (default 14); past three quarters of that
 * the query fails
 * (E = 0xFFFFFFFF in d_stats, ids 0xFFFFFFFF, its own serving request fails).
 * bits = 0: no device table, a query fails once its LDS table is 3/4 full
 * (for tests of that failure path); otherwise 8 <= bits <= 16.  No
 * counterpart in the reference, whose hash set grows without bound
 * (third/ngt/hashset.hpp:30-110). */
lqg_status lqg_h2_set_overflow_bits(lqg_index* index, uint32_t bits);

/* FastScan search kernel choice.  Indexes of degree 32 with D = 64, 96 or 128
 * run a kernel specialised for that record shape (offsets and the exact
 * distance's length fixed at compile time); generic = 1 runs the general
 * kernel for them too (same results; for comparisons).  Default 0.  No
 * counterpart in the reference. */
lqg_status lqg_h2_set_generic(lqg_index* index, int generic);

/* Read-only views of the loaded index (host copies). */
lqg_status lqg_info(const lqg_index* index, uint64_t* num_elements, uint64_t* dimension,
                    uint64_t* degree_bound, uint32_t* entry_point, uint32_t* ef);

/* Copies the neighbour table (N x deg uint32) to host memory. */
lqg_status lqg_get_graph(const lqg_index* index, uint32_t* neighbors);

#ifdef __cplusplus
}
#endif

#endif /* LITEQG_H */
