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
    LQG_ERR_UNSUPPORTED = 6
} lqg_status;

/* Output order of the k ids (see DESIGN: the reference returns the k farthest
 * entries of its final pool, qg.hpp:289-296,340-343). */
typedef enum {
    LQG_ORDER_REFERENCE = 0, /* k farthest of the ef pool, ascending (as shipped) */
    LQG_ORDER_NEAREST = 1    /* k nearest of the same pool, ascending            */
} lqg_order;

typedef struct lqg_index lqg_index;

/* Per-call counters (optional out-parameter of the batch entry points). */
typedef struct {
    uint64_t queries;
    uint64_t expansions;  /* E: nodes popped and expanded                       */
    uint64_t evaluations; /* distance evaluations done on the device            */
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
 * d_stats (optional) receives 2 x nq uint32 counters (E, V) per query. */
lqg_status lqg_search_batch_device(lqg_index* index, const float* d_queries, uint64_t nq,
                                   const uint32_t* d_ef, const uint32_t* d_k, uint32_t ef_max,
                                   uint32_t kmax, lqg_order order, uint32_t* d_results,
                                   float* d_dists, uint32_t* d_stats, void* stream);

/* Launch settings of the search kernel for one pool-capacity class
 * (ef <= 64, 128, 256, 512, 1024).  No counterpart in the reference (CPU
 * code has no launch geometry). */
typedef struct {
    int32_t waves_per_group; /* wavefronts (queries in flight) per workgroup, 1..4      */
    int32_t cache_bits;      /* log2 entries of the per-wave visited cache, 4..14       */
    int32_t occupancy;       /* 8: kernel capped at 64 VGPRs (8 waves/SIMD); 0: uncapped */
    int32_t groups;          /* workgroups of the persistent grid; 0 = fill the device  */
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

/* Read-only views of the loaded index (host copies). */
lqg_status lqg_info(const lqg_index* index, uint64_t* num_elements, uint64_t* dimension,
                    uint64_t* degree_bound, uint32_t* entry_point, uint32_t* ef);

/* Copies the neighbour table (N x deg uint32) to host memory. */
lqg_status lqg_get_graph(const lqg_index* index, uint32_t* neighbors);

#ifdef __cplusplus
}
#endif

#endif /* LITEQG_H */
