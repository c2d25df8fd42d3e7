// lqg_graph.cpp -- batch launches captured into HIP graphs (replayed with one
// hipGraphLaunch; no counterpart in the reference).
#include <hip/hip_runtime.h>

#include "lqg_index.h"

struct lqg_batch_graph {
    lqg_index* index = nullptr;
    lqg::Scratch scratch;  // its own work-queue counter and overflow tables
    uint64_t generation = 0;  // the index's generation at capture
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
};

using lqg::fail;

extern "C" {

lqg_status lqg_graph_create(lqg_index* h, const float* d_queries, uint64_t nq,
                            const uint32_t* d_ef, const uint32_t* d_k, uint32_t ef_max,
                            uint32_t kmax, lqg_order order, uint32_t* d_results, float* d_dists,
                            uint32_t* d_stats, lqg_batch_graph** out) {
    if (!h || !d_queries || !d_results || !out) return fail(LQG_ERR_ARG, "null argument");
    *out = nullptr;
    if (nq == 0 || nq > 0x7FFFFFFFull) return fail(LQG_ERR_ARG, "nq must be in [1, 2^31)");
    if (kmax == 0) return fail(LQG_ERR_ARG, "kmax must be positive");
    if (lqg::abi_enum(order) != LQG_ORDER_REFERENCE && lqg::abi_enum(order) != LQG_ORDER_NEAREST)
        return fail(LQG_ERR_ARG, "unknown order");
    if (ef_max == 0) {
        if (d_ef) return fail(LQG_ERR_ARG, "ef_max must bound the per-query ef array");
        if (!h->ef_set) return fail(LQG_ERR_STATE, "set_ef() or ef_max required");
        ef_max = h->ef;
    }
    if (lqg::ef_class(ef_max) == 0) return fail(LQG_ERR_UNSUPPORTED, "ef above 1024 is not supported");
    lqg_status st = lqg::ensure_device(h);
    if (st != LQG_OK) return st;
    hipStream_t cs = nullptr;
    LQG_HIP_TRY(hipStreamCreateWithFlags(&cs, hipStreamNonBlocking));
    // one uncaptured run first: every scratch buffer the launch needs is
    // allocated here, so the captured run allocates nothing
    lqg_batch_graph* g = new (std::nothrow) lqg_batch_graph();
    if (!g) {
        (void)hipStreamDestroy(cs);
        return fail(LQG_ERR_NOMEM, "allocation failed");
    }
    st = lqg::run_batch(h, d_queries, nq, d_ef, d_k, ef_max, kmax, order, d_results, d_dists,
                        d_stats, cs, false, &g->scratch);
    hipError_t e = st == LQG_OK ? hipStreamSynchronize(cs) : hipSuccess;
    if (st == LQG_OK && e == hipSuccess) {
        e = hipStreamBeginCapture(cs, hipStreamCaptureModeRelaxed);
        if (e == hipSuccess) {
            st = lqg::run_batch(h, d_queries, nq, d_ef, d_k, ef_max, kmax, order, d_results,
                                d_dists, d_stats, cs, false, &g->scratch);
            hipError_t e2 = hipStreamEndCapture(cs, &g->graph);
            if (st == LQG_OK) e = e2;
        }
        if (st == LQG_OK && e == hipSuccess)
            e = hipGraphInstantiate(&g->exec, g->graph, nullptr, nullptr, 0);
    }
    (void)hipStreamDestroy(cs);
    if (st != LQG_OK || e != hipSuccess) {
        if (g->exec) (void)hipGraphExecDestroy(g->exec);
        if (g->graph) (void)hipGraphDestroy(g->graph);
        g->scratch.release();
        delete g;
        return st != LQG_OK ? st : lqg::hip_fail(e, "graph capture");
    }
    g->index = h;
    g->generation = h->generation;
    *out = g;
    return LQG_OK;
}

lqg_status lqg_graph_launch(lqg_batch_graph* g, void* stream) {
    if (!g || !g->exec) return fail(LQG_ERR_ARG, "null graph");
    if (g->generation != g->index->generation)
        return fail(LQG_ERR_STATE, "the index was reloaded, rebuilt or given a new sidecar since "
                                   "this graph was captured: destroy it and capture again");
    hipStream_t s = stream ? static_cast<hipStream_t>(stream) : g->index->stream;
    LQG_HIP_TRY(hipGraphLaunch(g->exec, s));
    return LQG_OK;
}

lqg_status lqg_graph_destroy(lqg_batch_graph* g) {
    if (!g) return LQG_OK;
    if (g->exec) (void)hipGraphExecDestroy(g->exec);
    if (g->graph) (void)hipGraphDestroy(g->graph);
    g->scratch.release();
    delete g;
    return LQG_OK;
}

}  // extern "C"
