/*
 * lqg_oracle.c -- CPU restatement of LiteQG's query path (H1: SQ8 + exact L2 +
 * dual-heap beam search) used ONLY as a checker by tests/, the smoke check and
 * the CPU-baseline leg of bench.py.  Nothing in the shipped library links this
 * file.  Every function cites the reference file:line it restates.
 *
 * Pinning (how each function is checked, see tests/test_oracle_*.py):
 *   orc_sq8_dequantize  -- against the reference's own dequantised vectors
 *                          (tests/golden, written by oracle/_ref) and the
 *                          closed form min + q*scale.
 *   orc_l2_sqr          -- against the reference's l2_sqr on random vectors
 *                          (golden), and against an FP64 sum to 1e-6 rel.
 *   orc_search_heap     -- against the reference's search() ids on reference-
 *                          built index files (golden), bit-exact.
 *   orc_search_canon    -- against orc_search_heap on tie-free data (identical
 *                          ids) and against invariants (pool = ef smallest keys
 *                          of everything evaluated, every pool entry expanded).
 *   orc_bruteforce_fp64 -- against numpy argsort of FP64 distances.
 *   orc_knn_exact_f32   -- build-side helper, checked like the brute force.
 *
 * Compiled with -O2 -ffp-contract=off: every fused multiply-add is written
 * explicitly with fmaf(), so the rounding sequence is the one stated here.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#ifdef _OPENMP
#include <omp.h>
#endif

#define ORC_FILL 0xFFFFFFFFu

/* ------------------------------------------------------------------------ */
/* SQ8 dequantisation.  quantization/sq8.hpp:28-32:                         */
/*   data[i] = params[i].min + quantized[i] * params[i].scale               */
/* params is D x {float min, float scale} (sq8.hpp:6-9).  fused=1 computes  */
/* fmaf(q, scale, min) -- the form GCC emits for that expression under the  */
/* reference's -Ofast -march=native flags (python/setup.py:61-63); fused=0  */
/* rounds the product first.  Which one the reference build produces is     */
/* settled by the golden vectors, not assumed here.                         */
/* ------------------------------------------------------------------------ */
void orc_sq8_dequantize(const uint8_t* q, uint32_t dim, const float* params, float* out,
                        int fused) {
    for (uint32_t i = 0; i < dim; ++i) {
        float mn = params[2 * i], sc = params[2 * i + 1];
        float qf = (float)q[i];
        out[i] = fused ? fmaf(qf, sc, mn) : (mn + qf * sc);
    }
}

/* ------------------------------------------------------------------------ */
/* Squared L2, AVX-512 branch of space/l2.hpp:16-34.                        */
/* 16 running partial sums s[j] over dims j, j+16, j+32, ... each updated   */
/* with one FMA per step (l2.hpp:24-29), then _mm512_reduce_add_ps          */
/* (l2.hpp:30): pairwise sums (j, j+8), then (j, j+4), then (j, j+2), then  */
/* (0, 1).  A dimension tail (dim % 16, no workload has one) is added       */
/* sequentially afterwards (l2.hpp:31-34); the reference build contracts    */
/* that tail to an FMA, which is what is written here.                      */
/* ------------------------------------------------------------------------ */
float orc_l2_sqr(const float* a, const float* b, uint32_t dim) {
    float s[16];
    for (int j = 0; j < 16; ++j) s[j] = 0.0f;
    uint32_t mul16 = dim - (dim & 15u);
    uint32_t i = 0;
    for (; i < mul16; i += 16) {
        for (int j = 0; j < 16; ++j) {
            float t = a[i + j] - b[i + j];
            s[j] = fmaf(t, t, s[j]);
        }
    }
    float l1[8], l2[4];
    for (int j = 0; j < 8; ++j) l1[j] = s[j] + s[j + 8];
    for (int j = 0; j < 4; ++j) l2[j] = l1[j] + l1[j + 4];
    float c0 = l2[0] + l2[2];
    float c1 = l2[1] + l2[3];
    float r = c0 + c1;
    for (; i < dim; ++i) {
        float t = a[i] - b[i];
        r = fmaf(t, t, r);
    }
    return r;
}

/* ------------------------------------------------------------------------ */
/* Index view.  The file body is (qg.hpp:206-224, third/svs/array.hpp:      */
/* 109-116): u32 entry point | N x (D + deg) f32 rows, neighbour ids at     */
/* float offset D (qg.hpp:420-424) | N x D u8 SQ8 codes | D x {min,scale}.  */
/* The oracle takes the three arrays separately.                            */
/* ------------------------------------------------------------------------ */
typedef struct {
    uint32_t n, dim, deg, ep;
    const uint32_t* nbr; /* n x deg neighbour ids                         */
    const uint8_t* qdata; /* n x dim SQ8 codes                            */
    const float* params;  /* dim x {min, scale}                           */
    int fused;            /* dequantisation rounding (see above)          */
} orc_index;

typedef struct {
    uint32_t id;
    float d;
} orc_cand;

/* Exact visited set over [0, n): one bit per node.  HashBasedBooleanSet    */
/* (third/ngt/hashset.hpp:30-110) with its LazyCleanupSet overflow          */
/* (utils/lazy_cleanup_set.hpp:10-42) is an exact set; only membership      */
/* matters for the search.                                                  */
typedef struct {
    uint64_t* bits;
    uint32_t n;
} orc_bitset;

static int bs_get(const orc_bitset* s, uint32_t i) { return (int)((s->bits[i >> 6] >> (i & 63)) & 1u); }
static void bs_set(orc_bitset* s, uint32_t i) { s->bits[i >> 6] |= (uint64_t)1 << (i & 63); }

static float node_dist(const orc_index* ix, const float* query, uint32_t id, float* scratch) {
    orc_sq8_dequantize(ix->qdata + (size_t)id * ix->dim, ix->dim, ix->params, scratch, ix->fused);
    return orc_l2_sqr(query, scratch, ix->dim);
}

/* ------------------------------------------------------------------------ */
/* Binary max-heap by distance with the exact sift rules of libstdc++'s     */
/* std::push_heap / std::pop_heap (GCC 11 bits/stl_heap.h: __push_heap,    */
/* __adjust_heap), which std::priority_queue uses.  Comparator:             */
/* lhs.distance < rhs.distance (qg.hpp:289-296).  Reproducing the sift     */
/* rules is what fixes the order among equal distances.                     */
/* ------------------------------------------------------------------------ */
static void heap_sift_up(orc_cand* h, size_t hole, size_t top, orc_cand v) {
    while (hole > top) {
        size_t parent = (hole - 1) / 2;
        if (!(h[parent].d < v.d)) break;
        h[hole] = h[parent];
        hole = parent;
    }
    h[hole] = v;
}

static void heap_push(orc_cand* h, size_t* n, orc_cand v) {
    size_t hole = (*n)++;
    heap_sift_up(h, hole, 0, v);
}

static void heap_pop(orc_cand* h, size_t* n) {
    size_t len = *n;
    if (len > 1) {
        orc_cand v = h[len - 1];
        h[len - 1] = h[0];
        size_t L = len - 1; /* __adjust_heap(first, 0, L, v) */
        size_t hole = 0, child = 0;
        while (child < (L - 1) / 2) {
            child = 2 * (child + 1);
            if (h[child].d < h[child - 1].d) child--;
            h[hole] = h[child];
            hole = child;
        }
        if ((L & 1) == 0 && child == (L - 2) / 2) {
            child = 2 * (child + 1);
            h[hole] = h[child - 1];
            hole = child - 1;
        }
        heap_sift_up(h, hole, 0, v);
    }
    (*n)--;
}

/* ------------------------------------------------------------------------ */
/* H1 as shipped: QuantizedGraph::search_qg_with_heap, qg.hpp:298-344.      */
/*   results[k]      -- output exactly as the reference writes it: the k    */
/*                      entries popped from the top of res_pool (the k      */
/*                      FARTHEST of the final pool) into results[k-1..0]    */
/*                      (qg.hpp:340-343).  Slots the reference leaves       */
/*                      unwritten (pool smaller than k) are set to          */
/*                      0xFFFFFFFF here.                                    */
/*   pool_ids/pool_d -- optional (may be NULL), capacity ef: the whole final */
/*                      pool in the order the heap pops it, reversed        */
/*                      (ascending distance); *pool_n its size.             */
/*   stats[2]        -- optional: expansions E, distance evaluations V.     */
/* ------------------------------------------------------------------------ */
int orc_search_heap(const orc_index* ix, const float* query, uint32_t ef, uint32_t k,
                    uint32_t* results, uint32_t* pool_ids, float* pool_d, uint32_t* pool_n,
                    uint64_t* stats) {
    if (ef == 0 || k == 0) return -1;
    size_t cap_sp = (size_t)ix->n + 1;
    orc_cand* sp = (orc_cand*)malloc(sizeof(orc_cand) * cap_sp);      /* search_pool */
    orc_cand* rp = (orc_cand*)malloc(sizeof(orc_cand) * ((size_t)ef + 2)); /* res_pool */
    float* dec = (float*)malloc(sizeof(float) * ix->dim);
    orc_bitset vis = {(uint64_t*)calloc(((size_t)ix->n + 63) / 64, 8), ix->n};
    if (!sp || !rp || !dec || !vis.bits) {
        free(sp); free(rp); free(dec); free(vis.bits);
        return -2;
    }
    size_t nsp = 0, nrp = 0;
    uint64_t E = 0, V = 0;

    uint32_t cur = ix->ep;                                   /* qg.hpp:305-311 */
    float sqr_y = node_dist(ix, query, cur, dec);
    V++;
    float lower = sqr_y;
    heap_push(sp, &nsp, (orc_cand){cur, -sqr_y});
    heap_push(rp, &nrp, (orc_cand){cur, sqr_y});
    bs_set(&vis, cur);

    while (nsp > 0) {                                        /* qg.hpp:313-338 */
        orc_cand top = sp[0];
        cur = top.id;
        if (-top.d > lower) break;
        heap_pop(sp, &nsp);
        E++;
        const uint32_t* nb = ix->nbr + (size_t)cur * ix->deg;
        for (uint32_t i = 0; i < ix->deg; ++i) {
            uint32_t x = nb[i];
            if (bs_get(&vis, x)) continue;
            bs_set(&vis, x);
            sqr_y = node_dist(ix, query, x, dec);
            V++;
            if (nrp < ef || lower > sqr_y) {
                heap_push(sp, &nsp, (orc_cand){x, -sqr_y});
                heap_push(rp, &nrp, (orc_cand){x, sqr_y});
                if (nrp > ef) heap_pop(rp, &nrp);
                lower = rp[0].d;
            }
        }
    }

    /* drain the final pool from the top: farthest first (qg.hpp:340-343) */
    size_t total = nrp;
    for (uint32_t i = 0; i < k; ++i) results[i] = ORC_FILL;
    for (size_t j = 0; j < total; ++j) {
        orc_cand t = rp[0];
        if (j < k) results[k - 1 - j] = t.id;
        if (pool_ids) pool_ids[total - 1 - j] = t.id;
        if (pool_d) pool_d[total - 1 - j] = t.d;
        heap_pop(rp, &nrp);
    }
    if (pool_n) *pool_n = (uint32_t)total;
    if (stats) {
        stats[0] = E;
        stats[1] = V;
    }
    free(sp); free(rp); free(dec); free(vis.bits);
    return 0;
}

/* ------------------------------------------------------------------------ */
/* H1 with a total order: the same loop as qg.hpp:298-344, but every        */
/* comparison is on the key (distance, id) lexicographically, so no two     */
/* candidates compare equal.  With distinct keys the two heaps of the       */
/* reference reduce to one sorted pool of capacity ef whose entries are     */
/* expanded smallest-first until all are expanded (an entry pushed out of   */
/* res_pool always has a key above the current lowerBound, and the          */
/* reference stops at the first such top, qg.hpp:316-317).  This is the     */
/* semantics the device kernel implements; it equals the reference on any   */
/* input where no two evaluated distances tie.                              */
/*   order 0 = "reference": results as qg.hpp:340-343 (k farthest, asc.)    */
/*   order 1 = "nearest"  : the k nearest of the same pool, ascending.      */
/* Unfilled slots are 0xFFFFFFFF; dists (optional) get +inf there.          */
/* ------------------------------------------------------------------------ */
static int key_less(float da, uint32_t ia, float db, uint32_t ib) {
    return da < db || (da == db && ia < ib);
}

int orc_search_canon(const orc_index* ix, const float* query, uint32_t ef, uint32_t k,
                     int order, uint32_t* results, float* dists, uint32_t* pool_ids,
                     float* pool_d, uint32_t* pool_n, uint64_t* stats) {
    if (ef == 0 || k == 0) return -1;
    orc_cand* pool = (orc_cand*)malloc(sizeof(orc_cand) * ((size_t)ef + 1));
    uint8_t* done = (uint8_t*)malloc((size_t)ef + 1);
    float* dec = (float*)malloc(sizeof(float) * ix->dim);
    orc_bitset vis = {(uint64_t*)calloc(((size_t)ix->n + 63) / 64, 8), ix->n};
    if (!pool || !done || !dec || !vis.bits) {
        free(pool); free(done); free(dec); free(vis.bits);
        return -2;
    }
    uint64_t E = 0, V = 0;
    size_t size = 0;

    float d0 = node_dist(ix, query, ix->ep, dec);
    V++;
    pool[0] = (orc_cand){ix->ep, d0};
    done[0] = 0;
    size = 1;
    bs_set(&vis, ix->ep);

    for (;;) {
        size_t c = 0;
        while (c < size && done[c]) ++c;
        if (c == size) break;
        done[c] = 1;
        uint32_t cur = pool[c].id;
        E++;
        const uint32_t* nb = ix->nbr + (size_t)cur * ix->deg;
        for (uint32_t i = 0; i < ix->deg; ++i) {
            uint32_t x = nb[i];
            if (bs_get(&vis, x)) continue;
            bs_set(&vis, x);
            float d = node_dist(ix, query, x, dec);
            V++;
            if (size == ef && !key_less(d, x, pool[size - 1].d, pool[size - 1].id)) continue;
            size_t pos = size;
            while (pos > 0 && key_less(d, x, pool[pos - 1].d, pool[pos - 1].id)) {
                pool[pos] = pool[pos - 1];
                done[pos] = done[pos - 1];
                --pos;
            }
            pool[pos] = (orc_cand){x, d};
            done[pos] = 0;
            if (size < ef) ++size;
        }
    }

    for (uint32_t i = 0; i < k; ++i) {
        results[i] = ORC_FILL;
        if (dists) dists[i] = INFINITY;
    }
    if (order == 0) {
        size_t m = size < k ? size : k;
        for (size_t j = 0; j < m; ++j) {
            results[k - 1 - j] = pool[size - 1 - j].id;
            if (dists) dists[k - 1 - j] = pool[size - 1 - j].d;
        }
    } else {
        size_t m = size < k ? size : k;
        for (size_t j = 0; j < m; ++j) {
            results[j] = pool[j].id;
            if (dists) dists[j] = pool[j].d;
        }
    }
    if (pool_ids)
        for (size_t j = 0; j < size; ++j) pool_ids[j] = pool[j].id;
    if (pool_d)
        for (size_t j = 0; j < size; ++j) pool_d[j] = pool[j].d;
    if (pool_n) *pool_n = (uint32_t)size;
    if (stats) {
        stats[0] = E;
        stats[1] = V;
    }
    free(pool); free(done); free(dec); free(vis.bits);
    return 0;
}

/* Many queries, OpenMP over queries (the all-core CPU baseline).  results  */
/* is nq x k, dists nq x k (optional), stats nq x 2 (optional).             */
int orc_search_canon_batch(const orc_index* ix, const float* queries, uint64_t nq, uint32_t ef,
                           uint32_t k, int order, uint32_t* results, float* dists,
                           uint64_t* stats, int threads) {
    int err = 0;
#ifdef _OPENMP
    if (threads > 0) omp_set_num_threads(threads);
#pragma omp parallel for schedule(dynamic, 4) reduction(| : err)
#endif
    for (int64_t q = 0; q < (int64_t)nq; ++q) {
        int r = orc_search_canon(ix, queries + (size_t)q * ix->dim, ef, k, order,
                                 results + (size_t)q * k, dists ? dists + (size_t)q * k : NULL,
                                 NULL, NULL, NULL, stats ? stats + 2 * (size_t)q : NULL);
        err |= (r != 0);
    }
    (void)threads;
    return err ? -1 : 0;
}

int orc_search_heap_batch(const orc_index* ix, const float* queries, uint64_t nq, uint32_t ef,
                          uint32_t k, uint32_t* results, int threads) {
    int err = 0;
#ifdef _OPENMP
    if (threads > 0) omp_set_num_threads(threads);
#pragma omp parallel for schedule(dynamic, 4) reduction(| : err)
#endif
    for (int64_t q = 0; q < (int64_t)nq; ++q) {
        int r = orc_search_heap(ix, queries + (size_t)q * ix->dim, ef, k, results + (size_t)q * k,
                                NULL, NULL, NULL, NULL);
        err |= (r != 0);
    }
    (void)threads;
    return err ? -1 : 0;
}

/* ------------------------------------------------------------------------ */
/* Ground truth: FP64 squared L2 over the raw vectors, top-k ascending,     */
/* ties broken by ascending id.  (The reference computes its ground truth   */
/* in float32, reproduce/compute_gt.py:18-25; FP64 removes that rounding.)  */
/* ------------------------------------------------------------------------ */
int orc_bruteforce_fp64(const float* base, uint64_t n, uint32_t dim, const float* queries,
                        uint64_t nq, uint32_t k, uint32_t* ids, double* dists, int threads) {
    int err = 0;
#ifdef _OPENMP
    if (threads > 0) omp_set_num_threads(threads);
#pragma omp parallel for schedule(dynamic, 1) reduction(| : err)
#endif
    for (int64_t q = 0; q < (int64_t)nq; ++q) {
        const float* qv = queries + (size_t)q * dim;
        uint32_t* oi = ids + (size_t)q * k;
        double* od = (double*)malloc(sizeof(double) * k);
        if (!od) { err |= 1; continue; }
        uint32_t m = 0;
        for (uint64_t i = 0; i < n; ++i) {
            const float* x = base + (size_t)i * dim;
            double s = 0.0;
            for (uint32_t j = 0; j < dim; ++j) {
                double t = (double)qv[j] - (double)x[j];
                s += t * t;
            }
            if (m == k && !(s < od[m - 1])) continue; /* ties keep the smaller id */
            uint32_t pos = (m < k) ? m : k - 1;
            while (pos > 0 && s < od[pos - 1]) {
                od[pos] = od[pos - 1];
                oi[pos] = oi[pos - 1];
                --pos;
            }
            od[pos] = s;
            oi[pos] = (uint32_t)i;
            if (m < k) ++m;
        }
        for (uint32_t j = m; j < k; ++j) { oi[j] = ORC_FILL; od[j] = INFINITY; }
        if (dists) memcpy(dists + (size_t)q * k, od, sizeof(double) * k);
        free(od);
    }
    (void)threads;
    return err ? -1 : 0;
}
