/*
 * lqg_oracle.c -- CPU restatement of LiteQG's query paths (H1: SQ8 + exact L2 +
 * dual-heap beam search; H2: FastScan RaBitQ estimates + exact re-rank) used
 * ONLY as a checker by tests/, the smoke check and
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
 *
 * H2 (FastScan RaBitQ search_qg, qg.hpp:252-287), tests/test_oracle_h2.py:
 *   orc_buffers_replay  -- SearchBuffer/ResultBuffer against the reference's
 *                          own replay of a tie-heavy operation trace.
 *   orc_fht_rotate      -- bit-exact against the reference's rotated queries,
 *                          and against scipy's Sylvester Hadamard in FP64.
 *   orc_query_prepare   -- lo, width, sumq, LUT bytes against the reference;
 *                          closed forms (q range, LUT = subset sums).
 *   orc_scan_estimates  -- bit-exact against the reference's estimates; the
 *                          two code layouts give identical sums.
 *   orc_search_h2       -- ids against the reference's search_qg on rebuilt
 *                          rows (reference sidecar), every ef of the fixture.
 *   orc_h2_expansions_batch -- the same traversal's expansion order; pinned
 *                          through orc_search_h2 (its pops and expansions are
 *                          the ones that decide the pinned ids) and by the
 *                          invariants of its tests (the list's length is E,
 *                          distinct ids, the first is the entry point, every
 *                          later one is a neighbour of an earlier one).
 *   orc_sidecar_build   -- codes bit-exact against rabitq_codes; factors to
 *                          2e-6 of their magnitude (FP64 sums vs float).
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

/* ======================================================================== */
/* H2: the FastScan RaBitQ search, QuantizedGraph::search_qg (qg.hpp:252-   */
/* 287), dead code in LiteQG and restated here over an explicit sidecar.    */
/* ======================================================================== */

/* FHTRotator::rotate (utils/rotator.hpp:88-114): dst[i] = src[i] * s[i] for */
/* i < D (s = the +-1/sqrt(P) sign vector, mat_), zero padding to P, then   */
/* the unnormalised Walsh-Hadamard transform of FFHT (third/ffht/           */
/* fht_avx.hpp:223-1538): butterflies (a, b) -> (a + b, a - b) applied in   */
/* stride order 1, 2, 4, ..., P/2.                                           */
void orc_fht_rotate(const float* signs, const float* src, uint32_t dim, uint32_t p, float* dst) {
    for (uint32_t i = 0; i < dim; ++i) dst[i] = src[i] * signs[i];
    for (uint32_t i = dim; i < p; ++i) dst[i] = 0.0f;
    for (uint32_t h = 1; h < p; h <<= 1)
        for (uint32_t b = 0; b < p; b += 2 * h)
            for (uint32_t j = b; j < b + h; ++j) {
                float a = dst[j], c = dst[j + h];
                dst[j] = a + c;
                dst[j + h] = a - c;
            }
}

/* QGQuery::query_prepare (qg/qg_query.hpp:28-43):                           */
/*   data_range (utils/scalar_quantize.hpp:11-38): min and max seeded with  */
/*     +0 (the AVX-512 branch), so lo = min(0, min v), hi = max(0, max v);  */
/*   width = (hi - lo) / 63 (QG_BQUERY = 6, common.hpp:11);                  */
/*   quantize (scalar_quantize.hpp:40-58):                                   */
/*     q = lround(double((v - lo) * (1/width)) + 0.5), sumq = sum q;         */
/*   pack_lut_impl (quantization/fastscan_impl.hpp:258-270): LUT[c][j] =    */
/*     sum of q[4c + 3 - b] over the set bits b of j, in uint8 (wraps).      */
/* The reference is built with -Ofast (python/setup.py:61-63), which folds  */
/* the two constant divisions: width = (hi - lo) * fl(1/63) and 1/width =   */
/* fl(1/fl(1/63)) / (hi - lo) = 62.999996f / (hi - lo).  The latter puts   */
/* the maximum coordinate just below 63 instead of on it, so it decides q   */
/* there; the reference's own LUT and sumq outputs fix both forms.          */
/* A NaN product (hi == lo) quantises to 0, which is what the x86 lround    */
/* conversion of NaN yields after the uint8 cast.                            */
void orc_query_prepare(const float* signs, const float* query, uint32_t dim, uint32_t p,
                       float* rq, float* lo_out, float* width_out, int32_t* sumq_out,
                       uint8_t* qbytes, uint8_t* lut) {
    orc_fht_rotate(signs, query, dim, p, rq);
    float lo = 0.0f, hi = 0.0f;
    for (uint32_t i = 0; i < p; ++i) {
        float t = rq[i];
        lo = t < lo ? t : lo;
        hi = t > hi ? t : hi;
    }
    const float inv63 = 1.0f / 63.0f;
    float width = (hi - lo) * inv63;
    float ow = (1.0f / inv63) / (hi - lo);
    int32_t sum = 0;
    for (uint32_t i = 0; i < p; ++i) {
        float f = (rq[i] - lo) * ow;
        double x = (double)f + 0.5;
        uint8_t q = 0;
        if (x == x && x < 9.2e18 && x > -9.2e18) q = (uint8_t)lround(x);
        qbytes[i] = q;
        sum += q;
    }
    for (uint32_t c = 0; c < p / 4; ++c)
        for (uint32_t j = 0; j < 16; ++j) {
            uint32_t s = 0;
            for (uint32_t b = 0; b < 4; ++b)
                if (j & (1u << b)) s += qbytes[4 * c + 3 - b];
            lut[16 * c + j] = (uint8_t)s;
        }
    *lo_out = lo;
    *width_out = width;
    *sumq_out = sum;
}

/* Nibble of neighbour v for codebook c in the reference's FastScan block   */
/* layout (pack_codes / pack_codes_helper, fastscan_impl.hpp:67-124): block */
/* v/32 of 4P bytes; codebook pair c/2 is a 32-byte group whose byte j      */
/* holds vector kPerm0[j] (low nibble) and kPerm0[j]+16 (high nibble) of    */
/* codebook c (j < 16) or c+1 (j >= 16); kPerm0 = 0,8,1,9,...,7,15.          */
static uint32_t ref_nibble(const uint8_t* codes, uint32_t p, uint32_t v, uint32_t c) {
    uint32_t w = v & 31u, u = w & 15u;
    uint32_t j = u < 8 ? 2 * u : 2 * (u - 8) + 1;
    const uint8_t* g = codes + (size_t)(v >> 5) * 4 * p + (size_t)(c >> 1) * 32 + (c & 1u) * 16;
    uint8_t byte = g[j];
    return w < 16 ? (byte & 15u) : (byte >> 4);
}

/* Nibble of neighbour v for codebook c in this build's vector-major layout: */
/* P/8 bytes per vector, byte c/2, low nibble for even c, high for odd c.    */
static uint32_t vm_nibble(const uint8_t* codes, uint32_t p, uint32_t v, uint32_t c) {
    uint8_t byte = codes[(size_t)v * (p / 8) + c / 2];
    return (c & 1u) ? (byte >> 4) : (byte & 15u);
}

/* QGScanner::scan_neighbors (qg/qg_scanner.hpp:99-155) for deg neighbours: */
/* accumulate_impl's uint16 sums of LUT entries (fastscan_impl.hpp:127-255),*/
/* read back as int16 and turned into float(2r - sumq) (qg_scanner.hpp:    */
/* 117-135, _mm512_cvtepi16_epi32 sign-extends), then appro_dist_impl's     */
/* AVX-512 estimate (qg_scanner.hpp:23-48):                                  */
/*   est = (fac_dq * width) * f + fma(fac_vq, lo, triple_x + sqr_y),         */
/* which the -Ofast reference build reassociates to (f * fac_dq) * width;   */
/* the reference's own estimate outputs fix that order.                     */
/* layout 0 = reference block layout, 1 = vector-major layout.               */
void orc_scan_estimates(const uint8_t* codes, int layout, const float* factors, uint32_t deg,
                        uint32_t p, const uint8_t* lut, float sqr_y, float lo, float width,
                        int32_t sumq, uint16_t* sums_out, float* est) {
    const float* tx = factors;
    const float* dq = factors + deg;
    const float* vq = factors + 2 * deg;
    for (uint32_t v = 0; v < deg; ++v) {
        uint32_t r = 0;
        for (uint32_t c = 0; c < p / 4; ++c) {
            uint32_t nib = layout == 0 ? ref_nibble(codes, p, v, c) : vm_nibble(codes, p, v, c);
            r += lut[16 * c + nib];
        }
        uint16_t r16 = (uint16_t)(r & 0xFFFFu);
        if (sums_out) sums_out[v] = r16;
        float f = (float)((int32_t)(int16_t)r16 * 2 - sumq);
        float a = tx[v] + sqr_y;
        float b = fmaf(vq[v], lo, a);
        float c = (f * dq[v]) * width;
        est[v] = c + b;
    }
}

/* SearchBuffer (utils/buffer.hpp:11-78): sorted by distance, insert at the  */
/* first position whose distance is not below (before equal keys), the last */
/* entry dropped when full; pop returns the first unchecked entry and marks */
/* it checked.                                                               */
typedef struct {
    uint32_t* id;
    float* d;
    uint8_t* checked;
    size_t size, cur, cap;
} orc_sbuf;

static size_t first_not_below(const float* d, size_t n, float x) {
    size_t lo = 0, len = n;
    while (len > 0) {
        size_t h = len / 2;
        if (d[lo + h] < x) {
            lo += h + 1;
            len -= h + 1;
        } else {
            len = h;
        }
    }
    return lo;
}

static int sbuf_is_full(const orc_sbuf* s, float x) {
    return s->size == s->cap && x > s->d[s->size - 1];
}

static void sbuf_insert(orc_sbuf* s, uint32_t id, float x) {
    size_t lo = first_not_below(s->d, s->size, x);
    size_t n = s->size < s->cap ? s->size : s->cap - 1; /* entries that stay */
    for (size_t i = n; i > lo; --i) {
        s->id[i] = s->id[i - 1];
        s->d[i] = s->d[i - 1];
        s->checked[i] = s->checked[i - 1];
    }
    if (lo < s->cap) {
        s->id[lo] = id;
        s->d[lo] = x;
        s->checked[lo] = 0;
    }
    if (s->size < s->cap) s->size++;
    if (lo < s->cur) s->cur = lo;
}

static uint32_t sbuf_pop(orc_sbuf* s) {
    uint32_t id = s->id[s->cur];
    s->checked[s->cur] = 1;
    s->cur++;
    while (s->cur < s->size && s->checked[s->cur]) s->cur++;
    return id;
}

/* ResultBuffer (utils/buffer.hpp:81-120): capacity k, ids/distances zero-   */
/* initialised, insert skipped only when full and dist > last.               */
typedef struct {
    uint32_t* id; /* k + 1 */
    float* d;     /* k + 1 */
    size_t size, cap;
} orc_rbuf;

static void rbuf_insert(orc_rbuf* r, uint32_t id, float x) {
    if (r->size == r->cap && x > r->d[r->size - 1]) return;
    size_t lo = first_not_below(r->d, r->size, x);
    for (size_t i = r->size; i > lo; --i) {
        r->id[i] = r->id[i - 1];
        r->d[i] = r->d[i - 1];
    }
    r->id[lo] = id;
    r->d[lo] = x;
    if (r->size < r->cap) r->size++;
}

/* Replays buffer operations for the golden check against the reference's   */
/* classes: ops are (kind, id, dist-bits) with kind 0 = insert unless full,  */
/* 1 = pop (0xFFFFFFFF when empty), 2 = ResultBuffer insert.                 */
int orc_buffers_replay(const uint32_t* ops, uint64_t nops, uint32_t cap, uint32_t* popped,
                       uint32_t* npopped, uint32_t* results) {
    orc_sbuf s = {(uint32_t*)calloc(cap + 1, 4), (float*)calloc(cap + 1, 4),
                  (uint8_t*)calloc(cap + 1, 1), 0, 0, cap};
    orc_rbuf r = {(uint32_t*)calloc(cap + 1, 4), (float*)calloc(cap + 1, 4), 0, cap};
    uint32_t np = 0;
    for (uint64_t i = 0; i < nops; ++i) {
        uint32_t kind = ops[3 * i], id = ops[3 * i + 1];
        float x;
        memcpy(&x, &ops[3 * i + 2], 4);
        if (kind == 0) {
            if (!sbuf_is_full(&s, x)) sbuf_insert(&s, id, x);
        } else if (kind == 1) {
            popped[np++] = s.cur < s.size ? sbuf_pop(&s) : 0xFFFFFFFFu;
        } else {
            rbuf_insert(&r, id, x);
        }
    }
    *npopped = np;
    memcpy(results, r.id, 4 * (size_t)cap);
    free(s.id); free(s.d); free(s.checked); free(r.id); free(r.d);
    return 0;
}

/* H2 index view: raw float rows, neighbour ids, sidecar codes + factors.   */
typedef struct {
    uint32_t n, dim, deg, p, ep;
    const float* raw;       /* n x dim */
    const uint32_t* nbr;    /* n x deg */
    const uint8_t* codes;   /* n x (deg * p / 8) */
    const float* factors;   /* n x (3 * deg): triple_x, fac_dq, fac_vq */
    const float* signs;     /* p */
    int layout;             /* codes layout: 0 reference blocks, 1 vector-major */
} orc_h2;

/* search_qg (qg.hpp:252-287) with scan_neighbors (qg.hpp:348-393),          */
/* update_results (qg.hpp:395-418) and copy_results (buffer.hpp:102):        */
/* results are the k nearest expanded nodes by exact distance; slots never  */
/* filled hold 0, as the zero-initialised ResultBuffer copies them.         */
static int h2_search_impl(const orc_h2* ix, const float* query, uint32_t ef, uint32_t k,
                          uint32_t* results, float* dists, uint64_t* stats, uint32_t* exp,
                          uint32_t exp_cap);

int orc_search_h2(const orc_h2* ix, const float* query, uint32_t ef, uint32_t k,
                  uint32_t* results, float* dists, uint64_t* stats) {
    return h2_search_impl(ix, query, ef, k, results, dists, stats, NULL, 0);
}

/* find_candidates (qg.hpp:439-473): the same search from the entry point,    */
/* and the nodes it expands, in expansion order (the builder keeps them as a */
/* node's candidates, dropping the node itself).  out: nq x cap ids, the     */
/* first cap expansions of each query, 0xFFFFFFFF past its last.             */
int orc_h2_expansions_batch(const orc_h2* ix, const float* queries, uint64_t nq, uint32_t ef,
                            uint32_t cap, uint32_t* out, int threads) {
    int err = 0;
#ifdef _OPENMP
    if (threads > 0) omp_set_num_threads(threads);
#pragma omp parallel for schedule(dynamic, 4) reduction(| : err)
#endif
    for (int64_t q = 0; q < (int64_t)nq; ++q) {
        uint32_t one;
        uint32_t* o = out + (size_t)q * cap;
        for (uint32_t i = 0; i < cap; ++i) o[i] = 0xFFFFFFFFu;
        err |= h2_search_impl(ix, queries + (size_t)q * ix->dim, ef, 1, &one, NULL, NULL, o,
                              cap) != 0;
    }
    (void)threads;
    return err ? -1 : 0;
}

static int h2_search_impl(const orc_h2* ix, const float* query, uint32_t ef, uint32_t k,
                          uint32_t* results, float* dists, uint64_t* stats, uint32_t* exp,
                          uint32_t exp_cap) {
    if (ef == 0 || k == 0) return -1;
    uint32_t p = ix->p, deg = ix->deg, dim = ix->dim;
    float* rq = (float*)malloc(4 * (size_t)p);
    uint8_t* qb = (uint8_t*)malloc(p);
    uint8_t* lut = (uint8_t*)malloc(4 * (size_t)p);
    float* est = (float*)malloc(4 * (size_t)deg);
    orc_sbuf s = {(uint32_t*)calloc(ef + 1, 4), (float*)calloc(ef + 1, 4),
                  (uint8_t*)calloc(ef + 1, 1), 0, 0, ef};
    orc_rbuf r = {(uint32_t*)calloc(k + 1, 4), (float*)calloc(k + 1, 4), 0, k};
    orc_bitset vis = {(uint64_t*)calloc(((size_t)ix->n + 63) / 64, 8), ix->n};
    if (!rq || !qb || !lut || !est || !s.id || !s.d || !s.checked || !r.id || !r.d || !vis.bits)
        return -2;
    float lo, width;
    int32_t sumq;
    orc_query_prepare(ix->signs, query, dim, p, rq, &lo, &width, &sumq, qb, lut);
    uint64_t E = 0, pops = 0;

    sbuf_insert(&s, ix->ep, 3.402823466e+38f);                /* FLT_MAX, qg.hpp:260 */
    while (s.cur < s.size) {
        uint32_t cur = sbuf_pop(&s);
        pops++;
        if (bs_get(&vis, cur)) continue;
        bs_set(&vis, cur);
        if (exp && E < exp_cap) exp[E] = cur;
        E++;
        const float* row = ix->raw + (size_t)cur * dim;
        float sqr_y = orc_l2_sqr(query, row, dim);
        orc_scan_estimates(ix->codes + (size_t)cur * (deg * p / 8), ix->layout,
                           ix->factors + (size_t)cur * 3 * deg, deg, p, lut, sqr_y, lo, width,
                           sumq, NULL, est);
        const uint32_t* nb = ix->nbr + (size_t)cur * deg;
        for (uint32_t i = 0; i < deg; ++i) {
            if (sbuf_is_full(&s, est[i]) || bs_get(&vis, nb[i])) continue;
            sbuf_insert(&s, nb[i], est[i]);
        }
        rbuf_insert(&r, cur, sqr_y);
    }
    if (r.size < r.cap) {                                    /* update_results */
        uint32_t* ids = (uint32_t*)malloc(4 * ((size_t)k + 1));
        memcpy(ids, r.id, 4 * ((size_t)k + 1));
        for (uint32_t t = 0; t <= k; ++t) {
            const uint32_t* nb = ix->nbr + (size_t)ids[t] * deg;
            for (uint32_t i = 0; i < deg; ++i) {
                uint32_t x = nb[i];
                if (bs_get(&vis, x)) continue;
                bs_set(&vis, x);
                rbuf_insert(&r, x, orc_l2_sqr(query, ix->raw + (size_t)x * dim, dim));
            }
            if (r.size == r.cap) break;
        }
        free(ids);
    }
    for (uint32_t i = 0; i < k; ++i) {
        results[i] = r.id[i];
        if (dists) dists[i] = i < r.size ? r.d[i] : INFINITY;
    }
    if (stats) {
        stats[0] = E;
        stats[1] = pops;
    }
    free(rq); free(qb); free(lut); free(est);
    free(s.id); free(s.d); free(s.checked); free(r.id); free(r.d); free(vis.bits);
    return 0;
}

int orc_search_h2_batch(const orc_h2* ix, const float* queries, uint64_t nq, uint32_t ef,
                        uint32_t k, uint32_t* results, float* dists, uint64_t* stats,
                        int threads) {
    int err = 0;
#ifdef _OPENMP
    if (threads > 0) omp_set_num_threads(threads);
#pragma omp parallel for schedule(dynamic, 4) reduction(| : err)
#endif
    for (int64_t q = 0; q < (int64_t)nq; ++q) {
        int rc = orc_search_h2(ix, queries + (size_t)q * ix->dim, ef, k, results + (size_t)q * k,
                               dists ? dists + (size_t)q * k : NULL,
                               stats ? stats + 2 * (size_t)q : NULL);
        err |= (rc != 0);
    }
    (void)threads;
    return err ? -1 : 0;
}

/* Sidecar of one node (this build's definition, restating rabitq_codes /    */
/* rabitq_factors, quantization/rabitq.hpp:23-108): residuals               */
/* o_j = R x_nbr_j - R x_node (each rotated in float by orc_fht_rotate),     */
/* bits o > 0, codes vector-major (byte b: low nibble dims 8b..8b+3 with    */
/* bit 3 = dim 8b, high nibble dims 8b+4..8b+7), and the factors with every */
/* sum taken sequentially over the dims in double:                          */
/*   fn = float(1/sqrt(P));  nrm = float(sqrt(sum o^2));                      */
/*   x0 = float(float(sum |o| * fn) / nrm);  x1 = float(sum c*s * fn)        */
/*   (s = +-1 from the bits, c = R x_node);                                  */
/*   x_x0 = nrm / x0 (double);                                               */
/*   triple_x = float(nrm^2 + 2 x_x0 x1);  fac_dq = float(-2 x_x0 fn);       */
/*   fac_vq = float(-2 x_x0 fn (2 popcount - P)).                            */
/* An all-zero residual (a neighbour equal to the node) has no direction;   */
/* its factors are set to (0, 0, 0) so the estimate equals the node's exact */
/* distance (the reference divides by zero there and gets NaN).             */
void orc_sidecar_node(const orc_h2* ix, uint32_t node, uint8_t* codes, float* factors) {
    uint32_t p = ix->p, dim = ix->dim, deg = ix->deg;
    float* c = (float*)malloc(4 * (size_t)p);
    float* x = (float*)malloc(4 * (size_t)p);
    orc_fht_rotate(ix->signs, ix->raw + (size_t)node * dim, dim, p, c);
    double fn = (double)(1.0f / sqrtf((float)p));
    for (uint32_t j = 0; j < deg; ++j) {
        uint32_t v = ix->nbr[(size_t)node * deg + j];
        orc_fht_rotate(ix->signs, ix->raw + (size_t)v * dim, dim, p, x);
        uint8_t* code = codes + (size_t)j * (p / 8);
        memset(code, 0, p / 8);
        double s2 = 0.0, sabs = 0.0, sc = 0.0;
        uint32_t pop = 0;
        for (uint32_t i = 0; i < p; ++i) {
            float o = x[i] - c[i];
            int bit = o > 0.0f;
            if (bit) {
                uint32_t b = i >> 3, w = i & 7u;
                uint32_t nib_shift = w < 4 ? 0 : 4;
                code[b] |= (uint8_t)(1u << (nib_shift + 3 - (w & 3u)));
                pop++;
            }
            s2 += (double)o * (double)o;
            sabs += bit ? (double)o : -(double)o;
            sc += bit ? (double)c[i] : -(double)c[i];
        }
        float nrm = (float)sqrt(s2);
        float x0 = (float)((float)(sabs * fn) / nrm);
        float x1 = (float)(sc * fn);
        float* tx = factors;
        float* dq = factors + deg;
        float* vq = factors + 2 * deg;
        if (s2 == 0.0) {
            tx[j] = 0.0f;
            dq[j] = 0.0f;
            vq[j] = 0.0f;
            continue;
        }
        double x_x0 = (double)nrm / (double)x0;
        tx[j] = (float)((double)nrm * (double)nrm + 2.0 * x_x0 * (double)x1);
        dq[j] = (float)(-2.0 * x_x0 * fn);
        vq[j] = (float)(-2.0 * x_x0 * fn * (double)(2 * (int)pop - (int)p));
    }
    free(c);
    free(x);
}

int orc_sidecar_build(const orc_h2* ix, uint8_t* codes, float* factors, int threads) {
#ifdef _OPENMP
    if (threads > 0) omp_set_num_threads(threads);
#pragma omp parallel for schedule(dynamic, 64)
#endif
    for (int64_t u = 0; u < (int64_t)ix->n; ++u)
        orc_sidecar_node(ix, (uint32_t)u, codes + (size_t)u * (ix->deg * ix->p / 8),
                         factors + (size_t)u * 3 * ix->deg);
    (void)threads;
    return 0;
}

/* Converts codes of one node from the reference's block layout to the       */
/* vector-major layout (same bits).                                          */
void orc_codes_to_vector_major(const uint8_t* ref, uint32_t deg, uint32_t p, uint8_t* vm) {
    memset(vm, 0, (size_t)deg * (p / 8));
    for (uint32_t v = 0; v < deg; ++v)
        for (uint32_t c = 0; c < p / 4; ++c) {
            uint32_t nib = ref_nibble(ref, p, v, c);
            vm[(size_t)v * (p / 8) + c / 2] |= (uint8_t)((c & 1u) ? nib << 4 : nib);
        }
}
