// build_graph.hip -- device kernels of the graph builder behind
// Index.build_index (python/bindings.cpp:69-90).  The algorithm follows the
// reference's QGBuilder (qg/qg_builder.hpp): random initial graph
// (:263-291), then per iteration a beam search from the entry point for every
// node (:202-231, the search kernel in h1_search.hip), heuristic_prune of the
// candidates (:155-200), reverse edges (:233-261) and, in the last iteration,
// graph_refine's relaxed-angle fill (:113-153, :293-348).  Differences from the
// reference, all on purpose:
//   * distances are those the search itself uses: the node's float vector
//     against SQ8-decoded vectors, and decoded against decoded inside the
//     prune (the reference orders its build-time beam with a FastScan
//     estimator that reads codes LiteQG's rows do not hold, qg.hpp:348-369);
//   * reverse edges are gathered for all nodes and pruned once per node
//     instead of being inserted one at a time under per-node locks;
//   * random draws come from a seeded counter hash, and reverse edges are
//     ordered by a stable sort, so builds are repeatable: the same data and
//     parameters give the same graph, row for row.  (With atomics ordering
//     them, a node with more reverse edges than its candidate buffer holds
//     would keep whichever arrived first.)
//
// One wavefront processes one node.  The sequential prune rules (keep a
// candidate unless an already kept neighbour dominates it) are evaluated 64
// candidates at a time: each lane tests its candidate against the current
// selection, then the few survivors are admitted in order and the later
// survivors re-tested against each new member.  This yields exactly the
// sequential outcome.
#include <math.h>

#include <hipcub/hipcub.hpp>

#include "lqg_build.h"
#include "lqg_internal.h"

namespace lqg {
namespace {

constexpr uint32_t kCandMax = 1024;     // candidates held in LDS per node
constexpr uint32_t kDegMax = 64;        // degree bound handled by one wavefront
constexpr uint32_t kMaxCandPool = 750;  // max_candidate_pool_size_, qg_builder.hpp:30

__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ __forceinline__ uint64_t bkey(float d, uint32_t id) {
    return (static_cast<uint64_t>(__float_as_uint(d)) << 32) | id;
}
__device__ __forceinline__ uint32_t bkey_id(uint64_t k) { return static_cast<uint32_t>(k); }
__device__ __forceinline__ float bkey_d(uint64_t k) {
    return __uint_as_float(static_cast<uint32_t>(k >> 32));
}

__device__ __forceinline__ uint64_t mix64(uint64_t x) {  // splitmix64 finaliser
    x += 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
}

__device__ __forceinline__ uint32_t lanes_below(uint64_t mask) {
    return __builtin_amdgcn_mbcnt_hi(static_cast<uint32_t>(mask >> 32),
                                     __builtin_amdgcn_mbcnt_lo(static_cast<uint32_t>(mask), 0u));
}

// SQ8 rows and their decoding parameters (parameters in LDS).
struct Codes {
    const uint8_t* q;
    const float* mn;
    const float* sc;
    uint32_t dim;
};

__device__ __forceinline__ float dec(uint32_t byte, const float* mn, const float* sc, uint32_t j) {
    return fmaf(static_cast<float>(byte), sc[j], mn[j]);
}

__device__ __forceinline__ float reduce16(float* s) {
#pragma unroll
    for (int j = 0; j < 8; ++j) s[j] += s[j + 8];
#pragma unroll
    for (int j = 0; j < 4; ++j) s[j] += s[j + 4];
    return (s[0] + s[2]) + (s[1] + s[3]);
}

// ||x - dec(row)||^2, x a float vector in LDS, in l2_sqr's order (l2.hpp:16-34).
__device__ float l2_vec_codes(const float* x, const uint8_t* row, const Codes& C) {
    float s[16];
#pragma unroll
    for (int j = 0; j < 16; ++j) s[j] = 0.f;
    const uint32_t mul16 = C.dim & ~15u;
    uint32_t i = 0;
    if ((C.dim & 15u) == 0) {
        for (; i < mul16; i += 16) {
            const uint4 w = *reinterpret_cast<const uint4*>(row + i);
            const uint32_t ww[4] = {w.x, w.y, w.z, w.w};
#pragma unroll
            for (int v = 0; v < 4; ++v)
#pragma unroll
                for (int b = 0; b < 4; ++b) {
                    const uint32_t j = i + 4 * v + b;
                    const float t = x[j] - dec((ww[v] >> (8 * b)) & 0xFFu, C.mn, C.sc, j);
                    s[4 * v + b] = fmaf(t, t, s[4 * v + b]);
                }
        }
    } else {
        for (; i < mul16; i += 16)
#pragma unroll
            for (int j = 0; j < 16; ++j) {
                const float t = x[i + j] - dec(row[i + j], C.mn, C.sc, i + j);
                s[j] = fmaf(t, t, s[j]);
            }
    }
    float r = reduce16(s);
    for (; i < C.dim; ++i) {
        const float t = x[i] - dec(row[i], C.mn, C.sc, i);
        r = fmaf(t, t, r);
    }
    return r;
}

// ||dec(a) - dec(b)||^2 in l2_sqr's order.
__device__ float l2_codes(const uint8_t* a, const uint8_t* b, const Codes& C) {
    float s[16];
#pragma unroll
    for (int j = 0; j < 16; ++j) s[j] = 0.f;
    const uint32_t mul16 = C.dim & ~15u;
    uint32_t i = 0;
    if ((C.dim & 15u) == 0) {
        for (; i < mul16; i += 16) {
            const uint4 wa = *reinterpret_cast<const uint4*>(a + i);
            const uint4 wb = *reinterpret_cast<const uint4*>(b + i);
            const uint32_t xa[4] = {wa.x, wa.y, wa.z, wa.w};
            const uint32_t xb[4] = {wb.x, wb.y, wb.z, wb.w};
#pragma unroll
            for (int v = 0; v < 4; ++v)
#pragma unroll
                for (int bb = 0; bb < 4; ++bb) {
                    const uint32_t j = i + 4 * v + bb;
                    const float t = dec((xa[v] >> (8 * bb)) & 0xFFu, C.mn, C.sc, j) -
                                    dec((xb[v] >> (8 * bb)) & 0xFFu, C.mn, C.sc, j);
                    s[4 * v + bb] = fmaf(t, t, s[4 * v + bb]);
                }
        }
    } else {
        for (; i < mul16; i += 16)
#pragma unroll
            for (int j = 0; j < 16; ++j) {
                const float t = dec(a[i + j], C.mn, C.sc, i + j) - dec(b[i + j], C.mn, C.sc, i + j);
                s[j] = fmaf(t, t, s[j]);
            }
    }
    float r = reduce16(s);
    for (; i < C.dim; ++i) {
        const float t = dec(a[i], C.mn, C.sc, i) - dec(b[i], C.mn, C.sc, i);
        r = fmaf(t, t, r);
    }
    return r;
}

// Ascending bitonic sort of a[0..n) (n a power of two) by one wavefront.
__device__ void wave_sort(uint64_t* a, uint32_t n, uint32_t lane) {
    for (uint32_t k = 2; k <= n; k <<= 1) {
        for (uint32_t j = k >> 1; j > 0; j >>= 1) {
            for (uint32_t i = lane; i < n; i += kWave) {
                const uint32_t p = i ^ j;
                if (p > i) {
                    const uint64_t x = a[i], y = a[p];
                    const bool up = (i & k) == 0;
                    if ((x > y) == up) {
                        a[i] = y;
                        a[p] = x;
                    }
                }
            }
            wave_sync();
        }
    }
}

// Sorts keys[0..n) (n <= kCandMax), removes repeated ids (equal keys are
// adjacent) and the node itself, returns the new count.
__device__ uint32_t sort_unique(uint64_t* keys, uint32_t n, uint32_t self, uint32_t lane) {
    uint32_t p2 = 1;
    while (p2 < n) p2 <<= 1;
    for (uint32_t i = n + lane; i < p2; i += kWave) keys[i] = ~0ull;
    wave_sync();
    wave_sort(keys, p2, lane);
    uint32_t out = 0;
    for (uint32_t base = 0; base < n; base += kWave) {
        const uint32_t i = base + lane;
        uint64_t k = 0;
        bool keep = false;
        if (i < n) {
            k = keys[i];
            keep = bkey_id(k) != self && (i == 0 || keys[i - 1] != k);
        }
        const uint64_t m = __ballot(keep);
        wave_sync();
        if (keep) keys[out + lanes_below(m)] = k;
        out += static_cast<uint32_t>(__popcll(m));
        wave_sync();
    }
    return out;
}

// Does selected neighbour r (distance dr to the node) rule out candidate c
// (distance dc)?  mode 0: heuristic_prune's test d(r,c) < d(node,c)
// (qg_builder.hpp:187-190); mode 1: add_pruned_edges' test at cosine bound
// theta, applied to neighbours no farther than c (qg_builder.hpp:128-143).
__device__ __forceinline__ bool rules_out(const Codes& C, const uint8_t* crow, uint32_t c,
                                          float dc, uint32_t r, float dr, int mode, float theta) {
    if (mode == 0) return l2_codes(C.q + static_cast<size_t>(r) * C.dim, crow, C) < dc;
    if (r == c) return true;
    if (dr > dc) return false;
    const float djk = l2_codes(C.q + static_cast<size_t>(r) * C.dim, crow, C);
    const float cosine = (dc + dr - djk) / (2.0f * sqrtf(dr * dc));
    return cosine > theta;
}

// Walks the sorted candidates cand[0..nc) and appends to the selection
// (sel_ids/sel_d, ns entries) every candidate no selected neighbour rules out,
// until the selection holds deg.  Candidates ruled out go to pr (when given).
__device__ uint32_t select_neighbors(const Codes& C, const uint64_t* cand, uint32_t nc,
                                     uint32_t deg, uint32_t* sel_ids, float* sel_d, uint32_t ns,
                                     int mode, float theta, uint32_t* pr_ids, float* pr_d,
                                     uint32_t pr_cap, uint32_t* np_io, uint32_t lane) {
    uint32_t np = np_io ? *np_io : 0;
    for (uint32_t base = 0; base < nc && ns < deg; base += kWave) {
        const uint32_t j = base + lane;
        const bool valid = j < nc;
        uint32_t c = 0;
        float dc = 0.f;
        if (valid) {
            c = bkey_id(cand[j]);
            dc = bkey_d(cand[j]);
        }
        const uint8_t* crow = C.q + static_cast<size_t>(c) * C.dim;
        bool out = !valid;
        for (uint32_t r = 0; r < ns; ++r) {
            const uint32_t rid = sel_ids[r];
            const float rd = sel_d[r];
            if (!out) out = rules_out(C, crow, c, dc, rid, rd, mode, theta);
        }
        uint64_t surv = __ballot(!out);
        uint64_t ruled = __ballot(valid && out);
        while (surv != 0 && ns < deg) {
            const int s = __builtin_ctzll(surv);
            const uint32_t sid = __shfl(c, s);
            const float sd = __shfl(dc, s);
            if (lane == 0) {
                sel_ids[ns] = sid;
                sel_d[ns] = sd;
            }
            ++ns;
            surv &= surv - 1;
            bool o2 = false;
            if ((surv >> lane) & 1ull) o2 = rules_out(C, crow, c, dc, sid, sd, mode, theta);
            const uint64_t m2 = __ballot(o2);
            surv &= ~m2;
            ruled |= m2;
            wave_sync();
        }
        if (pr_ids && np < pr_cap) {
            const bool mine = (ruled >> lane) & 1ull;
            const uint32_t pos = np + lanes_below(ruled);
            if (mine && pos < pr_cap) {
                pr_ids[pos] = c;
                pr_d[pos] = dc;
            }
            np = min(pr_cap, np + static_cast<uint32_t>(__popcll(ruled)));
        }
        wave_sync();
    }
    if (np_io && lane == 0) *np_io = np;
    return ns;
}

struct BuildArgs {
    const float* raw;          // N x dim
    const uint8_t* qdata;      // N x dim
    const float* sq8_min;      // dim
    const float* sq8_scale;    // dim
    uint32_t n, dim, deg;
    uint32_t* graph;           // N x deg (read in phase A, written in phase B)
    const uint32_t* cand_ids;  // N x ncand, search output (nearest first)
    uint32_t ncand;
    uint32_t* kept_ids;        // N x deg
    float* kept_d;
    uint32_t* kept_n;          // N
    uint32_t* pr_ids;          // N x pr_cap
    float* pr_d;
    uint32_t* pr_n;            // N
    uint32_t pr_cap;
    const uint32_t* in_off;    // N + 1
    uint32_t* in_ids;          // N x deg total
    uint32_t* counter;
    uint64_t seed;
    int refine;
};

struct NodeLds {
    float* qv;        // dim
    uint64_t* cand;   // kCandMax
    uint32_t* k_ids;  // kDegMax
    float* k_d;       // kDegMax
    uint32_t* t_ids;  // kDegMax (fill scratch)
    float* t_d;       // kDegMax
};

__device__ NodeLds carve(unsigned char* lds, uint32_t dim_pad, uint32_t wave_bytes, uint32_t wid) {
    unsigned char* wb = lds + 8u * dim_pad + static_cast<size_t>(wid) * wave_bytes;
    NodeLds L;
    L.qv = reinterpret_cast<float*>(wb);
    L.cand = reinterpret_cast<uint64_t*>(wb + 4u * dim_pad);
    L.k_ids = reinterpret_cast<uint32_t*>(L.cand + kCandMax);
    L.k_d = reinterpret_cast<float*>(L.k_ids + kDegMax);
    L.t_ids = reinterpret_cast<uint32_t*>(L.k_d + kDegMax);
    L.t_d = reinterpret_cast<float*>(L.t_ids + kDegMax);
    return L;
}

__device__ Codes load_params(const BuildArgs& a, unsigned char* lds, uint32_t dim_pad) {
    float* mn = reinterpret_cast<float*>(lds);
    float* sc = mn + dim_pad;
    for (uint32_t i = threadIdx.x; i < a.dim; i += blockDim.x) {
        mn[i] = a.sq8_min[i];
        sc[i] = a.sq8_scale[i];
    }
    __syncthreads();
    Codes C;
    C.q = a.qdata;
    C.mn = mn;
    C.sc = sc;
    C.dim = a.dim;
    return C;
}

__device__ __forceinline__ uint32_t next_node(uint32_t* counter, uint32_t lane) {
    uint32_t v = 0;
    if (lane == 0) v = atomicAdd(counter, 1u);
    return __shfl(v, 0);
}

// Appends ids src[0..cnt) that are < n to L.cand (as bare ids), returns the count.
__device__ uint32_t gather_ids(uint64_t* cand, uint32_t nc, const uint32_t* src, uint32_t cnt,
                               uint32_t n, uint32_t lane) {
    for (uint32_t b = 0; b < cnt; b += kWave) {
        const uint32_t j = b + lane;
        const uint32_t c = j < cnt ? src[j] : kInvalid;
        const bool ok = c < n;
        const uint64_t m = __ballot(ok);
        if (ok) cand[nc + lanes_below(m)] = c;
        nc += static_cast<uint32_t>(__popcll(m));
    }
    return nc;
}

// Replaces the bare ids in cand[0..nc) by (distance to the node, id) keys.
__device__ void key_candidates(const Codes& C, const float* qv, uint64_t* cand, uint32_t nc,
                               uint32_t lane) {
    wave_sync();
    for (uint32_t j = lane; j < nc; j += kWave) {
        const uint32_t c = static_cast<uint32_t>(cand[j]);
        cand[j] = bkey(l2_vec_codes(qv, C.q + static_cast<size_t>(c) * C.dim, C), c);
    }
    wave_sync();
}

// Phase A: candidates of node i = its beam-search result plus its current
// neighbours (qg_builder.hpp:211-219), nearest 750 kept (:221-227),
// heuristic_prune (:229).
__global__ void __launch_bounds__(256) prune_search_candidates(BuildArgs a, uint32_t dim_pad,
                                                               uint32_t wave_bytes) {
    extern __shared__ __align__(16) unsigned char lds[];
    const uint32_t lane = threadIdx.x & (kWave - 1);
    const Codes C = load_params(a, lds, dim_pad);
    NodeLds L = carve(lds, dim_pad, wave_bytes, threadIdx.x / kWave);
    for (;;) {
        const uint32_t i = next_node(a.counter, lane);
        if (i >= a.n) break;
        const float* xi = a.raw + static_cast<size_t>(i) * a.dim;
        for (uint32_t t = lane; t < a.dim; t += kWave) L.qv[t] = xi[t];
        uint32_t nc = gather_ids(L.cand, 0, a.cand_ids + static_cast<size_t>(i) * a.ncand,
                                 min(a.ncand, kCandMax - a.deg), a.n, lane);
        nc = gather_ids(L.cand, nc, a.graph + static_cast<size_t>(i) * a.deg, a.deg, a.n, lane);
        key_candidates(C, L.qv, L.cand, nc, lane);
        nc = min(sort_unique(L.cand, nc, i, lane), kMaxCandPool);
        uint32_t nk;
        uint32_t* pr = a.pr_ids + static_cast<size_t>(i) * a.pr_cap;
        float* prd = a.pr_d + static_cast<size_t>(i) * a.pr_cap;
        if (nc <= a.deg) {  // a pool no larger than the bound is kept whole (:164-168)
            for (uint32_t j = lane; j < nc; j += kWave) {
                L.k_ids[j] = bkey_id(L.cand[j]);
                L.k_d[j] = bkey_d(L.cand[j]);
            }
            if (lane == 0) a.pr_n[i] = 0;
            nk = nc;
        } else {
            uint32_t np = 0;
            nk = select_neighbors(C, L.cand, nc, a.deg, L.k_ids, L.k_d, 0, 0, 0.f, pr, prd,
                                  a.pr_cap, &np, lane);
            if (lane == 0) a.pr_n[i] = np;
        }
        wave_sync();
        for (uint32_t j = lane; j < nk; j += kWave) {
            a.kept_ids[static_cast<size_t>(i) * a.deg + j] = L.k_ids[j];
            a.kept_d[static_cast<size_t>(i) * a.deg + j] = L.k_d[j];
        }
        if (lane == 0) a.kept_n[i] = nk;
        wave_sync();
    }
}

// Reverse edges, pass 1: in-degree of every node.
__global__ void count_reverse(BuildArgs a, uint32_t* in_cnt) {
    const size_t e = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (e >= static_cast<size_t>(a.n) * a.deg) return;
    const uint32_t i = static_cast<uint32_t>(e / a.deg), s = static_cast<uint32_t>(e % a.deg);
    if (s >= a.kept_n[i]) return;
    atomicAdd(&in_cnt[a.kept_ids[e]], 1u);
}

// Reverse edges, pass 2: one (target, source) pair per edge slot, the key N
// for empty slots.  Slots are in source order, so a stable sort by target
// gives every CSR row its sources in ascending order, the N keys last.
__global__ void emit_reverse(BuildArgs a, uint32_t* keys, uint32_t* vals) {
    const size_t e = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (e >= static_cast<size_t>(a.n) * a.deg) return;
    const uint32_t i = static_cast<uint32_t>(e / a.deg), s = static_cast<uint32_t>(e % a.deg);
    keys[e] = s < a.kept_n[i] ? a.kept_ids[e] : a.n;
    vals[e] = i;
}

// Key bits for targets 0..n (n = the empty-slot key).
uint32_t reverse_key_bits(uint32_t n) {
    uint32_t bits = 1;
    while (bits < 32 && (n >> bits) != 0) ++bits;
    return bits;
}

// Phase B: own kept list plus reverse edges, pruned again (add_reverse_edges,
// qg_builder.hpp:233-261); in the last iteration graph_refine (:293-348)
// fills short lists from the pruned candidates by a bisection on the cosine
// bound, then with random nodes.  Writes the node's graph row.
__global__ void __launch_bounds__(256) merge_reverse(BuildArgs a, uint32_t dim_pad,
                                                     uint32_t wave_bytes) {
    extern __shared__ __align__(16) unsigned char lds[];
    const uint32_t lane = threadIdx.x & (kWave - 1);
    const Codes C = load_params(a, lds, dim_pad);
    NodeLds L = carve(lds, dim_pad, wave_bytes, threadIdx.x / kWave);
    for (;;) {
        const uint32_t i = next_node(a.counter, lane);
        if (i >= a.n) break;
        const float* xi = a.raw + static_cast<size_t>(i) * a.dim;
        for (uint32_t t = lane; t < a.dim; t += kWave) L.qv[t] = xi[t];
        uint32_t nc = gather_ids(L.cand, 0, a.kept_ids + static_cast<size_t>(i) * a.deg,
                                 a.kept_n[i], a.n, lane);
        const uint32_t b0 = a.in_off[i], b1 = a.in_off[i + 1];
        nc = gather_ids(L.cand, nc, a.in_ids + b0, min(b1 - b0, kCandMax - nc), a.n, lane);
        key_candidates(C, L.qv, L.cand, nc, lane);
        nc = sort_unique(L.cand, nc, i, lane);
        uint32_t nk;
        if (nc <= a.deg) {
            for (uint32_t j = lane; j < nc; j += kWave) {
                L.k_ids[j] = bkey_id(L.cand[j]);
                L.k_d[j] = bkey_d(L.cand[j]);
            }
            nk = nc;
        } else {
            nk = select_neighbors(C, L.cand, nc, a.deg, L.k_ids, L.k_d, 0, 0, 0.f, nullptr,
                                  nullptr, 0, nullptr, lane);
        }
        wave_sync();
        uint32_t* out = a.graph + static_cast<size_t>(i) * a.deg;
        if (nk >= a.deg) {
            for (uint32_t j = lane; j < a.deg; j += kWave) out[j] = L.k_ids[j];
            wave_sync();
            continue;
        }
        // pruned candidates of phase A (sorted by distance) as keys
        const uint32_t np = a.pr_n[i];
        const uint32_t* pids = a.pr_ids + static_cast<size_t>(i) * a.pr_cap;
        const float* pd = a.pr_d + static_cast<size_t>(i) * a.pr_cap;
        for (uint32_t j = lane; j < np; j += kWave) L.cand[j] = bkey(pd[j], pids[j]);
        wave_sync();
        uint32_t nt;
        if (a.refine) {
            float left = 0.5f, right = 1.0f;  // bound of cosine (:308-322)
            for (int it = 0; it < 6; ++it) {
                const float theta = it < 5 ? (left + right) / 2 : right;
                for (uint32_t j = lane; j < nk; j += kWave) {
                    L.t_ids[j] = L.k_ids[j];
                    L.t_d[j] = L.k_d[j];
                }
                wave_sync();
                nt = select_neighbors(C, L.cand, np, a.deg, L.t_ids, L.t_d, nk, 1, theta, nullptr,
                                      nullptr, 0, nullptr, lane);
                if (it < 5) {
                    if (nt < a.deg) left = theta;
                    else right = theta;
                }
            }
        } else {
            // not the last iteration: nearest pruned candidates fill the row
            for (uint32_t j = lane; j < nk; j += kWave) L.t_ids[j] = L.k_ids[j];
            wave_sync();
            nt = nk;
            if (lane == 0) {
                for (uint32_t s = 0; s < np && nt < a.deg; ++s) {
                    const uint32_t c = bkey_id(L.cand[s]);
                    bool dup = false;
                    for (uint32_t r = 0; r < nt; ++r) dup |= L.t_ids[r] == c;
                    if (!dup) L.t_ids[nt++] = c;
                }
            }
            nt = __shfl(nt, 0);
        }
        // random fill (:326-342), seeded
        if (lane == 0) {
            uint64_t ctr = 0;
            while (nt < a.deg) {
                const uint32_t r = static_cast<uint32_t>(
                    mix64(a.seed ^ (static_cast<uint64_t>(i) << 20) ^ ctr++) % a.n);
                bool dup = r == i;
                for (uint32_t s = 0; s < nt && !dup; ++s) dup = L.t_ids[s] == r;
                if (!dup) L.t_ids[nt++] = r;
            }
        }
        wave_sync();
        for (uint32_t j = lane; j < a.deg; j += kWave) out[j] = L.t_ids[j];
        wave_sync();
    }
}

void node_geometry(uint32_t dim, uint32_t& dim_pad, uint32_t& wave_bytes) {
    dim_pad = (dim + 3u) & ~3u;
    size_t wb = 4u * dim_pad + 8u * kCandMax + 16u * kDegMax;
    wave_bytes = static_cast<uint32_t>((wb + 15u) & ~static_cast<size_t>(15u));
}

hipError_t launch_node_kernel(void (*kern)(BuildArgs, uint32_t, uint32_t), const BuildArgs& a,
                              hipStream_t s) {
    uint32_t dim_pad, wave_bytes;
    node_geometry(a.dim, dim_pad, wave_bytes);
    const int waves = 4;
    const size_t lds = 8u * dim_pad + static_cast<size_t>(waves) * wave_bytes;
    hipError_t e;
    if (lds > 64 * 1024) {
        e = hipFuncSetAttribute(reinterpret_cast<const void*>(kern),
                                hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(lds));
        if (e != hipSuccess) return e;
    }
    int dev = 0, cus = 0, per = 1;
    if ((e = hipGetDevice(&dev)) != hipSuccess) return e;
    if ((e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev)) != hipSuccess)
        return e;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, reinterpret_cast<const void*>(kern),
                                                     kWave * waves, lds) != hipSuccess || per < 1)
        per = 1;
    int groups = per * cus;
    const int need = static_cast<int>((a.n + waves - 1) / waves);
    if (groups > need) groups = need;
    hipLaunchKernelGGL(kern, dim3(groups), dim3(kWave * waves), lds, s, a, dim_pad, wave_bytes);
    return hipGetLastError();
}

}  // namespace

size_t scan_temp_bytes(uint32_t n) {
    size_t bytes = 0;
    if (hipcub::DeviceScan::ExclusiveSum(nullptr, bytes, static_cast<const uint32_t*>(nullptr),
                                         static_cast<uint32_t*>(nullptr),
                                         static_cast<int>(n)) != hipSuccess)
        return 0;
    return bytes;
}

size_t build_temp_bytes(uint32_t n, uint32_t deg) {
    size_t sort_bytes = 0;
    const int edges = static_cast<int>(static_cast<size_t>(n) * deg);
    if (hipcub::DeviceRadixSort::SortPairs(nullptr, sort_bytes, static_cast<const uint32_t*>(nullptr),
                                           static_cast<uint32_t*>(nullptr),
                                           static_cast<const uint32_t*>(nullptr),
                                           static_cast<uint32_t*>(nullptr), edges, 0,
                                           static_cast<int>(reverse_key_bits(n))) != hipSuccess)
        return 0;
    const size_t scan_bytes = scan_temp_bytes(n + 1);
    if (scan_bytes == 0) return 0;
    return sort_bytes > scan_bytes ? sort_bytes : scan_bytes;
}

hipError_t exclusive_scan_u32(const uint32_t* in, uint32_t* out, uint32_t n, void* tmp,
                              size_t tmp_bytes, hipStream_t s) {
    return hipcub::DeviceScan::ExclusiveSum(tmp, tmp_bytes, in, out, static_cast<int>(n), s);
}

hipError_t build_run_iteration(const BuildBuffers& b, const BuildParams& p, hipStream_t s) {
    BuildArgs a;
    a.raw = b.raw;
    a.qdata = b.qdata;
    a.sq8_min = b.sq8_min;
    a.sq8_scale = b.sq8_scale;
    a.n = p.n;
    a.dim = p.dim;
    a.deg = p.deg;
    a.graph = b.graph;
    a.cand_ids = b.cand_ids;
    a.ncand = p.ncand;
    a.kept_ids = b.kept_ids;
    a.kept_d = b.kept_d;
    a.kept_n = b.kept_n;
    a.pr_ids = b.pr_ids;
    a.pr_d = b.pr_d;
    a.pr_n = b.pr_n;
    a.pr_cap = p.pr_cap;
    a.in_off = b.in_off;
    a.in_ids = b.in_ids;
    a.counter = b.counter;
    a.seed = p.seed;
    a.refine = p.refine;
    hipError_t e;
    if ((e = launch_reset_counter(b.counter, s)) != hipSuccess) return e;
    if ((e = launch_node_kernel(prune_search_candidates, a, s)) != hipSuccess) return e;
    if ((e = hipMemsetAsync(b.in_cnt, 0, sizeof(uint32_t) * (p.n + 1), s)) != hipSuccess) return e;
    const size_t edges = static_cast<size_t>(p.n) * p.deg;
    const int blk = 256;
    const int grid = static_cast<int>((edges + blk - 1) / blk);
    hipLaunchKernelGGL(count_reverse, dim3(grid), dim3(blk), 0, s, a, b.in_cnt);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    if ((e = exclusive_scan_u32(b.in_cnt, b.in_off, p.n + 1, b.scan_tmp, b.scan_tmp_bytes, s)) !=
        hipSuccess)
        return e;
    hipLaunchKernelGGL(emit_reverse, dim3(grid), dim3(blk), 0, s, a, b.rev_keys, b.rev_vals);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    size_t sort_bytes = b.scan_tmp_bytes;
    if ((e = hipcub::DeviceRadixSort::SortPairs(b.scan_tmp, sort_bytes, b.rev_keys, b.rev_keys_out,
                                                b.rev_vals, b.in_ids, static_cast<int>(edges), 0,
                                                static_cast<int>(reverse_key_bits(p.n)), s)) !=
        hipSuccess)
        return e;
    if ((e = launch_reset_counter(b.counter, s)) != hipSuccess) return e;
    return launch_node_kernel(merge_reverse, a, s);
}

}  // namespace lqg
