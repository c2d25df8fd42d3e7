// h1_search.hip -- gfx950 kernel for the reference's live query path,
// QuantizedGraph::search_qg_with_heap (qg.hpp:298-344): SQ8-decoded exact L2
// distances (quantization/sq8.hpp:28-32, space/l2.hpp:16-34) driving a beam
// of width ef over the neighbour table.
//
// One wavefront walks one query.  Keys are (distance, id) packed into 64 bits
//   key = float_bits(d) << 32 | id << 1 | expanded
// so that comparing keys orders by distance and then id (distances are squared
// L2 >= +0, whose float bits are monotone).  Under this total order the
// reference's two heaps reduce to one sorted pool of capacity ef whose
// smallest unexpanded entry is expanded until none is left (an entry pushed
// out of res_pool always compares above the current lowerBound, and the
// reference stops at the first such top, qg.hpp:316-317).
//
// The pool is held in registers, lane-major: logical entry i = s*64 + lane is
// P[s] of that lane (EF_CAP/64 slots per lane), unused entries ~0.  Candidates
// of one expansion are admitted in one uniform pass that computes, with
// readlane broadcasts and ballots, each candidate's position, duplicate test
// and rank, and each pool entry's shift; one scatter/gather through LDS then
// applies the permutation.
//
// Visited set: a direct-mapped cache of ids in LDS (lossy).  A cache hit is a
// node evaluated before, so skipping it is what the reference does.  A miss
// may still be a node evaluated before; it is evaluated again and then
// dropped by the exact test against the pool: a node evaluated earlier is
// either still in the pool (same key -> found) or compares above the pool's
// current maximum (it was rejected or pushed out when the maximum was larger,
// and the maximum never grows).  The result is therefore exactly that of an
// exact visited set; only the number of evaluations can grow.  Repeated ids
// inside one row are masked out of the device copy of the neighbour table.
//
// Distances follow the reference's AVX-512 rounding order exactly: partial
// sum p (0..15) accumulates dims p, p+16, p+32, ... with one FMA each, then
// pairs (p, p+8), (p, p+4), (p, p+2), (0, 1) are added.  Neighbour rows are
// walked in chunks of 32 ids; a group of L lanes (L = 4 or 2, chosen so the
// fresh candidates of one chunk fit one pass) evaluates one candidate: lane t
// of the group holds partials t*16/L .. (t+1)*16/L - 1, i.e. 16/L consecutive
// bytes of each 16-byte chunk of the SQ8 row.  The candidate's key ends up in
// the group's first lane.
#include <math.h>

#include "lqg_internal.h"
#include "wave.h"

namespace lqg {
namespace {

__device__ __forceinline__ uint64_t make_key(float d, uint32_t id) {
    return (static_cast<uint64_t>(__float_as_uint(d)) << 32) | (static_cast<uint64_t>(id) << 1);
}
__device__ __forceinline__ uint32_t key_id(uint64_t k) {
    return static_cast<uint32_t>(k >> 1) & 0x7FFFFFFFu;
}
__device__ __forceinline__ float key_dist(uint64_t k) {
    return __uint_as_float(static_cast<uint32_t>(k >> 32));
}

// One SQ8 byte into partial s: x = min + q*scale (one FMA, as the reference
// build contracts it), t = query - x, s += t*t (one FMA).
__device__ __forceinline__ void acc1(uint32_t byte, float qq, float mn, float sc, float& s) {
    const float x = fmaf(static_cast<float>(byte), sc, mn);
    const float t = qq - x;
    s = fmaf(t, t, s);
}

// NB chunks of 16 dims from chunk `base` of an SQ8 row into a lane's PL
// partials (lane t of its group holds partials PL t .. PL t + PL - 1, i.e.
// bytes PL t .. of every chunk).  All NB loads are issued before any
// arithmetic, and the block carries no branch: a branch per chunk kept each
// chunk's LDS reads of the query and the SQ8 parameters from being issued
// ahead of the previous chunk's arithmetic, exposing their latency once per
// 16 dims.  MASK: chunks past the row (base + c >= nchunk) read the last
// chunk and add t = 0, i.e. fma(0, 0, s) = s exactly.
template <int PL, int NB, bool MASK>
__device__ __forceinline__ void chunk_block(const uint8_t* row, uint32_t t, const float* qv,
                                            const float* pmin, const float* pscale, uint32_t base,
                                            uint32_t nchunk, float (&s)[PL]) {
    constexpr int W = PL >= 4 ? PL / 4 : 1;  // 32-bit words per chunk and lane
    uint32_t wv[NB][W];
    uint32_t ch[NB];
#pragma unroll
    for (int c = 0; c < NB; ++c) {
        ch[c] = MASK ? min(base + c, nchunk - 1u) : base + c;
        const uint8_t* src = row + 16 * ch[c] + PL * t;
        if constexpr (PL == 1) {
            wv[c][0] = *src;
        } else if constexpr (PL == 2) {
            wv[c][0] = *reinterpret_cast<const uint16_t*>(src);
        } else if constexpr (PL == 8) {
            const uint2 v = *reinterpret_cast<const uint2*>(src);
            wv[c][0] = v.x;
            wv[c][1] = v.y;
        } else if constexpr (PL == 16) {
            const uint4 v = *reinterpret_cast<const uint4*>(src);
            wv[c][0] = v.x;
            wv[c][1] = v.y;
            wv[c][2] = v.z;
            wv[c][3] = v.w;
        } else {
            wv[c][0] = *reinterpret_cast<const uint32_t*>(src);
        }
    }
#pragma unroll
    for (int c = 0; c < NB; ++c) {
        const bool ok = !MASK || base + c < nchunk;
        const uint32_t b = 16 * ch[c] + PL * t;
        if constexpr (PL < 4) {
#pragma unroll
            for (int p = 0; p < PL; ++p) {
                const float x = fmaf(static_cast<float>((wv[c][0] >> (8 * p)) & 0xFFu),
                                     pscale[b + p], pmin[b + p]);
                float tt = qv[b + p] - x;
                if (MASK) tt = ok ? tt : 0.0f;
                s[p] = fmaf(tt, tt, s[p]);
            }
        } else {
#pragma unroll
            for (int v = 0; v < W; ++v) {
                const float4 qq = *reinterpret_cast<const float4*>(qv + b + 4 * v);
                const float4 mn = *reinterpret_cast<const float4*>(pmin + b + 4 * v);
                const float4 sc = *reinterpret_cast<const float4*>(pscale + b + 4 * v);
                const uint32_t x = wv[c][v];
                const float q4[4] = {qq.x, qq.y, qq.z, qq.w};
                const float m4[4] = {mn.x, mn.y, mn.z, mn.w};
                const float s4[4] = {sc.x, sc.y, sc.z, sc.w};
#pragma unroll
                for (int e = 0; e < 4; ++e) {
                    const float xd = fmaf(static_cast<float>((x >> (8 * e)) & 0xFFu), s4[e], m4[e]);
                    float tt = q4[e] - xd;
                    if (MASK) tt = ok ? tt : 0.0f;
                    s[4 * v + e] = fmaf(tt, tt, s[4 * v + e]);
                }
            }
        }
    }
}

// Squared L2 of the query to candidate `id` by a group of L lanes (t = lane
// index in the group).  Returns the full sum on every lane of the group.
template <int L, int NB>
__device__ __forceinline__ float group_l2(const DevIndex& ix, const float* qv, const float* pmin,
                                         const float* pscale, uint32_t id, bool act,
                                         uint32_t t) {
    constexpr int PL = 16 / L;  // partials (and bytes per chunk) per lane
    float s[PL];
#pragma unroll
    for (int p = 0; p < PL; ++p) s[p] = 0.f;
    const uint32_t nchunk = ix.dim >> 4;
    const uint8_t* row = ix.qdata + static_cast<size_t>(id) * ix.dim;
    // Chunks are taken in blocks of NB (8: a row of <= 128 bytes is one
    // memory round trip), then a block of the 6, 4 or 2 left, all without
    // branches (see chunk_block); other remainders take a block of 8, or of NB
    // when more than 8 are left, with the chunks past the row masked.
    if (act && (ix.dim & 15u) == 0) {
        uint32_t base = 0;
        for (; base + NB <= nchunk; base += NB)
            chunk_block<PL, NB, false>(row, t, qv, pmin, pscale, base, nchunk, s);
        const uint32_t rem = nchunk - base;
        if constexpr (NB > 8) {
            if (rem > 8u) {
                chunk_block<PL, NB, true>(row, t, qv, pmin, pscale, base, nchunk, s);
                base = nchunk;
            }
        }
        if (base == nchunk) {
        } else if (rem == 8u)
            chunk_block<PL, 8, false>(row, t, qv, pmin, pscale, base, nchunk, s);
        else if (rem == 6u)
            chunk_block<PL, 6, false>(row, t, qv, pmin, pscale, base, nchunk, s);
        else if (rem == 4u)
            chunk_block<PL, 4, false>(row, t, qv, pmin, pscale, base, nchunk, s);
        else if (rem == 2u)
            chunk_block<PL, 2, false>(row, t, qv, pmin, pscale, base, nchunk, s);
        else if (rem != 0u)
            chunk_block<PL, 8, true>(row, t, qv, pmin, pscale, base, nchunk, s);
    } else if (act) {
        for (uint32_t i = 0; i < nchunk; ++i)
#pragma unroll
            for (int p = 0; p < PL; ++p) {
                const uint32_t d = 16 * i + PL * t + p;
                acc1(row[d], qv[d], pmin[d], pscale[d], s[p]);
            }
    }
    float r;
    if constexpr (L == 16) {
        // lane t holds partial t: (p, p+8) across t^8, (p, p+4) across t^4, then
        // (b0 + b2) across t^2 and (b0 + b2) + (b1 + b3) across t^1
        s[0] += __shfl_xor(s[0], 8);
        s[0] += __shfl_xor(s[0], 4);
        s[0] += __shfl_xor(s[0], 2);
        s[0] += __shfl_xor(s[0], 1);
        r = s[0];
    } else if constexpr (L == 8) {
        // lane t holds partials 2t, 2t+1: (p, p+8) across t^4, (p, p+4) across
        // t^2; lane 0 then holds b0, b1 and lane 1 b2, b3
#pragma unroll
        for (int p = 0; p < 2; ++p) s[p] += __shfl_xor(s[p], 4);
#pragma unroll
        for (int p = 0; p < 2; ++p) s[p] += __shfl_xor(s[p], 2);
        const float c0 = s[0] + __shfl_xor(s[0], 1);
        const float c1 = s[1] + __shfl_xor(s[1], 1);
        r = c0 + c1;
    } else if constexpr (L == 4) {
        // lane t holds partials 4t..4t+3: (p, p+8) across t^2, (p, p+4) across t^1
#pragma unroll
        for (int p = 0; p < 4; ++p) s[p] += __shfl_xor(s[p], 2);
#pragma unroll
        for (int p = 0; p < 4; ++p) s[p] += __shfl_xor(s[p], 1);
        r = (s[0] + s[2]) + (s[1] + s[3]);
    } else if constexpr (L == 1) {
        // one lane holds all 16 partials: (p, p+8), (p, p+4), then
        // (b0 + b2) + (b1 + b3)
#pragma unroll
        for (int p = 0; p < 8; ++p) s[p] += s[p + 8];
#pragma unroll
        for (int p = 0; p < 4; ++p) s[p] += s[p + 4];
        r = (s[0] + s[2]) + (s[1] + s[3]);
    } else {
        static_assert(L == 2, "groups of 16, 8, 4, 2 or 1 lanes");
        // lane t holds partials 8t..8t+7: (p, p+8) across t^1, then in-lane
#pragma unroll
        for (int p = 0; p < 8; ++p) s[p] += __shfl_xor(s[p], 1);
#pragma unroll
        for (int p = 0; p < 4; ++p) s[p] += s[p + 4];
        r = (s[0] + s[2]) + (s[1] + s[3]);
    }
    if (act) {
        for (uint32_t d = nchunk << 4; d < ix.dim; ++d) acc1(row[d], qv[d], pmin[d], pscale[d], r);
    }
    return r;
}

// log2 of the lanes per candidate when a wave evaluates m fresh candidates:
// 64 / (the smallest of 4, 8, 16, 32, 64 that holds m).
__device__ __forceinline__ uint32_t fresh_lanes_log2(uint32_t m) {
    return m <= 4 ? 4u : m <= 8 ? 3u : m <= 16 ? 2u : m <= 32 ? 1u : 0u;
}

// Keys of the m fresh candidates bids[0..m) (m <= 32, or <= 64 with WIDE) in
// one pass: candidate c's key is returned on lane c*L, L = 1 <<
// fresh_lanes_log2(m), so that few candidates still spread over the whole
// wave (each lane then carries fewer of the 16 partials, and the dependent
// chains are shorter); other lanes return ~0.  All 64 lanes call it.  NB:
// chunks per block of row loads (8 in the batch kernels; more in the split
// kernel, whose waves hold few candidates each and have registers to spare).
// WIDE: one lane per candidate for 33..64 candidates (all 16 partials in the
// lane, 16-byte loads).
template <int NB = 8, bool WIDE = false>
__device__ __forceinline__ uint64_t eval_fresh(const DevIndex& ix, const float* qv,
                                               const float* pmin, const float* pscale,
                                               const uint32_t* bids, uint32_t m,
                                               uint32_t lane) {
    // blocks of NB chunks hold NB * max(1, PL / 4) row words per lane; two
    // lanes per candidate (8 partials, two words per chunk) take half as many
    // chunks, at most 16
    constexpr int NB2 = NB > 8 ? (NB / 2 > 16 ? 16 : NB / 2) : NB;
    float r;
    const uint32_t lg = fresh_lanes_log2(m);
    const uint32_t L = 1u << lg;
    const uint32_t c = lane >> lg;
    const bool act = c < m;
    const uint32_t id = act ? bids[c] : 0u;
    if (L == 16)
        r = group_l2<16, NB>(ix, qv, pmin, pscale, id, act, lane & 15u);
    else if (L == 8)
        r = group_l2<8, NB>(ix, qv, pmin, pscale, id, act, lane & 7u);
    else if (L == 4)
        r = group_l2<4, NB>(ix, qv, pmin, pscale, id, act, lane & 3u);
    else if (!WIDE || L == 2)
        r = group_l2<2, NB2>(ix, qv, pmin, pscale, id, act, lane & 1u);
    else
        r = group_l2<1, (NB > 8 ? 8 : NB)>(ix, qv, pmin, pscale, id, act, 0u);
    const bool leader = act && (lane & (L - 1)) == 0;
    return leader ? make_key(r, id) : ~0ull;
}

// Exact integer evaluation.  When every SQ8 parameter is min 0 / scale 1 the
// decode is the identity, and when the query is integer valued in [0, 255]
// every term of the reference's FP32 sum is an integer: t = q - c is exact,
// t*t <= 65025, and with dim <= 256 every partial and pairwise sum stays below
// 2^24, so each FP32 operation is exact and the reference's result is exactly
// sum (q_i - c_i)^2 whatever the order.  That integer is computed here as
// sum q^2 + sum c^2 - 2 sum q*c with v_dot4_u32_u8 and converted to float
// (exact below 2^24): bit-identical.  Candidate c's key is returned on lane 2c.
template <int NCH>
__device__ __forceinline__ uint64_t eval_fresh_int(const DevIndex& ix, const uint32_t* qp,
                                                   uint32_t qq, const uint32_t* bids, uint32_t m,
                                                   uint32_t lane) {
    // two lanes per candidate (m <= 32): lane 2c + h takes half h of the row's
    // 16-byte chunks (at most NCH: 8 of the 16 a 256-byte row has, or 4 for
    // rows of at most 128 bytes, which then need half the registers); the
    // integer partial sums are combined with one shuffle
    const uint32_t c = lane >> 1, h = lane & 1u;
    const bool act = c < m;
    const uint32_t n16 = ix.dim >> 4;
    const uint32_t per = (n16 + 1) >> 1;
    const uint32_t c0 = h * per;
    const uint32_t nch = h == 0 ? per : n16 - per;
    uint32_t qc = 0, cc = 0, id = 0;
    if (act) {
        id = bids[c];
        const uint4* row =
            reinterpret_cast<const uint4*>(ix.qdata + static_cast<size_t>(id) * ix.dim) + c0;
        uint4 w[NCH];
#pragma unroll
        for (uint32_t k = 0; k < NCH; ++k)
            if (k < nch) w[k] = row[k];
#pragma unroll
        for (uint32_t k = 0; k < NCH; ++k) {
            if (k < nch) {
                const uint4 qw = reinterpret_cast<const uint4*>(qp)[c0 + k];
                qc = __builtin_amdgcn_udot4(qw.x, w[k].x, qc, false);
                qc = __builtin_amdgcn_udot4(qw.y, w[k].y, qc, false);
                qc = __builtin_amdgcn_udot4(qw.z, w[k].z, qc, false);
                qc = __builtin_amdgcn_udot4(qw.w, w[k].w, qc, false);
                cc = __builtin_amdgcn_udot4(w[k].x, w[k].x, cc, false);
                cc = __builtin_amdgcn_udot4(w[k].y, w[k].y, cc, false);
                cc = __builtin_amdgcn_udot4(w[k].z, w[k].z, cc, false);
                cc = __builtin_amdgcn_udot4(w[k].w, w[k].w, cc, false);
            }
        }
    }
    qc += __shfl_xor(qc, 1);
    cc += __shfl_xor(cc, 1);
    if (!act || h) return ~0ull;
    return make_key(static_cast<float>(qq + cc - 2u * qc), id);
}

// eval_fresh_int for up to 64 candidates: candidates 32.. in a second call,
// whose keys (on even lanes) move to the odd lanes.
template <int NCH>
__device__ __forceinline__ uint64_t eval_fresh_int_wide(const DevIndex& ix, const uint32_t* qp,
                                                        uint32_t qq, const uint32_t* bids,
                                                        uint32_t m, uint32_t lane) {
    const uint64_t k1 = eval_fresh_int<NCH>(ix, qp, qq, bids, min(m, 32u), lane);
    if (m <= 32u) return k1;
    const uint64_t k2 = eval_fresh_int<NCH>(ix, qp, qq, bids + 32, m - 32u, lane);
    const int src = static_cast<int>(lane) - 1;
    const uint32_t lo = static_cast<uint32_t>(__shfl(static_cast<int>(static_cast<uint32_t>(k2)), src));
    const uint32_t hi =
        static_cast<uint32_t>(__shfl(static_cast<int>(static_cast<uint32_t>(k2 >> 32)), src));
    return (lane & 1u) ? ((static_cast<uint64_t>(hi) << 32) | lo) : k1;
}

struct WaveLds {
    float* qv;        // query (dim)
    uint64_t* stage;  // pool permutation staging (EF_CAP)
    uint32_t* bids;   // fresh candidate ids (64)
    uint32_t* cache;  // visited cache (1 << cache_bits)
    uint32_t* cnt;    // ranked admission: EF_CAP 16-bit counters (two per word),
                      // zero between passes
};

// Pool classes from this capacity on admit a pass's candidates by binary
// search and a prefix sum (search_one) instead of comparing each candidate
// with every register slot of the pool.  200k x 960 degree 64, 1,000 queries,
// unsplit kernel: ef U{64..1024} (16 slots) 11.3 -> 10.3 ms; ef U{64..512}
// (8 slots) 5.42 -> 5.68 ms, so the 512 class keeps the slot compares.  The
// counters are 16-bit (a pass admits at most 64 candidates): with 32-bit ones
// the split kernel's 960-d workgroup needed 64 bytes more than a quarter of
// the LDS and ran three to a CU (split 2: 10.7 -> 11.1 ms).
constexpr int kRankedSq8 = 1024;

// Split evaluation (h1_split_kernel): a query's workgroup of W waves shares
// each pass's distance evaluation.  Wave 0 walks the query; before each pass it
// posts the fresh candidates (WaveLds::bids) and their count here, every wave
// evaluates a share of them and writes their keys, and wave 0 takes the keys
// back.  go = 0 tells the other waves that the query is done.
struct SplitLds {
    uint32_t go;
    uint32_t m;
    uint32_t pad[2];
    uint64_t keys[kWave];  // one per fresh candidate (at most 64 per pass)
};

// Chunks per block of row loads in the split kernel: 32 (one 32-bit word per
// chunk and lane) for up to 4 partials per lane, 16 for 8 partials; half that
// for pools of 512 or more entries in the four-wave kernel, whose 128 VGPRs
// the pool shares.
template <int EF_CAP, int W>
constexpr int split_nb() {
    return W == 2 || EF_CAP < 512 ? 32 : 16;
}

// Wave wid's share of a pass of m fresh candidates: candidates [c0, c0 + mw)
// with c0 = wid * ceil(m / W), each key written to keys[c].
template <int W, int NB>
__device__ __forceinline__ void eval_share(const DevIndex& ix, const float* qv, const float* pmin,
                                           const float* pscale, const uint32_t* bids,
                                           uint32_t m, uint32_t wid, uint32_t lane,
                                           uint64_t* keys) {
    const uint32_t cpw = (m + W - 1) / W;
    const uint32_t c0 = wid * cpw;
    if (c0 >= m) return;
    const uint32_t mw = min(cpw, m - c0);
    const uint64_t key = eval_fresh<NB>(ix, qv, pmin, pscale, bids + c0, mw, lane);
    const uint32_t lg = fresh_lanes_log2(mw);
    if ((lane & ((1u << lg) - 1u)) == 0 && (lane >> lg) < mw) keys[c0 + (lane >> lg)] = key;
}

// Speculative evaluation (latency mode, depth 4): the second wave of a
// query's workgroup computes the keys of every neighbour of the best
// unexpanded pool entries before they are expanded, into these slots; wave 0
// takes a popped node's ids and keys from its slot instead of loading and
// evaluating them.  A slot's tag is the node it holds, kInvalid while the
// slot is written (a seqlock: wave 0 reads the tag, the row, then the tag
// again, and uses the row only if both reads give the node).
constexpr uint32_t kSpecSlots = 4;

struct SpecSlot {
    uint32_t tag;
    uint32_t pad[3];
    uint32_t ids[kWave];   // the node's neighbour ids (degree <= 64)
    uint64_t keys[kWave];  // make_key(distance, id) of each
};

struct SpecLds {
    uint32_t qready;  // 1 once wave 0 has prepared the query
    uint32_t qq;      // the integer path's sum of squares of the query
    uint32_t cur;     // the node wave 0 expands
    uint32_t pad;
    uint32_t bids[kWave];  // the speculating wave's candidate list
    SpecSlot slot[kSpecSlots];
};

// Marks node x in the counting bitmap; true when it was not marked before.
__device__ __forceinline__ bool first_visit(uint32_t* bits, uint32_t x) {
    return (atomicOr(&bits[x >> 5], 1u << (x & 31u)) & (1u << (x & 31u))) == 0;
}

// Evaluation modes of the kernel, fixed at compile time so each variant only
// holds the registers of its own distance code:
//   EVAL_INT   -- exact integer path only (eval_fresh_int); queries that do not
//                 qualify are left for a following EVAL_FLOAT launch
//   EVAL_FLOAT -- SQ8 float path; with job.skip_int, queries the EVAL_INT
//                 launch handled are skipped
constexpr int EVAL_FLOAT = 0;
constexpr int EVAL_INT = 1;
// EVAL_INT for rows of at most 128 bytes (D <= 128): the row buffer of the
// integer path is half as large
constexpr int EVAL_INT128 = 3;

// W > 1: wave 0 of a split-kernel workgroup, sharing every distance
// evaluation with the workgroup's other W - 1 waves through sp.
// SPEC: wave 0 of a speculative latency-mode workgroup (h1_spec_kernel),
// taking rows the other wave evaluated ahead from spec's slots.
template <int EF_CAP, int EVAL, int W = 1, bool SPEC = false>
__device__ __forceinline__ void search_one(const DevIndex& ix, const SearchJob& job, uint32_t q,
                           const float* pmin, const float* pscale, const WaveLds& w,
                           uint32_t cache_bits, uint32_t lane, uint32_t* vbits,
                           SplitLds* sp = nullptr, SpecLds* spec = nullptr) {
    constexpr int S = EF_CAP / kWave;
    // clamps keep register, LDS and output indexing in bounds whatever the arrays hold
    const uint32_t ef =
        max(1u, min(job.ef ? job.ef[q] : job.ef_all, static_cast<uint32_t>(EF_CAP)));
    const uint32_t kq = min(job.k ? job.k[q] : job.k_all, job.kstride);
    const float* qsrc = job.queries + static_cast<size_t>(q) * ix.dim;
    bool qint = true;
    for (uint32_t i = lane; i < ix.dim; i += kWave) {
        const float v = qsrc[i];
        w.qv[i] = v;
        qint = qint && v >= 0.f && v <= 255.f && v == rintf(v);
    }
    // exact integer path (see eval_fresh_int): identity SQ8, integer query
    const bool int_ok = ix.sq8_identity && (ix.dim & 15u) == 0 && ix.dim <= 256 &&
                        __ballot(!qint) == 0;
    constexpr bool int_path = EVAL == EVAL_INT || EVAL == EVAL_INT128;
    if (int_path ? !int_ok : (job.skip_int && int_ok)) {
        // the other launch of the pair handles this query
        if (int_path && job.left && lane == 0) atomicAdd(job.left, 1u);
        wave_sync();
        return;
    }
    uint32_t qq = 0;
    const uint32_t* qp = reinterpret_cast<const uint32_t*>(w.qv);
    if (int_path) {
        wave_sync();
        // pack the query to bytes in place: word j <- dims 4j..4j+3
        uint32_t word = 0;
        if (lane < ix.dim / 4) {
            word = static_cast<uint32_t>(w.qv[4 * lane]) |
                   (static_cast<uint32_t>(w.qv[4 * lane + 1]) << 8) |
                   (static_cast<uint32_t>(w.qv[4 * lane + 2]) << 16) |
                   (static_cast<uint32_t>(w.qv[4 * lane + 3]) << 24);
            qq = __builtin_amdgcn_udot4(word, word, 0u, false);
        }
        wave_sync();
        if (lane < ix.dim / 4) w.qv[lane] = __uint_as_float(word);
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) qq += __shfl_xor(qq, o);
    }
    if constexpr (SPEC) {
        // the query (packed in place for the integer path) is complete
        wave_sync();
        if (lane == 0) {
            spec->qq = qq;
            __hip_atomic_store(&spec->qready, 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_WORKGROUP);
        }
    }
    const uint32_t csize = 1u << cache_bits;
    for (uint32_t i = lane; i < csize; i += kWave) w.cache[i] = kInvalid;
    if constexpr (EF_CAP >= kRankedSq8)
        for (uint32_t i = lane; i < static_cast<uint32_t>(EF_CAP) / 2u; i += kWave) w.cnt[i] = 0u;
    if (lane == 0) {
        w.bids[0] = ix.ep;
        w.cache[(ix.ep * 2654435761u) >> (32 - cache_bits)] = ix.ep;
    }
    if (vbits) {
        for (uint32_t i = lane; i < job.visit_words; i += kWave) vbits[i] = 0u;
        __threadfence_block();
        if (lane == 0) first_visit(vbits, ix.ep);
    }
    wave_sync();

    // entry point: qg.hpp:305-311
    uint64_t P[S];
#pragma unroll
    for (int s = 0; s < S; ++s) P[s] = ~0ull;
    // Pools of 1024 entries (launches of a few high-ef queries) take the
    // neighbours of a node of degree <= 64 in one pass of 64 ids: one id load
    // and one evaluation instead of two.  Smaller pools keep passes of 32, and
    // the registers that go with them (the 512 class, which graph builds at
    // EF 500 run with a million queries, took 299 instead of 164 VGPRs).
    constexpr bool kWide = EF_CAP >= 1024;
    const uint32_t pass = kWide && ix.deg <= kWave ? kWave : 32u;
    auto evaluate = [&](uint32_t m) -> uint64_t {
        if constexpr (EVAL == EVAL_INT && kWide)
            return eval_fresh_int_wide<8>(ix, qp, qq, w.bids, m, lane);
        else if constexpr (EVAL == EVAL_INT128 && kWide)
            return eval_fresh_int_wide<4>(ix, qp, qq, w.bids, m, lane);
        else if constexpr (EVAL == EVAL_INT)
            return eval_fresh_int<8>(ix, qp, qq, w.bids, m, lane);
        else if constexpr (EVAL == EVAL_INT128)
            return eval_fresh_int<4>(ix, qp, qq, w.bids, m, lane);
        else if constexpr (W > 1) {
            if (lane == 0) {
                sp->m = m;
                sp->go = 1u;
            }
            __syncthreads();  // the other waves start their shares
            eval_share<W, split_nb<EF_CAP, W>()>(ix, w.qv, pmin, pscale, w.bids, m, 0u, lane,
                                              sp->keys);
            __syncthreads();  // every share's keys are written
            return lane < m ? sp->keys[lane] : ~0ull;
        } else {
            return eval_fresh<8, kWide>(ix, w.qv, pmin, pscale, w.bids, m, lane);
        }
    };
    {
        const uint64_t k0 = readlane64(evaluate(1), 0);
        if (lane == 0) {
            P[0] = k0;
            w.stage[0] = k0;  // the stage array mirrors the pool between merges
        }
        wave_sync();
    }
    uint32_t size = 1, n_exp = 0, n_eval = 1;

    for (;;) {
        // pop the smallest unexpanded entry (qg.hpp:314-318)
        uint64_t kc = 0;
        bool found = false;
#pragma unroll
        for (int s = 0; s < S; ++s) {
            if (!found) {
                const uint32_t i = static_cast<uint32_t>(s * kWave) + lane;
                const uint64_t m = __ballot(i < size && !(P[s] & 1ull));
                if (m) {
                    const int l = static_cast<int>(__builtin_ctzll(m));
                    kc = readlane64(P[s], l);
                    if (static_cast<int>(lane) == l) P[s] |= 1ull;
                    found = true;
                }
            }
        }
        if (!found) break;
        ++n_exp;
        const uint32_t* nrow = ix.nbr + static_cast<size_t>(key_id(kc)) * ix.deg;
        // the node's row from a speculation slot: lane l < 32 holds neighbours
        // l (s_id0, s_k0) and 32 + l (s_id1, s_k1)
        bool use_spec = false;
        uint32_t s_id0 = kInvalid, s_id1 = kInvalid;
        uint64_t s_k0 = ~0ull, s_k1 = ~0ull;
        // (with passes of 64, lane l holds neighbour l in s_id0, s_k0)
        if constexpr (SPEC) {
            const uint32_t x = key_id(kc);
            if (lane == 0) __hip_atomic_store(&spec->cur, x, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
            uint32_t t = kInvalid;
            if (lane < kSpecSlots)
                t = __hip_atomic_load(&spec->slot[lane].tag, __ATOMIC_ACQUIRE,
                                      __HIP_MEMORY_SCOPE_WORKGROUP);
            const uint64_t hit = __ballot(lane < kSpecSlots && t == x);
            if (hit && ix.deg <= kWave) {
                SpecSlot* sl = &spec->slot[__builtin_ctzll(hit)];
                if (lane < pass && lane < ix.deg) {
                    s_id0 = sl->ids[lane];
                    s_k0 = sl->keys[lane];
                }
                if (pass == 32u && lane < 32 && 32u + lane < ix.deg) {
                    s_id1 = sl->ids[32u + lane];
                    s_k1 = sl->keys[32u + lane];
                }
                __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
                const uint32_t t2 =
                    __hip_atomic_load(&sl->tag, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
                use_spec = uniform(t2) == x;
            }
        }

        for (uint32_t cb = 0; cb < ix.deg; cb += pass) {
            // neighbours not seen before (qg.hpp:320-324), `pass` ids at a time
            bool fresh = false;
            uint32_t nid = kInvalid;
            if (lane < pass && cb + lane < ix.deg) {
                nid = use_spec ? (cb == 0 ? s_id0 : s_id1) : nrow[cb + lane];
                if (nid < ix.n) {
                    const uint32_t slot = (nid * 2654435761u) >> (32 - cache_bits);
                    fresh = atomicExch(&w.cache[slot], nid) != nid;
                }
            }
            const uint64_t fm = __ballot(fresh);
            if (fm == 0) continue;
            const uint32_t m = static_cast<uint32_t>(__popcll(fm));
            if (!use_spec && fresh) w.bids[lanes_below(fm)] = nid;
            wave_sync();
            if (vbits) {
                const bool first = fresh && first_visit(vbits, nid);
                n_eval += static_cast<uint32_t>(__popcll(__ballot(first)));
            } else {
                n_eval += m;
            }
            const uint64_t key =
                use_spec ? (fresh ? (cb == 0 ? s_k0 : s_k1) : ~0ull) : evaluate(m);

            // admission (qg.hpp:329-335): pool not full, or below its maximum
            uint64_t last = 0;
#pragma unroll
            for (int s = 0; s < S; ++s)
                if (static_cast<uint32_t>(s) == (size - 1) / kWave)
                    last = readlane64(P[s], static_cast<int>((size - 1) % kWave));
            bool cand = key != ~0ull && (size < ef || key < last);
            uint64_t mm = __ballot(cand);
            if (mm == 0) continue;
            uint32_t lb = 0, rank = 0;
            uint32_t shift[S];
            if constexpr (EF_CAP >= kRankedSq8) {
                // lb by binary search over the stage array (keys ascending at
                // positions below size; only expanded bits, bit 0, may differ
                // from the registers, and no comparison with a candidate,
                // whose bit 0 is clear, depends on them), then the duplicate
                // check below, then a counter per position: entry i moves up
                // by the candidates with lb <= i (their prefix sum)
                if (cand) {
                    uint32_t pos = 0;
#pragma unroll
                    for (uint32_t step = static_cast<uint32_t>(EF_CAP); step > 0u; step >>= 1) {
                        const uint32_t t = pos + step;
                        if (t <= size && w.stage[t - 1u] < key) pos = t;
                    }
                    lb = pos;
                }
                cand = cand && !(lb < size && ((w.stage[lb] ^ key) >> 1) == 0);
                mm = __ballot(cand);
                if (mm == 0) continue;
                // 16-bit counters, position p in half p & 1 of word p >> 1;
                // no half passes 64, so a word-wide add never carries
                if (cand) atomicAdd(&w.cnt[lb >> 1], 1u << (16u * (lb & 1u)));  // lb < EF_CAP
                for (uint64_t rest = mm; rest; rest &= rest - 1)
                    rank += readlane64(key, static_cast<int>(__builtin_ctzll(rest))) < key ? 1u : 0u;
                wave_sync();
                constexpr int V = S / 8;  // lane l sums positions S l .. S l + S - 1
                uint4* run = reinterpret_cast<uint4*>(w.cnt) + V * lane;
                uint32_t c[4 * V];
                uint32_t acc = 0;
#pragma unroll
                for (int v = 0; v < V; ++v) {
                    const uint4 x = run[v];
                    const uint32_t xs[4] = {x.x, x.y, x.z, x.w};
#pragma unroll
                    for (int u = 0; u < 4; ++u) {
                        const uint32_t p0 = acc + (xs[u] & 0xFFFFu);
                        acc = p0 + (xs[u] >> 16);
                        c[4 * v + u] = p0 | (acc << 16);
                    }
                }
                uint32_t incl = acc;
#pragma unroll
                for (int d = 1; d < kWave; d <<= 1) {
                    const uint32_t t = static_cast<uint32_t>(__shfl_up(static_cast<int>(incl), d));
                    if (lane >= static_cast<uint32_t>(d)) incl += t;
                }
                const uint32_t off = (incl - acc) * 0x10001u;  // added to both halves
#pragma unroll
                for (int v = 0; v < V; ++v)
                    run[v] = make_uint4(c[4 * v] + off, c[4 * v + 1] + off, c[4 * v + 2] + off,
                                        c[4 * v + 3] + off);
                wave_sync();
                const uint16_t* c16 = reinterpret_cast<const uint16_t*>(w.cnt);
#pragma unroll
                for (int s = 0; s < S; ++s) shift[s] = c16[static_cast<uint32_t>(s * kWave) + lane];
                wave_sync();
#pragma unroll
                for (int v = 0; v < V; ++v) run[v] = make_uint4(0u, 0u, 0u, 0u);
            } else {
                // One pass over the candidates (broadcast from their lanes): each
                // candidate's position below the pool (lb, written into its lane),
                // its rank among the candidates, and every pool entry's shift (the
                // candidates below it).  Candidate keys have bit 0 clear, so these
                // plain comparisons order by (distance, id) as long as no candidate is
                // already in the pool -- a node evaluated before whose cache entry
                // was overwritten.  Such a candidate's pool copy is the entry at its
                // position lb; the stage array holds the pool as of the last merge
                // (only expanded bits may differ), so one LDS read per candidate
                // finds them, and only then is the pass redone with duplicates
                // dropped.
#pragma unroll
                for (int s = 0; s < S; ++s) shift[s] = 0;
                {
                    uint64_t rest = mm;
                    while (rest) {
                        const int j = static_cast<int>(__builtin_ctzll(rest));
                        rest &= rest - 1;
                        const uint64_t kj = readlane64(key, j);
                        uint32_t below = 0;
#pragma unroll
                        for (int s = 0; s < S; ++s) {
                            const bool lt = P[s] < kj;
                            below += static_cast<uint32_t>(__popcll(__ballot(lt)));
                            shift[s] += lt ? 0u : 1u;
                        }
                        rank += kj < key ? 1u : 0u;
                        if (static_cast<int>(lane) == j) lb = below;
                    }
                }
                bool drop = cand && lb < size && ((w.stage[lb] ^ key) >> 1) == 0;
                if (__ballot(drop)) {
                    // slow pass with the duplicates dropped
                    cand = cand && !drop;
                    mm = __ballot(cand);
                    rank = 0;
#pragma unroll
                    for (int s = 0; s < S; ++s) shift[s] = 0;
                    uint64_t rest = mm;
                    while (rest) {
                        const int j = static_cast<int>(__builtin_ctzll(rest));
                        rest &= rest - 1;
                        const uint64_t kj = readlane64(key, j);
                        rank += kj < key ? 1u : 0u;
#pragma unroll
                        for (int s = 0; s < S; ++s) shift[s] += kj < P[s] ? 1u : 0u;
                    }
                }
            }
            const uint32_t m3 = static_cast<uint32_t>(__popcll(__ballot(cand)));
            if (m3 == 0) continue;
            // merge: scatter to new positions, keep ef (res_pool.pop, qg.hpp:332-333)
            const uint32_t nsize = min(size + m3, ef);
#pragma unroll
            for (int s = 0; s < S; ++s) {
                const uint32_t i = static_cast<uint32_t>(s * kWave) + lane;
                const uint32_t np = i + shift[s];
                if (i < size && np < nsize) w.stage[np] = P[s];
            }
            if (cand && lb + rank < nsize) w.stage[lb + rank] = key;
            wave_sync();
#pragma unroll
            for (int s = 0; s < S; ++s) {
                const uint32_t i = static_cast<uint32_t>(s * kWave) + lane;
                P[s] = i < nsize ? w.stage[i] : ~0ull;
            }
            size = nsize;
            wave_sync();
        }
    }

    // results (qg.hpp:340-343 for the reference order)
    const uint32_t cnt = min(kq, size);
    uint32_t* oi = job.out_ids + static_cast<size_t>(q) * job.kstride;
    float* od = job.out_d ? job.out_d + static_cast<size_t>(q) * job.kstride : nullptr;
    for (uint32_t j = lane; j < job.kstride; j += kWave) {
        const bool fill = j >= kq || (job.order == 1 ? j >= cnt : j < kq - cnt);
        if (fill) {
            oi[j] = kInvalid;
            if (od) od[j] = INFINITY;
        }
    }
    const uint32_t first = job.order == 1 ? 0u : size - cnt;  // pool entries [first, first+cnt)
    const uint32_t dst0 = job.order == 1 ? 0u : kq - cnt;
#pragma unroll
    for (int s = 0; s < S; ++s) {
        const uint32_t i = static_cast<uint32_t>(s * kWave) + lane;
        if (i >= first && i < first + cnt) {
            oi[dst0 + i - first] = key_id(P[s]);
            if (od) od[dst0 + i - first] = key_dist(P[s]);
        }
    }
    if (job.out_stats && lane == 0) {
        job.out_stats[2 * static_cast<size_t>(q)] = n_exp;
        job.out_stats[2 * static_cast<size_t>(q) + 1] = n_eval;
    }
    if (job.done) signal_host(job.done + q, lane);
    wave_sync();
}

// Body shared by the kernel variants: SQ8 params to LDS, then a persistent
// loop pulling query indices from the work counter.
template <int EF_CAP, int EVAL>
__device__ __forceinline__ void search_kernel_body(const DevIndex& ix, const SearchJob& job,
                                                   uint32_t cache_bits, uint32_t dim_pad,
                                                   uint32_t wave_bytes) {
    extern __shared__ __align__(16) unsigned char lds[];
    const uint32_t lane = threadIdx.x & (kWave - 1);
    const uint32_t wid = threadIdx.x / kWave;
    // only the float path decodes with the SQ8 parameters; the integer
    // variants' workgroups leave that region out of their LDS
    constexpr bool kParams = EVAL == EVAL_FLOAT;
    float* pmin = nullptr;
    float* pscale = nullptr;
    if constexpr (kParams) {
        pmin = reinterpret_cast<float*>(lds);
        pscale = pmin + dim_pad;
        for (uint32_t i = threadIdx.x; i < ix.dim; i += blockDim.x) {
            pmin[i] = ix.sq8_min[i];
            pscale[i] = ix.sq8_scale[i];
        }
        __syncthreads();
    }

    unsigned char* wb =
        lds + (kParams ? 8u * dim_pad : 0u) + static_cast<size_t>(wid) * wave_bytes;
    WaveLds w;
    w.qv = reinterpret_cast<float*>(wb);
    w.stage = reinterpret_cast<uint64_t*>(wb + 4u * dim_pad);
    w.bids = reinterpret_cast<uint32_t*>(w.stage + EF_CAP);
    w.cache = w.bids + kWave;
    w.cnt = w.cache + (1u << cache_bits);

    // a float launch after an integer launch that left nothing: done
    if (EVAL == EVAL_FLOAT && job.skip_int && job.left && *job.left == 0u) return;
    // Work distribution: wave gw first takes query gw (no atomic, so the
    // launch does not start with one atomic per wave on a single counter),
    // then queries sw + atomicAdd(counter) until the batch is drained, where
    // sw is the number of waves that take part.  In counting launches a wave
    // without a bitmap takes no queries.
    const uint32_t wpg = blockDim.x / kWave;
    const uint32_t gw = blockIdx.x * wpg + wid;
    uint32_t sw = gridDim.x * wpg;
    if (job.visit_bits) sw = min(sw, job.visit_waves);
    if (gw >= sw) return;
    uint32_t* vbits =
        job.visit_bits ? job.visit_bits + static_cast<size_t>(gw) * job.visit_words : nullptr;
    // One queue for the grid (small grids, counting launches): wave gw takes
    // query gw, then sw + atomicAdd(counter).  Otherwise one queue per XCD.
    // Workgroup b runs on XCD b % 8 (round-robin dispatch); XCD x owns
    // queries [x nq / 8, (x + 1) nq / 8): its waves take the first ones
    // statically (wave j of the XCD takes q0 + j), the rest from the XCD's
    // own counter (kQueueStride words apart, so the eight counters sit on
    // separate cache lines and their atomics do not queue behind each other).
    // A wave whose range is drained moves on to the next XCD's counter, so the
    // ranges drain evenly and every query is taken exactly once (each XCD has
    // at least one workgroup, which takes its static part).  One call site of
    // search_one for both.
    const bool single = gridDim.x < kXcds || job.visit_bits;
    const uint32_t xcd = blockIdx.x % kXcds;
    const uint32_t lw = (blockIdx.x / kXcds) * wpg + wid;
    uint32_t r = 0;       // XCD queues visited
    uint32_t base, q1;    // the current queue: next = base + atomicAdd(ctr), up to q1
    uint32_t* ctr;
    uint32_t q;
    auto xcd_queue = [&](uint32_t x, uint32_t& q0) {
        const uint32_t nwx = ((gridDim.x + kXcds - 1u - x) / kXcds) * wpg;
        q0 = static_cast<uint32_t>((static_cast<uint64_t>(x) * job.nq) / kXcds);
        q1 = static_cast<uint32_t>((static_cast<uint64_t>(x + 1u) * job.nq) / kXcds);
        base = q0 + nwx;
        ctr = job.counter + kQueueStride * x;
    };
    if (single) {
        base = sw;
        q1 = job.nq;
        ctr = job.counter;
        q = gw;
    } else {
        uint32_t q0;
        xcd_queue(xcd, q0);
        q = q0 + lw;
    }
    for (;;) {
        q = uniform(q);
        if (q >= q1) {
            // this queue is drained: the next XCD's, until all are
            if (single || ++r == kXcds) break;
            uint32_t q0;
            xcd_queue((xcd + r) % kXcds, q0);
            uint32_t nx = 0;
            if (lane == 0) nx = base + atomicAdd(ctr, 1u);
            q = __shfl(nx, 0);
            continue;
        }
        search_one<EF_CAP, EVAL>(ix, job, q, pmin, pscale, w, cache_bits, lane, vbits);
        uint32_t nx = 0;
        if (lane == 0) nx = base + atomicAdd(ctr, 1u);
        q = __shfl(nx, 0);
    }
}

// Register allocation left to the compiler.
template <int EF_CAP, int EVAL>
__global__ void __launch_bounds__(256) h1_search_kernel(DevIndex ix, SearchJob job,
                                                         uint32_t cache_bits, uint32_t dim_pad,
                                                         uint32_t wave_bytes) {
    search_kernel_body<EF_CAP, EVAL>(ix, job, cache_bits, dim_pad, wave_bytes);
}

// The same with at least 4 waves per SIMD (at most 128 VGPRs): the default
// for pools of up to 256 entries.  Left alone, the float-path kernel of the
// 256 class takes 131 VGPRs and drops to 3 waves per SIMD.
template <int EF_CAP, int EVAL>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4)))
h1_search_kernel_w4(DevIndex ix, SearchJob job, uint32_t cache_bits, uint32_t dim_pad,
                    uint32_t wave_bytes) {
    search_kernel_body<EF_CAP, EVAL>(ix, job, cache_bits, dim_pad, wave_bytes);
}

// Register file capped at 64 VGPRs so 8 waves fit per SIMD (32 per CU): more
// queries in flight to hide memory latency, at the price of spills where the
// code needs more.
template <int EF_CAP, int EVAL>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(8, 8)))
h1_search_kernel_w8(DevIndex ix, SearchJob job, uint32_t cache_bits, uint32_t dim_pad,
                    uint32_t wave_bytes) {
    search_kernel_body<EF_CAP, EVAL>(ix, job, cache_bits, dim_pad, wave_bytes);
}

// Register file capped at 80 VGPRs: 6 waves per SIMD.
template <int EF_CAP, int EVAL>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(6, 6)))
h1_search_kernel_w6(DevIndex ix, SearchJob job, uint32_t cache_bits, uint32_t dim_pad,
                    uint32_t wave_bytes) {
    search_kernel_body<EF_CAP, EVAL>(ix, job, cache_bits, dim_pad, wave_bytes);
}

// ---------------------------------------------------------------------------
// Latency mode, for launches of a few queries (a lone Index.search call): one
// query per workgroup of two waves.  Wave 0 walks the query with search_one,
// the code every batch kernel runs, so results are the same.  Wave 1 only
// loads: it watches wave 0's pool mirror in LDS (the stage array, rewritten
// at every merge) and fetches ahead what the next expansions will read, so
// that wave 0 finds it in the L2 instead of waiting on HBM twice per
// expansion (the neighbour-id row, then the candidates' SQ8 rows):
//   depth 1: the id rows of the two best unexpanded pool entries;
//   depth 2: also every 128-byte line of the SQ8 rows of their neighbours;
//   depth 3: the same for the four best unexpanded entries.
// The node wave 0 is expanding is one of the two (the mirror shows it
// unexpanded until the next merge); the other is the runner-up, which is the
// next node expanded unless the current expansion admits a closer one.
// Nothing wave 1 does reaches wave 0 except through the caches.  It stops
// when wave 0 raises its done flag, or after kPrefetchRounds rounds.
constexpr uint32_t kPrefetchRounds = 1u << 16;

__device__ void prefetch_ahead(const DevIndex& ix, const uint64_t* stage, uint32_t ef_cap,
                               const uint32_t* done, uint32_t depth, uint32_t lane,
                               uint32_t* sink_out) {
    uint32_t sink = 0;
    // the last eight nodes fetched (scalar registers)
    uint32_t hist[8];
#pragma unroll
    for (int i = 0; i < 8; ++i) hist[i] = kInvalid;
    const uint32_t deg = min(ix.deg, 64u);
    // nodes fetched per round: 2, or 4 at depth 3; a node's id row is read by
    // 64 / T lanes, each taking ids c, c + 64 / T, ...
    const uint32_t tcount = depth >= 3 ? 4u : 2u;
    const uint32_t per = kWave / tcount;
    for (uint32_t r = 0; r < kPrefetchRounds; ++r) {
        if (__hip_atomic_load(done, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP) != 0u) break;
        uint64_t key = ~0ull;
        if (lane < ef_cap)
            key = __hip_atomic_load(stage + lane, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
        const uint32_t id = key_id(key);
        bool open = (key & 1ull) == 0ull && id < ix.n;
#pragma unroll
        for (int i = 0; i < 8; ++i) open = open && id != hist[i];
        uint64_t m = __ballot(open);
        if (m == 0) {
            __builtin_amdgcn_s_sleep(1);
            continue;
        }
        // the first tcount open entries, best first; this lane's node
        uint32_t node = kInvalid;
        for (uint32_t t = 0; t < tcount && m != 0; ++t) {
            const uint32_t x = uniform(__shfl(id, static_cast<int>(__builtin_ctzll(m))));
            m &= m - 1;
#pragma unroll
            for (int i = 7; i > 0; --i) hist[i] = hist[i - 1];
            hist[0] = x;
            if (lane / per == t) node = x;
        }
        uint32_t nid[4] = {kInvalid, kInvalid, kInvalid, kInvalid};
        if (node < ix.n) {
            const uint32_t* row = ix.nbr + static_cast<size_t>(node) * ix.deg;
#pragma unroll
            for (uint32_t j = 0; j < 4; ++j) {
                const uint32_t c = lane % per + j * per;
                if (c < deg) nid[j] = row[c];
            }
        }
        sink += nid[0] ^ nid[1] ^ nid[2] ^ nid[3];
        if (depth >= 2) {
            uint32_t v = 0;
            for (uint32_t o = 0; o < ix.dim; o += 128u)
#pragma unroll
                for (uint32_t j = 0; j < 4; ++j)
                    if (nid[j] < ix.n) v += ix.qdata[static_cast<size_t>(nid[j]) * ix.dim + o];
            sink += v;
        }
    }
    // keeps the loads: the condition depends on every value they returned
    if (sink == 0x9E3779B9u) *sink_out = sink;
}

template <int EF_CAP, int EVAL>
__global__ void __launch_bounds__(128) h1_latency_kernel(DevIndex ix, SearchJob job,
                                                         uint32_t cache_bits, uint32_t dim_pad,
                                                         uint32_t wave_bytes, uint32_t depth) {
    extern __shared__ __align__(16) unsigned char lds[];
    const uint32_t lane = threadIdx.x & (kWave - 1);
    const uint32_t wid = threadIdx.x / kWave;
    constexpr bool kParams = EVAL == EVAL_FLOAT;
    float* pmin = nullptr;
    float* pscale = nullptr;
    if constexpr (kParams) {
        pmin = reinterpret_cast<float*>(lds);
        pscale = pmin + dim_pad;
        for (uint32_t i = threadIdx.x; i < ix.dim; i += blockDim.x) {
            pmin[i] = ix.sq8_min[i];
            pscale[i] = ix.sq8_scale[i];
        }
    }
    unsigned char* wb = lds + (kParams ? 8u * dim_pad : 0u);
    WaveLds w;
    w.qv = reinterpret_cast<float*>(wb);
    w.stage = reinterpret_cast<uint64_t*>(wb + 4u * dim_pad);
    w.bids = reinterpret_cast<uint32_t*>(w.stage + EF_CAP);
    w.cache = w.bids + kWave;
    w.cnt = w.cache + (1u << cache_bits);
    uint32_t* done = reinterpret_cast<uint32_t*>(wb + wave_bytes);
    // the mirror starts empty (keys ~0 read as expanded)
    for (uint32_t i = threadIdx.x; i < static_cast<uint32_t>(EF_CAP); i += blockDim.x)
        w.stage[i] = ~0ull;
    if (threadIdx.x == 0) {
        done[0] = 0u;
        done[1] = 0u;
    }
    __syncthreads();
    const uint32_t q = blockIdx.x;
    if (wid == 0) {
        if (q < job.nq)
            search_one<EF_CAP, EVAL>(ix, job, q, pmin, pscale, w, cache_bits, lane, nullptr);
        if (lane == 0) __hip_atomic_store(done, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    } else {
        prefetch_ahead(ix, w.stage, static_cast<uint32_t>(EF_CAP), done, depth, lane, done + 1);
    }
}

// ---------------------------------------------------------------------------
// Split evaluation, for the SQ8 float path at high dimension: one query per
// workgroup of W waves.  Wave 0 walks the query with search_one; at every
// pass all W waves evaluate a share of the fresh candidates (each candidate
// by the same group_l2 code as the batch kernels, so results are the same),
// with blocks of split_nb chunks per row load.  At D = 960 a row is then two
// memory round trips instead of eight, and a pass's arithmetic is spread over
// W SIMDs.  The waves meet at two barriers per pass (SplitLds).
// Registers: at most 128 VGPRs for W = 4, so that four waves per SIMD, a
// 1,000-query launch's workgroups, are resident at once; W = 2 needs half as
// many waves and is left at 256.
template <int EF_CAP, int W>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(W == 4 ? 4 : 2)))
h1_split_kernel(DevIndex ix, SearchJob job,
                                                       uint32_t cache_bits, uint32_t dim_pad,
                                                       uint32_t wave_bytes) {
    extern __shared__ __align__(16) unsigned char lds[];
    const uint32_t lane = threadIdx.x & (kWave - 1);
    const uint32_t wid = uniform(threadIdx.x / kWave);
    float* pmin = reinterpret_cast<float*>(lds);
    float* pscale = pmin + dim_pad;
    for (uint32_t i = threadIdx.x; i < ix.dim; i += blockDim.x) {
        pmin[i] = ix.sq8_min[i];
        pscale[i] = ix.sq8_scale[i];
    }
    unsigned char* wb = lds + 8u * dim_pad;
    WaveLds w;
    w.qv = reinterpret_cast<float*>(wb);
    w.stage = reinterpret_cast<uint64_t*>(wb + 4u * dim_pad);
    w.bids = reinterpret_cast<uint32_t*>(w.stage + EF_CAP);
    w.cache = w.bids + kWave;
    w.cnt = w.cache + (1u << cache_bits);
    SplitLds* sp = reinterpret_cast<SplitLds*>(wb + wave_bytes);
    __syncthreads();
    // a float launch after an integer launch that left nothing: done (the
    // same value for every wave of the workgroup)
    if (job.skip_int && job.left && uniform(*job.left) == 0u) return;
    const uint32_t q = blockIdx.x;
    if (wid == 0) {
        if (q < job.nq)
            search_one<EF_CAP, EVAL_FLOAT, W>(ix, job, q, pmin, pscale, w, cache_bits, lane,
                                              nullptr, sp);
        if (lane == 0) sp->go = 0u;
        __syncthreads();
    } else {
        for (;;) {
            __syncthreads();
            if (uniform(sp->go) == 0u) break;
            eval_share<W, split_nb<EF_CAP, W>()>(ix, w.qv, pmin, pscale, w.bids, uniform(sp->m),
                                              wid, lane, sp->keys);
            __syncthreads();
        }
    }
}

// Latency mode at depth 4: wave 1 computes, for the best unexpanded pool
// entry that is neither the node wave 0 expands nor held by a slot, the keys
// of all of its neighbours (the batch kernels' evaluation code, so the same
// bits), and publishes them in the next slot (round robin).  Wave 0 never
// waits for it.  Stops on wave 0's done flag or after kPrefetchRounds rounds.
template <int EVAL>
__device__ void spec_ahead(const DevIndex& ix, const uint64_t* stage, uint32_t ef_cap,
                           const uint32_t* done, SpecLds* sp, const float* qv,
                           const float* pmin, const float* pscale, uint32_t lane) {
    const uint32_t deg = ix.deg;
    if (deg > kWave) return;
    uint32_t r = 0;
    for (; r < kPrefetchRounds; ++r) {
        if (__hip_atomic_load(done, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP) != 0u) return;
        if (__hip_atomic_load(&sp->qready, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_WORKGROUP) != 0u)
            break;
        __builtin_amdgcn_s_sleep(1);
    }
    const uint32_t qq = uniform(sp->qq);
    const uint32_t* qp = reinterpret_cast<const uint32_t*>(qv);
    uint32_t next = 0;
    for (; r < kPrefetchRounds; ++r) {
        if (__hip_atomic_load(done, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP) != 0u) break;
        uint64_t key = ~0ull;
        if (lane < ef_cap)
            key = __hip_atomic_load(stage + lane, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
        const uint32_t id = key_id(key);
        const uint32_t cur =
            uniform(__hip_atomic_load(&sp->cur, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP));
        bool open = (key & 1ull) == 0ull && id < ix.n && id != cur;
#pragma unroll
        for (uint32_t j = 0; j < kSpecSlots; ++j)
            open = open && id != uniform(sp->slot[j].tag);
        const uint64_t m = __ballot(open);
        if (m == 0) {
            __builtin_amdgcn_s_sleep(1);
            continue;
        }
        const uint32_t x = uniform(__shfl(id, static_cast<int>(__builtin_ctzll(m))));
        SpecSlot* sl = &sp->slot[next];
        next = (next + 1u) % kSpecSlots;
        if (lane == 0) __hip_atomic_store(&sl->tag, kInvalid, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
        uint32_t nid = kInvalid;
        if (lane < deg) {
            nid = ix.nbr[static_cast<size_t>(x) * deg + lane];
            sl->ids[lane] = nid;
        }
        for (uint32_t cb = 0; cb < deg; cb += 32) {
            // rows of ids past the index evaluate node x instead (never used)
            if (lane >= cb && lane < cb + 32u) sp->bids[lane - cb] = nid < ix.n ? nid : x;
            wave_sync();
            uint64_t k;
            uint32_t lg;
            if constexpr (EVAL == EVAL_INT) {
                k = eval_fresh_int<8>(ix, qp, qq, sp->bids, 32u, lane);
                lg = 1u;
            } else if constexpr (EVAL == EVAL_INT128) {
                k = eval_fresh_int<4>(ix, qp, qq, sp->bids, 32u, lane);
                lg = 1u;
            } else {
                k = eval_fresh(ix, qv, pmin, pscale, sp->bids, 32u, lane);
                lg = fresh_lanes_log2(32u);
            }
            if ((lane & ((1u << lg) - 1u)) == 0) sl->keys[cb + (lane >> lg)] = k;
            wave_sync();
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
        if (lane == 0) __hip_atomic_store(&sl->tag, x, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    }
}

template <int EF_CAP, int EVAL>
__global__ void __launch_bounds__(128) h1_spec_kernel(DevIndex ix, SearchJob job,
                                                      uint32_t cache_bits, uint32_t dim_pad,
                                                      uint32_t wave_bytes, uint32_t depth) {
    extern __shared__ __align__(16) unsigned char lds[];
    const uint32_t lane = threadIdx.x & (kWave - 1);
    const uint32_t wid = uniform(threadIdx.x / kWave);
    constexpr bool kParams = EVAL == EVAL_FLOAT;
    float* pmin = nullptr;
    float* pscale = nullptr;
    if constexpr (kParams) {
        pmin = reinterpret_cast<float*>(lds);
        pscale = pmin + dim_pad;
        for (uint32_t i = threadIdx.x; i < ix.dim; i += blockDim.x) {
            pmin[i] = ix.sq8_min[i];
            pscale[i] = ix.sq8_scale[i];
        }
    }
    unsigned char* wb = lds + (kParams ? 8u * dim_pad : 0u);
    WaveLds w;
    w.qv = reinterpret_cast<float*>(wb);
    w.stage = reinterpret_cast<uint64_t*>(wb + 4u * dim_pad);
    w.bids = reinterpret_cast<uint32_t*>(w.stage + EF_CAP);
    w.cache = w.bids + kWave;
    w.cnt = w.cache + (1u << cache_bits);
    uint32_t* done = reinterpret_cast<uint32_t*>(wb + wave_bytes);
    SpecLds* sp = reinterpret_cast<SpecLds*>(wb + wave_bytes + 16u);
    for (uint32_t i = threadIdx.x; i < static_cast<uint32_t>(EF_CAP); i += blockDim.x)
        w.stage[i] = ~0ull;
    if (threadIdx.x < kSpecSlots) sp->slot[threadIdx.x].tag = kInvalid;
    if (threadIdx.x == 0) {
        done[0] = 0u;
        sp->qready = 0u;
        sp->cur = kInvalid;
    }
    __syncthreads();
    (void)depth;
    const uint32_t q = blockIdx.x;
    if (wid == 0) {
        if (q < job.nq)
            search_one<EF_CAP, EVAL, 1, true>(ix, job, q, pmin, pscale, w, cache_bits, lane,
                                              nullptr, nullptr, sp);
        if (lane == 0) __hip_atomic_store(done, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    } else {
        spec_ahead<EVAL>(ix, w.stage, static_cast<uint32_t>(EF_CAP), done, sp, w.qv, pmin, pscale,
                         lane);
    }
}

__global__ void reset_counter_kernel(uint32_t* counter, uint32_t n) {
    for (uint32_t i = threadIdx.x; i < n; i += blockDim.x) counter[i] = 0u;
}

using KernelFn = void (*)(DevIndex, SearchJob, uint32_t, uint32_t, uint32_t);

// occ: 8 or 6 waves per SIMD (64 / 80 VGPRs); -1 uncapped; 0 the default, at
// least 4 waves per SIMD (128 VGPRs) up to the 256 class and uncapped above.
template <int EF_CAP, int EVAL>
KernelFn kernel_for(int occ) {
    if (occ == 8) return &h1_search_kernel_w8<EF_CAP, EVAL>;
    if (occ == 6) return &h1_search_kernel_w6<EF_CAP, EVAL>;
    if (occ == -1) return &h1_search_kernel<EF_CAP, EVAL>;
    if constexpr (EF_CAP <= 256)
        return &h1_search_kernel_w4<EF_CAP, EVAL>;
    else
        return &h1_search_kernel<EF_CAP, EVAL>;
}

template <int EVAL>
KernelFn select_for_eval(int ef_cap, int occ) {
    switch (ef_cap) {
        case 64: return kernel_for<64, EVAL>(occ);
        case 128: return kernel_for<128, EVAL>(occ);
        case 192: return kernel_for<192, EVAL>(occ);
        case 256: return kernel_for<256, EVAL>(occ);
        case 512: return kernel_for<512, EVAL>(occ);
        case 1024: return kernel_for<1024, EVAL>(occ);
        default: return nullptr;
    }
}

using LatencyFn = void (*)(DevIndex, SearchJob, uint32_t, uint32_t, uint32_t, uint32_t);

template <int EVAL>
LatencyFn latency_for_eval(int ef_cap) {
    switch (ef_cap) {
        case 64: return &h1_latency_kernel<64, EVAL>;
        case 128: return &h1_latency_kernel<128, EVAL>;
        case 192: return &h1_latency_kernel<192, EVAL>;
        case 256: return &h1_latency_kernel<256, EVAL>;
        case 512: return &h1_latency_kernel<512, EVAL>;
        case 1024: return &h1_latency_kernel<1024, EVAL>;
        default: return nullptr;
    }
}

template <int EVAL>
LatencyFn spec_for_eval(int ef_cap) {
    switch (ef_cap) {
        case 64: return &h1_spec_kernel<64, EVAL>;
        case 128: return &h1_spec_kernel<128, EVAL>;
        case 192: return &h1_spec_kernel<192, EVAL>;
        case 256: return &h1_spec_kernel<256, EVAL>;
        case 512: return &h1_spec_kernel<512, EVAL>;
        case 1024: return &h1_spec_kernel<1024, EVAL>;
        default: return nullptr;
    }
}

LatencyFn select_latency(const LaunchCfg& cfg) {
    if (cfg.prefetch == 4) {
        if (cfg.eval == EVAL_INT128) return spec_for_eval<EVAL_INT128>(cfg.ef_cap);
        return cfg.eval == EVAL_INT ? spec_for_eval<EVAL_INT>(cfg.ef_cap)
                                    : spec_for_eval<EVAL_FLOAT>(cfg.ef_cap);
    }
    if (cfg.eval == EVAL_INT128) return latency_for_eval<EVAL_INT128>(cfg.ef_cap);
    return cfg.eval == EVAL_INT ? latency_for_eval<EVAL_INT>(cfg.ef_cap)
                                : latency_for_eval<EVAL_FLOAT>(cfg.ef_cap);
}

template <int W>
KernelFn split_for_w(int ef_cap) {
    switch (ef_cap) {
        case 64: return &h1_split_kernel<64, W>;
        case 128: return &h1_split_kernel<128, W>;
        case 192: return &h1_split_kernel<192, W>;
        case 256: return &h1_split_kernel<256, W>;
        case 512: return &h1_split_kernel<512, W>;
        case 1024: return &h1_split_kernel<1024, W>;
        default: return nullptr;
    }
}

KernelFn select_split(const LaunchCfg& cfg) {
    if (cfg.eval != EVAL_FLOAT) return nullptr;
    if (cfg.split == 2) return split_for_w<2>(cfg.ef_cap);
    if (cfg.split == 4) return split_for_w<4>(cfg.ef_cap);
    return nullptr;
}

KernelFn select_kernel(const LaunchCfg& cfg) {
    if (cfg.eval == EVAL_INT128) return select_for_eval<EVAL_INT128>(cfg.ef_cap, cfg.occupancy);
    return cfg.eval == EVAL_INT ? select_for_eval<EVAL_INT>(cfg.ef_cap, cfg.occupancy)
                                : select_for_eval<EVAL_FLOAT>(cfg.ef_cap, cfg.occupancy);
}

void geometry(const DevIndex& ix, const LaunchCfg& cfg, uint32_t& dim_pad, uint32_t& wave_bytes) {
    dim_pad = (ix.dim + 3u) & ~3u;
    size_t wb = 4u * dim_pad + 8u * static_cast<size_t>(cfg.ef_cap) + 4u * kWave +
                (4u << cfg.cache_bits) +
                (cfg.ef_cap >= kRankedSq8 ? 2u * static_cast<size_t>(cfg.ef_cap) : 0u);
    wave_bytes = static_cast<uint32_t>((wb + 15u) & ~static_cast<size_t>(15u));
}

}  // namespace

size_t search_lds_bytes(const DevIndex& ix, const LaunchCfg& cfg) {
    uint32_t dim_pad, wave_bytes;
    geometry(ix, cfg, dim_pad, wave_bytes);
    return (cfg.eval == EVAL_FLOAT ? 8u * static_cast<size_t>(dim_pad) : 0u) +
           static_cast<size_t>(cfg.waves_per_group) * wave_bytes;
}

int search_full_grid(const DevIndex& ix, const LaunchCfg& cfg) {
    int dev = 0, cus = 0;
    if (hipGetDevice(&dev) != hipSuccess) return 1;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
        return 1;
    KernelFn k = select_kernel(cfg);
    if (!k) return 1;
    int per_cu = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, reinterpret_cast<const void*>(k),
                                                     kWave * cfg.waves_per_group,
                                                     search_lds_bytes(ix, cfg)) != hipSuccess ||
        per_cu < 1)
        per_cu = 1;
    return per_cu * cus;
}

int search_groups(const DevIndex& ix, const SearchJob& job, const LaunchCfg& cfg) {
    int groups = cfg.groups > 0 ? cfg.groups : search_full_grid(ix, cfg);
    const int need = static_cast<int>((job.nq + cfg.waves_per_group - 1) / cfg.waves_per_group);
    if (groups > need) groups = need > 0 ? need : 1;
    return groups;
}

hipError_t launch_search(const DevIndex& ix, const SearchJob& job, const LaunchCfg& cfg,
                         hipStream_t stream) {
    if (cfg.split > 1 && cfg.eval == EVAL_FLOAT && !job.visit_bits) {
        // split evaluation: one query per workgroup of cfg.split waves
        KernelFn k = select_split(cfg);
        if (!k || cfg.cache_bits < 4 || cfg.cache_bits > 14) return hipErrorInvalidValue;
        uint32_t dim_pad, wave_bytes;
        geometry(ix, cfg, dim_pad, wave_bytes);
        const size_t lds = 8u * static_cast<size_t>(dim_pad) + wave_bytes + sizeof(SplitLds);
        if (lds > 64 * 1024) {
            hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(k),
                                               hipFuncAttributeMaxDynamicSharedMemorySize,
                                               static_cast<int>(lds));
            if (e != hipSuccess) return e;
        }
        if (job.nq == 0) return hipSuccess;
        hipLaunchKernelGGL(k, dim3(job.nq), dim3(kWave * cfg.split), lds, stream, ix, job,
                           static_cast<uint32_t>(cfg.cache_bits), dim_pad, wave_bytes);
        return hipGetLastError();
    }
    if (cfg.prefetch > 0 && job.nq <= sq8_latency_max(ix.dim) && !job.visit_bits) {
        // latency mode: one query per workgroup (search wave + prefetch wave)
        LatencyFn k = select_latency(cfg);
        if (!k || cfg.cache_bits < 4 || cfg.cache_bits > 14 || cfg.prefetch > 4)
            return hipErrorInvalidValue;
        uint32_t dim_pad, wave_bytes;
        geometry(ix, cfg, dim_pad, wave_bytes);
        const size_t lds = (cfg.eval == EVAL_FLOAT ? 8u * static_cast<size_t>(dim_pad) : 0u) +
                           wave_bytes + 16u + (cfg.prefetch == 4 ? sizeof(SpecLds) : 0u);
        if (lds > 64 * 1024) {
            hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(k),
                                               hipFuncAttributeMaxDynamicSharedMemorySize,
                                               static_cast<int>(lds));
            if (e != hipSuccess) return e;
        }
        hipLaunchKernelGGL(k, dim3(job.nq), dim3(2 * kWave), lds, stream, ix, job,
                           static_cast<uint32_t>(cfg.cache_bits), dim_pad, wave_bytes,
                           static_cast<uint32_t>(cfg.prefetch));
        return hipGetLastError();
    }
    KernelFn k = select_kernel(cfg);
    if (!k || cfg.cache_bits < 4 || cfg.cache_bits > 14 || cfg.waves_per_group < 1 ||
        cfg.waves_per_group > 4)
        return hipErrorInvalidValue;
    uint32_t dim_pad, wave_bytes;
    geometry(ix, cfg, dim_pad, wave_bytes);
    const size_t lds = search_lds_bytes(ix, cfg);
    if (lds > 64 * 1024) {
        hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(k),
                                           hipFuncAttributeMaxDynamicSharedMemorySize,
                                           static_cast<int>(lds));
        if (e != hipSuccess) return e;
    }
    const int groups = search_groups(ix, job, cfg);
    hipLaunchKernelGGL(k, dim3(groups), dim3(kWave * cfg.waves_per_group), lds, stream, ix, job,
                       static_cast<uint32_t>(cfg.cache_bits), dim_pad, wave_bytes);
    return hipGetLastError();
}

hipError_t launch_reset_counter(uint32_t* counter, hipStream_t stream, uint32_t n) {
    hipLaunchKernelGGL(reset_counter_kernel, dim3(1), dim3(64), 0, stream, counter, n);
    return hipGetLastError();
}

}  // namespace lqg
