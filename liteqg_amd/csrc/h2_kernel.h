// h2_kernel.h -- device code of the reference's FastScan RaBitQ query
// path, QuantizedGraph::search_qg (qg.hpp:252-287): query preparation
// (QGQuery::query_prepare, qg_query.hpp:28-43), FastScan estimates of every
// neighbour of an expanded node (QGScanner::scan_neighbors,
// qg_scanner.hpp:99-155), the exact L2 distance of the node itself
// (space/l2.hpp:16-34) kept in a top-k ResultBuffer, update_results
// (qg.hpp:395-418), and the sidecar build that fills the codes and factors
// the scan reads (quantization/rabitq.hpp:23-108).
//
// One wavefront walks one query.
//
// Query preparation runs in LDS: the FHT rotation (sign multiply, then the
// unnormalised Walsh-Hadamard butterflies with strides 1, 2, ..., P/2, the
// operation order of utils/rotator.hpp:88-114 and FFHT), the [lo, hi] range
// seeded with 0, 6-bit quantisation and the 16-entry uint8 LUT of each
// 4-dimension codebook.  The two constant divisions are evaluated as the
// reference's -Ofast build folds them: width = (hi - lo) * fl(1/63) and
// 1/width = fl(1/fl(1/63)) / (hi - lo); q = lround(double(f) + 0.5).
//
// SearchBuffer (utils/buffer.hpp:11-78) semantics: entries sorted by
// estimate, a new entry placed before equal ones, the last dropped when full,
// an entry skipped when the buffer is full and its estimate is above the
// last, duplicates of an id allowed, pop = first unchecked entry.  Inserting
// the neighbours of one node one by one under these rules leaves the ef
// smallest entries of (buffer + all inserted) in the order
//   (estimate ascending, later insertion first)
// so the kernel keeps the buffer as a register pool sorted by the 64-bit key
//   sortable(estimate) << 32 | (2^31 - 1 - seq) << 1 | checked
// (seq = insertion number) and merges each node's neighbours in one pass, as
// the SQ8 kernel does.  The visited set (nodes expanded, plus the nodes
// update_results adds) is exact: an LDS open-addressing table, moved to a
// per-wave table in device memory if it passes three quarters of its capacity.
//
// FastScan: a lane holds the codes of one neighbour (deg >= 64) or half of
// them (deg = 32) and adds LUT bytes looked up in LDS; sums are taken mod
// 2^16 and read back as int16, as the AVX-512 code does (qg_scanner.hpp:
// 117-135); est = (f * fac_dq) * width + fma(fac_vq, lo, triple_x + sqr_y), the
// operation order of the reference build (qg_scanner.hpp:23-48 as compiled).
#pragma once

#include <math.h>

#include "lqg_internal.h"
#include "wave.h"

namespace lqg {

struct Layout {
    uint32_t qv, tb, qb, lut, pl, st, hv, cnt, wave_bytes;
    int has_qv;
};

// Pool classes from this capacity on admit a pass's candidates by binary
// search and a prefix sum (h2_search_one) instead of comparing each candidate
// with every register slot of the pool.  200k x 960 degree 64, 1,000 queries:
// ef U{64..1024} 4.57 -> 3.72 ms, ef U{64..512} 2.74 -> 2.49 ms.
constexpr int kRankedPool = 512;

__host__ __device__ inline uint32_t align16(uint32_t x) { return (x + 15u) & ~15u; }

// Bytes of a query's LUT in LDS (layout: lut_pos below).
__host__ __device__ inline uint32_t lut_bytes(uint32_t p) { return 4u * p + 32u; }

// Per-wave LDS layout.  with_qv: keep the query in LDS (the probe kernel, and
// searches with dim > 16 * kRegJ, whose exact distances read dims past the
// registers' from there).
__host__ __device__ inline Layout h2_layout(uint32_t dim, uint32_t p, uint32_t ef_cap,
                                            uint32_t hash_bits, bool with_qv) {
    Layout l;
    uint32_t o = 0;
    l.has_qv = with_qv ? 1 : 0;
    l.qv = o;   o += with_qv ? align16(4 * dim) : 0u;
    l.lut = o;  o += align16(lut_bytes(p));
    l.pl = o;   o += align16(p);  // bit planes: 8 words per 32-bit code word
    // one region for the preparation buffers (tb, qb) and, once the LUT is
    // built, the pool merge staging (ef_cap records and a spare)
    const uint32_t prep = align16(4 * p) + align16(p);
    const uint32_t stage = 16u * (ef_cap + 1);
    l.tb = o;
    l.qb = o + align16(4 * p);
    l.st = o;
    o += prep > stage ? prep : stage;
    l.hv = o;   o += 4u << hash_bits;
    // ranked admission (pool classes from kRankedPool on): one counter per
    // pool position
    l.cnt = o;  o += ef_cap >= static_cast<uint32_t>(kRankedPool) ? 4u * ef_cap : 0u;
    l.wave_bytes = o;
    return l;
}

namespace {

constexpr uint32_t kEmpty = 0xFFFFFFFFu;
constexpr float kInv63 = 1.0f / 63.0f;  // width = (hi - lo) * fl(1/63)
constexpr float kC63 = 1.0f / kInv63;   // 1/width = fl(1/fl(1/63)) / (hi - lo)

// Estimate bits that sort like the floats (-0 taken as +0, as a float
// comparison does).
__device__ __forceinline__ uint32_t sortable(float x) {
    // x + 0 turns -0 into +0 and leaves every other value as it is
    const uint32_t u = __float_as_uint(x + 0.0f);
    const uint32_t m = static_cast<uint32_t>(static_cast<int32_t>(u) >> 31);
    return u ^ (m | 0x80000000u);
}

__device__ __forceinline__ uint64_t sb_key(float est, uint32_t seq) {
    return (static_cast<uint64_t>(sortable(est)) << 32) |
           (static_cast<uint64_t>(0x7FFFFFFFu - seq) << 1);
}

struct H2Lds {
    float* qv;        // query (dim), only where dim > 16 * kRegJ (the rest sits in registers)
    float* tb;        // rotated query during preparation; q - x in exact_l2 (p floats)
    uint8_t* qb;      // quantised query (p bytes)
    uint8_t* lut;     // lut_bytes(p): 4p bytes and 2 x 16 bytes of bank skew
    uint4* pl;        // query bit planes (see prepare): 2 uint4 per 32-bit code word
    uint4* st;        // pool merge staging, once the LUT is built (shares tb's memory):
                      // EF_CAP + 1 records {key lo, key hi, id, 0}
    uint32_t* hv;     // visited table (1 << hash_bits)
    uint32_t* cnt;    // ranked admission: EF_CAP counters, zero between passes
};

__device__ __forceinline__ H2Lds carve(unsigned char* base, const Layout& l) {
    H2Lds w;
    w.qv = l.has_qv ? reinterpret_cast<float*>(base + l.qv) : nullptr;
    w.tb = reinterpret_cast<float*>(base + l.tb);
    w.qb = base + l.qb;
    w.lut = base + l.lut;
    w.pl = reinterpret_cast<uint4*>(base + l.pl);
    w.st = reinterpret_cast<uint4*>(base + l.st);
    w.hv = reinterpret_cast<uint32_t*>(base + l.hv);
    w.cnt = reinterpret_cast<uint32_t*>(base + l.cnt);
    return w;
}

// ---------------------------------------------------------------------------
// LUT layout for the scan.  A code byte holds codebook 2g (low nibble) and
// 2g + 1 (high nibble) of codebook pair g.  Even codebooks: 16 bytes per pair,
// row after row (pair g's entry j at 16g + j); odd codebooks: the same rows
// from 2p + 16 on.  With two lanes per vector (lpv = 2) each lane scans half
// of the pairs, and the rows of the second half sit 16 bytes further on.  The
// LDS serves a wavefront in two 32-lane passes, and two dwords 32 dwords
// apart share a bank: without the skew, the rows the two lanes of a vector
// read at once (pairs g and g + p/16, 2p / 16 bytes apart) share banks, which
// an LDS probe of exactly these lookups measured at 1.9 conflict cycles per
// lookup, against none with the skew.  4p + 32 bytes.
__host__ __device__ inline uint32_t lut_pos(uint32_t p, uint32_t c, uint32_t j, uint32_t lpv) {
    const uint32_t g = c >> 1;
    const uint32_t skew = (lpv == 2u && g >= p / 16u) ? 16u : 0u;
    return (c & 1u) * (2u * p + 16u) + 16u * g + skew + j;
}

// Query preparation (qg_query.hpp:28-43) into LDS: qv (raw query), lut.
struct Prep {
    float lo, width;
    int32_t sumq;
    // 1: the bit-plane sum (plane_sum) equals the LUT scan for this query
    int planes;
};

// In-place unnormalised Walsh-Hadamard transform of t[0..p) in LDS by one
// wave, butterflies (a, b) -> (a + b, a - b) with strides 1, 2, ..., p/2.
__device__ void fht_lds(float* t, uint32_t p, uint32_t lane) {
    for (uint32_t lh = 0; (1u << lh) < p; ++lh) {
        const uint32_t h = 1u << lh;
        for (uint32_t b = lane; b < p / 2; b += kWave) {
            const uint32_t j = ((b >> lh) << (lh + 1)) | (b & (h - 1u));
            const float a = t[j], c = t[j + h];
            t[j] = a + c;
            t[j + h] = a - c;
        }
        wave_sync();
    }
}

// rsrc: the query already rotated (p floats, the GEMM rotation), or null to
// rotate here with the FHT.
__device__ Prep prepare(const H2Dev& ix, const float* qsrc, const float* rsrc, const H2Lds& w,
                        uint32_t lpv, uint32_t lane, bool want_planes = true) {
    const uint32_t dim = ix.dim, p = ix.p;
    if (w.qv)
        for (uint32_t i = lane; i < dim; i += kWave) w.qv[i] = qsrc[i];
    if (rsrc) {
        for (uint32_t i = lane; i < p; i += kWave) w.tb[i] = rsrc[i];
        wave_sync();
    } else {
        for (uint32_t i = lane; i < p; i += kWave) w.tb[i] = i < dim ? qsrc[i] * ix.signs[i] : 0.0f;
        wave_sync();
        fht_lds(w.tb, p, lane);
    }
    // data_range (utils/scalar_quantize.hpp:11-38): min and max seeded with 0
    float lo = 0.0f, hi = 0.0f;
    for (uint32_t i = lane; i < p; i += kWave) {
        const float t = w.tb[i];
        lo = t < lo ? t : lo;
        hi = t > hi ? t : hi;
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        const float a = __shfl_xor(lo, o), b = __shfl_xor(hi, o);
        lo = a < lo ? a : lo;
        hi = b > hi ? b : hi;
    }
    const float width = (hi - lo) * kInv63;
    const float ow = kC63 / (hi - lo);
    // quantize (scalar_quantize.hpp:40-58): q = lround(double(f) + 0.5); for
    // f >= 0 that is floor(f + 1) (exact in double).  A NaN or infinite f
    // (hi == lo) gives 0, the low byte of the x86 conversion's result.
    int32_t sq = 0;
    for (uint32_t i = lane; i < p; i += kWave) {
        const float f = (w.tb[i] - lo) * ow;
        uint32_t q = 0;
        if (f <= 1.0e9f) q = static_cast<uint32_t>(floor(static_cast<double>(f) + 1.0)) & 0xFFu;
        w.qb[i] = static_cast<uint8_t>(q);
        sq += static_cast<int32_t>(q);
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) sq += __shfl_xor(sq, o);
    wave_sync();
    // pack_lut_impl (fastscan_impl.hpp:258-270): entry j of codebook c is the
    // uint8 sum of q[4c + 3 - b] over the set bits b of j.  Stored at
    // lut_pos(p, c, j, lpv) (the scan's layout).
    const uint32_t ncb = p / 4;
    for (uint32_t idx = lane; idx < 16u * ncb; idx += kWave) {
        const uint32_t c = idx >> 4, j = idx & 15u;
        uint32_t s = 0;
#pragma unroll
        for (uint32_t b = 0; b < 4; ++b)
            if ((j >> b) & 1u) s += w.qb[4 * c + 3 - b];
        w.lut[lut_pos(p, c, j, lpv)] = static_cast<uint8_t>(s & 0xFFu);
    }
    // Bit planes.  Code bit i of 32-bit code word W (little-endian bytes
    // 4W .. 4W + 3) stands for dim d(W, i) = 8B + r, B = 4W + i / 8, with
    // r = 3 - m for bit m = i % 8 below 4 (codebook 2B, bit 3 = dim 8B) and
    // r = 11 - m above (codebook 2B + 1).  Plane k of word W holds bit k of
    // q[d(W, i)] at bit i, at w.pl word 8W + k (k < 7; q <= 64).  Then the
    // scan's sum, the LUT entries of a code added up, is
    //   sum_k 2^k popcount(code word AND plane k word), over the words,
    // as long as no LUT entry wrapped: an entry is the uint8 sum of up to
    // four q, so only a codebook whose four q add up to 256 or more (all four
    // 64) differs; such a query keeps the LUT scan.  Lane l takes code bit l
    // of the word pair (2t, 2t + 1): dim 64t + 8 (l / 8) + r(l % 8); one ballot
    // per plane gives the pair's two words.
    int planes = 0;
    if (want_planes) {
        bool wrap = false;
        const uint32_t m = lane & 7u;
        const uint32_t r = m < 4u ? 3u - m : 11u - m;
        for (uint32_t t = 0; 64u * t < p; ++t) {
            const uint32_t q = w.qb[64u * t + (lane & ~7u) + r];
            // the codebook of this bit: lanes 8g .. 8g + 3 or 8g + 4 .. 8g + 7
            uint32_t cs = q + static_cast<uint32_t>(__shfl_xor(static_cast<int>(q), 1));
            cs += static_cast<uint32_t>(__shfl_xor(static_cast<int>(cs), 2));
            wrap = wrap || cs > 255u || q > 127u;
            uint32_t* pw = reinterpret_cast<uint32_t*>(w.pl) + 16u * t;
#pragma unroll
            for (uint32_t k = 0; k < 7u; ++k) {
                const uint64_t b = __ballot((q >> k) & 1u);
                if (lane == 0) {
                    pw[k] = static_cast<uint32_t>(b);
                    pw[8u + k] = static_cast<uint32_t>(b >> 32);
                }
            }
            if (lane == 0) {
                pw[7] = 0u;
                pw[15] = 0u;
            }
        }
        planes = __ballot(wrap) == 0 ? 1 : 0;
        wave_sync();
    }
    Prep res;
    res.lo = lo;
    res.width = width;
    res.sumq = sq;
    res.planes = planes;
    return res;
}

// The same distance with the partials in registers: lane p < 16 runs the FMA
// chain over dims p, p + 16, ... (the first kRegJ of them from xr / qr, which
// the caller loaded with load_partials, the rest from the row and w.qv), the
// tree is taken with DPP row shifts (lane p adds lane p + 8, p + 4, p + 2,
// p + 1 of its row: the pairs of reduce_l2), and lane 0's sum is broadcast.
// Lanes 16..63 repeat the work of lane (lane & 15), so no load or step needs
// a lane mask; only lanes 0..15 feed the tree.  Registers past the chain
// length hold 0, and fma(0, 0, s) = s, so their steps change nothing.
constexpr int kRegJ = 8;

__device__ __forceinline__ void load_partials(float (&r)[kRegJ], const float* src, uint32_t dim,
                                              uint32_t lane) {
    const uint32_t nj = dim / 16, p = lane & 15u;
    const gptr<float> g = uniform_gptr<float>(src);
#pragma unroll
    for (int u = 0; u < kRegJ; ++u) {
        r[u] = 0.0f;
        if (static_cast<uint32_t>(u) < nj) r[u] = g[p + 16u * static_cast<uint32_t>(u)];
    }
}

__device__ __forceinline__ float dpp_shl(float x, int ctrl_n) {
    // row_shl:n, lanes past the row end read 0
    int r = 0;
    switch (ctrl_n) {
        case 8: r = __builtin_amdgcn_update_dpp(0, __float_as_int(x), 0x108, 0xF, 0xF, false); break;
        case 4: r = __builtin_amdgcn_update_dpp(0, __float_as_int(x), 0x104, 0xF, 0xF, false); break;
        case 2: r = __builtin_amdgcn_update_dpp(0, __float_as_int(x), 0x102, 0xF, 0xF, false); break;
        default: r = __builtin_amdgcn_update_dpp(0, __float_as_int(x), 0x101, 0xF, 0xF, false); break;
    }
    return __int_as_float(r);
}

__device__ __forceinline__ float exact_l2_regs(const float (&xr)[kRegJ], const float (&qr)[kRegJ],
                                               const float* row, const H2Lds& w, uint32_t dim,
                                               uint32_t lane) {
    const uint32_t nj = dim / 16;  // chain length of every partial
    const uint32_t p = lane & 15u;
    float s = 0.0f;
#pragma unroll
    for (int u = 0; u < kRegJ; ++u) {
        const float t = qr[u] - xr[u];
        s = fmaf(t, t, s);
    }
    // Dims past the registers' (D > 16 kRegJ), in blocks of 16 per partial.
    // Up to four blocks per round trip: lane row h (lanes 16h .. 16h + 15)
    // loads block h of the round, so all four are in flight together; block
    // b then reaches lane p from lane p + 16b (ds_bpermute), and the FMAs run
    // in dim order.  Steps past the chain (a last block's tail) take t = 0:
    // fma(0, 0, s) = s.
    const uint32_t h = lane >> 4;
    for (uint32_t j0 = kRegJ; j0 < nj; j0 += 64u) {
        const uint32_t jb = j0 + 16u * h;
        float xv[16];
#pragma unroll
        for (uint32_t u = 0; u < 16u; ++u) xv[u] = jb + u < nj ? row[p + 16u * (jb + u)] : 0.0f;
#pragma unroll
        for (uint32_t b = 0; b < 4u; ++b) {
            const uint32_t jj = j0 + 16u * b;
            if (jj >= nj) break;
            const int src = static_cast<int>(4u * (p + 16u * b));
#pragma unroll
            for (uint32_t u = 0; u < 16u; ++u) {
                const float x = b == 0 ? xv[u]
                                       : __int_as_float(__builtin_amdgcn_ds_bpermute(
                                             src, __float_as_int(xv[u])));
                const bool ok = jj + u < nj;
                const float t = w.qv[p + 16u * min(jj + u, nj - 1u)] - x;
                s = fmaf(ok ? t : 0.0f, ok ? t : 0.0f, s);
            }
        }
    }
    s = s + dpp_shl(s, 8);
    s = s + dpp_shl(s, 4);
    s = s + dpp_shl(s, 2);
    s = s + dpp_shl(s, 1);
    float r = __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(s)));
    for (uint32_t i = nj * 16; i < dim; ++i) {
        const float t = w.qv[i] - row[i];
        r = fmaf(t, t, r);
    }
    return r;
}

// ---------------------------------------------------------------------------
// Exact visited set (nodes expanded, and the nodes update_results adds):
// open addressing over 2^bits slots (empty = 0xFFFFFFFF), in LDS or, for a
// query whose set would pass vis_limit of its LDS table, in a per-wave table in
// device memory (the query is then searched again from the start with that
// table: GLOBAL = true).  An id's probe sequence starts at the first slot of
// its 4-slot bucket and runs through consecutive slots.  Nothing is ever
// removed, so the filled slots of a bucket are a prefix of it: a lane tests a
// whole bucket with one 16-byte LDS read and stops at a bucket whose last
// slot is empty, and a uniform probe reads 64 consecutive slots in one
// instruction, one per lane.  The device table is accessed with atomics, so
// every lane sees every insertion.
template <bool GLOBAL>
struct Visited {
    uint32_t* tab;  // 2^bits slots (LDS, or this wave's device table)
    uint32_t bits, count;
};

__device__ __forceinline__ uint32_t vbucket(uint32_t id, uint32_t bits) {
    return ((id * 2654435761u) >> (32u - bits)) & ~3u;
}

__device__ __forceinline__ uint32_t g_load(const uint32_t* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

enum : uint32_t { kVisNew = 0, kVisSeen = 1, kVisFull = 2 };

// Most ids a table of 2^bits slots takes: three quarters of it, so that every
// probe meets an empty slot within a few buckets.  (At half, the 2^11-slot
// table of the 1,024-entry pool class turned away the queries at ef near
// 1,024, which expand about 1,030 nodes, and they were searched twice.)
__host__ __device__ inline uint32_t vis_limit(uint32_t bits) { return 3u << (bits - 2u); }

// For an id equal on every lane: kVisSeen if it is in the set; otherwise
// inserts it (kVisNew), or, if the set would pass vis_limit, inserts nothing
// (kVisFull).
template <bool GLOBAL>
__device__ uint32_t v_test_set(Visited<GLOBAL>& v, uint32_t id, uint32_t lane) {
    const uint32_t mask = (1u << v.bits) - 1u;
    uint32_t h = vbucket(id, v.bits);
    uint32_t fe = 0, slot = 0;
    for (;;) {
        slot = (h + lane) & mask;
        const uint32_t e = GLOBAL ? g_load(&v.tab[slot]) : v.tab[slot];
        const uint64_t hit = __ballot(e == id), emp = __ballot(e == kEmpty);
        if (hit | emp) {
            const uint32_t fh = hit ? static_cast<uint32_t>(__builtin_ctzll(hit)) : 64u;
            fe = emp ? static_cast<uint32_t>(__builtin_ctzll(emp)) : 64u;
            if (fh < fe) return kVisSeen;
            break;
        }
        h = (h + kWave) & mask;
    }
    if (v.count + 1u > vis_limit(v.bits)) return kVisFull;
    if (GLOBAL) {
        // the first empty slot may have been taken meanwhile by no one else
        // (the table is this wave's), so the CAS lands there
        if (lane == fe) atomicCAS(&v.tab[slot], kEmpty, id);
        __threadfence();
    } else {
        if (lane == fe) v.tab[slot] = id;
    }
    wave_sync();
    ++v.count;
    return kVisNew;
}

// For each lane with `active` set: whether its id is in the set (false on
// the other lanes).  Every lane reads each round (inactive lanes an in-range
// bucket), so the loop has no divergent branch.
template <bool GLOBAL>
__device__ __forceinline__ bool v_contains_lanes(const Visited<GLOBAL>& v, uint32_t id,
                                                 bool active) {
    bool found = false;
    const uint32_t bmask = (1u << (v.bits - 2u)) - 1u;
    uint32_t b = vbucket(id, v.bits) >> 2;
    for (;;) {
        uint4 e;
        if (GLOBAL) {
            const uint32_t* t = v.tab + 4u * b;
            e = make_uint4(g_load(t), g_load(t + 1), g_load(t + 2), g_load(t + 3));
        } else {
            e = reinterpret_cast<const uint4*>(v.tab)[b];
        }
        const bool f = active && (e.x == id || e.y == id || e.z == id || e.w == id);
        found = found || f;
        active = active && !f && e.w != kEmpty;
        if (__ballot(active) == 0) break;
        b = (b + 1u) & bmask;
    }
    return found;
}

// ---------------------------------------------------------------------------
// ResultBuffer (utils/buffer.hpp:81-120) in registers: entry j is slot j / 64
// of lane j % 64 of rd / ri, KS slots (k <= 64 * KS), sorted by exact
// distance, empty entries +inf.  insert (buffer.hpp:84-96): placed before
// equal distances, skipped when full and above the last; the entry pushed
// past k is dropped.
template <int KS>
struct ResBuf {
    float rd[KS];
    uint32_t ri[KS];
    uint32_t size;
    float tail;  // distance of entry k - 1 once full
};

template <int KS>
__device__ __forceinline__ void rb_init(ResBuf<KS>& r) {
#pragma unroll
    for (int s = 0; s < KS; ++s) {
        r.rd[s] = INFINITY;
        r.ri[s] = 0u;
    }
    r.size = 0;
    r.tail = INFINITY;
}

__device__ __forceinline__ float wave_shr1(float x) {
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(x), 0x138, 0xF, 0xF, false));
}
__device__ __forceinline__ uint32_t wave_shr1(uint32_t x) {
    return static_cast<uint32_t>(
        __builtin_amdgcn_update_dpp(0, static_cast<int>(x), 0x138, 0xF, 0xF, false));
}

template <int KS>
__device__ __forceinline__ void rb_insert(ResBuf<KS>& r, uint32_t kq, uint32_t id, float x,
                                          uint32_t lane) {
    if (kq == 0 || (r.size == kq && x > r.tail)) return;
    // entries below x among the first k (entries past k are stale copies)
    uint32_t pos = 0;
#pragma unroll
    for (int s = 0; s < KS; ++s) {
        const uint32_t j = static_cast<uint32_t>(s * kWave) + lane;
        pos += static_cast<uint32_t>(__popcll(__ballot(j < kq && r.rd[s] < x)));
    }
    // entries at or after pos move up one place (lane 63 of a slot into lane
    // 0 of the next), then x goes to pos
#pragma unroll
    for (int s = KS - 1; s >= 0; --s) {
        const uint32_t j = static_cast<uint32_t>(s * kWave) + lane;
        float sd = wave_shr1(r.rd[s]);
        uint32_t si = wave_shr1(r.ri[s]);
        if (s > 0) {
            const float cd = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(r.rd[s - 1]), 63));
            const uint32_t ci = static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(r.ri[s - 1]), 63));
            sd = lane == 0 ? cd : sd;
            si = lane == 0 ? ci : si;
        }
        r.rd[s] = j > pos ? sd : (j == pos ? x : r.rd[s]);
        r.ri[s] = j > pos ? si : (j == pos ? id : r.ri[s]);
    }
    r.size = min(r.size + 1u, kq);
    if (r.size == kq) {
        const uint32_t t = kq - 1u;
        float tl = r.rd[0];
#pragma unroll
        for (int s = 0; s < KS; ++s)
            if (static_cast<uint32_t>(s) == t / kWave) tl = r.rd[s];
        r.tail = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(tl), static_cast<int>(t % kWave)));
    }
}

// Id of entry t (t < size; uniform t).
template <int KS>
__device__ __forceinline__ uint32_t rb_id(const uint32_t (&ri)[KS], uint32_t t) {
    uint32_t v = ri[0];
#pragma unroll
    for (int s = 0; s < KS; ++s)
        if (static_cast<uint32_t>(s) == t / kWave) v = ri[s];
    return static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(v), static_cast<int>(t % kWave)));
}

// ---------------------------------------------------------------------------
// base + byte J of t in one instruction (SDWA source select).
template <int J>
__device__ __forceinline__ uint32_t add_byte(uint32_t base, uint32_t t) {
    uint32_t r;
    if constexpr (J == 0)
        asm("v_add_u32_sdwa %0, %1, %2 dst_sel:DWORD dst_unused:UNUSED_PAD src0_sel:DWORD "
            "src1_sel:BYTE_0" : "=v"(r) : "v"(base), "v"(t));
    else if constexpr (J == 1)
        asm("v_add_u32_sdwa %0, %1, %2 dst_sel:DWORD dst_unused:UNUSED_PAD src0_sel:DWORD "
            "src1_sel:BYTE_1" : "=v"(r) : "v"(base), "v"(t));
    else if constexpr (J == 2)
        asm("v_add_u32_sdwa %0, %1, %2 dst_sel:DWORD dst_unused:UNUSED_PAD src0_sel:DWORD "
            "src1_sel:BYTE_2" : "=v"(r) : "v"(base), "v"(t));
    else
        asm("v_add_u32_sdwa %0, %1, %2 dst_sel:DWORD dst_unused:UNUSED_PAD src0_sel:DWORD "
            "src1_sel:BYTE_3" : "=v"(r) : "v"(base), "v"(t));
    return r;
}

typedef __attribute__((address_space(3))) const uint8_t lds_u8;

// The two LDS base addresses of a lane's LUT lookups (lut_pos for the lane's
// first codebook pair, 4 CW pairs per lane, `part` of lpv lanes per vector;
// part 1 only where lpv = 2, and then its pairs are the skewed half).
struct LutBase {
    uint32_t even, odd;
};

template <int CW>
__device__ __forceinline__ LutBase lut_base(const uint8_t* lut, uint32_t p, uint32_t part) {
    // the LDS offset of the generic pointer (an address-space cast)
    const uint32_t l = static_cast<uint32_t>(reinterpret_cast<uintptr_t>((lds_u8*)lut));
    const uint32_t first = 16u * (4u * CW * part) + (part ? 16u : 0u);
    return LutBase{l + first, l + 2u * p + 16u + first};
}

// FastScan sum of one lane: CW code words, pairs g0 .. g0 + 4 CW - 1 of the
// lane's LUT bases: per word, the low nibbles index the even rows and the
// high nibbles, shifted down, the odd rows.
template <int CW>
__device__ __forceinline__ uint32_t scan_sum(const uint32_t (&cw)[CW], LutBase lb) {
    uint32_t s = 0;
#pragma unroll
    for (int k = 0; k < CW; ++k) {
        const uint32_t t = cw[k] & 0x0F0F0F0Fu, v = (cw[k] >> 4) & 0x0F0F0F0Fu;
#define LQG_SCAN_BYTE(J)                                                                     \
    {                                                                                       \
        const uint32_t b = 4u * static_cast<uint32_t>(k) + (J);                             \
        s += reinterpret_cast<lds_u8*>(static_cast<uintptr_t>(add_byte<J>(lb.even, t)))[16u * b]; \
        s += reinterpret_cast<lds_u8*>(static_cast<uintptr_t>(add_byte<J>(lb.odd, v)))[16u * b]; \
    }
        LQG_SCAN_BYTE(0)
        LQG_SCAN_BYTE(1)
        LQG_SCAN_BYTE(2)
        LQG_SCAN_BYTE(3)
#undef LQG_SCAN_BYTE
    }
    return s;
}

// CW code words from word index w of a global array (w a multiple of
// min(CW, 4): the loads are 4, 8 or 16 bytes wide and aligned).
typedef uint32_t u32x2 __attribute__((ext_vector_type(2)));
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));

template <int CW>
__device__ __forceinline__ void load_codes(uint32_t (&cw)[CW], gptr<uint32_t> base, uint32_t w) {
    if constexpr (CW == 1) {
        cw[0] = base[w];
    } else if constexpr (CW == 2) {
        const u32x2 v = ((gptr<u32x2>)base)[w / 2];
        cw[0] = v.x;
        cw[1] = v.y;
    } else {
#pragma unroll
        for (int k = 0; k < CW; k += 4) {
            const u32x4 v = ((gptr<u32x4>)base)[w / 4 + k / 4];
            cw[k] = v.x;
            cw[k + 1] = v.y;
            cw[k + 2] = v.z;
            cw[k + 3] = v.w;
        }
    }
}

// The lane's FastScan sum by the query's bit planes (prepare): per code
// word, seven ANDs with the plane words (one uniform LDS address per
// instruction for one lane per neighbour, two for two) and seven population
// counts, weighted 2^k.  pl: the plane words of the lane's first code word.
template <int CW>
__device__ __forceinline__ uint32_t plane_sum(const uint32_t (&cw)[CW], const uint4* pl) {
    uint32_t a[7] = {0u, 0u, 0u, 0u, 0u, 0u, 0u};
#pragma unroll
    for (int k = 0; k < CW; ++k) {
        const uint4 x = pl[2 * k], y = pl[2 * k + 1];
        a[0] += static_cast<uint32_t>(__builtin_popcount(cw[k] & x.x));
        a[1] += static_cast<uint32_t>(__builtin_popcount(cw[k] & x.y));
        a[2] += static_cast<uint32_t>(__builtin_popcount(cw[k] & x.z));
        a[3] += static_cast<uint32_t>(__builtin_popcount(cw[k] & x.w));
        a[4] += static_cast<uint32_t>(__builtin_popcount(cw[k] & y.x));
        a[5] += static_cast<uint32_t>(__builtin_popcount(cw[k] & y.y));
        a[6] += static_cast<uint32_t>(__builtin_popcount(cw[k] & y.z));
    }
    return a[0] + 2u * a[1] + 4u * a[2] + 8u * a[3] + 16u * a[4] + 32u * a[5] + 64u * a[6];
}

// Estimates of the neighbours a lane holds for one node (pass of VP
// neighbours starting at vb).  Returns the estimate on every lane of the
// neighbour's lanes.  pl: the lane's plane words (used when pr.planes).
template <int CW>
__device__ __forceinline__ float lane_estimate(const uint32_t (&cw)[CW], LutBase lb,
                                               const uint4* pl, uint32_t lpv, const Prep& pr,
                                               float sqr_y, float tx, float dq, float vq) {
    uint32_t s = pr.planes ? plane_sum<CW>(cw, pl) : scan_sum<CW>(cw, lb);
    // the two lanes of a neighbour: DPP quad_perm [1, 0, 3, 2]
    if (lpv == 2)
        s += static_cast<uint32_t>(
            __builtin_amdgcn_update_dpp(0, static_cast<int>(s), 0xB1, 0xF, 0xF, false));
    const int32_t r = static_cast<int32_t>(static_cast<int16_t>(static_cast<uint16_t>(s & 0xFFFFu)));
    const float f = static_cast<float>(r * 2 - pr.sumq);
    const float b = fmaf(vq, pr.lo, tx + sqr_y);
    const float c = (f * dq) * pr.width;
    return c + b;
}

// ---------------------------------------------------------------------------
// One query.  GLOBAL = false (the search kernel): the visited set in the
// wave's LDS table; if it would pass vis_limit, the query is appended to
// the launch's overflow list and false returned with nothing written (the
// overflow kernel searches it again from the start), or, where the launch
// has no overflow list, the query fails.  GLOBAL = true (the overflow
// kernel): the set in the wave's device table gtab.  A failed query (past
// vis_limit of its last table, or no device table) has E = 0xFFFFFFFF in the
// stats and ids 0xFFFFFFFF, sets job.err and its kFlagFailed flag.
// Appends query q to the launch's overflow list (ovf[0] = count, ovf[2 + i]
// = entries); returns false.
__device__ __forceinline__ bool h2_defer(const H2Job& job, uint32_t q, uint32_t lane) {
    if (lane == 0) job.ovf[2 + atomicAdd(&job.ovf[0], 1u)] = q;
    wave_sync();
    return false;
}

template <int EF_CAP, int CW, int KS, bool GLOBAL>
__device__ bool h2_search_one(const H2Dev& ix, const H2Job& job, uint32_t q, const H2Lds& w,
                              uint32_t hash_bits, uint32_t* gtab, uint32_t lane) {
    constexpr int S = EF_CAP / kWave;
    const uint32_t ef =
        uniform(max(1u, min(job.ef ? job.ef[q] : job.ef_all, static_cast<uint32_t>(EF_CAP))));
    const uint32_t kq = uniform(min(job.k ? job.k[q] : job.k_all, job.kstride));
    const uint32_t deg = ix.deg, dim = ix.dim;
    const uint32_t lpv = deg >= 64 ? 1u : 2u;
    const uint32_t vp = kWave / lpv;              // neighbours per pass
    const uint32_t code_bytes = ix.p / 8;         // per neighbour
    const uint32_t part = lane % lpv;
    const LutBase lutb = lut_base<CW>(w.lut, ix.p, part);
    const uint4* plv = w.pl + 2u * CW * part;

    // bit planes from P = 256 on: at P = 128 the specialised kernel ran
    // slower with them (0.559 -> 0.574 ms per 10,000-query launch), its LUT
    // scan being 16 lookups per lane
    const Prep pr = prepare(ix, job.queries + static_cast<size_t>(q) * dim,
                            job.rot ? job.rot + static_cast<size_t>(q) * ix.p : nullptr, w, lpv,
                            lane, ix.p >= 256u);

    Visited<GLOBAL> vis;
    vis.tab = GLOBAL ? gtab : w.hv;
    vis.bits = GLOBAL ? job.gvis_bits : hash_bits;
    vis.count = 0;
    bool failed = GLOBAL && gtab == nullptr;
    if (!failed) {
        for (uint32_t i = lane; i < (1u << vis.bits); i += kWave) vis.tab[i] = kEmpty;
        if (GLOBAL) __threadfence();
    }
    // the merge staging reads back every position below EF_CAP; positions the
    // pool has not reached yet must read as empty entries
    for (uint32_t i = lane; i <= static_cast<uint32_t>(EF_CAP); i += kWave)
        w.st[i] = make_uint4(kEmpty, kEmpty, kEmpty, 0u);
    if constexpr (EF_CAP >= kRankedPool) {
        // ranked admission searches the staging copy of the pool, so it must
        // hold the entry point from the start
        if (lane == 0) {
            const uint64_t k0 = sb_key(3.402823466e+38f, 0);
            w.st[0] = make_uint4(static_cast<uint32_t>(k0), static_cast<uint32_t>(k0 >> 32), ix.ep, 0u);
        }
        for (uint32_t i = lane; i < static_cast<uint32_t>(EF_CAP); i += kWave) w.cnt[i] = 0u;
    }
    wave_sync();

    // search_pool_.insert(entry_point_, FLT_MAX) (qg.hpp:260); empty entries
    // are ~0, whose bit 0 (checked) keeps them from being popped
    uint64_t P[S];
    uint32_t I[S];
#pragma unroll
    for (int s = 0; s < S; ++s) {
        P[s] = ~0ull;
        I[s] = kEmpty;
    }
    if (lane == 0) {
        P[0] = sb_key(3.402823466e+38f, 0);
        I[0] = ix.ep;
    }
    uint32_t size = 1, seq = 1, n_exp = 0, n_pop = 0;
    ResBuf<KS> rb;
    rb_init(rb);
    // query dims (lane & 15) + 16u (u < kRegJ) for the partials of exact_l2_regs
    float qr[kRegJ];
    load_partials(qr, job.queries + static_cast<size_t>(q) * dim, dim, lane);
    // per-lane offsets inside a node record for the neighbour this lane scans
    // in a pass (the pass's first neighbour added per pass); every load of a
    // pass is then a uniform base plus one of these
    const uint32_t lv = lane / lpv;
    const uint32_t code_off = ix.off_codes + lv * code_bytes + part * (code_bytes / lpv);
    const uint32_t lv4 = 4u * lv, lv2 = 2u * lv;

    while (!failed) {
        // pop: first unchecked entry (buffer.hpp:52-60)
        uint64_t pm = 0;
        int fs = S;
#pragma unroll
        for (int s = S - 1; s >= 0; --s) {
            const uint64_t m = __ballot((static_cast<uint32_t>(P[s]) & 1u) == 0u);
            if (m) {
                pm = m;
                fs = s;
            }
        }
        if (fs == S) break;
        const int pl = static_cast<int>(__builtin_ctzll(pm));
        uint32_t cur = 0;
#pragma unroll
        for (int s = 0; s < S; ++s) {
            if (s == fs) {
                cur = readlane32(I[s], pl);
                P[s] |= static_cast<int>(lane) == pl ? 1ull : 0ull;
            }
        }
        ++n_pop;
        const uint32_t vs = v_test_set(vis, cur, lane);
        if (vs == kVisSeen) continue;
        if (vs == kVisFull) {
            if (!GLOBAL && job.ovf) return h2_defer(job, q, lane);
            failed = true;
            break;
        }
        if (job.exp_ids && n_exp < job.exp_cap && lane == 0)
            job.exp_ids[static_cast<size_t>(q) * job.exp_cap + n_exp] = cur;
        ++n_exp;

        const unsigned char* rec = ix.rec + static_cast<size_t>(cur) * ix.stride;
        const float* raw = reinterpret_cast<const float*>(rec);
        const float* fac = reinterpret_cast<const float*>(rec + ix.off_fac);
        const uint32_t* ids = reinterpret_cast<const uint32_t*>(rec + ix.off_ids);
        // the node's raw row for the exact distance: dims (lane & 15) + 16u
        float xr[kRegJ];
        load_partials(xr, raw, dim, lane);
        float sqr_y = 0.0f;

        for (uint32_t vb = 0; vb < deg; vb += vp) {
            // the pass's loads first (codes, factors, id).  Lanes past the
            // last neighbour (a pass of a degree that is not a multiple of
            // the pass width) read bytes further on in the record, or in the
            // spare record after the last one, and take no part.
            const uint32_t vi = vb + lv;
            const bool act = vi < deg;
            uint32_t cw[CW];
            load_codes<CW>(cw, uniform_gptr<uint32_t>(rec + vb * code_bytes), code_off / 4u);
            const uint32_t nid = uniform_gptr<uint32_t>(ids + vb)[lv];
            const float tx = uniform_gptr<float>(fac + vb)[lv];
            const float dq = uniform_gptr<float>(fac + deg + vb)[lv];
            const float vq = uniform_gptr<float>(fac + 2 * deg + vb)[lv];
            // scan_neighbors (qg.hpp:355): exact distance to the node
            if (vb == 0) sqr_y = exact_l2_regs(xr, qr, raw, w, dim, lane);
            const float est = lane_estimate<CW>(cw, lutb, plv, lpv, pr, sqr_y, tx, dq, vq);
            // is_full / visited_.get (qg.hpp:375-377); candidate i gets insertion
            // number seq + i.  The visited table is probed only by the lanes the
            // pool would take.
            // sb_key(est, seq + vi), the low word a uniform base minus 2 lv
            const uint32_t klo = ((0x7FFFFFFFu - seq - vb) << 1) - lv2;
            const uint64_t key = (static_cast<uint64_t>(sortable(est)) << 32) | klo;
            bool cand = act && part == 0;
            if (size == ef) {
                const uint32_t t = size - 1u;
                uint64_t lastv = P[0];
#pragma unroll
                for (int s = 0; s < S; ++s)
                    if (static_cast<uint32_t>(s) == t / kWave) lastv = P[s];
                const uint64_t last = readlane64(lastv, static_cast<int>(t % kWave));
                cand = cand && key < last;
            }
            if (__ballot(cand) == 0) continue;
            cand = cand && !v_contains_lanes(vis, nid, cand);
            uint64_t mm = __ballot(cand);
            if (mm == 0) continue;
            const uint32_t m3 = static_cast<uint32_t>(__popcll(mm));
            // per candidate j: lb_j = pool entries below it and rank = the
            // candidates below it; per pool entry i: the candidates below it,
            // #{j : lb_j <= i}, by which it moves up
            uint32_t lb = 0, rank = 0;
            uint32_t up[S];
            if constexpr (EF_CAP >= kRankedPool) {
                // lb by binary search over the staging copy of the pool (as of
                // the last merge: keys ascending at positions below size; only
                // checked bits, bit 0, may differ from the registers, and keys
                // differ above bit 0), then a counter per position
                if (cand) {
                    uint32_t pos = 0;
#pragma unroll
                    for (uint32_t step = static_cast<uint32_t>(EF_CAP); step > 0u; step >>= 1) {
                        const uint32_t t = pos + step;
                        if (t <= size) {
                            const uint2 e = *reinterpret_cast<const uint2*>(&w.st[t - 1u]);
                            if (((static_cast<uint64_t>(e.y) << 32) | e.x) < key) pos = t;
                        }
                    }
                    lb = pos;  // below ef - 1 < EF_CAP: the pool's bound admits it
                    atomicAdd(&w.cnt[lb], 1u);
                }
                while (mm) {
                    const int j = static_cast<int>(__builtin_ctzll(mm));
                    mm &= mm - 1;
                    rank += readlane64(key, j) < key ? 1u : 0u;
                }
                wave_sync();
                // inclusive prefix sums of the counters: lane l sums positions
                // S l .. S l + S - 1 in place, then adds the lanes below
                constexpr int V = S / 4;
                uint4* run = reinterpret_cast<uint4*>(w.cnt) + V * lane;
                uint4 c[V];
                uint32_t acc = 0;
#pragma unroll
                for (int v = 0; v < V; ++v) {
                    c[v] = run[v];
                    c[v].x += acc;
                    c[v].y += c[v].x;
                    c[v].z += c[v].y;
                    c[v].w += c[v].z;
                    acc = c[v].w;
                }
                uint32_t incl = acc;
#pragma unroll
                for (int d = 1; d < kWave; d <<= 1) {
                    const uint32_t t = static_cast<uint32_t>(__shfl_up(static_cast<int>(incl), d));
                    if (lane >= static_cast<uint32_t>(d)) incl += t;
                }
                const uint32_t off = incl - acc;
#pragma unroll
                for (int v = 0; v < V; ++v)
                    run[v] = make_uint4(c[v].x + off, c[v].y + off, c[v].z + off, c[v].w + off);
                wave_sync();
#pragma unroll
                for (int s = 0; s < S; ++s) up[s] = w.cnt[static_cast<uint32_t>(s * kWave) + lane];
                wave_sync();
#pragma unroll
                for (int v = 0; v < V; ++v) run[v] = make_uint4(0u, 0u, 0u, 0u);
            } else {
                uint32_t nlt[S];  // candidates above each pool entry
#pragma unroll
                for (int s = 0; s < S; ++s) nlt[s] = 0;
                while (mm) {
                    const int j = static_cast<int>(__builtin_ctzll(mm));
                    mm &= mm - 1;
                    const uint64_t kj = readlane64(key, j);
                    uint32_t below = 0;
#pragma unroll
                    for (int s = 0; s < S; ++s) {
                        const bool lt = P[s] < kj;
                        below += static_cast<uint32_t>(__popcll(__ballot(lt)));
                        nlt[s] += lt ? 1u : 0u;
                    }
                    lb = static_cast<int>(lane) == j ? below : lb;
                    rank += kj < key ? 1u : 0u;
                }
#pragma unroll
                for (int s = 0; s < S; ++s) up[s] = m3 - nlt[s];
            }
            const uint32_t nsize = min(size + m3, ef);
            // merge through LDS: each entry and candidate goes to its new
            // position (those that fall off to the spare slot EF_CAP); every
            // position below nsize is written, those above still hold the
            // empty entries they were given at the start
#pragma unroll
            for (int s = 0; s < S; ++s) {
                const uint32_t np = static_cast<uint32_t>(s * kWave) + lane + up[s];
                const uint32_t dst = np < nsize ? np : static_cast<uint32_t>(EF_CAP);
                w.st[dst] = make_uint4(static_cast<uint32_t>(P[s]), static_cast<uint32_t>(P[s] >> 32),
                                       I[s], 0u);
            }
            {
                const uint32_t np = lb + rank;
                const uint32_t dst = (cand && np < nsize) ? np : static_cast<uint32_t>(EF_CAP);
                w.st[dst] = make_uint4(static_cast<uint32_t>(key), static_cast<uint32_t>(key >> 32),
                                       nid, 0u);
            }
            wave_sync();
#pragma unroll
            for (int s = 0; s < S; ++s) {
                const uint4 e = w.st[static_cast<uint32_t>(s * kWave) + lane];
                P[s] = (static_cast<uint64_t>(e.y) << 32) | e.x;
                I[s] = e.z;
            }
            size = nsize;
            wave_sync();
        }
        seq += deg;
        // res_pool.insert (qg.hpp:282)
        rb_insert<KS>(rb, kq, cur, sqr_y, lane);
    }

    // update_results (qg.hpp:395-418): fewer than k nodes expanded.  The ids
    // snapshot holds k + 1 entries, those past the buffer's size 0 (the
    // ResultBuffer is zero-initialised, buffer.hpp:81-102), so node 0 may be
    // probed.
    if (!failed && kq > 0 && rb.size < kq) {
        uint32_t snap[KS];
#pragma unroll
        for (int s = 0; s < KS; ++s) snap[s] = rb.ri[s];
        const uint32_t snap_size = rb.size;
        for (uint32_t t = 0; t <= kq && !failed; ++t) {
            const uint32_t node = t < snap_size ? rb_id<KS>(snap, t) : 0u;
            const unsigned char* rec = ix.rec + static_cast<size_t>(node) * ix.stride;
            const uint32_t* nb = reinterpret_cast<const uint32_t*>(rec + ix.off_ids);
            for (uint32_t i = 0; i < deg; ++i) {
                const uint32_t x = uniform(nb[i]);
                const uint32_t xs = v_test_set(vis, x, lane);
                if (xs == kVisSeen) continue;
                if (xs == kVisFull) {
                    if (!GLOBAL && job.ovf) return h2_defer(job, q, lane);
                    failed = true;
                    break;
                }
                const float* row =
                    reinterpret_cast<const float*>(ix.rec + static_cast<size_t>(x) * ix.stride);
                float xr[kRegJ];
                load_partials(xr, row, dim, lane);
                rb_insert<KS>(rb, kq, x, exact_l2_regs(xr, qr, row, w, dim, lane), lane);
            }
            if (rb.size == kq) break;
        }
    }

    // copy_results (buffer.hpp:102): unfilled slots hold id 0; a query whose
    // visited set overflowed returns no ids (0xFFFFFFFF everywhere)
    uint32_t* oi = job.out_ids + static_cast<size_t>(q) * job.kstride;
    float* od = job.out_d ? job.out_d + static_cast<size_t>(q) * job.kstride : nullptr;
#pragma unroll
    for (int s = 0; s < KS; ++s) {
        const uint32_t j = static_cast<uint32_t>(s * kWave) + lane;
        if (j < job.kstride) {
            uint32_t id = kInvalid;
            float d = INFINITY;
            if (j < kq && !failed) {
                id = j < rb.size ? rb.ri[s] : 0u;
                if (j < rb.size) d = rb.rd[s];
            }
            oi[j] = id;
            if (od) od[j] = d;
        }
    }
    if (lane == 0) {
        if (job.out_stats) {
            job.out_stats[2 * static_cast<size_t>(q)] = failed ? kInvalid : n_exp;
            job.out_stats[2 * static_cast<size_t>(q) + 1] = n_pop;
        }
        if (failed) atomicOr(job.err, 1u);
    }
    if (job.done) signal_host(job.done + q, lane, failed ? kFlagFailed : kFlagDone);
    wave_sync();
    return true;
}

template <int EF_CAP, int CW, int KS>
__device__ __forceinline__ void h2_search_body(const H2Dev& ix, const H2Job& job,
                                               uint32_t hash_bits, const Layout& lay) {
    extern __shared__ __align__(16) unsigned char lds[];
    const uint32_t lane = threadIdx.x & (kWave - 1);
    const uint32_t wid = threadIdx.x / kWave;
    const H2Lds w = carve(lds + static_cast<size_t>(wid) * lay.wave_bytes, lay);
    const uint32_t gwave = blockIdx.x * (blockDim.x / kWave) + wid;
    // wave gwave first takes query gwave, then sw + atomicAdd(counter) (sw =
    // waves in the grid), so the launch does not start with one atomic per
    // wave on a single counter
    const uint32_t sw = gridDim.x * (blockDim.x / kWave);
    uint32_t q = gwave;
    for (;;) {
        q = uniform(q);
        if (q >= job.nq) break;
        (void)h2_search_one<EF_CAP, CW, KS, false>(ix, job, q, w, hash_bits, nullptr, lane);
        uint32_t nx = 0;
        if (lane == 0) nx = sw + atomicAdd(job.counter, 1u);
        q = __shfl(nx, 0);
    }
}

// ---------------------------------------------------------------------------
// Latency mode (launches of a few queries: a lone Index.search): one query
// per workgroup of two waves.  Wave 0 runs the search of the batch kernels
// (same results).  Wave 1 only loads: it reads wave 0's pool staging records
// in LDS (rewritten at every merge; an entry is unexpanded when bit 0 of its
// key and the top bit of its id are clear, which covers both kernels' flag
// conventions) and fetches every 128-byte line of the records of the first
// `targets` unexpanded entries, so that the next expansion finds its record
// in the L2 instead of waiting on HBM.  Nothing it does reaches wave 0 except
// through the caches; it stops on wave 0's done flag in LDS, or after
// kH2PrefetchRounds rounds.
constexpr uint32_t kH2PrefetchRounds = 1u << 16;

__device__ void h2_prefetch_ahead(const H2Dev& ix, const uint4* st, uint32_t ef_cap,
                                  const uint32_t* done, uint32_t targets, uint32_t lane,
                                  uint32_t* sink_out) {
    uint32_t sink = 0;
    uint32_t hist[8];  // the last eight nodes fetched (scalar registers)
#pragma unroll
    for (int i = 0; i < 8; ++i) hist[i] = kEmpty;
    targets = min(max(targets, 1u), 8u);
    const uint32_t lines = (ix.stride + 127u) / 128u;
    const uint32_t* stw = reinterpret_cast<const uint32_t*>(st);
    for (uint32_t r = 0; r < kH2PrefetchRounds; ++r) {
        if (__hip_atomic_load(done, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP) != 0u) break;
        uint32_t klo = 1u, id = kEmpty;
        if (lane < ef_cap) {
            klo = __hip_atomic_load(stw + 4u * lane, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
            id = __hip_atomic_load(stw + 4u * lane + 2u, __ATOMIC_RELAXED,
                                   __HIP_MEMORY_SCOPE_WORKGROUP);
        }
        bool open = (klo & 1u) == 0u && id < ix.n;
#pragma unroll
        for (int i = 0; i < 8; ++i) open = open && id != hist[i];
        uint64_t m = __ballot(open);
        if (m == 0) {
            __builtin_amdgcn_s_sleep(1);
            continue;
        }
        uint32_t tg[8];
        uint32_t cnt = 0;
#pragma unroll
        for (uint32_t t = 0; t < 8u; ++t) {
            tg[t] = kEmpty;
            if (t < targets && m != 0) {
                const uint32_t x = uniform(__shfl(id, static_cast<int>(__builtin_ctzll(m))));
                m &= m - 1;
#pragma unroll
                for (int i = 7; i > 0; --i) hist[i] = hist[i - 1];
                hist[0] = x;
                tg[t] = x;
                cnt = t + 1u;
            }
        }
        for (uint32_t i = lane; i < cnt * lines; i += kWave) {
            const uint32_t t = i / lines, line = i - t * lines;
            uint32_t node = tg[0];
#pragma unroll
            for (uint32_t u = 1; u < 8u; ++u) node = t == u ? tg[u] : node;
            sink += *reinterpret_cast<const uint32_t*>(ix.rec + static_cast<size_t>(node) * ix.stride +
                                                       128u * line);
        }
    }
    // keeps the loads: the condition depends on every value they returned
    if (sink == 0x9E3779B9u) *sink_out = sink;
}

template <int EF_CAP, int CW, int KS>
__global__ void __launch_bounds__(128) h2_latency_kernel(H2Dev ix, H2Job job, uint32_t hash_bits,
                                                          Layout lay) {
    extern __shared__ __align__(16) unsigned char lds[];
    const uint32_t lane = threadIdx.x & (kWave - 1);
    const uint32_t wid = threadIdx.x / kWave;
    const H2Lds w = carve(lds, lay);
    uint32_t* flag = reinterpret_cast<uint32_t*>(lds + lay.wave_bytes);
    if (threadIdx.x == 0) {
        flag[0] = 0u;
        flag[1] = 0u;
    }
    __syncthreads();
    if (wid == 0) {
        if (blockIdx.x < job.nq)
            (void)h2_search_one<EF_CAP, CW, KS, false>(ix, job, blockIdx.x, w, hash_bits, nullptr,
                                                       lane);
        if (lane == 0) __hip_atomic_store(flag, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    } else {
        h2_prefetch_ahead(ix, w.st, static_cast<uint32_t>(EF_CAP), flag, job.prefetch_targets,
                          lane, flag + 1);
    }
}

// The overflow kernel: searches again, with the visited set in a per-wave
// device table, the queries the search kernel put on the overflow list
// (run after it on the same stream; it finds an empty list at once).
template <int EF_CAP, int CW, int KS>
__global__ void __launch_bounds__(64) h2_overflow_kernel(H2Dev ix, H2Job job, uint32_t hash_bits,
                                                          Layout lay) {
    extern __shared__ __align__(16) unsigned char lds[];
    const uint32_t lane = threadIdx.x & (kWave - 1);
    const H2Lds w = carve(lds, lay);
    const uint32_t n = uniform(__hip_atomic_load(&job.ovf[0], __ATOMIC_RELAXED,
                                                 __HIP_MEMORY_SCOPE_AGENT));
    uint32_t* gtab = job.gvis ? job.gvis + (static_cast<size_t>(blockIdx.x) << job.gvis_bits)
                              : nullptr;
    for (;;) {
        uint32_t i = 0;
        if (lane == 0) i = atomicAdd(&job.ovf[1], 1u);
        i = uniform(__shfl(i, 0));
        if (i >= n) break;
        const uint32_t q = uniform(job.ovf[2 + i]);
        (void)h2_search_one<EF_CAP, CW, KS, true>(ix, job, q, w, hash_bits, gtab, lane);
    }
}

// Occupancy variants: uncapped registers, and capped for 6 or 8 waves per
// SIMD.
template <int EF_CAP, int CW, int KS>
__global__ void __launch_bounds__(256) h2_search_kernel(H2Dev ix, H2Job job, uint32_t hash_bits,
                                                         Layout lay) {
    h2_search_body<EF_CAP, CW, KS>(ix, job, hash_bits, lay);
}

template <int EF_CAP, int CW, int KS>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(6, 6)))
h2_search_kernel_w6(H2Dev ix, H2Job job, uint32_t hash_bits, Layout lay) {
    h2_search_body<EF_CAP, CW, KS>(ix, job, hash_bits, lay);
}

template <int EF_CAP, int CW, int KS>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(8, 8)))
h2_search_kernel_w8(H2Dev ix, H2Job job, uint32_t hash_bits, Layout lay) {
    h2_search_body<EF_CAP, CW, KS>(ix, job, hash_bits, lay);
}

}  // namespace
}  // namespace lqg
