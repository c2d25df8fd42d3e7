// h2_search.hip -- gfx950 kernels for the reference's FastScan RaBitQ query
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
// per-wave table in device memory if it passes half its capacity.
//
// FastScan: a lane holds the codes of one neighbour (deg >= 64) or half of
// them (deg = 32) and adds LUT bytes looked up in LDS; sums are taken mod
// 2^16 and read back as int16, as the AVX-512 code does (qg_scanner.hpp:
// 117-135); est = (f * fac_dq) * width + fma(fac_vq, lo, triple_x + sqr_y), the
// operation order of the reference build (qg_scanner.hpp:23-48 as compiled).
#include <math.h>

#include "lqg_internal.h"
#include "wave.h"

namespace lqg {
namespace {

constexpr uint32_t kEmpty = 0xFFFFFFFFu;
constexpr float kInv63 = 1.0f / 63.0f;  // width = (hi - lo) * fl(1/63)
constexpr float kC63 = 1.0f / kInv63;   // 1/width = fl(1/fl(1/63)) / (hi - lo)

// Estimate bits that sort like the floats (-0 taken as +0, as a float
// comparison does).
__device__ __forceinline__ uint32_t sortable(float x) {
    const uint32_t u = __float_as_uint(x == 0.0f ? 0.0f : x);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}

__device__ __forceinline__ uint64_t sb_key(float est, uint32_t seq) {
    return (static_cast<uint64_t>(sortable(est)) << 32) |
           (static_cast<uint64_t>(0x7FFFFFFFu - seq) << 1);
}

struct H2Lds {
    float* qv;        // query (dim)
    float* tb;        // rotated query during preparation, q - x in exact_l2 (p floats)
    uint8_t* qb;      // quantised query (p bytes)
    uint8_t* lut;     // 4p bytes (+16 bytes of bank skew for 2 lanes per vector)
    uint64_t* sk;     // pool merge staging: keys (EF_CAP + 1; shares tb's memory)
    uint32_t* si;     // pool merge staging: ids (EF_CAP + 1)
    uint32_t* hv;     // visited table (1 << hash_bits)
    float* rd;        // ResultBuffer distances (kcap)
    uint32_t* ri;     // ResultBuffer ids (kcap)
    uint32_t* snap;   // ids snapshot for update_results (kcap)
};

struct Layout {
    uint32_t qv, tb, qb, lut, sk, si, hv, rd, ri, snap, wave_bytes;
};

__host__ __device__ inline uint32_t align16(uint32_t x) { return (x + 15u) & ~15u; }

__host__ __device__ inline Layout h2_layout(uint32_t dim, uint32_t p, uint32_t ef_cap,
                                            uint32_t hash_bits, uint32_t kmax) {
    Layout l;
    const uint32_t kcap = kmax + 1;
    uint32_t o = 0;
    l.qv = o;   o += align16(4 * dim);
    l.lut = o;  o += align16(4 * p + 16);
    // one region for the preparation buffers (tb, qb) and, once the LUT is
    // built, the pool merge staging (sk, si: ef_cap entries and a spare)
    const uint32_t prep = align16(4 * p) + align16(p);
    const uint32_t stage = align16(8 * (ef_cap + 1)) + align16(4 * (ef_cap + 1));
    l.tb = o;
    l.qb = o + align16(4 * p);
    l.sk = o;
    l.si = o + align16(8 * (ef_cap + 1));
    o += prep > stage ? prep : stage;
    l.hv = o;   o += 4u << hash_bits;
    l.rd = o;   o += align16(4 * kcap);
    l.ri = o;   o += align16(4 * kcap);
    l.snap = o; o += align16(4 * kcap);
    l.wave_bytes = o;
    return l;
}

__device__ __forceinline__ H2Lds carve(unsigned char* base, const Layout& l) {
    H2Lds w;
    w.qv = reinterpret_cast<float*>(base + l.qv);
    w.tb = reinterpret_cast<float*>(base + l.tb);
    w.qb = base + l.qb;
    w.lut = base + l.lut;
    w.sk = reinterpret_cast<uint64_t*>(base + l.sk);
    w.si = reinterpret_cast<uint32_t*>(base + l.si);
    w.hv = reinterpret_cast<uint32_t*>(base + l.hv);
    w.rd = reinterpret_cast<float*>(base + l.rd);
    w.ri = reinterpret_cast<uint32_t*>(base + l.ri);
    w.snap = reinterpret_cast<uint32_t*>(base + l.snap);
    return w;
}

// ---------------------------------------------------------------------------
// Query preparation (qg_query.hpp:28-43) into LDS: qv (raw query), lut.
struct Prep {
    float lo, width;
    int32_t sumq;
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
                        uint32_t lpv, uint32_t lane) {
    const uint32_t dim = ix.dim, p = ix.p;
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
    // uint8 sum of q[4c + 3 - b] over the set bits b of j.  With two lanes per
    // neighbour the second half of the codebooks sits 16 bytes further, so the
    // two halves of a wavefront read different LDS banks.
    const uint32_t ncb = p / 4, cpl = ncb / lpv;
    for (uint32_t idx = lane; idx < 16u * ncb; idx += kWave) {
        const uint32_t c = idx >> 4, j = idx & 15u;
        uint32_t s = 0;
#pragma unroll
        for (uint32_t b = 0; b < 4; ++b)
            if ((j >> b) & 1u) s += w.qb[4 * c + 3 - b];
        w.lut[16 * c + (c >= cpl ? 16u : 0u) + j] = static_cast<uint8_t>(s & 0xFFu);
    }
    wave_sync();
    Prep r;
    r.lo = lo;
    r.width = width;
    r.sumq = sq;
    return r;
}

// Squared L2 from w.tb = q - x (dim floats, synchronised): the reference's
// 16 FMA partials (dims p, p+16, ...), the reduction tree (p, p+8), (p, p+4),
// (p, p+2), (0, 1), then an FMA tail for dims beyond the last multiple of 16.
__device__ __forceinline__ float reduce_l2(const float* tb, uint32_t dim, uint32_t lane) {
    const uint32_t mul16 = dim & ~15u;
    float s = 0.0f;
    if (lane < 16)
        for (uint32_t j = lane; j < mul16; j += 16) {
            const float t = tb[j];
            s = fmaf(t, t, s);
        }
    s = s + __shfl_down(s, 8);
    s = s + __shfl_down(s, 4);
    s = s + __shfl_down(s, 2);
    s = s + __shfl_down(s, 1);
    if (lane == 0)
        for (uint32_t i = mul16; i < dim; ++i) {
            const float t = tb[i];
            s = fmaf(t, t, s);
        }
    return __shfl(s, 0);
}

// Exact distance of the query (w.qv) to a raw row in device memory.
__device__ float exact_l2(const float* row, const H2Lds& w, uint32_t dim, uint32_t lane) {
    for (uint32_t i = lane; i < dim; i += kWave) w.tb[i] = w.qv[i] - row[i];
    wave_sync();
    const float r = reduce_l2(w.tb, dim, lane);
    wave_sync();
    return r;
}

// The same distance with the partials in registers: lane p < 16 runs the FMA
// chain over dims p, p + 16, ... (the first kRegJ of them from xr / qr, which
// the caller loaded, the rest from the row and w.qv), the tree is taken with
// DPP row shifts (lane p adds lane p + 8, p + 4, p + 2, p + 1 of its row:
// the pairs of reduce_l2), and lane 0's sum is broadcast.
constexpr int kRegJ = 8;

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
    float s = 0.0f;
#pragma unroll
    for (int u = 0; u < kRegJ; ++u) {
        if (static_cast<uint32_t>(u) < nj) {
            const float t = qr[u] - xr[u];
            s = fmaf(t, t, s);
        }
    }
    if (nj > static_cast<uint32_t>(kRegJ) && lane < 16)
        for (uint32_t j = kRegJ; j < nj; ++j) {
            const float t = w.qv[lane + 16 * j] - row[lane + 16 * j];
            s = fmaf(t, t, s);
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
// Exact visited set.  LDS table while it is at most half full, then a table
// of (1 << gbits) entries in device memory (atomics, so every lane sees every
// insertion).
struct Visited {
    uint32_t* lds;
    uint32_t* glb;
    uint32_t lbits, gbits, count;
    bool global;
    bool failed;
};

__device__ __forceinline__ uint32_t vhash(uint32_t id, uint32_t bits) {
    return (id * 2654435761u) >> (32u - bits);
}

__device__ __forceinline__ bool v_contains(const Visited& v, uint32_t id) {
    if (!v.global) {
        const uint32_t mask = (1u << v.lbits) - 1u;
        uint32_t h = vhash(id, v.lbits);
        for (;;) {
            const uint32_t e = v.lds[h];
            if (e == id) return true;
            if (e == kEmpty) return false;
            h = (h + 1u) & mask;
        }
    }
    const uint32_t mask = (1u << v.gbits) - 1u;
    uint32_t h = vhash(id, v.gbits);
    for (;;) {
        const uint32_t e = __hip_atomic_load(&v.glb[h], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (e == id) return true;
        if (e == kEmpty) return false;
        h = (h + 1u) & mask;
    }
}

// v_contains on the lanes where `active` is set (false on the others).  The
// LDS probe runs without divergent branches: every lane reads a slot of the
// table each round (inactive ones a harmless in-range slot) until every
// active lane has met its id or an empty slot.
__device__ __forceinline__ bool v_contains_lanes(const Visited& v, uint32_t id, bool active) {
    if (v.global) return active && v_contains(v, id);
    const uint32_t mask = (1u << v.lbits) - 1u;
    uint32_t h = vhash(id, v.lbits);
    bool found = false;
    while (__ballot(active)) {
        const uint32_t e = v.lds[h];
        found = found || (active && e == id);
        active = active && e != id && e != kEmpty;
        h = (h + 1u) & mask;
    }
    return found;
}

// v_contains for an id that is the same on every lane: the probe loop runs on
// scalar values.
__device__ __forceinline__ bool v_contains_u(const Visited& v, uint32_t id) {
    const uint32_t bits = v.global ? v.gbits : v.lbits;
    const uint32_t mask = (1u << bits) - 1u;
    uint32_t h = vhash(id, bits);
    for (;;) {
        const uint32_t raw = v.global
                                 ? __hip_atomic_load(&v.glb[h], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)
                                 : v.lds[h];
        const uint32_t e = __builtin_amdgcn_readfirstlane(raw);
        if (e == id) return true;
        if (e == kEmpty) return false;
        h = (h + 1u) & mask;
    }
}

__device__ __forceinline__ void g_insert(uint32_t* g, uint32_t gbits, uint32_t id) {
    const uint32_t mask = (1u << gbits) - 1u;
    uint32_t h = vhash(id, gbits);
    for (;;) {
        const uint32_t prev = atomicCAS(&g[h], kEmpty, id);
        if (prev == kEmpty || prev == id) return;
        h = (h + 1u) & mask;
    }
}

// Inserts an id known to be absent (uniform across the wave).  Sets
// v.failed if both tables would pass half capacity.
__device__ void v_insert(Visited& v, uint32_t id, uint32_t lane) {
    if (!v.global && v.count + 1u > (1u << v.lbits) / 2u) {
        if (!v.glb) {
            v.failed = true;
            return;
        }
        for (uint32_t i = lane; i < (1u << v.gbits); i += kWave) v.glb[i] = kEmpty;
        __threadfence();
        wave_sync();
        for (uint32_t i = lane; i < (1u << v.lbits); i += kWave) {
            const uint32_t e = v.lds[i];
            if (e != kEmpty) g_insert(v.glb, v.gbits, e);
        }
        __threadfence();
        wave_sync();
        v.global = true;
    }
    if (v.global && v.count + 1u > (1u << v.gbits) / 2u) {
        v.failed = true;
        return;
    }
    if (!v.global) {
        const uint32_t mask = (1u << v.lbits) - 1u;
        uint32_t h = vhash(id, v.lbits);
        while (__builtin_amdgcn_readfirstlane(v.lds[h]) != kEmpty) h = (h + 1u) & mask;
        if (lane == 0) v.lds[h] = id;
    } else if (lane == 0) {
        g_insert(v.glb, v.gbits, id);
    }
    if (v.global) __threadfence();
    wave_sync();
    ++v.count;
}

// ---------------------------------------------------------------------------
// ResultBuffer::insert (utils/buffer.hpp:84-96): kept in LDS, sorted by exact
// distance, a new entry placed before equal ones, skipped when full and above
// the last.
__device__ void rb_insert(const H2Lds& w, uint32_t& rsize, uint32_t kq, uint32_t id, float x,
                          uint32_t lane) {
    if (kq == 0 || (rsize == kq && x > w.rd[rsize - 1])) return;
    uint32_t pos = 0;
    for (uint32_t base = 0; base < rsize; base += kWave) {
        const uint32_t i = base + lane;
        pos += static_cast<uint32_t>(__popcll(__ballot(i < rsize && w.rd[i] < x)));
    }
    const uint32_t n = min(rsize, kq - 1u);  // entries [pos, n) move up by one
    if (n > pos) {
        for (int base = static_cast<int>(((n - 1u) / kWave) * kWave);
             base >= static_cast<int>((pos / kWave) * kWave); base -= kWave) {
            const uint32_t i = static_cast<uint32_t>(base) + lane;
            const bool mv = i >= pos && i < n;
            float d = 0.0f;
            uint32_t v = 0;
            if (mv) {
                d = w.rd[i];
                v = w.ri[i];
            }
            wave_sync();
            if (mv) {
                w.rd[i + 1] = d;
                w.ri[i + 1] = v;
            }
            wave_sync();
        }
    }
    if (lane == 0) {
        w.rd[pos] = x;
        w.ri[pos] = id;
    }
    rsize = min(rsize + 1u, kq);
    wave_sync();
}

// ---------------------------------------------------------------------------
// FastScan sum of one lane: CW code words (8 codebooks each) starting at
// codebook c0, LUT base already offset for the lane's half.
template <int CW>
__device__ __forceinline__ uint32_t scan_sum(const uint32_t (&cw)[CW], const uint8_t* lut) {
    uint32_t s = 0;
#pragma unroll
    for (int k = 0; k < CW; ++k) {
        const uint32_t word = cw[k];
#pragma unroll
        for (int n = 0; n < 8; ++n) s += lut[16 * (8 * k + n) + ((word >> (4 * n)) & 15u)];
    }
    return s;
}

template <int CW>
__device__ __forceinline__ void load_codes(uint32_t (&cw)[CW], const uint8_t* src) {
    if constexpr (CW == 1) {
        cw[0] = *reinterpret_cast<const uint32_t*>(src);
    } else if constexpr (CW == 2) {
        const uint2 v = *reinterpret_cast<const uint2*>(src);
        cw[0] = v.x;
        cw[1] = v.y;
    } else {
#pragma unroll
        for (int k = 0; k < CW; k += 4) {
            const uint4 v = *reinterpret_cast<const uint4*>(src + 4 * k);
            cw[k] = v.x;
            cw[k + 1] = v.y;
            cw[k + 2] = v.z;
            cw[k + 3] = v.w;
        }
    }
}

// Estimates of the neighbours a lane holds for one node (pass of VP
// neighbours starting at vb).  Returns the estimate on every lane of the
// neighbour's lanes.
template <int CW>
__device__ __forceinline__ float lane_estimate(const uint32_t (&cw)[CW], const uint8_t* lut_part,
                                               uint32_t lpv, const Prep& pr, float sqr_y,
                                               float tx, float dq, float vq) {
    uint32_t s = scan_sum<CW>(cw, lut_part);
    if (lpv == 2) s += __shfl_xor(s, 1);
    const int32_t r = static_cast<int32_t>(static_cast<int16_t>(static_cast<uint16_t>(s & 0xFFFFu)));
    const float f = static_cast<float>(r * 2 - pr.sumq);
    const float b = fmaf(vq, pr.lo, tx + sqr_y);
    const float c = (f * dq) * pr.width;
    return c + b;
}

// ---------------------------------------------------------------------------
template <int EF_CAP, int CW>
__device__ void h2_search_one(const H2Dev& ix, const H2Job& job, uint32_t q, const H2Lds& w,
                              uint32_t hash_bits, uint32_t* gtab, uint32_t lane) {
    constexpr int S = EF_CAP / kWave;
    const uint32_t ef =
        max(1u, min(job.ef ? job.ef[q] : job.ef_all, static_cast<uint32_t>(EF_CAP)));
    const uint32_t kq = min(job.k ? job.k[q] : job.k_all, job.kstride);
    const uint32_t deg = ix.deg, dim = ix.dim;
    const uint32_t lpv = deg >= 64 ? 1u : 2u;
    const uint32_t vp = kWave / lpv;              // neighbours per pass
    const uint32_t cpl = (ix.p / 4) / lpv;        // codebooks per lane
    const uint32_t code_bytes = ix.p / 8;         // per neighbour
    const uint32_t part = lane % lpv;
    const uint8_t* lut_part = w.lut + 16u * cpl * part + (part ? 16u : 0u);

    const Prep pr = prepare(ix, job.queries + static_cast<size_t>(q) * dim,
                            job.rot ? job.rot + static_cast<size_t>(q) * ix.p : nullptr, w, lpv,
                            lane);

    Visited vis;
    vis.lds = w.hv;
    vis.glb = gtab;
    vis.lbits = hash_bits;
    vis.gbits = job.gvis_bits;
    vis.count = 0;
    vis.global = false;
    vis.failed = false;
    for (uint32_t i = lane; i < (1u << hash_bits); i += kWave) w.hv[i] = kEmpty;
    wave_sync();

    // search_pool_.insert(entry_point_, FLT_MAX) (qg.hpp:260)
    uint64_t P[S];
    uint32_t I[S];
#pragma unroll
    for (int s = 0; s < S; ++s) {
        P[s] = ~0ull;
        I[s] = 0;
    }
    if (lane == 0) {
        P[0] = sb_key(3.402823466e+38f, 0);
        I[0] = ix.ep;
    }
    uint32_t size = 1, seq = 1, n_exp = 0, n_pop = 0, rsize = 0;
    float rtail = INFINITY;  // ResultBuffer's last distance once it holds kq entries
    // query dims lane + 16u (u < kRegJ) for the partials of exact_l2_regs
    const uint32_t mul16 = dim & ~15u;
    float qr[kRegJ];
#pragma unroll
    for (int u = 0; u < kRegJ; ++u) {
        const uint32_t d = lane + 16u * static_cast<uint32_t>(u);
        qr[u] = (lane < 16 && d < mul16) ? w.qv[d] : 0.0f;
    }

    for (;;) {
        // pop: first unchecked entry (buffer.hpp:52-60)
        uint64_t pm = 0;
        int fs = S;
#pragma unroll
        for (int s = S - 1; s >= 0; --s) {
            const uint32_t i = static_cast<uint32_t>(s * kWave) + lane;
            const uint64_t m = __ballot(i < size && !(P[s] & 1ull));
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
            const uint32_t c = readlane32(I[s], pl);
            if (s == fs) cur = c;
            P[s] |= (s == fs && static_cast<int>(lane) == pl) ? 1ull : 0ull;
        }
        ++n_pop;
        if (v_contains_u(vis, cur)) continue;
        v_insert(vis, cur, lane);
        if (vis.failed) break;
        ++n_exp;

        const unsigned char* rec = ix.rec + static_cast<size_t>(cur) * ix.stride;
        const float* raw = reinterpret_cast<const float*>(rec);
        const unsigned char* codes = rec + ix.off_codes;
        const float* fac = reinterpret_cast<const float*>(rec + ix.off_fac);
        const uint32_t* ids = reinterpret_cast<const uint32_t*>(rec + ix.off_ids);
        // the node's raw row for the exact distance: dims lane + 16u
        float xr[kRegJ];
#pragma unroll
        for (int u = 0; u < kRegJ; ++u) {
            const uint32_t d = lane + 16u * static_cast<uint32_t>(u);
            xr[u] = (lane < 16 && d < mul16) ? raw[d] : 0.0f;
        }
        float sqr_y = 0.0f;

        for (uint32_t vb = 0; vb < deg; vb += vp) {
            // the pass's loads first (codes, factors, id)
            const uint32_t vi = vb + lane / lpv;
            const bool act = vi < deg;
            uint32_t cw[CW];
            uint32_t nid = kInvalid;
            float tx = 0.0f, dq = 0.0f, vq = 0.0f;
            if (act) {
                load_codes<CW>(cw, codes + static_cast<size_t>(vi) * code_bytes +
                                       static_cast<size_t>(part) * (code_bytes / lpv));
                nid = ids[vi];
                tx = fac[vi];
                dq = fac[deg + vi];
                vq = fac[2 * deg + vi];
            } else {
#pragma unroll
                for (int k = 0; k < CW; ++k) cw[k] = 0;
            }
            // scan_neighbors (qg.hpp:355): exact distance to the node
            if (vb == 0) sqr_y = exact_l2_regs(xr, qr, raw, w, dim, lane);
            const float est = lane_estimate<CW>(cw, lut_part, lpv, pr, sqr_y, tx, dq, vq);
            // is_full / visited_.get (qg.hpp:375-377); candidate i gets insertion
            // number seq + i.  The visited table is probed only by the lanes the
            // pool would take.
            const uint64_t key = sb_key(est, seq + vi);
            uint64_t last = 0;
#pragma unroll
            for (int s = 0; s < S; ++s)
                if (static_cast<uint32_t>(s) == (size - 1) / kWave)
                    last = readlane64(P[s], static_cast<int>((size - 1) % kWave));
            bool cand = act && part == 0 && (size < ef || key < last);
            if (__ballot(cand) == 0) continue;
            cand = cand && !v_contains_lanes(vis, nid, cand);
            uint64_t mm = __ballot(cand);
            if (mm == 0) continue;
            const uint32_t m3 = static_cast<uint32_t>(__popcll(mm));
            // per candidate j: lb_j = pool entries below it; per pool entry:
            // nlt = candidates above it, so it moves up by m3 - nlt
            uint32_t lb = 0, rank = 0;
            uint32_t nlt[S];
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
            const uint32_t nsize = min(size + m3, ef);
            uint32_t shift[S];
#pragma unroll
            for (int s = 0; s < S; ++s) shift[s] = m3 - nlt[s];
            // merge through LDS; entries that fall off the end go to the spare
            // slot EF_CAP, so the stores need no branch
#pragma unroll
            for (int s = 0; s < S; ++s) {
                const uint32_t i = static_cast<uint32_t>(s * kWave) + lane;
                const uint32_t np = i + shift[s];
                const uint32_t dst = (i < size && np < nsize) ? np : static_cast<uint32_t>(EF_CAP);
                w.sk[dst] = P[s];
                w.si[dst] = I[s];
            }
            {
                const uint32_t np = lb + rank;
                const uint32_t dst = (cand && np < nsize) ? np : static_cast<uint32_t>(EF_CAP);
                w.sk[dst] = key;
                w.si[dst] = nid;
            }
            wave_sync();
#pragma unroll
            for (int s = 0; s < S; ++s) {
                const uint32_t i = static_cast<uint32_t>(s * kWave) + lane;
                const uint64_t kv = w.sk[i];
                const uint32_t iv = w.si[i];
                P[s] = i < nsize ? kv : ~0ull;
                I[s] = i < nsize ? iv : 0u;
            }
            size = nsize;
            wave_sync();
        }
        seq += deg;
        // res_pool.insert (qg.hpp:282); skipped without a look at LDS when the
        // buffer is full and the distance is above its last
        if (kq != 0 && !(rsize == kq && sqr_y > rtail)) {
            rb_insert(w, rsize, kq, cur, sqr_y, lane);
            if (rsize == kq) rtail = w.rd[kq - 1];
        }
    }

    // update_results (qg.hpp:395-418): fewer than k nodes expanded
    if (!vis.failed && kq > 0 && rsize < kq) {
        for (uint32_t t = lane; t <= kq; t += kWave) w.snap[t] = t < rsize ? w.ri[t] : 0u;
        wave_sync();
        for (uint32_t t = 0; t <= kq && !vis.failed; ++t) {
            const uint32_t node = w.snap[t];
            const uint32_t* nb = reinterpret_cast<const uint32_t*>(
                ix.rec + static_cast<size_t>(node) * ix.stride + ix.off_ids);
            for (uint32_t i = 0; i < deg; ++i) {
                const uint32_t x = nb[i];
                if (v_contains(vis, x)) continue;
                v_insert(vis, x, lane);
                if (vis.failed) break;
                const float d = exact_l2(
                    reinterpret_cast<const float*>(ix.rec + static_cast<size_t>(x) * ix.stride), w,
                    dim, lane);
                rb_insert(w, rsize, kq, x, d, lane);
            }
            if (rsize == kq) break;
        }
    }

    // copy_results (buffer.hpp:102): unfilled slots hold id 0; a query whose
    // visited set overflowed returns no ids (0xFFFFFFFF everywhere)
    uint32_t* oi = job.out_ids + static_cast<size_t>(q) * job.kstride;
    float* od = job.out_d ? job.out_d + static_cast<size_t>(q) * job.kstride : nullptr;
    for (uint32_t j = lane; j < job.kstride; j += kWave) {
        uint32_t id = kInvalid;
        float d = INFINITY;
        if (j < kq && !vis.failed) {
            id = j < rsize ? w.ri[j] : 0u;
            if (j < rsize) d = w.rd[j];
        }
        oi[j] = id;
        if (od) od[j] = d;
    }
    if (lane == 0) {
        if (job.out_stats) {
            job.out_stats[2 * static_cast<size_t>(q)] = vis.failed ? kInvalid : n_exp;
            job.out_stats[2 * static_cast<size_t>(q) + 1] = n_pop;
        }
        if (vis.failed) atomicOr(job.err, 1u);
    }
    if (job.done) signal_host(job.done + q, lane, vis.failed ? kFlagFailed : kFlagDone);
    wave_sync();
}

template <int EF_CAP, int CW>
__device__ __forceinline__ void h2_search_body(const H2Dev& ix, const H2Job& job,
                                               uint32_t hash_bits, const Layout& lay) {
    extern __shared__ __align__(16) unsigned char lds[];
    const uint32_t lane = threadIdx.x & (kWave - 1);
    const uint32_t wid = threadIdx.x / kWave;
    const H2Lds w = carve(lds + static_cast<size_t>(wid) * lay.wave_bytes, lay);
    const uint32_t gwave = blockIdx.x * (blockDim.x / kWave) + wid;
    uint32_t* gtab = job.gvis ? job.gvis + (static_cast<size_t>(gwave) << job.gvis_bits) : nullptr;
    // wave gwave first takes query gwave, then sw + atomicAdd(counter) (sw =
    // waves in the grid), so the launch does not start with one atomic per
    // wave on a single counter
    const uint32_t sw = gridDim.x * (blockDim.x / kWave);
    uint32_t q = gwave;
    for (;;) {
        q = uniform(q);
        if (q >= job.nq) break;
        h2_search_one<EF_CAP, CW>(ix, job, q, w, hash_bits, gtab, lane);
        uint32_t nx = 0;
        if (lane == 0) nx = sw + atomicAdd(job.counter, 1u);
        q = __shfl(nx, 0);
    }
}

// Occupancy variants: uncapped registers, and capped for 5 or 6 waves per
// SIMD (a few spills traded for more waves to hide the record loads).
template <int EF_CAP, int CW>
__global__ void __launch_bounds__(256) h2_search_kernel(H2Dev ix, H2Job job, uint32_t hash_bits,
                                                         Layout lay) {
    h2_search_body<EF_CAP, CW>(ix, job, hash_bits, lay);
}

template <int EF_CAP, int CW>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(5, 5)))
h2_search_kernel_w5(H2Dev ix, H2Job job, uint32_t hash_bits, Layout lay) {
    h2_search_body<EF_CAP, CW>(ix, job, hash_bits, lay);
}

template <int EF_CAP, int CW>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(6, 6)))
h2_search_kernel_w6(H2Dev ix, H2Job job, uint32_t hash_bits, Layout lay) {
    h2_search_body<EF_CAP, CW>(ix, job, hash_bits, lay);
}

// Preparation (and one node's scan) per query, for parity checks.
template <int CW>
__global__ void __launch_bounds__(64) h2_probe_kernel(H2Dev ix, const float* queries,
                                                       const uint32_t* nodes, uint8_t* lut_out,
                                                       float* params, float* sqr_out, float* est_out,
                                                       Layout lay) {
    extern __shared__ __align__(16) unsigned char lds[];
    const uint32_t lane = threadIdx.x;
    const uint32_t q = blockIdx.x;
    const H2Lds w = carve(lds, lay);
    const uint32_t deg = ix.deg;
    const uint32_t lpv = deg >= 64 ? 1u : 2u;
    const Prep pr = prepare(ix, queries + static_cast<size_t>(q) * ix.dim, nullptr, w, lpv, lane);
    const uint32_t cpl = (ix.p / 4) / lpv;
    // LUT in the reference's layout (no bank skew)
    for (uint32_t idx = lane; idx < 4 * ix.p; idx += kWave) {
        const uint32_t c = idx >> 4;
        lut_out[static_cast<size_t>(q) * 4 * ix.p + idx] =
            w.lut[idx + (lpv == 2 && c >= cpl ? 16u : 0u)];
    }
    if (lane == 0) {
        params[3 * q] = pr.lo;
        params[3 * q + 1] = pr.width;
        params[3 * q + 2] = __int_as_float(pr.sumq);
    }
    if (!nodes) return;
    const uint32_t node = nodes[q];
    const unsigned char* rec = ix.rec + static_cast<size_t>(node) * ix.stride;
    const float sqr_y = exact_l2(reinterpret_cast<const float*>(rec), w, ix.dim, lane);
    if (lane == 0) sqr_out[q] = sqr_y;
    const uint32_t part = lane % lpv;
    const uint8_t* lut_part = w.lut + 16u * cpl * part + (part ? 16u : 0u);
    const float* fac = reinterpret_cast<const float*>(rec + ix.off_fac);
    const uint32_t code_bytes = ix.p / 8;
    for (uint32_t vb = 0; vb < deg; vb += kWave / lpv) {
        const uint32_t vi = vb + lane / lpv;
        const bool act = vi < deg;
        uint32_t cw[CW];
        float tx = 0.0f, dq = 0.0f, vq = 0.0f;
        if (act) {
            load_codes<CW>(cw, rec + ix.off_codes + static_cast<size_t>(vi) * code_bytes +
                                   static_cast<size_t>(part) * (code_bytes / lpv));
            tx = fac[vi];
            dq = fac[deg + vi];
            vq = fac[2 * deg + vi];
        } else {
#pragma unroll
            for (int k = 0; k < CW; ++k) cw[k] = 0;
        }
        const float est = lane_estimate<CW>(cw, lut_part, lpv, pr, sqr_y, tx, dq, vq);
        if (act && part == 0) est_out[static_cast<size_t>(q) * deg + vi] = est;
    }
}

// ---------------------------------------------------------------------------
// Sidecar build.  rot[u] = FHT rotation of raw[u] (one wave per node, LDS).
__global__ void __launch_bounds__(256) h2_rotate_kernel(H2Dev ix, float* rot) {
    extern __shared__ __align__(16) unsigned char lds[];
    const uint32_t lane = threadIdx.x & (kWave - 1);
    const uint32_t wid = threadIdx.x / kWave;
    const uint32_t u = blockIdx.x * (blockDim.x / kWave) + wid;
    if (u >= ix.n) return;
    float* t = reinterpret_cast<float*>(lds) + static_cast<size_t>(wid) * ix.p;
    const float* src = reinterpret_cast<const float*>(ix.rec + static_cast<size_t>(u) * ix.stride);
    for (uint32_t i = lane; i < ix.p; i += kWave) t[i] = i < ix.dim ? src[i] * ix.signs[i] : 0.0f;
    wave_sync();
    fht_lds(t, ix.p, lane);
    for (uint32_t i = lane; i < ix.p; i += kWave) rot[static_cast<size_t>(u) * ix.p + i] = t[i];
}

// Queries rotated by the FHT (one wave per query): what prepare computes.
__global__ void __launch_bounds__(256) h2_rotq_fht_kernel(H2Dev ix, const float* queries,
                                                           uint32_t nq, float* out) {
    extern __shared__ __align__(16) unsigned char lds[];
    const uint32_t lane = threadIdx.x & (kWave - 1);
    const uint32_t wid = threadIdx.x / kWave;
    const uint32_t q = blockIdx.x * (blockDim.x / kWave) + wid;
    if (q >= nq) return;
    float* t = reinterpret_cast<float*>(lds) + static_cast<size_t>(wid) * ix.p;
    const float* src = queries + static_cast<size_t>(q) * ix.dim;
    for (uint32_t i = lane; i < ix.p; i += kWave) t[i] = i < ix.dim ? src[i] * ix.signs[i] : 0.0f;
    wave_sync();
    fht_lds(t, ix.p, lane);
    for (uint32_t i = lane; i < ix.p; i += kWave) out[static_cast<size_t>(q) * ix.p + i] = t[i];
}

// Queries rotated by one FP32 GEMM on the matrix cores: out[nq x p] =
// Q[nq x dim] M[dim x p] with M[i][j] = signs[i] (-1)^popcount(i & j), the
// FHT's matrix (H_p diag(signs), Sylvester order), generated in registers.
// Each wave computes a 16-query x 64-column block with
// v_mfma_f32_16x16x4_f32: lane l supplies A[l % 16][k0 + l / 16] and
// B[k0 + l / 16][l % 16] per step of 4 and holds D[4 (l / 16) + r][l % 16],
// r = 0..3.  Same products as the FHT, different summation order: the
// rounding differs from the reference's rotation by a few ulps.
typedef float floatx4 __attribute__((ext_vector_type(4)));

__global__ void __launch_bounds__(256) h2_rotq_gemm_kernel(H2Dev ix, const float* queries,
                                                            uint32_t nq, float* out) {
    const uint32_t lane = threadIdx.x & (kWave - 1);
    const uint32_t gw = blockIdx.x * (blockDim.x / kWave) + threadIdx.x / kWave;
    const uint32_t ct_n = ix.p / 64;
    const uint32_t qt = gw / ct_n, ct = gw % ct_n;
    if (qt * 16 >= nq) return;
    const uint32_t li = lane & 15u, lk = lane >> 4;
    const uint32_t row = qt * 16 + li;
    const float* qrow = queries + static_cast<size_t>(row < nq ? row : 0) * ix.dim;
    floatx4 acc[4];
#pragma unroll
    for (int t = 0; t < 4; ++t) acc[t] = floatx4{0.0f, 0.0f, 0.0f, 0.0f};
    for (uint32_t k0 = 0; k0 < ix.dim; k0 += 4) {
        const uint32_t k = k0 + lk;
        const bool in = k < ix.dim;
        const float a = in && row < nq ? qrow[k] : 0.0f;
        const float sk = in ? ix.signs[k] : 0.0f;
#pragma unroll
        for (int t = 0; t < 4; ++t) {
            const uint32_t j = ct * 64 + 16 * static_cast<uint32_t>(t) + li;
            const float b = (__popc(k & j) & 1u) ? -sk : sk;
            acc[t] = __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, acc[t], 0, 0, 0);
        }
    }
#pragma unroll
    for (int t = 0; t < 4; ++t)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const uint32_t orow = qt * 16 + 4 * lk + static_cast<uint32_t>(r);
            const uint32_t col = ct * 64 + 16 * static_cast<uint32_t>(t) + li;
            if (orow < nq) out[static_cast<size_t>(orow) * ix.p + col] = acc[t][r];
        }
}

// One thread per (node u, neighbour slot j): residual o = R x_j - R x_u, sign
// codes, and the three factors with every sum taken over the dims in order,
// in double:
//   fn = float(1/sqrt(P)); nrm = float(sqrt(sum o^2));
//   x0 = float(float(sum |o| * fn) / nrm); x1 = float(sum +-c * fn);
//   x_x0 = nrm / x0; triple_x = float(nrm^2 + 2 x_x0 x1);
//   fac_dq = float(-2 x_x0 fn); fac_vq = float(-2 x_x0 fn (2 popcount - P))
// (rabitq.hpp:61-108).  An all-zero residual gets factors (0, 0, 0): its
// estimate is then the node's exact distance.
__global__ void __launch_bounds__(256) h2_sidecar_kernel(H2Dev ix, const float* rot, double fn) {
    const uint64_t t = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (t >= static_cast<uint64_t>(ix.n) * ix.deg) return;
    const uint32_t u = static_cast<uint32_t>(t / ix.deg), j = static_cast<uint32_t>(t % ix.deg);
    unsigned char* rec = const_cast<unsigned char*>(ix.rec) + static_cast<size_t>(u) * ix.stride;
    const uint32_t v = reinterpret_cast<const uint32_t*>(rec + ix.off_ids)[j];
    const float* c = rot + static_cast<size_t>(u) * ix.p;
    const float* x = rot + static_cast<size_t>(v) * ix.p;
    uint32_t* code = reinterpret_cast<uint32_t*>(rec + ix.off_codes + static_cast<size_t>(j) * (ix.p / 8));
    double s2 = 0.0, sabs = 0.0, sc = 0.0;
    uint32_t pop = 0;
    for (uint32_t i0 = 0; i0 < ix.p; i0 += 8) {
        // 8 dims -> one code byte: low nibble dims i0..i0+3 (bit 3 = i0), high
        // nibble dims i0+4..i0+7
        const float4 xa = *reinterpret_cast<const float4*>(x + i0);
        const float4 xb = *reinterpret_cast<const float4*>(x + i0 + 4);
        const float4 ca = *reinterpret_cast<const float4*>(c + i0);
        const float4 cb = *reinterpret_cast<const float4*>(c + i0 + 4);
        const float xs[8] = {xa.x, xa.y, xa.z, xa.w, xb.x, xb.y, xb.z, xb.w};
        const float cs[8] = {ca.x, ca.y, ca.z, ca.w, cb.x, cb.y, cb.z, cb.w};
        uint32_t byte = 0;
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            const float o = xs[k] - cs[k];
            const bool bit = o > 0.0f;
            if (bit) {
                byte |= 1u << ((k < 4 ? 0 : 4) + 3 - (k & 3));
                ++pop;
            }
            s2 += static_cast<double>(o) * static_cast<double>(o);
            sabs += bit ? static_cast<double>(o) : -static_cast<double>(o);
            sc += bit ? static_cast<double>(cs[k]) : -static_cast<double>(cs[k]);
        }
        // 4 bytes per word; i0 / 8 = byte index
        const uint32_t bi = i0 / 8;
        if ((bi & 3u) == 0) code[bi / 4] = 0u;
        code[bi / 4] |= byte << (8 * (bi & 3u));
    }
    float* tx = reinterpret_cast<float*>(rec + ix.off_fac);
    float* dq = tx + ix.deg;
    float* vq = tx + 2 * ix.deg;
    if (s2 == 0.0) {
        tx[j] = 0.0f;
        dq[j] = 0.0f;
        vq[j] = 0.0f;
        return;
    }
    const float nrm = static_cast<float>(sqrt(s2));
    const float x0 = static_cast<float>(static_cast<float>(sabs * fn) / nrm);
    const float x1 = static_cast<float>(sc * fn);
    const double x_x0 = static_cast<double>(nrm) / static_cast<double>(x0);
    tx[j] = static_cast<float>(static_cast<double>(nrm) * static_cast<double>(nrm) +
                               2.0 * x_x0 * static_cast<double>(x1));
    dq[j] = static_cast<float>(-2.0 * x_x0 * fn);
    vq[j] = static_cast<float>(-2.0 * x_x0 * fn *
                               static_cast<double>(2 * static_cast<int>(pop) - static_cast<int>(ix.p)));
}

// Copies w bytes per node (w a multiple of 4) between a contiguous n x w
// array and offset `off` of every record; to_rec selects the direction.
__global__ void __launch_bounds__(256) h2_rows_kernel(unsigned char* rec, uint32_t stride,
                                                       uint32_t off, unsigned char* buf,
                                                       uint32_t w, uint32_t n, int to_rec) {
    const uint32_t words = w / 4;
    const uint64_t t = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (t >= static_cast<uint64_t>(n) * words) return;
    const uint64_t u = t / words, i = t % words;
    uint32_t* r = reinterpret_cast<uint32_t*>(rec + u * stride + off) + i;
    uint32_t* b = reinterpret_cast<uint32_t*>(buf + u * w) + i;
    if (to_rec)
        *r = *b;
    else
        *b = *r;
}

// ---------------------------------------------------------------------------
using SearchFn = void (*)(H2Dev, H2Job, uint32_t, Layout);
using ProbeFn = void (*)(H2Dev, const float*, const uint32_t*, uint8_t*, float*, float*, float*,
                         Layout);

// Code words (4 bytes) a lane reads per neighbour pass.
int code_words(const H2Dev& ix) {
    const uint32_t lpv = ix.deg >= 64 ? 1u : 2u;
    return static_cast<int>(ix.p / 8 / lpv / 4);
}

template <int EF_CAP>
SearchFn search_for_cw(int cw, int occ) {
    if (occ == 5) {
        switch (cw) {
            case 1: return &h2_search_kernel_w5<EF_CAP, 1>;
            case 2: return &h2_search_kernel_w5<EF_CAP, 2>;
            case 4: return &h2_search_kernel_w5<EF_CAP, 4>;
            default: return nullptr;
        }
    }
    if (occ == 6) {
        switch (cw) {
            case 1: return &h2_search_kernel_w6<EF_CAP, 1>;
            case 2: return &h2_search_kernel_w6<EF_CAP, 2>;
            case 4: return &h2_search_kernel_w6<EF_CAP, 4>;
            default: return nullptr;
        }
    }
    if (occ != 0) return nullptr;
    switch (cw) {
        case 1: return &h2_search_kernel<EF_CAP, 1>;
        case 2: return &h2_search_kernel<EF_CAP, 2>;
        case 4: return &h2_search_kernel<EF_CAP, 4>;
        case 8: return &h2_search_kernel<EF_CAP, 8>;
        case 16: return &h2_search_kernel<EF_CAP, 16>;
        case 32: return &h2_search_kernel<EF_CAP, 32>;
        default: return nullptr;
    }
}

SearchFn select_search(int ef_cap, int cw, int occ) {
    switch (ef_cap) {
        case 64: return search_for_cw<64>(cw, occ);
        case 128: return search_for_cw<128>(cw, occ);
        case 192: return search_for_cw<192>(cw, occ);
        case 256: return search_for_cw<256>(cw, occ);
        case 512: return search_for_cw<512>(cw, occ);
        case 1024: return search_for_cw<1024>(cw, occ);
        default: return nullptr;
    }
}

ProbeFn select_probe(int cw) {
    switch (cw) {
        case 1: return &h2_probe_kernel<1>;
        case 2: return &h2_probe_kernel<2>;
        case 4: return &h2_probe_kernel<4>;
        case 8: return &h2_probe_kernel<8>;
        case 16: return &h2_probe_kernel<16>;
        case 32: return &h2_probe_kernel<32>;
        default: return nullptr;
    }
}

}  // namespace

bool h2_occupancy_ok(const H2Dev& ix) { return code_words(ix) <= 4; }

size_t h2_lds_bytes(const H2Dev& ix, const H2Cfg& cfg, uint32_t kmax) {
    const Layout l = h2_layout(ix.dim, ix.p, static_cast<uint32_t>(cfg.ef_cap),
                               static_cast<uint32_t>(cfg.hash_bits), kmax);
    return static_cast<size_t>(cfg.waves_per_group) * l.wave_bytes;
}

int h2_groups(const H2Dev& ix, const H2Cfg& cfg, uint32_t kmax, uint32_t nq) {
    const int need = static_cast<int>((nq + cfg.waves_per_group - 1) / cfg.waves_per_group);
    int groups = cfg.groups;
    if (groups <= 0) {
        SearchFn k = select_search(cfg.ef_cap, code_words(ix), cfg.occupancy);
        int dev = 0, cus = 0, per_cu = 0;
        if (!k || hipGetDevice(&dev) != hipSuccess ||
            hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
            return 0;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, reinterpret_cast<const void*>(k),
                                                         kWave * cfg.waves_per_group,
                                                         h2_lds_bytes(ix, cfg, kmax)) != hipSuccess ||
            per_cu < 1)
            per_cu = 1;
        groups = per_cu * cus;
    }
    if (groups > need) groups = need > 0 ? need : 1;
    return groups;
}

hipError_t h2_launch_search(const H2Dev& ix, const H2Job& job, const H2Cfg& cfg,
                            hipStream_t stream) {
    SearchFn k = select_search(cfg.ef_cap, code_words(ix), cfg.occupancy);
    if (!k || cfg.waves_per_group < 1 || cfg.waves_per_group > 4 || cfg.hash_bits < 6 ||
        cfg.hash_bits > 14 || job.kstride == 0)
        return hipErrorInvalidValue;
    const Layout lay = h2_layout(ix.dim, ix.p, static_cast<uint32_t>(cfg.ef_cap),
                                 static_cast<uint32_t>(cfg.hash_bits), job.kstride);
    const size_t lds = static_cast<size_t>(cfg.waves_per_group) * lay.wave_bytes;
    if (lds > 160 * 1024) return hipErrorInvalidValue;
    if (lds > 64 * 1024) {
        hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(k),
                                           hipFuncAttributeMaxDynamicSharedMemorySize,
                                           static_cast<int>(lds));
        if (e != hipSuccess) return e;
    }
    const int groups = h2_groups(ix, cfg, job.kstride, job.nq);
    if (groups <= 0) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k, dim3(groups), dim3(kWave * cfg.waves_per_group), lds, stream, ix, job,
                       static_cast<uint32_t>(cfg.hash_bits), lay);
    return hipGetLastError();
}

hipError_t h2_launch_probe(const H2Dev& ix, const float* queries, uint32_t nq,
                           const uint32_t* nodes, uint8_t* lut, float* params, float* sqr_y,
                           float* est, hipStream_t stream) {
    ProbeFn k = select_probe(code_words(ix));
    if (!k || nq == 0) return hipErrorInvalidValue;
    const Layout lay = h2_layout(ix.dim, ix.p, 64, 6, 1);
    hipLaunchKernelGGL(k, dim3(nq), dim3(kWave), lay.wave_bytes, stream, ix, queries, nodes, lut,
                       params, sqr_y, est, lay);
    return hipGetLastError();
}

hipError_t h2_launch_sidecar(const H2Dev& ix, float* rot, hipStream_t stream) {
    const uint32_t waves = 4;
    hipLaunchKernelGGL(h2_rotate_kernel, dim3((ix.n + waves - 1) / waves), dim3(kWave * waves),
                       static_cast<size_t>(waves) * ix.p * 4, stream, ix, rot);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    const double fn = static_cast<double>(1.0f / sqrtf(static_cast<float>(ix.p)));
    const uint64_t threads = static_cast<uint64_t>(ix.n) * ix.deg;
    hipLaunchKernelGGL(h2_sidecar_kernel, dim3(static_cast<uint32_t>((threads + 255) / 256)),
                       dim3(256), 0, stream, ix, rot, fn);
    return hipGetLastError();
}

hipError_t h2_launch_rotate_queries(const H2Dev& ix, const float* queries, uint32_t nq,
                                    int method, float* out, hipStream_t stream) {
    if (nq == 0 || ix.p < 64 || (method != 0 && method != 1)) return hipErrorInvalidValue;
    const uint32_t waves = 4;
    if (method == 0) {
        hipLaunchKernelGGL(h2_rotq_fht_kernel, dim3((nq + waves - 1) / waves), dim3(kWave * waves),
                           static_cast<size_t>(waves) * ix.p * 4, stream, ix, queries, nq, out);
    } else {
        const uint64_t tiles = static_cast<uint64_t>((nq + 15) / 16) * (ix.p / 64);
        hipLaunchKernelGGL(h2_rotq_gemm_kernel, dim3(static_cast<uint32_t>((tiles + waves - 1) / waves)),
                           dim3(kWave * waves), 0, stream, ix, queries, nq, out);
    }
    return hipGetLastError();
}

hipError_t h2_launch_rows(const H2Dev& ix, uint32_t off, void* buf, uint32_t w, bool to_rec,
                          hipStream_t stream) {
    if (w % 4 != 0 || off % 4 != 0) return hipErrorInvalidValue;
    const uint64_t threads = static_cast<uint64_t>(ix.n) * (w / 4);
    hipLaunchKernelGGL(h2_rows_kernel, dim3(static_cast<uint32_t>((threads + 255) / 256)),
                       dim3(256), 0, stream, const_cast<unsigned char*>(ix.rec), ix.stride, off,
                       static_cast<unsigned char*>(buf), w, ix.n, to_rec ? 1 : 0);
    return hipGetLastError();
}

}  // namespace lqg
