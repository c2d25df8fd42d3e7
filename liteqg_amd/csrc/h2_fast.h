// h2_fast.h -- the FastScan search of h2_kernel.h (QuantizedGraph::search_qg,
// qg.hpp:252-287) specialised for one record shape: D = 16 NJ dimensions
// (NJ = 4, 6 or 8, so P = 64 or 128) and degree 32 (two lanes per
// neighbour, all 32 neighbours of a node in one pass).  Same results as
// h2_search_one, with less work per expansion:
//   - every record offset is a compile-time immediate, so each load is one
//     instruction on a scalar base (the node's record) and a per-lane offset;
//   - the exact distance's 16 partial chains sit in registers whole (no
//     dimension is left for LDS);
//   - a pool entry's checked flag is the top bit of its id word (ids are
//     below 2^31), and the unchecked entries of each register slot are kept
//     as a scalar lane mask, so a pop is scalar work plus one lane write;
//   - the key of entry ef - 1 is kept in scalar registers once the pool is
//     full, so admission is one 64-bit compare per neighbour;
//   - the rank of each admitted neighbour is counted with scalar popcounts and
//     written to its lane, and the merge writes only the entries that stay.
#pragma once

#include "h2_kernel.h"

namespace lqg {

// Record shape of the specialised kernel (the layout set_layout gives these
// shapes: raw row, codes, three factor arrays, ids, 128-byte stride).
template <int NJ>
struct FastShape {
    static constexpr uint32_t kDim = 16u * NJ;
    static constexpr uint32_t kP = kDim <= 64u ? 64u : 128u;  // 1 << ceil(log2 D)
    static constexpr uint32_t kDeg = 32;
    static constexpr uint32_t kCodeBytes = kP / 8u;             // per neighbour
    static constexpr int kCW = static_cast<int>(kP / 64u);      // code words per lane
    static constexpr uint32_t kOffCodes = 4u * kDim;
    static constexpr uint32_t kOffFac = kOffCodes + kDeg * kCodeBytes;
    static constexpr uint32_t kOffIds = kOffFac + 12u * kDeg;
    static constexpr uint32_t kStride = (kOffIds + 4u * kDeg + 127u) / 128u * 128u;
};

// NJ of the specialised kernel for this index, or 0 if it has another shape.
__host__ inline int h2_fast_nj(const H2Dev& ix) {
    if (ix.deg != 32 || ix.n >= 0x80000000u) return 0;
    int nj = 0;
    switch (ix.dim) {
        case 64: nj = 4; break;
        case 96: nj = 6; break;
        case 128: nj = 8; break;
        default: return 0;
    }
    uint32_t p = 0, oc = 0, of = 0, oi = 0, st = 0;
    switch (nj) {
        case 4: p = FastShape<4>::kP; oc = FastShape<4>::kOffCodes; of = FastShape<4>::kOffFac;
                oi = FastShape<4>::kOffIds; st = FastShape<4>::kStride; break;
        case 6: p = FastShape<6>::kP; oc = FastShape<6>::kOffCodes; of = FastShape<6>::kOffFac;
                oi = FastShape<6>::kOffIds; st = FastShape<6>::kStride; break;
        default: p = FastShape<8>::kP; oc = FastShape<8>::kOffCodes; of = FastShape<8>::kOffFac;
                oi = FastShape<8>::kOffIds; st = FastShape<8>::kStride; break;
    }
    if (ix.p != p || ix.off_codes != oc || ix.off_fac != of || ix.off_ids != oi || ix.stride != st)
        return 0;
    return nj;
}

namespace {

constexpr uint32_t kChecked = 0x80000000u;

// old with lane `l` replaced by v.  The lane select goes through M0 (one
// SGPR operand per VALU instruction); the s_nop covers a lane select that a
// VALU instruction wrote.
__device__ __forceinline__ uint32_t write_lane(uint32_t old, uint32_t v, uint32_t l) {
    asm("s_nop 3\n\tv_writelane_b32 %0, %1, m0" : "+v"(old) : "s"(uniform(v)), "{m0}"(uniform(l)));
    return old;
}

// m with bit b cleared (one scalar instruction).
__device__ __forceinline__ uint64_t clear_bit(uint64_t m, uint32_t b) {
    m = (static_cast<uint64_t>(uniform(static_cast<uint32_t>(m >> 32))) << 32) |
        uniform(static_cast<uint32_t>(m));
    asm("s_bitset0_b64 %0, %1" : "+s"(m) : "s"(uniform(b)));
    return m;
}

// Pops the first unchecked entry of the pool (slot s on): its id, marked
// checked in I and cleared from U; kEmpty if there is none.
template <int S, int s = 0>
__device__ __forceinline__ uint32_t pop_first(uint64_t (&U)[S], uint32_t (&I)[S]) {
    if constexpr (s == S) {
        return kEmpty;
    } else {
        if (U[s]) {
            const uint32_t pl = static_cast<uint32_t>(__builtin_ctzll(U[s]));
            U[s] = clear_bit(U[s], pl);
            const uint32_t cur = readlane32(I[s], static_cast<int>(pl));
            I[s] = write_lane(I[s], cur | kChecked, pl);
            return cur;
        }
        return pop_first<S, s + 1>(U, I);
    }
}

// scan_sum with the eight lookups of a code word issued before their sums.
template <int CW>
__device__ __forceinline__ uint32_t scan_sum_grouped(const uint32_t (&cw)[CW], LutBase lb) {
    uint32_t s = 0;
#pragma unroll
    for (int k = 0; k < CW; ++k) {
        const uint32_t t = cw[k] & 0x0F0F0F0Fu, v = (cw[k] >> 4) & 0x0F0F0F0Fu;
        uint32_t e[4], o[4];
#define LQG_SCAN_BYTE2(J)                                                                   \
    {                                                                                       \
        const uint32_t b = 4u * static_cast<uint32_t>(k) + (J);                             \
        e[J] = reinterpret_cast<lds_u8*>(static_cast<uintptr_t>(add_byte<J>(lb.even, t)))[16u * b]; \
        o[J] = reinterpret_cast<lds_u8*>(static_cast<uintptr_t>(add_byte<J>(lb.odd, v)))[16u * b]; \
    }
        LQG_SCAN_BYTE2(0)
        LQG_SCAN_BYTE2(1)
        LQG_SCAN_BYTE2(2)
        LQG_SCAN_BYTE2(3)
#undef LQG_SCAN_BYTE2
        __builtin_amdgcn_sched_barrier(0);
        s += (e[0] + o[0]) + (e[1] + o[1]) + (e[2] + o[2]) + (e[3] + o[3]);
    }
    return s;
}

// lane_estimate for two lanes per neighbour, the pair sum one DPP add.
template <int CW>
__device__ __forceinline__ float estimate2(const uint32_t (&cw)[CW], LutBase lb, const Prep& pr,
                                           float sqr_y, float tx, float dq, float vq) {
    uint32_t s = scan_sum_grouped<CW>(cw, lb);
    // s + the other lane's s (quad_perm [1, 0, 3, 2]); the s_nop covers the
    // VALU write of s before a DPP read
    asm("s_nop 1\n\tv_add_u32_dpp %0, %0, %0 quad_perm:[1,0,3,2] row_mask:0xf bank_mask:0xf"
        : "+v"(s));
    const int32_t r = static_cast<int32_t>(static_cast<int16_t>(static_cast<uint16_t>(s & 0xFFFFu)));
    const float f = static_cast<float>(r * 2 - pr.sumq);
    const float b = fmaf(vq, pr.lo, tx + sqr_y);
    const float c = (f * dq) * pr.width;
    return c + b;
}

// visited_.get for the lanes in `active` (the LDS table): the lanes whose id
// is in the set.  Every lane reads its first bucket; further buckets are read
// only while some active lane has not met its id or an empty slot.
__device__ __forceinline__ uint64_t v_seen_lanes(const Visited<false>& v, uint32_t id,
                                                 uint64_t active) {
    const uint32_t bmask = (1u << (v.bits - 2u)) - 1u;
    uint32_t b = vbucket(id, v.bits) >> 2;
    const uint4* t4 = reinterpret_cast<const uint4*>(v.tab);
    uint4 e = t4[b];
    uint64_t found =
        (__ballot(e.x == id) | __ballot(e.y == id) | __ballot(e.z == id) | __ballot(e.w == id)) &
        active;
    uint64_t more = active & ~found & __ballot(e.w != kEmpty);
    while (more) {
        b = (b + 1u) & bmask;
        e = t4[b];
        const uint64_t f =
            (__ballot(e.x == id) | __ballot(e.y == id) | __ballot(e.z == id) | __ballot(e.w == id)) &
            more;
        found |= f;
        more &= ~f & __ballot(e.w != kEmpty);
    }
    return found;
}

// v_test_set on the LDS table with the first-hit / first-empty comparison done
// on the lane masks.
__device__ __forceinline__ uint32_t v_test_set_fast(Visited<false>& v, uint32_t id, uint32_t lane) {
    const uint32_t mask = (1u << v.bits) - 1u;
    uint32_t h = vbucket(id, v.bits);
    uint32_t slot = 0;
    uint64_t hit = 0, emp = 0;
    for (;;) {
        slot = (h + lane) & mask;
        const uint32_t e = v.tab[slot];
        hit = __ballot(e == id);
        emp = __ballot(e == kEmpty);
        if (hit | emp) break;
        h = (h + kWave) & mask;
    }
    // a hit before the first empty slot (all of them when there is none)
    const uint64_t first_emp = emp & (0ull - emp);
    if (hit & (first_emp - 1ull)) return kVisSeen;
    if (v.count + 1u > vis_limit(v.bits)) return kVisFull;
    if (lane == static_cast<uint32_t>(__builtin_ctzll(emp))) v.tab[slot] = id;
    wave_sync();
    ++v.count;
    return kVisNew;
}

// res_pool.insert (buffer.hpp:84-96) for x >= 0 (a sum of squares, or +inf):
// as rb_insert, with the skip test on the float bits and the position counted
// over all entries (entries past k are stale copies pushed out, never below
// the tail, so never below an x that is inserted).
template <int KS>
__device__ __forceinline__ void rb_insert_fast(ResBuf<KS>& r, uint32_t kq, uint32_t id, float x,
                                               uint32_t lane) {
    if (r.size == kq && __float_as_uint(x) > __float_as_uint(r.tail)) return;
    uint32_t pos = 0;
#pragma unroll
    for (int s = 0; s < KS; ++s) pos += static_cast<uint32_t>(__popcll(__ballot(r.rd[s] < x)));
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

// Exact distance (space/l2.hpp:16-34) with all NJ partials of lane (lane & 15)
// in registers: the chain of exact_l2_regs, then its DPP tree.
template <int NJ>
__device__ __forceinline__ float exact_fast(const float (&xr)[NJ], const float (&qr)[NJ]) {
    float s = 0.0f;
#pragma unroll
    for (int u = 0; u < NJ; ++u) {
        const float t = qr[u] - xr[u];
        s = fmaf(t, t, s);
    }
    s = s + dpp_shl(s, 8);
    s = s + dpp_shl(s, 4);
    s = s + dpp_shl(s, 2);
    s = s + dpp_shl(s, 1);
    return __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(s)));
}

template <int NJ>
__device__ __forceinline__ void load_row(float (&r)[NJ], gptr<float> row, uint32_t lane) {
#pragma unroll
    for (int u = 0; u < NJ; ++u) r[u] = row[(lane & 15u) + 16u * static_cast<uint32_t>(u)];
}

// One query, the visited set in the wave's LDS table (the search kernel's
// GLOBAL = false case of h2_search_one, with the same overflow deferral).
template <int EF_CAP, int NJ, int KS>
__device__ bool h2_fast_one(const H2Dev& ix, const H2Job& job, uint32_t q, const H2Lds& w,
                            uint32_t hash_bits, uint32_t lane) {
    using Sh = FastShape<NJ>;
    constexpr int S = EF_CAP / kWave;
    constexpr int CW = Sh::kCW;
    const uint32_t ef =
        uniform(max(1u, min(job.ef ? job.ef[q] : job.ef_all, static_cast<uint32_t>(EF_CAP))));
    const uint32_t kq = uniform(min(job.k ? job.k[q] : job.k_all, job.kstride));
    const uint32_t part = lane & 1u;
    const uint32_t lv = lane >> 1;  // the lane's neighbour
    const LutBase lutb = lut_base<CW>(w.lut, Sh::kP, part);
    const float* qsrc = job.queries + static_cast<size_t>(q) * Sh::kDim;
    Prep pr = prepare(ix, qsrc, job.rot ? job.rot + static_cast<size_t>(q) * Sh::kP : nullptr, w,
                      2u, lane, false);
    pr.lo = __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(pr.lo)));
    pr.width = __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(pr.width)));
    pr.sumq = __builtin_amdgcn_readfirstlane(pr.sumq);

    Visited<false> vis;
    vis.tab = w.hv;
    vis.bits = hash_bits;
    vis.count = 0;
    for (uint32_t i = lane; i < (1u << vis.bits); i += kWave) vis.tab[i] = kEmpty;
    // positions the pool has not reached read back as empty entries
    for (uint32_t i = lane; i < static_cast<uint32_t>(EF_CAP); i += kWave)
        w.st[i] = make_uint4(kEmpty, kEmpty, kEmpty, 0u);
    wave_sync();

    // search_pool_.insert(entry_point_, FLT_MAX) (qg.hpp:260)
    uint64_t P[S];
    uint32_t I[S];
    uint64_t U[S];  // per slot: lanes holding an unchecked entry
#pragma unroll
    for (int s = 0; s < S; ++s) {
        P[s] = ~0ull;
        I[s] = kEmpty;
        U[s] = 0;
    }
    if (lane == 0) {
        P[0] = sb_key(3.402823466e+38f, 0);
        I[0] = ix.ep;
    }
    U[0] = 1;
    uint64_t last = ~0ull;  // key of entry ef - 1 once the pool is full
    uint32_t size = 1, seq = 1, n_exp = 0, n_pop = 0;
    bool full = false;  // the visited set would pass vis_limit of its table
    ResBuf<KS> rb;
    rb_init(rb);
    float qr[NJ];
    load_row<NJ>(qr, uniform_gptr<float>(qsrc), lane);
    const uint32_t code_w = (Sh::kOffCodes + lv * Sh::kCodeBytes + part * (Sh::kCodeBytes / 2)) / 4u;
    const uint32_t lv2 = 2u * lv;
    constexpr uint64_t kEvenLanes = 0x5555555555555555ull;  // a neighbour's first lane

    for (;;) {
        // pop: first unchecked entry (buffer.hpp:52-60)
        const uint32_t cur = pop_first<S>(U, I);
        if (cur == kEmpty) break;  // ids are below 2^31
        ++n_pop;
        // the node's record (raw row partials, codes, factors, ids), asked for
        // before the visited test so that the two latencies overlap; a node
        // already expanded (about one pop in eight) leaves it unused
        const unsigned char* rec = ix.rec + static_cast<size_t>(cur) * Sh::kStride;
        const gptr<float> rf = uniform_gptr<float>(rec);
        const gptr<uint32_t> ru = uniform_gptr<uint32_t>(rec);
        float xr[NJ];
        load_row<NJ>(xr, rf, lane);
        uint32_t cw[CW];
        load_codes<CW>(cw, ru, code_w);
        const uint32_t nid = ru[Sh::kOffIds / 4u + lv];
        const float tx = rf[Sh::kOffFac / 4u + lv];
        const float dq = rf[Sh::kOffFac / 4u + Sh::kDeg + lv];
        const float vq = rf[Sh::kOffFac / 4u + 2u * Sh::kDeg + lv];
        const uint32_t vs = v_test_set_fast(vis, cur, lane);
        if (vs != kVisNew) {
            if (vs == kVisSeen) continue;
            full = true;
            break;
        }
        ++n_exp;
        // scan_neighbors (qg.hpp:355): exact distance to the node, estimates
        const float sqr_y = exact_fast<NJ>(xr, qr);
        const float est = estimate2<CW>(cw, lutb, pr, sqr_y, tx, dq, vq);
        // sb_key(est, seq + lv)
        const uint32_t klo = ((0x7FFFFFFFu - seq) << 1) - lv2;
        const uint64_t key = (static_cast<uint64_t>(sortable(est)) << 32) | klo;
        // is_full / visited_.get (qg.hpp:375-377)
        uint64_t cm = __ballot(key < last) & kEvenLanes;
        if (cm) cm &= ~v_seen_lanes(vis, nid, cm);
        if (cm) {
            const uint32_t m3 = static_cast<uint32_t>(__popcll(cm));
            // per admitted neighbour j: its new position (pool entries and
            // admitted neighbours below it), written to its lane; per pool
            // entry: nlt = admitted neighbours above it
            uint32_t nlt[S];
#pragma unroll
            for (int s = 0; s < S; ++s) nlt[s] = 0;
            uint32_t fp = 0xFFFFu;  // lanes not admitted keep a position past the pool
            // two neighbours per step (independent chains); a step with one
            // left takes key 0 for the second, which is above nothing
            uint64_t mm = cm;
            while (mm) {
                const uint32_t j1 = static_cast<uint32_t>(__builtin_ctzll(mm));
                mm = clear_bit(mm, j1);
                const bool two = mm != 0;
                const uint32_t j2 = two ? static_cast<uint32_t>(__builtin_ctzll(mm)) : j1;
                mm = clear_bit(mm, j2);
                const uint64_t k1 = readlane64(key, static_cast<int>(j1));
                const uint64_t k2 = two ? readlane64(key, static_cast<int>(j2)) : 0ull;
                // every compare first, each into its own lane mask, then the
                // counts (no compare waits for the scalar unit)
                const uint64_t r1 = __ballot(key < k1), r2 = __ballot(key < k2);
                bool l1[S], l2[S];
                uint64_t a1[S], a2[S];
#pragma unroll
                for (int s = 0; s < S; ++s) {
                    l1[s] = P[s] < k1;
                    l2[s] = P[s] < k2;
                    a1[s] = __ballot(l1[s]);
                    a2[s] = __ballot(l2[s]);
                }
#pragma unroll
                for (int s = 0; s < S; ++s) nlt[s] += (l1[s] ? 1u : 0u) + (l2[s] ? 1u : 0u);
                uint32_t b1 = static_cast<uint32_t>(__popcll(r1 & cm));
                uint32_t b2 = static_cast<uint32_t>(__popcll(r2 & cm));
#pragma unroll
                for (int s = 0; s < S; ++s) {
                    b1 += static_cast<uint32_t>(__popcll(a1[s]));
                    b2 += static_cast<uint32_t>(__popcll(a2[s]));
                }
                fp = write_lane(fp, b2, j2);  // j2 == j1 with one left: rewritten next
                fp = write_lane(fp, b1, j1);
            }
            const uint32_t nsize = min(size + m3, ef);
            // merge through LDS: every entry that stays goes to its new position
            // (positions from nsize on keep the empty entries they hold)
#pragma unroll
            for (int s = 0; s < S; ++s) {
                const uint32_t np = static_cast<uint32_t>(s * kWave) + lane + m3 - nlt[s];
                if (np < nsize)
                    w.st[np] = make_uint4(static_cast<uint32_t>(P[s]),
                                          static_cast<uint32_t>(P[s] >> 32), I[s], 0u);
            }
            if (fp < nsize)
                w.st[fp] = make_uint4(static_cast<uint32_t>(key), static_cast<uint32_t>(key >> 32),
                                      nid, 0u);
            wave_sync();
#pragma unroll
            for (int s = 0; s < S; ++s) {
                const uint4 e = w.st[static_cast<uint32_t>(s * kWave) + lane];
                P[s] = (static_cast<uint64_t>(e.y) << 32) | e.x;
                I[s] = e.z;
                U[s] = __ballot(static_cast<int32_t>(e.z) >= 0);
            }
            size = nsize;
            if (size == ef) {
                const uint32_t t = ef - 1u;
#pragma unroll
                for (int s = 0; s < S; ++s)
                    if (static_cast<uint32_t>(s) == t / kWave)
                        last = readlane64(P[s], static_cast<int>(t % kWave));
            }
            wave_sync();
        }
        seq += Sh::kDeg;
        // res_pool.insert (qg.hpp:282)
        if (kq > 0) rb_insert_fast<KS>(rb, kq, cur, sqr_y, lane);
    }

    // update_results (qg.hpp:395-418), as in h2_search_one
    if (!full && kq > 0 && rb.size < kq) {
        uint32_t snap[KS];
#pragma unroll
        for (int s = 0; s < KS; ++s) snap[s] = rb.ri[s];
        const uint32_t snap_size = rb.size;
        for (uint32_t t = 0; t <= kq && !full; ++t) {
            const uint32_t node = t < snap_size ? rb_id<KS>(snap, t) : 0u;
            const gptr<uint32_t> nb = uniform_gptr<uint32_t>(
                ix.rec + static_cast<size_t>(node) * Sh::kStride + Sh::kOffIds);
            for (uint32_t i = 0; i < Sh::kDeg; ++i) {
                const uint32_t x = uniform(nb[i]);
                const uint32_t xs = v_test_set_fast(vis, x, lane);
                if (xs == kVisSeen) continue;
                if (xs == kVisFull) {
                    full = true;
                    break;
                }
                float xr[NJ];
                load_row<NJ>(xr, uniform_gptr<float>(ix.rec + static_cast<size_t>(x) * Sh::kStride),
                             lane);
                rb_insert_fast<KS>(rb, kq, x, exact_fast<NJ>(xr, qr), lane);
            }
            if (rb.size == kq) break;
        }
    }

    // a query whose visited set would pass vis_limit of the LDS table goes to the
    // overflow list, or fails where the launch has none
    if (full && job.ovf) return h2_defer(job, q, lane);
    const bool failed = full;
    // copy_results (buffer.hpp:102), as in h2_search_one
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

template <int EF_CAP, int NJ, int KS>
__device__ __forceinline__ void h2_fast_body(const H2Dev& ix, const H2Job& job,
                                             uint32_t hash_bits, const Layout& lay) {
    extern __shared__ __align__(16) unsigned char lds[];
    const uint32_t lane = threadIdx.x & (kWave - 1);
    const uint32_t wid = threadIdx.x / kWave;
    const H2Lds w = carve(lds + static_cast<size_t>(wid) * lay.wave_bytes, lay);
    const uint32_t gwave = blockIdx.x * (blockDim.x / kWave) + wid;
    const uint32_t sw = gridDim.x * (blockDim.x / kWave);
    uint32_t q = gwave;
    for (;;) {
        q = uniform(q);
        if (q >= job.nq) break;
        (void)h2_fast_one<EF_CAP, NJ, KS>(ix, job, q, w, hash_bits, lane);
        uint32_t nx = 0;
        if (lane == 0) nx = sw + atomicAdd(job.counter, 1u);
        q = __shfl(nx, 0);
    }
}

// Latency mode of the specialised kernel (as h2_latency_kernel).
template <int EF_CAP, int NJ, int KS>
__global__ void __launch_bounds__(128) h2_fast_latency_kernel(H2Dev ix, H2Job job,
                                                               uint32_t hash_bits, Layout lay) {
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
            (void)h2_fast_one<EF_CAP, NJ, KS>(ix, job, blockIdx.x, w, hash_bits, lane);
        if (lane == 0) __hip_atomic_store(flag, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    } else {
        h2_prefetch_ahead(ix, w.st, static_cast<uint32_t>(EF_CAP), flag, job.prefetch_targets,
                          lane, flag + 1);
    }
}

template <int EF_CAP, int NJ, int KS>
__global__ void __launch_bounds__(256) h2_fast_kernel(H2Dev ix, H2Job job, uint32_t hash_bits,
                                                       Layout lay) {
    h2_fast_body<EF_CAP, NJ, KS>(ix, job, hash_bits, lay);
}

template <int EF_CAP, int NJ, int KS>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(6, 6)))
h2_fast_kernel_w6(H2Dev ix, H2Job job, uint32_t hash_bits, Layout lay) {
    h2_fast_body<EF_CAP, NJ, KS>(ix, job, hash_bits, lay);
}

template <int EF_CAP, int NJ, int KS>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(8, 8)))
h2_fast_kernel_w8(H2Dev ix, H2Job job, uint32_t hash_bits, Layout lay) {
    h2_fast_body<EF_CAP, NJ, KS>(ix, job, hash_bits, lay);
}

}  // namespace
}  // namespace lqg
