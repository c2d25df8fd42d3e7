// h1_search.hip -- gfx950 kernel for the reference's live query path,
// QuantizedGraph::search_qg_with_heap (qg.hpp:298-344): SQ8-decoded exact L2
// distances (quantization/sq8.hpp:28-32, space/l2.hpp:16-34) driving a beam
// of width ef over the neighbour table.
//
// One wavefront walks one query.  Keys are (distance, id) packed into 64 bits
//   key = float_bits(d) << 32 | id << 1 | expanded
// so that comparing key >> 1 orders by distance and then id (distances are
// squared L2 >= +0, whose float bits are monotone).  Under this total order
// the reference's two heaps reduce to one sorted pool of capacity ef whose
// smallest unexpanded entry is expanded until none is left (an entry pushed
// out of res_pool always compares above the current lowerBound, and the
// reference stops at the first such top, qg.hpp:316-317).
//
// Visited set: a direct-mapped cache of ids in LDS (lossy).  A cache hit is a
// node evaluated before, so skipping it is what the reference does.  A miss
// may still be a node evaluated before; it is evaluated again and then
// dropped by the exact test against the pool: a node evaluated earlier is
// either still in the pool (same key -> found) or compares above the pool's
// current maximum (it was rejected or pushed out when the maximum was larger,
// and the maximum never grows).  The result is therefore exactly that of an
// exact visited set; only the number of evaluations can grow.
//
// Distances follow the reference's AVX-512 rounding order exactly: partial
// sum p (0..15) accumulates dims p, p+16, p+32, ... with one FMA each, then
// pairs (p, p+8), (p, p+4), (p, p+2), (0, 1) are added.  A quad of lanes
// evaluates one candidate: lane t holds partials 4t..4t+3, i.e. the dword at
// byte 16i+4t of each 16-byte chunk i of the SQ8 row.
#include <math.h>

#include "lqg_internal.h"

namespace lqg {
namespace {

__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ __forceinline__ uint32_t lanes_below(uint64_t mask) {
    return __builtin_amdgcn_mbcnt_hi(static_cast<uint32_t>(mask >> 32),
                                     __builtin_amdgcn_mbcnt_lo(static_cast<uint32_t>(mask), 0u));
}

__device__ __forceinline__ uint64_t make_key(float d, uint32_t id) {
    return (static_cast<uint64_t>(__float_as_uint(d)) << 32) | (static_cast<uint64_t>(id) << 1);
}
__device__ __forceinline__ uint64_t kcmp(uint64_t k) { return k >> 1; }
__device__ __forceinline__ uint32_t key_id(uint64_t k) {
    return static_cast<uint32_t>(k >> 1) & 0x7FFFFFFFu;
}
__device__ __forceinline__ float key_dist(uint64_t k) {
    return __uint_as_float(static_cast<uint32_t>(k >> 32));
}

// Number of entries of the sorted LDS array a[0..n) whose key is below x.
__device__ __forceinline__ uint32_t lower_bound(const uint64_t* a, uint32_t n, uint64_t x) {
    const uint64_t xk = kcmp(x);
    uint32_t lo = 0;
    while (n > 0) {
        const uint32_t h = n >> 1;
        if (kcmp(a[lo + h]) < xk) {
            lo += h + 1;
            n -= h + 1;
        } else {
            n = h;
        }
    }
    return lo;
}

__device__ __forceinline__ void acc4(uint32_t w, const float4& qq, const float4& mn,
                                     const float4& sc, float& s0, float& s1, float& s2,
                                     float& s3) {
    // x = min + q * scale, contracted to one FMA as the reference build does
    float x, t;
    x = fmaf(static_cast<float>(w & 0xFFu), sc.x, mn.x);
    t = qq.x - x;
    s0 = fmaf(t, t, s0);
    x = fmaf(static_cast<float>((w >> 8) & 0xFFu), sc.y, mn.y);
    t = qq.y - x;
    s1 = fmaf(t, t, s1);
    x = fmaf(static_cast<float>((w >> 16) & 0xFFu), sc.z, mn.z);
    t = qq.z - x;
    s2 = fmaf(t, t, s2);
    x = fmaf(static_cast<float>(w >> 24), sc.w, mn.w);
    t = qq.w - x;
    s3 = fmaf(t, t, s3);
}

// Keys of the m candidates bids[0..m) into bkeys[0..m).  All 64 lanes call it.
__device__ void eval_batch(const DevIndex& ix, const float* qv, const float* pmin,
                           const float* pscale, const uint32_t* bids, uint64_t* bkeys,
                           uint32_t m, uint32_t lane) {
    const uint32_t t = lane & 3u;
    const uint32_t nchunk = ix.dim >> 4;
    const uint32_t tail0 = nchunk << 4;
    for (uint32_t base = 0; base < m; base += 16) {
        const uint32_t c = base + (lane >> 2);
        const bool act = c < m;
        float s0 = 0.f, s1 = 0.f, s2 = 0.f, s3 = 0.f;
        uint32_t id = 0;
        const uint8_t* row = nullptr;
        if (act) {
            id = bids[c];
            row = ix.qdata + static_cast<size_t>(id) * ix.dim;
            if ((ix.dim & 3u) == 0) {
                const uint32_t* rw = reinterpret_cast<const uint32_t*>(row) + t;
#pragma unroll 4
                for (uint32_t i = 0; i < nchunk; ++i) {
                    const uint32_t w = rw[4 * i];
                    const uint32_t b = 16 * i + 4 * t;
                    acc4(w, *reinterpret_cast<const float4*>(qv + b),
                         *reinterpret_cast<const float4*>(pmin + b),
                         *reinterpret_cast<const float4*>(pscale + b), s0, s1, s2, s3);
                }
            } else {
                for (uint32_t i = 0; i < nchunk; ++i) {
                    const uint32_t b = 16 * i + 4 * t;
                    const uint32_t w = static_cast<uint32_t>(row[b]) |
                                       (static_cast<uint32_t>(row[b + 1]) << 8) |
                                       (static_cast<uint32_t>(row[b + 2]) << 16) |
                                       (static_cast<uint32_t>(row[b + 3]) << 24);
                    const float4 qq = make_float4(qv[b], qv[b + 1], qv[b + 2], qv[b + 3]);
                    const float4 mn =
                        make_float4(pmin[b], pmin[b + 1], pmin[b + 2], pmin[b + 3]);
                    const float4 sc =
                        make_float4(pscale[b], pscale[b + 1], pscale[b + 2], pscale[b + 3]);
                    acc4(w, qq, mn, sc, s0, s1, s2, s3);
                }
            }
        }
        // (p, p+8): lanes t and t^2;  (p, p+4): lanes t and t^1
        s0 += __shfl_xor(s0, 2);
        s1 += __shfl_xor(s1, 2);
        s2 += __shfl_xor(s2, 2);
        s3 += __shfl_xor(s3, 2);
        s0 += __shfl_xor(s0, 1);
        s1 += __shfl_xor(s1, 1);
        s2 += __shfl_xor(s2, 1);
        s3 += __shfl_xor(s3, 1);
        float r = (s0 + s2) + (s1 + s3);  // (0,2),(1,3) then (0,1)
        if (act) {
            for (uint32_t d = tail0; d < ix.dim; ++d) {  // l2.hpp:31-34
                const float x = fmaf(static_cast<float>(row[d]), pscale[d], pmin[d]);
                const float tt = qv[d] - x;
                r = fmaf(tt, tt, r);
            }
            if (t == 0) bkeys[c] = make_key(r, id);
        }
    }
}

struct WaveLds {
    float* qv;
    uint64_t* pool;
    uint64_t* bkeys;
    uint64_t* bsorted;
    uint32_t* bids;
    uint32_t* cache;
};

template <int EF_CAP>
__device__ void search_one(const DevIndex& ix, const SearchJob& job, uint32_t q,
                           const float* pmin, const float* pscale, const WaveLds& w,
                           uint32_t cache_bits, uint32_t lane) {
    // clamps keep LDS and output writes in bounds whatever the arrays hold
    const uint32_t ef = min(job.ef ? job.ef[q] : job.ef_all, static_cast<uint32_t>(EF_CAP));
    const uint32_t kq = min(job.k ? job.k[q] : job.k_all, job.kstride);
    const float* qsrc = job.queries + static_cast<size_t>(q) * ix.dim;
    for (uint32_t i = lane; i < ix.dim; i += kWave) w.qv[i] = qsrc[i];
    const uint32_t csize = 1u << cache_bits;
    for (uint32_t i = lane; i < csize; i += kWave) w.cache[i] = kInvalid;
    if (lane == 0) w.bids[0] = ix.ep;
    wave_sync();

    // entry point: qg.hpp:305-311
    eval_batch(ix, w.qv, pmin, pscale, w.bids, w.bkeys, 1, lane);
    wave_sync();
    if (lane == 0) {
        w.pool[0] = w.bkeys[0];
        w.cache[(ix.ep * 2654435761u) >> (32 - cache_bits)] = ix.ep;
    }
    wave_sync();
    uint32_t size = 1, cursor = 0, n_exp = 0, n_eval = 1;

    for (;;) {
        // pop the smallest unexpanded entry (qg.hpp:314-318)
        uint32_t c = size;
        for (uint32_t b = cursor; b < size; b += kWave) {
            const uint32_t i = b + lane;
            const bool un = (i < size) && !(w.pool[i] & 1ull);
            const uint64_t mk = __ballot(un);
            if (mk) {
                c = b + static_cast<uint32_t>(__builtin_ctzll(mk));
                break;
            }
        }
        if (c >= size) break;
        const uint64_t kc = w.pool[c];
        wave_sync();
        if (lane == 0) w.pool[c] = kc | 1ull;
        cursor = c + 1;
        ++n_exp;
        const uint32_t* nrow = ix.nbr + static_cast<size_t>(key_id(kc)) * ix.deg;

        for (uint32_t cb = 0; cb < ix.deg; cb += kWave) {
            // neighbours not seen before (qg.hpp:320-324)
            uint32_t nid = kInvalid;
            bool fresh = false;
            if (cb + lane < ix.deg) {
                nid = nrow[cb + lane];
                if (nid < ix.n) {
                    const uint32_t slot = (nid * 2654435761u) >> (32 - cache_bits);
                    fresh = atomicExch(&w.cache[slot], nid) != nid;
                }
            }
            const uint64_t fm = __ballot(fresh);
            if (fm == 0) continue;
            const uint32_t fp = lanes_below(fm);
            wave_sync();
            if (fresh) w.bids[fp] = nid;
            wave_sync();
            // repeated ids in one list: the first occurrence is the one evaluated
            bool keep = fresh;
            if (fresh) {
                for (uint32_t i = 0; i < fp; ++i) {
                    if (w.bids[i] == nid) {
                        keep = false;
                        break;
                    }
                }
            }
            const uint64_t km = __ballot(keep);
            const uint32_t m = static_cast<uint32_t>(__popcll(km));
            if (km != fm) {
                wave_sync();
                if (keep) w.bids[lanes_below(km)] = nid;
            }
            wave_sync();
            n_eval += m;
            eval_batch(ix, w.qv, pmin, pscale, w.bids, w.bkeys, m, lane);
            wave_sync();

            // admission (qg.hpp:329-335): below the pool maximum, or pool not full;
            // and not already in the pool
            bool cand = lane < m;
            const uint64_t key = cand ? w.bkeys[lane] : ~0ull;
            const uint64_t last = w.pool[size - 1];
            if (cand && size >= ef && !(kcmp(key) < kcmp(last))) cand = false;
            uint32_t lb = 0;
            if (cand) {
                lb = lower_bound(w.pool, size, key);
                if (lb < size && kcmp(w.pool[lb]) == kcmp(key)) cand = false;
            }
            const uint64_t cm = __ballot(cand);
            if (cm == 0) continue;
            const uint32_t m3 = static_cast<uint32_t>(__popcll(cm));
            wave_sync();
            if (cand) w.bkeys[lanes_below(cm)] = key;
            wave_sync();
            uint32_t rank = 0;
            if (cand) {
                for (uint32_t i = 0; i < m3; ++i) rank += (kcmp(w.bkeys[i]) < kcmp(key)) ? 1u : 0u;
                w.bsorted[rank] = key;
            }
            const uint32_t npos = lb + rank;
            const uint64_t r0 = __ballot(cand && rank == 0);
            const uint32_t first_new = __shfl(npos, static_cast<int>(__builtin_ctzll(r0)));
            wave_sync();

            // merge: an entry at i moves to i + (new keys below it); keep ef
            const uint32_t nsize = min(size + m3, ef);
            uint64_t pv[EF_CAP / kWave];
            uint32_t pn[EF_CAP / kWave];
#pragma unroll
            for (int j = 0; j < EF_CAP / kWave; ++j) {
                const uint32_t i = static_cast<uint32_t>(j * kWave) + lane;
                pn[j] = kInvalid;
                pv[j] = 0;
                if (i >= first_new && i < size) {
                    pv[j] = w.pool[i];
                    pn[j] = i + lower_bound(w.bsorted, m3, pv[j]);
                }
            }
            wave_sync();
#pragma unroll
            for (int j = 0; j < EF_CAP / kWave; ++j)
                if (pn[j] < nsize) w.pool[pn[j]] = pv[j];
            if (cand && npos < nsize) w.pool[npos] = key;
            size = nsize;
            cursor = min(cursor, first_new);
            wave_sync();
        }
    }

    // results (qg.hpp:340-343 for the reference order)
    const uint32_t cnt = min(kq, size);
    uint32_t* oi = job.out_ids + static_cast<size_t>(q) * job.kstride;
    float* od = job.out_d ? job.out_d + static_cast<size_t>(q) * job.kstride : nullptr;
    for (uint32_t j = lane; j < job.kstride; j += kWave) {
        uint32_t id = kInvalid;
        float d = INFINITY;
        int src = -1;
        if (j < kq) {
            if (job.order == 1) {
                if (j < cnt) src = static_cast<int>(j);
            } else if (j >= kq - cnt) {
                src = static_cast<int>(size - kq + j);
            }
        }
        if (src >= 0) {
            const uint64_t kk = w.pool[src];
            id = key_id(kk);
            d = key_dist(kk);
        }
        oi[j] = id;
        if (od) od[j] = d;
    }
    if (job.out_stats && lane == 0) {
        job.out_stats[2 * static_cast<size_t>(q)] = n_exp;
        job.out_stats[2 * static_cast<size_t>(q) + 1] = n_eval;
    }
    wave_sync();
}

template <int EF_CAP>
__global__ void __launch_bounds__(256) h1_search_kernel(DevIndex ix, SearchJob job,
                                                         uint32_t cache_bits, uint32_t dim_pad,
                                                         uint32_t wave_bytes) {
    extern __shared__ __align__(16) unsigned char lds[];
    const uint32_t lane = threadIdx.x & (kWave - 1);
    const uint32_t wid = threadIdx.x / kWave;
    float* pmin = reinterpret_cast<float*>(lds);
    float* pscale = pmin + dim_pad;
    for (uint32_t i = threadIdx.x; i < ix.dim; i += blockDim.x) {
        pmin[i] = ix.sq8_min[i];
        pscale[i] = ix.sq8_scale[i];
    }
    __syncthreads();

    unsigned char* wb = lds + 8u * dim_pad + static_cast<size_t>(wid) * wave_bytes;
    WaveLds w;
    w.qv = reinterpret_cast<float*>(wb);
    w.pool = reinterpret_cast<uint64_t*>(wb + 4u * dim_pad);
    w.bkeys = w.pool + EF_CAP;
    w.bsorted = w.bkeys + kWave;
    w.bids = reinterpret_cast<uint32_t*>(w.bsorted + kWave);
    w.cache = w.bids + kWave;

    for (;;) {
        uint32_t q = 0;
        if (lane == 0) q = atomicAdd(job.counter, 1u);
        q = __shfl(q, 0);
        if (q >= job.nq) break;
        search_one<EF_CAP>(ix, job, q, pmin, pscale, w, cache_bits, lane);
    }
}

__global__ void reset_counter_kernel(uint32_t* counter) { *counter = 0u; }

template <int EF_CAP>
hipError_t launch_t(const DevIndex& ix, const SearchJob& job, const LaunchCfg& cfg,
                    hipStream_t stream, size_t lds, uint32_t dim_pad, uint32_t wave_bytes,
                    int groups) {
    static size_t attr_bytes = 64 * 1024;
    if (lds > attr_bytes) {
        hipError_t e = hipFuncSetAttribute(
            reinterpret_cast<const void*>(&h1_search_kernel<EF_CAP>),
            hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(lds));
        if (e != hipSuccess) return e;
        attr_bytes = lds;
    }
    hipLaunchKernelGGL(h1_search_kernel<EF_CAP>, dim3(groups), dim3(kWave * cfg.waves_per_group),
                       lds, stream, ix, job, static_cast<uint32_t>(cfg.cache_bits), dim_pad,
                       wave_bytes);
    return hipGetLastError();
}

template <int EF_CAP>
int occupancy_t(int threads, size_t lds) {
    int nb = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(
            &nb, reinterpret_cast<const void*>(&h1_search_kernel<EF_CAP>), threads, lds) !=
        hipSuccess)
        return 1;
    return nb > 0 ? nb : 1;
}

void geometry(const DevIndex& ix, const LaunchCfg& cfg, uint32_t& dim_pad, uint32_t& wave_bytes) {
    dim_pad = (ix.dim + 3u) & ~3u;
    size_t wb = 4u * dim_pad + 8u * static_cast<size_t>(cfg.ef_cap) + 8u * kWave * 2 +
                4u * kWave + (4u << cfg.cache_bits);
    wave_bytes = static_cast<uint32_t>((wb + 15u) & ~static_cast<size_t>(15u));
}

}  // namespace

size_t search_lds_bytes(const DevIndex& ix, const LaunchCfg& cfg) {
    uint32_t dim_pad, wave_bytes;
    geometry(ix, cfg, dim_pad, wave_bytes);
    return 8u * static_cast<size_t>(dim_pad) +
           static_cast<size_t>(cfg.waves_per_group) * wave_bytes;
}

int search_full_grid(const DevIndex& ix, const LaunchCfg& cfg) {
    int dev = 0, cus = 0;
    if (hipGetDevice(&dev) != hipSuccess) return 1;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
        return 1;
    const size_t lds = search_lds_bytes(ix, cfg);
    const int threads = kWave * cfg.waves_per_group;
    int per_cu = 1;
    switch (cfg.ef_cap) {
        case 64: per_cu = occupancy_t<64>(threads, lds); break;
        case 128: per_cu = occupancy_t<128>(threads, lds); break;
        case 256: per_cu = occupancy_t<256>(threads, lds); break;
        case 512: per_cu = occupancy_t<512>(threads, lds); break;
        default: per_cu = occupancy_t<1024>(threads, lds); break;
    }
    return per_cu * cus;
}

hipError_t launch_search(const DevIndex& ix, const SearchJob& job, const LaunchCfg& cfg,
                         hipStream_t stream) {
    uint32_t dim_pad, wave_bytes;
    geometry(ix, cfg, dim_pad, wave_bytes);
    const size_t lds = search_lds_bytes(ix, cfg);
    int groups = cfg.groups > 0 ? cfg.groups : search_full_grid(ix, cfg);
    const int need = static_cast<int>((job.nq + cfg.waves_per_group - 1) / cfg.waves_per_group);
    if (groups > need) groups = need > 0 ? need : 1;
    switch (cfg.ef_cap) {
        case 64: return launch_t<64>(ix, job, cfg, stream, lds, dim_pad, wave_bytes, groups);
        case 128: return launch_t<128>(ix, job, cfg, stream, lds, dim_pad, wave_bytes, groups);
        case 256: return launch_t<256>(ix, job, cfg, stream, lds, dim_pad, wave_bytes, groups);
        case 512: return launch_t<512>(ix, job, cfg, stream, lds, dim_pad, wave_bytes, groups);
        case 1024: return launch_t<1024>(ix, job, cfg, stream, lds, dim_pad, wave_bytes, groups);
        default: return hipErrorInvalidValue;
    }
}

hipError_t launch_reset_counter(uint32_t* counter, hipStream_t stream) {
    hipLaunchKernelGGL(reset_counter_kernel, dim3(1), dim3(1), 0, stream, counter);
    return hipGetLastError();
}

}  // namespace lqg
