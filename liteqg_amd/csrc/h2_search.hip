// h2_search.hip -- host launchers of the FastScan RaBitQ path (search_qg,
// qg.hpp:252-287) and its smaller kernels: the parity probe, the sidecar
// build (quantization/rabitq.hpp:23-108), the query rotation by the FHT or
// one FP32 GEMM on the matrix cores, and record row copies.  The search and
// overflow kernels themselves are in h2_kernel.h, instantiated by
// h2_kern_*.hip (general) and h2_fast.h, instantiated by h2_fast_n*.hip
// (specialised for D = 64, 96, 128 at degree 32).
#include "h2_fast.h"

namespace lqg {

// Kernel selectors of the instantiation files (null where no variant exists).
using SearchFn = void (*)(H2Dev, H2Job, uint32_t, Layout);
SearchFn h2_select_ks1(int ef_cap, int cw, int occ);
SearchFn h2_select_ks2(int ef_cap, int cw, int occ);
SearchFn h2_select_overflow(int ef_cap, int cw, int ks);
SearchFn h2_select_fast_n4(int ef_cap, int ks, int occ);
SearchFn h2_select_fast_n6(int ef_cap, int ks, int occ);
SearchFn h2_select_fast_n8(int ef_cap, int ks, int occ);

namespace {

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
    // LUT in the reference's layout (entry j of codebook c at 16 c + j)
    for (uint32_t idx = lane; idx < 4 * ix.p; idx += kWave)
        lut_out[static_cast<size_t>(q) * 4 * ix.p + idx] = w.lut[lut_pos(ix.p, idx >> 4, idx & 15u, lpv)];
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
    const LutBase lutb = lut_base<CW>(w.lut, ix.p, part);
    const float* fac = reinterpret_cast<const float*>(rec + ix.off_fac);
    const uint32_t code_bytes = ix.p / 8;
    for (uint32_t vb = 0; vb < deg; vb += kWave / lpv) {
        const uint32_t vi = vb + lane / lpv;
        const bool act = vi < deg;
        uint32_t cw[CW];
        float tx = 0.0f, dq = 0.0f, vq = 0.0f;
        if (act) {
            load_codes<CW>(cw, uniform_gptr<uint32_t>(rec),
                           (ix.off_codes + vi * code_bytes + part * (code_bytes / lpv)) / 4u);
            tx = fac[vi];
            dq = fac[deg + vi];
            vq = fac[2 * deg + vi];
        } else {
#pragma unroll
            for (int k = 0; k < CW; ++k) cw[k] = 0;
        }
        const float est = lane_estimate<CW>(cw, lutb, w.pl + 2u * CW * part, lpv, pr, sqr_y, tx,
                                            dq, vq);
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
using ProbeFn = void (*)(H2Dev, const float*, const uint32_t*, uint8_t*, float*, float*, float*,
                         Layout);

// Code words (4 bytes) a lane reads per neighbour pass.
int code_words(const H2Dev& ix) {
    const uint32_t lpv = ix.deg >= 64 ? 1u : 2u;
    return static_cast<int>(ix.p / 8 / lpv / 4);
}

// Result-buffer register slots for k (0: k above 128, not supported).
int rb_slots(uint32_t kmax) { return kmax <= 64 ? 1 : kmax <= 128 ? 2 : 0; }

// The search kernel: the one specialised for the index's record shape where
// there is one (unless ix.generic), else the general one.
SearchFn select_search(const H2Dev& ix, int ef_cap, int occ, uint32_t kmax) {
    const int ks = rb_slots(kmax);
    if (ks == 0) return nullptr;
    if (!ix.generic) {
        switch (h2_fast_nj(ix)) {
            case 4: return h2_select_fast_n4(ef_cap, ks, occ);
            case 6: return h2_select_fast_n6(ef_cap, ks, occ);
            case 8: return h2_select_fast_n8(ef_cap, ks, occ);
            default: break;
        }
    }
    const int cw = code_words(ix);
    return ks == 1 ? h2_select_ks1(ef_cap, cw, occ) : h2_select_ks2(ef_cap, cw, occ);
}

// The search's LDS layout: the query stays in LDS only for dims the
// registers do not hold (dim > 16 * kRegJ, or a tail past a multiple of 16).
Layout search_layout(const H2Dev& ix, const H2Cfg& cfg) {
    const bool with_qv = ix.dim > 16u * kRegJ || ix.dim % 16u != 0;
    return h2_layout(ix.dim, ix.p, static_cast<uint32_t>(cfg.ef_cap),
                     static_cast<uint32_t>(cfg.hash_bits), with_qv);
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

bool h2_kmax_ok(uint32_t kmax) { return rb_slots(kmax) != 0; }

bool h2_variant_ok(const H2Dev& ix, int occupancy, uint32_t kmax) {
    return select_search(ix, 128, occupancy, kmax) != nullptr;
}

size_t h2_lds_bytes(const H2Dev& ix, const H2Cfg& cfg, uint32_t kmax) {
    (void)kmax;
    return static_cast<size_t>(cfg.waves_per_group) * search_layout(ix, cfg).wave_bytes;
}

int h2_groups(const H2Dev& ix, const H2Cfg& cfg, uint32_t kmax, uint32_t nq) {
    const int need = static_cast<int>((nq + cfg.waves_per_group - 1) / cfg.waves_per_group);
    int groups = cfg.groups;
    if (groups <= 0) {
        SearchFn k = select_search(ix, cfg.ef_cap, cfg.occupancy, kmax);
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
    if (cfg.prefetch > 0 && job.nq <= kLatencyMaxQueries && job.prefetch_targets > 0 &&
        !job.exp_ids) {
        // latency mode: one query per workgroup (search wave + prefetch wave)
        SearchFn k = select_search(ix, cfg.ef_cap, kOccLatency, job.kstride);
        if (!k || cfg.hash_bits < 6 || cfg.hash_bits > 14 || job.kstride == 0)
            return hipErrorInvalidValue;
        const Layout lay = search_layout(ix, cfg);
        const size_t lds = lay.wave_bytes + 16u;
        if (lds > 160 * 1024) return hipErrorInvalidValue;
        if (lds > 64 * 1024) {
            hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(k),
                                               hipFuncAttributeMaxDynamicSharedMemorySize,
                                               static_cast<int>(lds));
            if (e != hipSuccess) return e;
        }
        hipLaunchKernelGGL(k, dim3(job.nq), dim3(2 * kWave), lds, stream, ix, job,
                           static_cast<uint32_t>(cfg.hash_bits), lay);
        return hipGetLastError();
    }
    SearchFn k = select_search(ix, cfg.ef_cap, cfg.occupancy, job.kstride);
    if (!k || cfg.waves_per_group < 1 || cfg.waves_per_group > 4 || cfg.hash_bits < 6 ||
        cfg.hash_bits > 14 || job.kstride == 0)
        return hipErrorInvalidValue;
    const Layout lay = search_layout(ix, cfg);
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

hipError_t h2_launch_overflow(const H2Dev& ix, const H2Job& job, const H2Cfg& cfg,
                              uint32_t groups, hipStream_t stream) {
    SearchFn k = h2_select_overflow(cfg.ef_cap, code_words(ix), rb_slots(job.kstride));
    if (!k || groups == 0 || !job.ovf) return hipErrorInvalidValue;
    H2Cfg c = cfg;
    c.hash_bits = 6;  // the LDS visited table is not used there
    const Layout lay = search_layout(ix, c);
    if (lay.wave_bytes > 64 * 1024) {
        hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(k),
                                           hipFuncAttributeMaxDynamicSharedMemorySize,
                                           static_cast<int>(lay.wave_bytes));
        if (e != hipSuccess) return e;
    }
    hipLaunchKernelGGL(k, dim3(groups), dim3(kWave), lay.wave_bytes, stream, ix, job, 6u, lay);
    return hipGetLastError();
}

hipError_t h2_launch_probe(const H2Dev& ix, const float* queries, uint32_t nq,
                           const uint32_t* nodes, uint8_t* lut, float* params, float* sqr_y,
                           float* est, hipStream_t stream) {
    ProbeFn k = select_probe(code_words(ix));
    if (!k || nq == 0) return hipErrorInvalidValue;
    const Layout lay = h2_layout(ix.dim, ix.p, 64, 6, true);
    hipLaunchKernelGGL(k, dim3(nq), dim3(kWave), lay.wave_bytes, stream, ix, queries, nodes, lut,
                       params, sqr_y, est, lay);
    return hipGetLastError();
}

hipError_t h2_launch_rotate_rows(const H2Dev& ix, float* rot, hipStream_t stream) {
    const uint32_t waves = 4;
    hipLaunchKernelGGL(h2_rotate_kernel, dim3((ix.n + waves - 1) / waves), dim3(kWave * waves),
                       static_cast<size_t>(waves) * ix.p * 4, stream, ix, rot);
    return hipGetLastError();
}

hipError_t h2_launch_codes(const H2Dev& ix, const float* rot, hipStream_t stream) {
    const double fn = static_cast<double>(1.0f / sqrtf(static_cast<float>(ix.p)));
    const uint64_t threads = static_cast<uint64_t>(ix.n) * ix.deg;
    hipLaunchKernelGGL(h2_sidecar_kernel, dim3(static_cast<uint32_t>((threads + 255) / 256)),
                       dim3(256), 0, stream, ix, rot, fn);
    return hipGetLastError();
}

hipError_t h2_launch_sidecar(const H2Dev& ix, float* rot, hipStream_t stream) {
    hipError_t e = h2_launch_rotate_rows(ix, rot, stream);
    if (e != hipSuccess) return e;
    return h2_launch_codes(ix, rot, stream);
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
