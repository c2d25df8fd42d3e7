// h2_kern_variants.h -- the table of FastScan search-kernel variants one
// instantiation file compiles (h2_kern_ks1.hip, h2_kern_ks2.hip): pool
// capacity class x code words per lane x register cap, for one result-buffer
// size.  Register caps (6 or 8 waves per SIMD) exist for at most 4 code words
// per lane; the 8-wave cap only for k <= 64.
#pragma once

#include "h2_kernel.h"

namespace lqg {
namespace {

using SearchFn = void (*)(H2Dev, H2Job, uint32_t, Layout);

template <int EF_CAP, int KS>
SearchFn variant_for_cw(int cw, int occ) {
    if (occ == kOccLatency) {
        switch (cw) {
            case 1: return &h2_latency_kernel<EF_CAP, 1, KS>;
            case 2: return &h2_latency_kernel<EF_CAP, 2, KS>;
            case 4: return &h2_latency_kernel<EF_CAP, 4, KS>;
            case 8: return &h2_latency_kernel<EF_CAP, 8, KS>;
            case 16: return &h2_latency_kernel<EF_CAP, 16, KS>;
            case 32: return &h2_latency_kernel<EF_CAP, 32, KS>;
            default: return nullptr;
        }
    }
    if (occ == 6) {
        switch (cw) {
            case 1: return &h2_search_kernel_w6<EF_CAP, 1, KS>;
            case 2: return &h2_search_kernel_w6<EF_CAP, 2, KS>;
            case 4: return &h2_search_kernel_w6<EF_CAP, 4, KS>;
            default: return nullptr;
        }
    }
    if constexpr (KS == 1) {
        if (occ == 8) {
            switch (cw) {
                case 1: return &h2_search_kernel_w8<EF_CAP, 1, KS>;
                case 2: return &h2_search_kernel_w8<EF_CAP, 2, KS>;
                case 4: return &h2_search_kernel_w8<EF_CAP, 4, KS>;
                default: return nullptr;
            }
        }
    }
    if (occ != 0) return nullptr;
    switch (cw) {
        case 1: return &h2_search_kernel<EF_CAP, 1, KS>;
        case 2: return &h2_search_kernel<EF_CAP, 2, KS>;
        case 4: return &h2_search_kernel<EF_CAP, 4, KS>;
        case 8: return &h2_search_kernel<EF_CAP, 8, KS>;
        case 16: return &h2_search_kernel<EF_CAP, 16, KS>;
        case 32: return &h2_search_kernel<EF_CAP, 32, KS>;
        default: return nullptr;
    }
}

template <int KS>
SearchFn variant_for(int ef_cap, int cw, int occ) {
    switch (ef_cap) {
        case 64: return variant_for_cw<64, KS>(cw, occ);
        case 128: return variant_for_cw<128, KS>(cw, occ);
        case 192: return variant_for_cw<192, KS>(cw, occ);
        case 256: return variant_for_cw<256, KS>(cw, occ);
        case 512: return variant_for_cw<512, KS>(cw, occ);
        case 1024: return variant_for_cw<1024, KS>(cw, occ);
        default: return nullptr;
    }
}

}  // namespace
}  // namespace lqg
