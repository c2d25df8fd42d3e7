// h2_fast_variants.h -- the table of specialised FastScan search kernels for
// one record shape (NJ = D / 16): pool capacity class x result-buffer size x
// register cap (uncapped, 6 or 8 waves per SIMD; the 8-wave cap for k <= 64).
#pragma once

#include "h2_fast.h"

namespace lqg {
namespace {

using SearchFn = void (*)(H2Dev, H2Job, uint32_t, Layout);

template <int EF_CAP, int NJ>
SearchFn fast_for(int ks, int occ) {
    if (occ == kOccLatency) {
        if (ks == 1) return &h2_fast_latency_kernel<EF_CAP, NJ, 1>;
        if (ks == 2) return &h2_fast_latency_kernel<EF_CAP, NJ, 2>;
        return nullptr;
    }
    if (ks == 1) {
        switch (occ) {
            case 0: return &h2_fast_kernel<EF_CAP, NJ, 1>;
            case 6: return &h2_fast_kernel_w6<EF_CAP, NJ, 1>;
            case 8: return &h2_fast_kernel_w8<EF_CAP, NJ, 1>;
            default: return nullptr;
        }
    }
    if (ks == 2) {
        switch (occ) {
            case 0: return &h2_fast_kernel<EF_CAP, NJ, 2>;
            case 6: return &h2_fast_kernel_w6<EF_CAP, NJ, 2>;
            default: return nullptr;
        }
    }
    return nullptr;
}

template <int NJ>
SearchFn fast_variant(int ef_cap, int ks, int occ) {
    switch (ef_cap) {
        case 64: return fast_for<64, NJ>(ks, occ);
        case 128: return fast_for<128, NJ>(ks, occ);
        case 192: return fast_for<192, NJ>(ks, occ);
        case 256: return fast_for<256, NJ>(ks, occ);
        case 512: return fast_for<512, NJ>(ks, occ);
        case 1024: return fast_for<1024, NJ>(ks, occ);
        default: return nullptr;
    }
}

}  // namespace
}  // namespace lqg
