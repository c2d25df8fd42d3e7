// h2_kern_ovf.hip -- FastScan overflow kernels: the queries whose visited set
// outgrew their LDS table, searched again with it in device memory (one
// variant per pool class x code words per lane x result-buffer size).
#include "h2_kernel.h"

namespace lqg {
namespace {

using SearchFn = void (*)(H2Dev, H2Job, uint32_t, Layout);

template <int EF_CAP, int KS>
SearchFn overflow_for_cw(int cw) {
    switch (cw) {
        case 1: return &h2_overflow_kernel<EF_CAP, 1, KS>;
        case 2: return &h2_overflow_kernel<EF_CAP, 2, KS>;
        case 4: return &h2_overflow_kernel<EF_CAP, 4, KS>;
        case 8: return &h2_overflow_kernel<EF_CAP, 8, KS>;
        case 16: return &h2_overflow_kernel<EF_CAP, 16, KS>;
        case 32: return &h2_overflow_kernel<EF_CAP, 32, KS>;
        default: return nullptr;
    }
}

template <int KS>
SearchFn overflow_for(int ef_cap, int cw) {
    switch (ef_cap) {
        case 64: return overflow_for_cw<64, KS>(cw);
        case 128: return overflow_for_cw<128, KS>(cw);
        case 192: return overflow_for_cw<192, KS>(cw);
        case 256: return overflow_for_cw<256, KS>(cw);
        case 512: return overflow_for_cw<512, KS>(cw);
        case 1024: return overflow_for_cw<1024, KS>(cw);
        default: return nullptr;
    }
}

}  // namespace

SearchFn h2_select_overflow(int ef_cap, int cw, int ks) {
    return ks == 1 ? overflow_for<1>(ef_cap, cw) : ks == 2 ? overflow_for<2>(ef_cap, cw) : nullptr;
}

}  // namespace lqg
