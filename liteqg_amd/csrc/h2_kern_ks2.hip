// h2_kern_ks2.hip -- FastScan search kernels for 64 < k <= 128 (two
// result-buffer register slots per lane).
#include "h2_kern_variants.h"

namespace lqg {

SearchFn h2_select_ks2(int ef_cap, int cw, int occ) { return variant_for<2>(ef_cap, cw, occ); }

}  // namespace lqg
