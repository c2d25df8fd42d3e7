// h2_kern_ks1.hip -- FastScan search kernels for k <= 64 (one result-buffer
// register slot per lane).
#include "h2_kern_variants.h"

namespace lqg {

SearchFn h2_select_ks1(int ef_cap, int cw, int occ) { return variant_for<1>(ef_cap, cw, occ); }

}  // namespace lqg
