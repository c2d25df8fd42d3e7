// h2_fast_n4.hip -- FastScan search kernels specialised for D = 64, degree 32.
#include "h2_fast_variants.h"

namespace lqg {

SearchFn h2_select_fast_n4(int ef_cap, int ks, int occ) { return fast_variant<4>(ef_cap, ks, occ); }

}  // namespace lqg
