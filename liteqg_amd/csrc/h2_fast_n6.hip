// h2_fast_n6.hip -- FastScan search kernels specialised for D = 96, degree 32.
#include "h2_fast_variants.h"

namespace lqg {

SearchFn h2_select_fast_n6(int ef_cap, int ks, int occ) { return fast_variant<6>(ef_cap, ks, occ); }

}  // namespace lqg
