// h2_fast_n8.hip -- FastScan search kernels specialised for D = 128, degree 32.
#include "h2_fast_variants.h"

namespace lqg {

SearchFn h2_select_fast_n8(int ef_cap, int ks, int occ) { return fast_variant<8>(ef_cap, ks, occ); }

}  // namespace lqg
