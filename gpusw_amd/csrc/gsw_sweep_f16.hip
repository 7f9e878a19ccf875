// The half2 instantiations of the batched kernels (gsw_kernels.h): every tile, local and
// global/semiglobal, in a translation unit of its own so that it compiles beside the
// others.  The host merges these rows into variant_table().
#include "gsw_kernels.h"

namespace gsw {

std::vector<Variant> half_variants() { return {GSW_TILE_ROWS(h2, KIND_F16)}; }

}  // namespace gsw
