// The half2 instantiations of the batched kernels (gsw_kernels.h): every tile, local and
// global/semiglobal, in a translation unit of its own so that it compiles beside the
// others.  The host merges these rows into variant_table().
#include "gsw_kernels.h"

namespace gsw {

// Local cross-product tiles between the 16-lane widths of 64-column steps (384..768):
// queries of 384-768 columns then pad to 32 columns at most.
std::vector<Variant> half_variants() {
    return {GSW_TILE_ROWS(h2, KIND_F16),
            GSW_XVARIANT(h2, KIND_F16, 16, 26, true), GSW_XVARIANT(h2, KIND_F16, 16, 30, true),
            GSW_XVARIANT(h2, KIND_F16, 16, 34, true), GSW_XVARIANT(h2, KIND_F16, 16, 38, true),
            GSW_XVARIANT(h2, KIND_F16, 16, 42, true), GSW_XVARIANT(h2, KIND_F16, 16, 46, true)};
}

}  // namespace gsw
