// The half2 instantiations of the batched kernels (gsw_kernels.h): every tile, local and
// global/semiglobal, in a translation unit of its own so that it compiles beside the
// others.  The host merges these rows into variant_table().
#include "gsw_kernels.h"

namespace gsw {

// Local cross-product tiles between the 16-lane widths of 64-column steps (384..768):
// queries of 384-768 columns then pad to 32 columns at most.  Local stripes of 224, 240 and
// 288 columns beside those of 160, 192 and 256: a query past the widest tile pads its last
// stripe less (about 4 % of the striped cells instead of 7 % for the protein top-k).
std::vector<Variant> half_variants() {
    std::vector<Variant> rows = {
        GSW_TILE_ROWS(h2, KIND_F16),
        GSW_XVARIANT(h2, KIND_F16, 16, 26, true), GSW_XVARIANT(h2, KIND_F16, 16, 30, true),
        GSW_XVARIANT(h2, KIND_F16, 16, 34, true), GSW_XVARIANT(h2, KIND_F16, 16, 38, true),
        GSW_XVARIANT(h2, KIND_F16, 16, 42, true), GSW_XVARIANT(h2, KIND_F16, 16, 46, true),
        GSW_XSVARIANT(h2, KIND_F16, 8, 28, true)};
    for (Variant &v : rows) {
        if (!v.local || v.G != 8) continue;
        if (v.R == 30) {
            v.stripe_fn = reinterpret_cast<const void *>(&stripe_kernel<h2, 8, 30, true, CROSS_WAVES>);
            v.stripe = &launch_stripe<h2, 8, 30, true>;
        } else if (v.R == 36) {
            v.stripe_fn = reinterpret_cast<const void *>(&stripe_kernel<h2, 8, 36, true, CROSS_WAVES>);
            v.stripe = &launch_stripe<h2, 8, 36, true>;
        }
    }
    return rows;
}

}  // namespace gsw
