// The int32 instantiations of the batched kernels (gsw_kernels.h): every tile, local and
// global/semiglobal, in a translation unit of its own so that it compiles beside the
// others.  The host merges these rows into variant_table().
#include "gsw_kernels.h"

namespace gsw {

// Stripe kernels also on the 16- and 32-lane (R = 32) int32 tiles: explicit pairs longer
// than every tile on both sides run on int32 stripes, one pair per lane group, and a call
// of few such pairs spreads each one over more lanes on these (512 and 1,024 columns per
// stripe) than on the 8-lane ones.
template <int G, bool L> static void attach_stripes(Variant &v) {
    v.stripe_fn = reinterpret_cast<const void *>(&stripe_kernel<int, G, 32, L, CROSS_WAVES>);
    v.stripe = &launch_stripe<int, G, 32, L>;
    v.stripe1 = &launch_stripe1<int, G, 32, L>;
}

std::vector<Variant> int32_variants() {
    std::vector<Variant> rows = {GSW_TILE_ROWS(int, KIND_I32)};
    for (Variant &v : rows) {
        if (v.R != 32) continue;
        if (v.G == 16) v.local ? attach_stripes<16, true>(v) : attach_stripes<16, false>(v);
        if (v.G == 32) v.local ? attach_stripes<32, true>(v) : attach_stripes<32, false>(v);
    }
    return rows;
}

}  // namespace gsw
