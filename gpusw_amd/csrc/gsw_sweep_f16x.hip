// More local half2 tiles (gsw_kernels.h), in a translation unit of their own so that they
// compile beside the other element types: cross-product widths between those of the full
// tile set, and stripe kernels for more of the 8-lane local tiles.  The host merges these
// rows into variant_table() with the rest of the half2 rows.
#include "gsw_kernels.h"

namespace gsw {

// 8-lane local tiles of 176, 208, 304, 336 and 368 columns: a protein query of up to 384
// columns then pads its tile by at most 16 columns, not 32 (measured on the 1,000 x 100,000
// BLOSUM62 top-k: its queries of up to 384 columns ran at 0.875 of the VALU roofline
// against 0.859).  The 208-column tile also runs stripes, as do the 272-, 320- and
// 352-column tiles (attach_half_stripes): the last stripe of a longer query pads less
// (stripes at 0.803 against 0.797).  16-lane tiles at odd R between the even ones were
// measured too and dropped: queries of 385-768 columns ran at 0.839 on them against 0.851.
std::vector<Variant> half_extra_variants() {
    return {GSW_XVARIANT(h2, KIND_F16, 8, 22, true), GSW_XSVARIANT(h2, KIND_F16, 8, 26, true),
            GSW_XVARIANT(h2, KIND_F16, 8, 38, true), GSW_XVARIANT(h2, KIND_F16, 8, 42, true),
            GSW_XVARIANT(h2, KIND_F16, 8, 46, true)};
}

// Stripe kernels for the local half2 rows of the (8, 34), (8, 40) and (8, 44) tiles.
void attach_half_stripes(std::vector<Variant> &rows) {
    for (Variant &v : rows) {
        if (v.kind != KIND_F16 || !v.local || v.G != 8) continue;
        if (v.R == 34) {
            v.stripe_fn = reinterpret_cast<const void *>(&stripe_kernel<h2, 8, 34, true, CROSS_WAVES>);
            v.stripe = &launch_stripe<h2, 8, 34, true>;
        } else if (v.R == 40) {
            v.stripe_fn = reinterpret_cast<const void *>(&stripe_kernel<h2, 8, 40, true, CROSS_WAVES>);
            v.stripe = &launch_stripe<h2, 8, 40, true>;
        } else if (v.R == 44) {
            v.stripe_fn = reinterpret_cast<const void *>(&stripe_kernel<h2, 8, 44, true, CROSS_WAVES>);
            v.stripe = &launch_stripe<h2, 8, 44, true>;
        }
    }
}

// Byte-select pairs kernels (pairs_tab_kernel PERM) of the 8-lane half2 tiles that explicit
// pairs of up to 192 columns run on (150-nt reads take the (8, 19) one).
void attach_half_perm(std::vector<Variant> &rows) {
#define GSW_PERM_CASE(RR)                                                                             \
    if (v.R == RR)                                                                                    \
        v.pairs_perm_fn = v.local ? reinterpret_cast<const void *>(&pairs_tab_kernel<h2, 8, RR, true, PAIRS_TAB_WAVES, true>) \
                                  : reinterpret_cast<const void *>(&pairs_tab_kernel<h2, 8, RR, false, PAIRS_TAB_WAVES, true>);
    for (Variant &v : rows) {
        if (v.kind != KIND_F16 || v.G != 8 || !v.pairs_tab_fn) continue;
        GSW_PERM_CASE(8)
        GSW_PERM_CASE(12)
        GSW_PERM_CASE(16)
        GSW_PERM_CASE(19)
        GSW_PERM_CASE(20)
        GSW_PERM_CASE(24)
    }
#undef GSW_PERM_CASE
}

}  // namespace gsw
