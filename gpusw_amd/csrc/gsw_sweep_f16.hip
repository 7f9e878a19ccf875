// The half2 instantiations of the batched kernels (gsw_kernels.h): every tile of the
// int16x2 table, local and global/semiglobal, in a translation unit of its own so that it
// compiles beside gsw_sweep.hip.  The host merges these rows into variant_table().
#include "gsw_kernels.h"

namespace gsw {

std::vector<Variant> half_variants() {
    return {
        GSW_HVARIANT2(8, 8),   GSW_HVARIANT2(8, 12),  GSW_HVARIANT2(8, 16),  GSW_HVARIANT2(8, 19),
        GSW_HSVARIANT2(8, 20), GSW_HSVARIANT2(8, 24), GSW_HVARIANT2(8, 30),  GSW_HVARIANT2(8, 31),
        GSW_HSVARIANT2(8, 32), GSW_HVARIANT2(8, 33),  GSW_HVARIANT2(8, 34),  GSW_HVARIANT2(8, 36),
        GSW_HVARIANT2(8, 40),  GSW_HVARIANT2(8, 44),  GSW_HVARIANT2(8, 48),  GSW_HVARIANT2(16, 24),
        GSW_HVARIANT2(16, 28), GSW_HVARIANT2(16, 32), GSW_HVARIANT2(16, 36), GSW_HVARIANT2(16, 40),
        GSW_HVARIANT2(16, 44), GSW_HVARIANT2(16, 48), GSW_HVARIANT2(32, 32), GSW_HVARIANT2(32, 48),
        GSW_HVARIANT2(64, 32),
    };
}

}  // namespace gsw
