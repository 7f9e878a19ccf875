// Internal interface of the batched wavefront kernels (shared by the one-shot
// scoring calls and the serving loop).
#pragma once

#include "gsw_host.h"
#include "gsw_internal.h"
#include "gsw_sweep.h"

namespace gsw {

// Work item of the cross kernel: one query pair (or single query) against the
// length-sorted references [rbeg, rend).
struct CrossItem {
    int32_t qsa, lqa;        // query A: start in the query code buffer, length
    int32_t qsb, lqb;        // query B (lqb < 0: single query)
    int64_t outa, outb;      // output offset of reference 0 for A / B (outb < 0: none)
    int32_t rbeg, rend;      // range in the sorted reference order
    int32_t shift;           // front padding columns (semiglobal, free query ends)
    int32_t nstr;            // stripes of G*R columns (stripe kernel; 1 otherwise)
};

// An int16 result that reached 32767 (local mode under GSW_DTYPE_INT16): rescored in int32.
struct SatEntry {
    int32_t q;               // query index (-1: the output row names it, cross products)
    int32_t r;               // reference index
    int64_t o;               // output index
};

struct KArgs {
    const uint8_t *qcodes;
    const int32_t *qoff;
    const uint8_t *rcodes;
    const int32_t *roff;
    const int32_t *rorder;
    const int32_t *sub;
    int32_t ns;
    SweepConst k;
    int32_t *out;
    int64_t ldo;
    SatEntry *sat;
    int32_t *sat_count;
    int32_t sat_cap;
    int32_t sat_check;
    const void *items;
    int32_t nitems;
    int32_t *counter;
    const int32_t *nitems_ptr;   // pairs kernel: task count read from device memory
    uint32_t *bnd;               // stripe kernel: boundary regions, one per workgroup,
    int64_t bnd_stride;          //   each an H plane then an E plane of bnd_stride words (>= item rows)
    const int64_t *bpos;         //   row offsets of the sorted references (n + 1)
    const int64_t *oidx;         //   explicit pairs: output index of each position (else null)
    int32_t tab_copies;          // pairs table kernel: interleaved copies of the table
    int32_t prow;                // pairs kernel: last profile row (rows 0..prow are built)
};


constexpr int CROSS_WAVES = 8;
constexpr int PAIRS_WAVES = 8;   // most waves of a pairs workgroup (LDS decides the actual count)
constexpr int PAIRS_TAB_WAVES = 4;   // 4-wave workgroups: three fit a CU at 3 waves per SIMD

typedef void (*LaunchFn)(const KArgs &, int grid, size_t lds, hipStream_t);

const char *kind_name(int kind);   // "int32", "int16x2", "half2"

struct Variant {
    int id, G, R;
    bool packed, local;                             // packed: two alignments per lane
    const void *cross_fn, *pairs_fn, *stripe_fn;   // stripe_fn null: no stripe kernel
    LaunchFn cross, pairs, stripe;
    const void *pairs_tab_fn;                       // pairs with the substitution table in LDS
    int kind;                                       // KIND_*
    LaunchFn stripe1 = nullptr;                     // the stripe kernel in one-wave workgroups
    const void *pairs_perm_fn = nullptr;            // half2 pairs on byte-select tables (or null)
};
// The half2 and int32 rows of the variant table (gsw_sweep_f16.hip, gsw_sweep_i32.hip), and
// the further local half2 tiles (gsw_sweep_f16x.hip), which also attaches stripe kernels to
// local half2 rows of the full tile set.
std::vector<Variant> half_variants();
std::vector<Variant> int32_variants();
std::vector<Variant> half_extra_variants();
void attach_half_stripes(std::vector<Variant> &rows);
// The byte-select pairs kernels of the 8-lane half2 tiles up to 192 columns.
void attach_half_perm(std::vector<Variant> &rows);
// Byte-select pairs tables (binary16 entries, pairs_tab_kernel PERM) fit the scheme: at most
// 6 codes and every profile entry's low byte zero; their LDS bytes.
bool half_perm_ok(const SchemeDev &s);
inline size_t pairs_perm_lds_bytes(int ns) { return 8 * (size_t)ns; }
// LDS of the pairs table kernel: ns rows of (ns+2)^2 packed words (int32 tiles: ns+2),
// times pairs_tab_copies interleaved copies (32 when that stays within 48 KiB, else 1).
int pairs_tab_copies(const Variant &v, int ns);
size_t pairs_tab_lds_bytes(const Variant &v, int ns);

std::vector<Variant> &variant_table();
const Variant *pick_variant(int lq, double mean_lr, int kind, bool local, int ns, size_t lds_limit,
                            bool for_pairs);
int max_width();
// Tile of an explicit pair with query length lq; *tab: run it with the table kernel, else
// with the profile kernel (profiles of prof_rows rows; prof_rows <= 0: table kernel only).
const Variant *pick_pairs_tile(int lq, double mean_lr, int kind, bool local, int ns, int prof_rows,
                               size_t lds_limit, bool *tab);
size_t cross_lds_bytes(const Variant &v, int ns);
size_t pairs_wave_bytes(const Variant &v, int ns);               // one wave's private profiles
int pairs_waves(const Variant &v, int rows, size_t lds_limit);    // waves per workgroup (0: none fit)
int occupancy_grid(DeviceCtx *ctx, const void *fn, int threads, size_t lds, int *grid);

// A query scored against a whole reference set: code start/length in the query code
// buffer, output offset of its row, and an id (query index) for the comparator route.
struct CrossQuery {
    int32_t start, len;
    int64_t out;
    int32_t id;
};
struct CrossUnit {
    const Variant *v;
    CrossQuery a, b;
    bool has_b;
    int shift;
    int nstr;   // > 1: the queries run as nstr stripes of v's width (stripe kernel)
};

// Route queries to tiles, pairing packed int16 queries of the same tile (and of the same
// last-column register when semiglobal tracks the last column).  Queries no tile covers
// are appended to *plain.  Units come out sorted by tile.
int plan_cross_units(const SchemeDev &s, int dtype_policy, const std::vector<CrossQuery> &queries,
                     int max_lr, double mean_lr, size_t lds_limit, std::vector<CrossUnit> *units,
                     std::vector<CrossQuery> *plain);
// Tile for a cross query of length lq: the cheapest compiled tile covering it, or a tile
// with a stripe kernel run as *nstr > 1 stripes when that is cheaper or nothing else fits.
const Variant *pick_cross_tile(int lq, double mean_lr, int kind, bool local, int ns, size_t lds_limit,
                               bool track_last_col, int *nstr);
// One item per (unit, run of whole blocks of gpw sorted references holding about
// target_rows rows); the appended items are ordered longest first.
void expand_cross_items(const CrossUnit *u0, const CrossUnit *u1, const SeqSet &rs, int64_t target_rows,
                        int gpw, std::vector<CrossItem> *items);
// The same for one persistent launch of `grid` workgroups: per_wg x grid items in all (at
// least one per unit, rounded up to a multiple of grid, none under min_rows where the
// unit allows), spread evenly over the units and cut at
// equal row fractions, so every workgroup gets per_wg items of about the same work and the
// launch has no partly filled last round.  Sorted longest first.
void expand_cross_items_balanced(const CrossUnit *u0, const CrossUnit *u1, const SeqSet &rs, int grid,
                                 int per_wg, int64_t min_rows, int gpw, std::vector<CrossItem> *items);
// Striped units: with grid == 0, one item per (unit, run of sorted references of at most
// max_rows total length, or one reference); with grid > 0, a whole number (at least two)
// of items per workgroup of the launch, each unit cut at equal row fractions into pieces of
// about max_rows or less and, where the unit allows, at least min_rows.  Returns the
// largest item's row count.
int64_t stripe_item_rows(double mean_lr, int max_lr);
int64_t expand_stripe_items(const CrossUnit *u0, const CrossUnit *u1, const SeqSet &rs, int64_t max_rows,
                            std::vector<CrossItem> *items, int grid = 0, int64_t min_rows = 0);

// out[idx[i]] = val[i] on the context's stream (synchronous).
int scatter_values(DeviceCtx *ctx, int32_t *d_out, const std::vector<int64_t> &idx,
                   const std::vector<int32_t> &val);

// Device top-k of nrows rows of d_scores (row stride ncols) into device idx/sc (nrows x k).
void topk_launch(const int32_t *d_scores, int32_t nrows, int32_t ncols, int32_t k, int32_t *d_idx,
                 int32_t *d_sc, hipStream_t stream);

}  // namespace gsw
