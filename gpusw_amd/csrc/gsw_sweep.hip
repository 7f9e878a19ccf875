// Batched scoring on gfx950: persistent wavefront kernels fed from atomic work queues,
// plus their host-side scheduling (length binning, query pairing, int16/int32 routing,
// saturation rescoring).  Replaces the reference's one-launch-per-batch path
// (src/gpusw/backends/cuda.py:47-85, src/gpusw/aligner.py:95-167).
#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <memory>
#include <numeric>
#include <set>
#include <tuple>
#include <atomic>
#include <condition_variable>
#include <functional>
#include <thread>
#include <string>

#include "gsw_internal.h"
#include "gsw_sweep.h"
#include "gsw_batch.h"
#include "gsw_kernels.h"

namespace gsw {

// ---------------------------------------------------------------- edge cells
// Closed-form scores of pairs with an empty side, written to out[o[k]].
__global__ void scatter_kernel(int32_t *out, const int64_t *idx, const int32_t *val, int64_t n) {
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n) out[idx[k]] = val[k];
}


std::vector<Variant> &variant_table() {
    static std::vector<Variant> table = [] {
        // int16x2 rows here, int32 and half2 rows from their own units; the table is ordered
        // by tile (G, then R), each tile's rows int16x2, int32, half2
        const std::vector<Variant> v = {GSW_TILE_ROWS(s2, KIND_I16)};
        const std::vector<Variant> w = int32_variants();
        const std::vector<Variant> h = [] {
            std::vector<Variant> rows = half_variants();
            for (const Variant &x : half_extra_variants()) rows.push_back(x);
            attach_half_stripes(rows);
            attach_half_perm(rows);
            return rows;
        }();
        std::set<std::pair<int, int>> tiles;
        for (const std::vector<Variant> *o : {&v, &w, &h})
            for (const Variant &x : *o) tiles.insert({x.G, x.R});
        std::vector<Variant> all;
        for (const auto &gr : tiles)
            for (const std::vector<Variant> *o : {&v, &w, &h})
                for (const Variant &x : *o)
                    if (x.G == gr.first && x.R == gr.second) all.push_back(x);
        for (size_t i = 0; i < all.size(); ++i) all[i].id = (int)i;
        return all;
    }();
    return table;
}

const char *kind_name(int kind) { return kind == KIND_F16 ? "half2" : kind == KIND_I16 ? "int16x2" : "int32"; }

// Forced variant (G, R) for tuning; 0 = automatic.
// Packed (G << 16 | R) in one atomic: planner threads read it while a caller may set it.
static std::atomic<int> g_force{0};
static inline void forced_tile(int *g, int *r) {
    const int f = g_force.load(std::memory_order_relaxed);
    *g = f >> 16;
    *r = f & 0xFFFF;
}

// Measured tile table (from the autotuner): for a query length in [lo, hi] and the given
// element type / mode, use tile (G, R) when it is compiled, covers the query and fits.
struct TileRule { int lo, hi, packed, local, G, R; };
static std::mutex g_rules_mu;
static std::vector<TileRule> g_rules;

// Estimated relative cost per cell of a (G, R) tile for a query of length lq against
// references of mean length lr: padded columns, per-step overhead and pipeline fill.
// Per-step cost of a tile's columns and its fixed work: half2 7.5 lane-slots per column
// and about 6 per step; int16/int32 9 and 14.
static double step_factor(int kind, int R, double extra) {
    const double col = kind == KIND_F16 ? 7.5 : 9.0, fixed = kind == KIND_F16 ? 6.0 : 14.0;
    return (col * R + fixed + extra) / (col * R);
}
static double tile_cost(int G, int R, int lq, double lr, int kind) {
    const double width = (double)G * R;
    const double pad = width / std::max(lq, 1);
    const double fill = (lr + G - 1) / std::max(lr, 1.0);
    return pad * step_factor(kind, R, 0.0) * fill;
}

size_t cross_lds_bytes(const Variant &v, int ns);

// Occupancy of a cross or stripe tile relative to the 4 waves per SIMD of R <= 32 (local
// G = 8: R <= 34): wider tiles run 3 waves up to R 44, R 48 two, and a profile above 80 KiB
// leaves one workgroup per CU.  (Measured on BLOSUM62 top-k, 1,000 x 100,000: padded-column
// rates 5-8 % below the 4-wave tiles.)
static double tile_occupancy_cost(const Variant &v, int ns) {
    const int r4 = v.local && v.G <= 8 ? 34 : 32;
    double f = v.R > 44 ? 1.08 : v.R > r4 ? 1.05 : 1.0;
    if (cross_lds_bytes(v, ns) > 80 * 1024) f *= 1.05;
    return f;
}


// Cheapest compiled tile covering a query of length lq whose LDS profile fits
// (cross: one profile per workgroup; pairs: one per lane group).
const Variant *pick_variant(int lq, double mean_lr, int kind, bool local, int ns,
                                   size_t lds_limit, bool for_pairs) {
    const bool packed = kind != KIND_I32;
    const Variant *best = nullptr;
    double best_cost = 1e30;
    int g_force_g, g_force_r;
    forced_tile(&g_force_g, &g_force_r);
    if (!g_force_g) {
        std::lock_guard<std::mutex> lk(g_rules_mu);
        for (const TileRule &r : g_rules) {
            if (lq < r.lo || lq > r.hi || r.packed != (int)packed || r.local != (int)local) continue;
            for (const Variant &v : variant_table())
                if (v.G == r.G && v.R == r.R && v.kind == kind && v.local == local && v.G * v.R >= lq &&
                    (for_pairs ? v.pairs_fn && pairs_waves(v, ns, lds_limit) > 0 : cross_lds_bytes(v, ns) <= lds_limit))
                    return &v;
        }
    }
    for (const Variant &v : variant_table()) {
        if (v.kind != kind || v.local != local) continue;
        if (v.G * v.R < lq) continue;
        if (for_pairs ? !v.pairs_fn || pairs_waves(v, ns, lds_limit) == 0 : cross_lds_bytes(v, ns) > lds_limit) continue;
        if (g_force_g && (v.G != g_force_g || v.R != g_force_r)) continue;
        const double c = tile_cost(v.G, v.R, lq, mean_lr, kind) * (for_pairs ? 1.0 : tile_occupancy_cost(v, ns));
        if (c < best_cost) {
            best_cost = c;
            best = &v;
        }
    }
    return best;
}

// Relative cost per cell of running a query of length lq as nstr stripes of a (G, R)
// tile: padding of the last stripe, per-step overhead including the boundary load and
// store, pipeline fill per stripe, and the per-stripe profile build, boundary traffic
// and barrier (measured: about 12 % more per padded column than a single tile).
static double stripe_cost(int G, int R, int lq, double lr, int nstr, int kind) {
    const double pad = (double)nstr * G * R / std::max(lq, 1);
    const double fill = (lr + G - 1) / std::max(lr, 1.0);
    return pad * step_factor(kind, R, 3.0) * fill * 1.10;
}

static int stripes_for(const Variant &v, int lq, bool track_last_col) {
    const int shift = track_last_col ? (v.R - lq % v.R) % v.R : 0;
    const int w = v.G * v.R;
    return (lq + shift + w - 1) / w;
}

const Variant *pick_cross_tile(int lq, double mean_lr, int kind, bool local, int ns, size_t lds_limit,
                               bool track_last_col, int *nstr) {
    *nstr = 1;
    const bool packed = kind != KIND_I32;
    int g_force_g, g_force_r;
    forced_tile(&g_force_g, &g_force_r);
    auto usable = [&](const Variant &v, int G, int R) {
        return v.G == G && v.R == R && v.kind == kind && v.local == local && cross_lds_bytes(v, ns) <= lds_limit;
    };
    auto take = [&](const Variant &v) -> const Variant * {
        if (v.G * v.R >= lq) return &v;
        if (!v.stripe_fn) return nullptr;
        *nstr = stripes_for(v, lq, track_last_col);
        return &v;
    };
    if (g_force_g) {
        for (const Variant &v : variant_table())
            if (usable(v, g_force_g, g_force_r)) return take(v);
        return nullptr;
    }
    {
        std::lock_guard<std::mutex> lk(g_rules_mu);
        for (const TileRule &r : g_rules) {
            if (lq < r.lo || lq > r.hi || r.packed != (int)packed || r.local != (int)local) continue;
            for (const Variant &v : variant_table())
                if (usable(v, r.G, r.R) && (v.G * v.R >= lq || v.stripe_fn)) return take(v);
        }
    }
    const Variant *best = pick_variant(lq, mean_lr, kind, local, ns, lds_limit, false);
    double best_cost = best ? tile_cost(best->G, best->R, lq, mean_lr, kind) * tile_occupancy_cost(*best, ns) : 1e30;
    for (const Variant &v : variant_table()) {
        if (!v.stripe_fn || v.kind != kind || v.local != local || v.G * v.R >= lq) continue;
        if (cross_lds_bytes(v, ns) > lds_limit) continue;
        const int n = stripes_for(v, lq, track_last_col);
        const double c = stripe_cost(v.G, v.R, lq, mean_lr, n, kind) * tile_occupancy_cost(v, ns);
        if (c < best_cost) {
            best_cost = c;
            best = &v;
            *nstr = n;
        }
    }
    if (best && best->G * best->R >= lq) *nstr = 1;
    return best;
}

int max_width() {
    int w = 0;
    for (const Variant &v : variant_table()) w = std::max(w, v.G * v.R);
    return w;
}

// Cross: one shared profile (16/G copies when G < 16).  Pairs: 64/G private profiles per
// wave, interleaved in blocks of 16/G.  See ProfLayout in gsw_sweep.h.
size_t cross_lds_bytes(const Variant &v, int ns) {
    const size_t ch = (size_t)std::max(v.G, 16) * 16;
    return (size_t)ns * ((v.R + 3) / 4) * ch;
}
size_t pairs_wave_bytes(const Variant &v, int rows) { return (size_t)rows * ((v.R + 3) / 4) * 64 * 16; }
int pairs_waves(const Variant &v, int rows, size_t lds_limit) {
    return (int)std::min<size_t>(PAIRS_WAVES, lds_limit / pairs_wave_bytes(v, rows));
}
int pairs_tab_copies(const Variant &v, int ns) {
    const size_t nq = (size_t)ns + 2;
    const size_t one = 4 * (size_t)ns * (v.packed ? nq * nq : nq);
    return 32 * one <= 48 * 1024 ? 32 : 1;
}
size_t pairs_tab_lds_bytes(const Variant &v, int ns) {
    const size_t nq = (size_t)ns + 2;
    return 4 * (size_t)ns * (v.packed ? nq * nq : nq) * pairs_tab_copies(v, ns);
}

// The scheme admits byte-select tables: at most 6 codes (8 bytes per row with the two
// padding codes), and every binary16 profile entry -- s, or s + 2D in the diagonal-offset
// form, and the front padding's 0 or 2D -- an integer of at most 2048 whose binary16 low
// byte is zero (at most three significant bits: 0-8, 10, 12, 14, 16, 20, 24, ...).
bool half_perm_ok(const SchemeDev &s) {
    if (s.ns > 6 || !half_entries(s)) return false;
    const int64_t d2 = s.mode == GSW_MODE_LOCAL ? 0 : -2 * (int64_t)s.ge;
    auto high_byte_only = [](int64_t v) {
        const int64_t a = v < 0 ? -v : v;
        if (a > 2048) return false;
        if (a < 8) return true;
        const int e = 63 - __builtin_clzll((unsigned long long)a);
        return (a & ((INT64_C(1) << (e - 2)) - 1)) == 0;
    };
    if (!high_byte_only(d2)) return false;
    for (int32_t v : s.h_sub)
        if (!high_byte_only(v + d2)) return false;
    return true;
}

// A table-kernel batch runs the byte-select kernel where the tile has one and the scheme
// allows it (perm), else the LDS-table kernel.
static const void *tab_kernel_fn(const Variant &v, bool perm) {
    return perm && v.pairs_perm_fn ? v.pairs_perm_fn : v.pairs_tab_fn;
}
static size_t tab_kernel_lds(const Variant &v, bool perm, int ns) {
    return perm && v.pairs_perm_fn ? pairs_perm_lds_bytes(ns) : pairs_tab_lds_bytes(v, ns);
}

// Kernel of explicit pairs: by default the substitution-table kernel (three waves per SIMD,
// one address add per column), else per-lane profiles where the table does not fit.  The
// profile kernel issues 11 % fewer VALU instructions on the 1,000,000 read pairs (178 against
// 201 per step of 20 columns) but runs two waves per SIMD, whose memory waits cost as much:
// both take 1.00-1.02 ms per batch of 131,072 pairs.  GSW_PAIRS_TAB=0 runs pairs on profiles
// wherever they fit, GSW_PAIRS_TAB=1 on the table only.  Returns 1 (table), 0 (profiles) or
// -1 (default).
static int pairs_kernel_pref() {
    const char *e = std::getenv("GSW_PAIRS_TAB");   // read per call: a caller may change it
    return e && e[0] == '1' ? 1 : e && e[0] == '0' ? 0 : -1;
}

const Variant *pick_pairs_tile(int lq, double mean_lr, int kind, bool local, int ns, int prof_rows,
                               size_t lds_limit, bool *tab) {
    const int pref = prof_rows <= 0 ? 1 : pairs_kernel_pref();
    const Variant *best = nullptr;
    bool best_tab = false;
    double best_cost = 1e30;
    int g_force_g, g_force_r;
    forced_tile(&g_force_g, &g_force_r);
    for (const Variant &v : variant_table()) {
        if (v.kind != kind || v.local != local || v.G * v.R < lq || !v.pairs_fn) continue;
        if (g_force_g && (v.G != g_force_g || v.R != g_force_r)) continue;
        const int pw = prof_rows > 0 ? pairs_waves(v, prof_rows, lds_limit) : 0;
        const bool tab_ok = v.pairs_tab_fn && pairs_tab_lds_bytes(v, ns) <= lds_limit;
        bool use_tab;
        if (pref == 1) {
            if (!tab_ok) continue;
            use_tab = true;
        } else if (pref == 0) {
            if (pw == 0 && !tab_ok) continue;
            use_tab = pw == 0;
        } else {
            if (pw == 0 && !tab_ok) continue;
            use_tab = tab_ok;
        }
        const double c = tile_cost(v.G, v.R, lq, mean_lr, kind);
        if (c < best_cost) {
            best_cost = c;
            best = &v;
            best_tab = use_tab;
        }
    }
    *tab = best_tab;
    return best;
}

// Persistent grid of a kernel: resident workgroups per CU x CUs.  The kernel's dynamic-LDS
// limit is raised only when a launch needs more than it was set to, and occupancies are
// cached per (device, kernel, threads, LDS bytes): the runtime queries cost tens of
// microseconds per launch group, which small calls notice.
int occupancy_grid(DeviceCtx *ctx, const void *fn, int threads, size_t lds, int *grid) {
    static std::mutex mu;
    static std::map<std::pair<int, const void *>, size_t> lds_set;
    static std::map<std::tuple<int, const void *, int, size_t>, int> occ;
    std::lock_guard<std::mutex> lk(mu);
    auto oit = occ.find(std::make_tuple(ctx->device, fn, threads, lds));
    int per_cu = 0;
    if (oit != occ.end()) {
        per_cu = oit->second;
    } else {
        size_t &set = lds_set[std::make_pair(ctx->device, fn)];
        if (lds > set) {
            GSW_HIP_TRY(hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
            set = lds;
        }
        GSW_HIP_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, threads, lds));
        occ[std::make_tuple(ctx->device, fn, threads, lds)] = per_cu;
    }
    if (per_cu <= 0) return fail(GSW_ERR_LIMIT, "kernel does not fit on a CU (lds %zu bytes)", lds);
    *grid = per_cu * ctx->num_cus;
    return GSW_OK;
}

// ---------------------------------------------------------------- shared helpers
struct SatBuffers {
    SatEntry *d_sat = nullptr;
    int32_t cap = 0;
};

static int prepare_sat(DeviceCtx *ctx, SatBuffers *sb, int32_t cap) {
    static thread_local void *buf = nullptr;
    static thread_local size_t bytes = 0;
    static thread_local int dev = -1;
    if (dev != ctx->device) {
        buf = nullptr;
        bytes = 0;
        dev = ctx->device;
    }
    GSW_TRY(ensure_buffer(&buf, &bytes, sizeof(SatEntry) * (size_t)cap));
    sb->d_sat = static_cast<SatEntry *>(buf);
    sb->cap = cap;
    GSW_HIP_TRY(hipMemsetAsync(ctx->counters + 1, 0, sizeof(int32_t), ctx->stream));
    return GSW_OK;
}

// The saturated results listed by the launches since prepare_sat.  *overflowed: more results
// saturated than the list holds (then none is returned and the caller rescores all of its
// packed local work).
static int read_saturated(DeviceCtx *ctx, const SatBuffers &sb, std::vector<SatEntry> *out, bool *overflowed) {
    int32_t count = 0;
    GSW_HIP_TRY(hipMemcpyAsync(&count, ctx->counters + 1, sizeof(int32_t), hipMemcpyDeviceToHost, ctx->stream));
    GSW_HIP_TRY(hipStreamSynchronize(ctx->stream));
    *overflowed = count > sb.cap;
    out->clear();
    if (count == 0 || *overflowed) return GSW_OK;
    out->resize((size_t)count);
    GSW_HIP_TRY(hipMemcpy(out->data(), sb.d_sat, sizeof(SatEntry) * (size_t)count, hipMemcpyDeviceToHost));
    return GSW_OK;
}

// Rescore explicit pairs (qi[i], rj[i]) into d_out[oi[i]] with the batched int32 tiles.
static int rescore_pairs_int32(DeviceCtx *ctx, const SchemeDev &s, const SeqSet &qs, const SeqSet &rs,
                               const std::vector<int32_t> &qi, const std::vector<int32_t> &rj,
                               const std::vector<int64_t> &oi, int32_t *d_out, BatchResult *res) {
    if (qi.empty()) return GSW_OK;
    BatchResult r;
    GSW_TRY(batch_pairs(ctx, s, GSW_DTYPE_INT32, qs, rs, qi.data(), rj.data(), (int64_t)qi.size(), d_out, &r,
                        oi.data()));
    res->kernel_ms += r.kernel_ms;
    res->launches += r.launches;
    res->plain_pairs += r.plain_pairs;
    res->rescored += (int64_t)qi.size();
    return GSW_OK;
}

int scatter_values(DeviceCtx *ctx, int32_t *d_out, const std::vector<int64_t> &idx,
                          const std::vector<int32_t> &val) {
    const int64_t n = (int64_t)idx.size();
    if (n == 0) return GSW_OK;
    GSW_TRY(ensure_buffer(&ctx->meta, &ctx->meta_bytes, n * 12 + 64));
    int64_t *d_idx = static_cast<int64_t *>(ctx->meta);
    int32_t *d_val = reinterpret_cast<int32_t *>(d_idx + n);
    GSW_HIP_TRY(hipMemcpyAsync(d_idx, idx.data(), sizeof(int64_t) * n, hipMemcpyHostToDevice, ctx->stream));
    GSW_HIP_TRY(hipMemcpyAsync(d_val, val.data(), sizeof(int32_t) * n, hipMemcpyHostToDevice, ctx->stream));
    const int threads = 256;
    hipLaunchKernelGGL(scatter_kernel, dim3((unsigned)((n + threads - 1) / threads)), dim3(threads), 0,
                       ctx->stream, d_out, d_idx, d_val, n);
    GSW_HIP_TRY(hipGetLastError());
    GSW_HIP_TRY(hipStreamSynchronize(ctx->stream));
    return GSW_OK;
}

static KArgs base_args(const SchemeDev &s, const SeqSet &qs, const SeqSet &rs, int32_t *d_out) {
    KArgs a{};
    a.qcodes = qs.d_codes;
    a.qoff = qs.d_off;
    a.rcodes = rs.d_codes;
    a.roff = rs.d_off;
    a.rorder = rs.d_order_desc;
    a.sub = s.d_sub;
    a.ns = s.ns;
    a.k = SweepConst{s.mode, s.go, s.ge, s.free_q, s.free_r, s.ns};
    a.out = d_out;
    return a;
}

// ---------------------------------------------------------------- cross planning
int plan_cross_units(const SchemeDev &s, int dtype_policy, const std::vector<CrossQuery> &queries,
                     int max_lr, double mean_lr, size_t lds_limit, std::vector<CrossUnit> *units,
                     std::vector<CrossQuery> *plain) {
    const bool local = s.mode == GSW_MODE_LOCAL;
    const int wmax = max_width();
    // semiglobal with free query ends tracks the last query column: front-pad the query so
    // that column lands on a lane's last register (see sweep_group)
    const bool track_last_col = s.mode == GSW_MODE_SEMIGLOBAL && s.free_q;
    (void)wmax;
    struct QInfo { CrossQuery q; const Variant *v; int shift, nstr; };
    std::vector<QInfo> packed_q, single_q;
    std::vector<CrossQuery> packed_free[3];   // packed queries (by kind) paired before their tile is chosen
    const bool half = half_enabled();
    for (const CrossQuery &q : queries) {
        if (!fits_int32(s, q.len, max_lr))
            return fail(GSW_ERR_OVERFLOW, "score range of query %d exceeds the int32 kernels", q.id);
        const int kind = pair_kind(s, dtype_policy, sweep_fits_int16(s, q.len, max_lr), sweep_fits_half(s, q.len, max_lr),
                                  half);
        if (kind != KIND_I32 && !track_last_col) {
            packed_free[kind].push_back(q);
            continue;
        }
        int nstr = 1;
        const Variant *v = pick_cross_tile(q.len, mean_lr, kind, local, s.ns, lds_limit, track_last_col, &nstr);
        if (!v) {
            plain->push_back(q);
            continue;
        }
        (kind != KIND_I32 ? packed_q : single_q)
            .push_back(QInfo{q, v, track_last_col ? (v->R - q.len % v->R) % v->R : 0, nstr});
    }
    // Without last-column tracking any two packed queries of a kind can share a unit: pair
    // neighbours in length order and give the pair the tile of the longer one (the shorter
    // one's extra columns are padding), so every tile and stripe count is left with no lone
    // query.
    const CrossQuery none{0, 0, -1, -1};
    for (int kind = KIND_I16; kind <= KIND_F16; ++kind) {
        std::vector<CrossQuery> &pf = packed_free[kind];
        std::stable_sort(pf.begin(), pf.end(), [](const CrossQuery &x, const CrossQuery &y) { return x.len < y.len; });
        for (size_t i = 0; i < pf.size(); i += 2) {
            const CrossQuery &a = pf[i];
            const bool has_b = i + 1 < pf.size();
            const CrossQuery &b = has_b ? pf[i + 1] : none;
            const int lq = std::max(a.len, has_b ? b.len : 0);
            int nstr = 1;
            const Variant *v = pick_cross_tile(lq, mean_lr, kind, local, s.ns, lds_limit, false, &nstr);
            if (!v) {
                plain->push_back(a);
                if (has_b) plain->push_back(b);
                continue;
            }
            units->push_back(CrossUnit{v, has_b ? b : a, has_b ? a : none, has_b, 0, nstr});
        }
    }
    // pair packed queries of the same tile, stripe count and shift (wave-uniform)
    auto key = [&](const QInfo &x) { return std::make_tuple(x.v->id, x.nstr, track_last_col ? x.shift : 0, x.q.len); };
    auto by_key = [&](const QInfo &x, const QInfo &y) { return key(x) < key(y); };
    std::stable_sort(packed_q.begin(), packed_q.end(), by_key);
    std::stable_sort(single_q.begin(), single_q.end(), by_key);
    for (size_t i = 0; i < packed_q.size();) {
        const QInfo &x = packed_q[i];
        if (i + 1 < packed_q.size() && packed_q[i + 1].v == x.v && packed_q[i + 1].nstr == x.nstr &&
            (!track_last_col || packed_q[i + 1].shift == x.shift)) {
            units->push_back(CrossUnit{x.v, x.q, packed_q[i + 1].q, true, x.shift, x.nstr});
            i += 2;
        } else {
            units->push_back(CrossUnit{x.v, x.q, none, false, x.shift, x.nstr});
            i += 1;
        }
    }
    for (const QInfo &x : single_q) units->push_back(CrossUnit{x.v, x.q, none, false, x.shift, x.nstr});
    // grouped by kernel: striped units of a tile form their own group, launched first
    std::stable_sort(units->begin(), units->end(), [](const CrossUnit &x, const CrossUnit &y) {
        return std::make_pair(x.nstr == 1, x.v->id) < std::make_pair(y.nstr == 1, y.v->id);
    });
    return GSW_OK;
}

void expand_cross_items(const CrossUnit *u0, const CrossUnit *u1, const SeqSet &rs, int64_t target_rows,
                        int gpw, std::vector<CrossItem> *items) {
    const int64_t nref = rs.n_nonempty;
    const size_t first = items->size();
    for (const CrossUnit *u = u0; u < u1; ++u) {
        for (int64_t rb = 0; rb < nref;) {
            // whole blocks of gpw references until the item holds target_rows rows
            int64_t re = std::min(nref, rb + gpw);
            while (re < nref && rs.bpos[re] - rs.bpos[rb] < target_rows) re = std::min(nref, re + gpw);
            items->push_back(CrossItem{u->a.start, u->a.len, u->has_b ? u->b.start : 0, u->has_b ? u->b.len : -1,
                                       u->a.out, u->has_b ? u->b.out : -1, (int32_t)rb, (int32_t)re, u->shift,
                                       u->nstr});
            rb = re;
        }
    }
    // longest first: the work queue then ends on the smallest items
    std::stable_sort(items->begin() + first, items->end(), [&](const CrossItem &x, const CrossItem &y) {
        return rs.bpos[x.rend] - rs.bpos[x.rbeg] > rs.bpos[y.rend] - rs.bpos[y.rbeg];
    });
}

void expand_cross_items_balanced(const CrossUnit *u0, const CrossUnit *u1, const SeqSet &rs, int grid,
                                 int per_wg, int64_t min_rows, int gpw, std::vector<CrossItem> *items) {
    const int64_t nref = rs.n_nonempty;
    const int64_t nu = (int64_t)(u1 - u0);
    const int64_t unit_rows = rs.bpos[nref];
    const int64_t g = std::max(grid, 1);
    const int64_t want = (std::max<int64_t>(nu, (int64_t)per_wg * g) + g - 1) / g * g;
    const int64_t max_pieces = std::max<int64_t>(1, unit_rows / std::max<int64_t>(min_rows, 1));
    const size_t first = items->size();
    for (int64_t i = 0; i < nu; ++i) {
        const CrossUnit *u = u0 + i;
        // `want` pieces over the units, the first `want % nu` units one more
        const int64_t pieces = std::min<int64_t>(max_pieces, want / nu + (i < want % nu ? 1 : 0));
        int64_t rb = 0;
        for (int64_t j = 1; j <= pieces && rb < nref; ++j) {
            // cut at the first whole block of gpw references reaching j/pieces of the rows
            const int64_t goal = unit_rows * j / pieces;
            int64_t re = std::min<int64_t>(nref, rb + gpw);
            while (re < nref && rs.bpos[re] < goal) re = std::min<int64_t>(nref, re + gpw);
            if (j == pieces) re = nref;
            items->push_back(CrossItem{u->a.start, u->a.len, u->has_b ? u->b.start : 0, u->has_b ? u->b.len : -1,
                                       u->a.out, u->has_b ? u->b.out : -1, (int32_t)rb, (int32_t)re, u->shift,
                                       u->nstr});
            rb = re;
        }
    }
    std::stable_sort(items->begin() + first, items->end(), [&](const CrossItem &x, const CrossItem &y) {
        return rs.bpos[x.rend] - rs.bpos[x.rbeg] > rs.bpos[y.rend] - rs.bpos[y.rbeg];
    });
}

int64_t stripe_item_rows(double mean_lr, int max_lr) {
    // about 16 blocks of 64/G references per wave and stripe, so the waves reach a stripe's
    // barrier within about one block of each other; at most 2^19 rows (4 MiB of boundaries
    // per workgroup) unless a single reference is longer
    const int64_t want = std::min<int64_t>((int64_t)(1024.0 * mean_lr), (int64_t)1 << 19);
    return std::max<int64_t>({want, (int64_t)1 << 17, (int64_t)max_lr});
}

int64_t expand_stripe_items(const CrossUnit *u0, const CrossUnit *u1, const SeqSet &rs, int64_t max_rows,
                            std::vector<CrossItem> *items, int grid, int64_t min_rows) {
    const int64_t nref = rs.n_nonempty;
    const int64_t nu = (int64_t)(u1 - u0);
    const int64_t unit_rows = rs.bpos[nref];
    const size_t first = items->size();
    int64_t widest = 0;
    auto push = [&](const CrossUnit *u, int64_t rb, int64_t re) {
        widest = std::max(widest, rs.bpos[re] - rs.bpos[rb]);
        items->push_back(CrossItem{u->a.start, u->a.len, u->has_b ? u->b.start : 0, u->has_b ? u->b.len : -1,
                                   u->a.out, u->has_b ? u->b.out : -1, (int32_t)rb, (int32_t)re, u->shift,
                                   u->nstr});
    };
    if (grid <= 0) {
        // runs of at most max_rows rows (or one reference)
        for (const CrossUnit *u = u0; u < u1; ++u) {
            for (int64_t rb = 0; rb < nref;) {
                int64_t re = rb + 1;
                while (re < nref && rs.bpos[re + 1] - rs.bpos[rb] <= max_rows) ++re;
                push(u, rb, re);
                rb = re;
            }
        }
    } else {
        // a whole number of items per workgroup (at least two); unit i gets a share of them
        // in proportion to its stripes, so items carry about equal rows x stripes, cut at
        // equal row fractions, enough pieces for max_rows and pieces of at least min_rows
        const int64_t pmin = (unit_rows + max_rows - 1) / std::max<int64_t>(max_rows, 1);
        const int64_t pmax = std::max(pmin, unit_rows / std::max<int64_t>(min_rows, 1));
        int64_t sum_str = 0;
        for (const CrossUnit *u = u0; u < u1; ++u) sum_str += u->nstr;
        int64_t want = std::max<int64_t>(2 * (int64_t)grid, nu * pmin);
        want = (want + grid - 1) / grid * grid;
        double carry = 0;   // running remainder of the proportional shares
        for (int64_t i = 0; i < nu; ++i) {
            const CrossUnit *u = u0 + i;
            const double share = (double)want * u->nstr / (double)sum_str + carry;
            int64_t quota = (int64_t)std::floor(share + 1e-9);
            carry = share - (double)quota;
            const int64_t pieces = std::min(pmax, std::max(pmin, std::max<int64_t>(quota, 1)));
            int64_t rb = 0;
            for (int64_t j = 1; j <= pieces && rb < nref; ++j) {
                int64_t re = nref;
                if (j < pieces) {   // first reference boundary at or past j/pieces of the rows
                    const int64_t goal = unit_rows * j / pieces;
                    re = std::lower_bound(rs.bpos.begin() + rb + 1, rs.bpos.begin() + nref, goal) - rs.bpos.begin();
                }
                push(u, rb, re);
                rb = re;
            }
        }
    }
    // most work (rows x stripes) first
    std::stable_sort(items->begin() + first, items->end(), [&](const CrossItem &x, const CrossItem &y) {
        return (rs.bpos[x.rend] - rs.bpos[x.rbeg]) * x.nstr > (rs.bpos[y.rend] - rs.bpos[y.rbeg]) * y.nstr;
    });
    return widest;
}

// ---------------------------------------------------------------- cross
int batch_cross(DeviceCtx *ctx, const SchemeDev &s, int dtype_policy, const SeqSet &qs,
                const int32_t *h_qlist, int32_t nq, const SeqSet &rs, int32_t *d_out,
                BatchResult *res, const int32_t *h_outrow) {
    const int32_t nr = rs.n;
    auto orow = [&](int a) -> int64_t { return h_outrow ? (int64_t)h_outrow[a] : (int64_t)a; };
    if (nq == 0 || nr == 0) return GSW_OK;
    const bool local = s.mode == GSW_MODE_LOCAL;
    double mean_lr = rs.n_nonempty ? (double)rs.total / rs.n_nonempty : 1.0;
    using clk = std::chrono::steady_clock;
    const clk::time_point t_in = clk::now();

    // -- empty rows/columns: closed forms
    std::vector<int64_t> e_idx;
    std::vector<int32_t> e_val;
    for (int a = 0; a < nq; ++a) {
        const int q = h_qlist[a];
        if (q < 0 || q >= qs.n) return fail(GSW_ERR_INVALID, "query index %d out of range", q);
        const int lq = qs.len[q];
        if (lq == 0) {
            for (int b = 0; b < nr; ++b) {
                e_idx.push_back(orow(a) * nr + b);
                e_val.push_back(empty_edge_score(s, 0, rs.len[b]));
            }
        } else {
            for (int p = rs.n_nonempty; p < nr; ++p) {
                const int b = rs.order_desc[p];
                e_idx.push_back(orow(a) * nr + b);
                e_val.push_back(empty_edge_score(s, lq, 0));
            }
        }
    }

    // -- route each query: batched (int16 packed / int32) or plain; pair packed queries
    std::vector<CrossQuery> cq;
    for (int a = 0; a < nq; ++a) {
        const int q = h_qlist[a];
        if (qs.len[q] == 0 || rs.n_nonempty == 0) continue;
        cq.push_back(CrossQuery{qs.h_off[q], qs.len[q], orow(a) * nr, q});
    }
    std::vector<CrossUnit> units;
    std::vector<CrossQuery> plain_cq;
    GSW_TRY(plan_cross_units(s, dtype_policy, cq, rs.max_len, mean_lr, (size_t)ctx->max_smem, &units, &plain_cq));
    // Small calls: below ~1e10 cells a call is bound by launch latency (each tile's launch
    // is a few short items), so the single-stripe units of each element type take the
    // widest tile chosen for that type (it covers them all) and the call has one launch per
    // type.  Not with last-column tracking, whose front padding depends on R.
    if (!(s.mode == GSW_MODE_SEMIGLOBAL && s.free_q)) {
        double cells = 0;
        const Variant *wide[3] = {nullptr, nullptr, nullptr};   // by kind
        std::set<const Variant *> used;
        for (const CrossUnit &u : units) {
            cells += (double)(u.a.len + (u.has_b ? u.b.len : 0)) * (double)rs.total;
            if (u.nstr != 1) continue;
            used.insert(u.v);
            const Variant *&w = wide[u.v->kind];
            if (!w || u.v->G * u.v->R > w->G * w->R) w = u.v;
        }
        const size_t nkinds = (wide[0] != nullptr) + (wide[1] != nullptr) + (wide[2] != nullptr);
        if (cells < 1e10 && used.size() > nkinds) {
            for (CrossUnit &u : units)
                if (u.nstr == 1) u.v = wide[u.v->kind];
            std::stable_sort(units.begin(), units.end(), [](const CrossUnit &x, const CrossUnit &y) {
                return std::make_pair(x.nstr == 1, x.v->id) < std::make_pair(y.nstr == 1, y.v->id);
            });
        }
    }
    std::vector<int32_t> plain_rows, plain_q;
    for (const CrossQuery &x : plain_cq) {
        plain_rows.push_back((int32_t)(x.out / nr));
        plain_q.push_back(x.id);
    }

    // -- one persistent launch per tile, all queued back to back: the items of every
    //    tile are uploaded once and each launch has its own work counter
    SatBuffers sb;
    const int32_t sat_cap = 1 << 20;
    GSW_TRY(prepare_sat(ctx, &sb, sat_cap));
    struct Group { size_t u0, u1, i0, i1; int grid; size_t lds; bool striped; };
    std::vector<Group> groups;
    std::vector<CrossItem> items;
    int dominant = -1;
    int64_t dominant_units = 0;
    int64_t bnd_entries = 0;   // boundary entries of the widest striped launch
    for (size_t u0 = 0; u0 < units.size();) {
        size_t u1 = u0;
        const bool striped = units[u0].nstr > 1;
        while (u1 < units.size() && units[u1].v == units[u0].v && (units[u1].nstr > 1) == striped) ++u1;
        const Variant &v = *units[u0].v;
        if ((int64_t)(u1 - u0) > dominant_units) {
            dominant_units = (int64_t)(u1 - u0);
            dominant = v.id;
        }
        Group g{u0, u1, items.size(), 0, 0, cross_lds_bytes(v, s.ns), striped};
        GSW_TRY(occupancy_grid(ctx, striped ? v.stripe_fn : v.cross_fn, CROSS_WAVES * 64, g.lds, &g.grid));
        const int gpw = 64 / v.G;
        const int64_t nref = rs.n_nonempty;
        if (striped) {
            const int64_t rows = expand_stripe_items(units.data() + u0, units.data() + u1, rs,
                                                     stripe_item_rows(mean_lr, rs.max_len), &items, g.grid,
                                                     (int64_t)(4.0 * gpw * CROSS_WAVES * mean_lr));
            bnd_entries = std::max(bnd_entries, rows * std::min<int64_t>(g.grid, (int64_t)(items.size() - g.i0)));
        } else {
            // 8 items per workgroup of about equal rows, each at least one block per wave
            expand_cross_items_balanced(units.data() + u0, units.data() + u1, rs, g.grid, 8,
                                        (int64_t)(gpw * CROSS_WAVES * mean_lr), gpw, &items);
            (void)nref;
        }
        g.i1 = items.size();
        groups.push_back(g);
        u0 = u1;
    }
    // two boundary regions: consecutive striped launches may run at the same time
    if (bnd_entries > 0) GSW_TRY(ensure_buffer(&ctx->bnd, &ctx->bnd_bytes, 4 * sizeof(uint32_t) * (size_t)bnd_entries));
    if (groups.size() > (size_t)MAX_GROUPS) return fail(GSW_ERR_LIMIT, "too many tile groups (%zu)", groups.size());
    // Launches alternate between the library stream and the second stream, so each one's
    // workgroups take the CUs as the previous launch's last workgroups finish instead of
    // waiting for all of them.  Every launch is timed by events right before and after it
    // on its stream (the launch with the most cells is reported).  GSW_TRACE_LAUNCHES=1
    // also prints tile, cells and GCUPS of each launch, and the call's host phases, to stderr.
    static const bool trace = [] { const char *e = std::getenv("GSW_TRACE_LAUNCHES"); return e && *e == '1'; }();
    const clk::time_point t_planned = clk::now();
    while (ctx->marks.size() < 2 * groups.size() + 2) {
        hipEvent_t e;
        GSW_HIP_TRY(hipEventCreate(&e));
        ctx->marks.push_back(e);
    }
    hipEvent_t fork = ctx->marks[2 * groups.size()], join = ctx->marks[2 * groups.size() + 1];
    const hipStream_t streams[2] = {ctx->stream, ctx->stream2};
    GSW_HIP_TRY(hipEventRecord(ctx->ev0, ctx->stream));
    if (!groups.empty()) {
        GSW_TRY(ensure_buffer(&ctx->meta2, &ctx->meta2_bytes, sizeof(CrossItem) * items.size()));
        GSW_HIP_TRY(hipMemcpyAsync(ctx->meta2, items.data(), sizeof(CrossItem) * items.size(),
                                   hipMemcpyHostToDevice, ctx->stream));
        GSW_HIP_TRY(hipMemsetAsync(ctx->counters + 2, 0, sizeof(int32_t) * groups.size(), ctx->stream));
        GSW_HIP_TRY(hipEventRecord(fork, ctx->stream));
        GSW_HIP_TRY(hipStreamWaitEvent(ctx->stream2, fork, 0));
    }
    for (size_t gi = 0; gi < groups.size(); ++gi) {
        const Group &g = groups[gi];
        const hipStream_t st = streams[gi & 1];
        const Variant &v = *units[g.u0].v;
        KArgs a = base_args(s, qs, rs, d_out);
        a.ldo = nr;
        a.sat = sb.d_sat;
        a.sat_count = ctx->counters + 1;
        a.sat_cap = sb.cap;
        a.sat_check = (v.packed && local) ? 1 : 0;
        a.items = static_cast<const CrossItem *>(ctx->meta2) + g.i0;
        a.nitems = (int32_t)(g.i1 - g.i0);
        a.counter = ctx->counters + 2 + gi;
        const int grid = std::min<int>(g.grid, a.nitems);
        GSW_HIP_TRY(hipEventRecord(ctx->marks[2 * gi], st));
        if (g.striped) {
            // region gi & 1: the launches of one stream run one after the other
            a.bnd = static_cast<uint32_t *>(ctx->bnd) + (gi & 1) * 2 * (size_t)bnd_entries;
            a.bnd_stride = bnd_entries / grid;
            a.bpos = rs.d_bpos;
            v.stripe(a, grid, g.lds, st);
        } else {
            v.cross(a, grid, g.lds, st);
        }
        GSW_HIP_TRY(hipGetLastError());
        res->launches++;
        GSW_HIP_TRY(hipEventRecord(ctx->marks[2 * gi + 1], st));
    }
    if (groups.size() > 1) {   // the library stream waits for the second one
        GSW_HIP_TRY(hipEventRecord(join, ctx->stream2));
        GSW_HIP_TRY(hipStreamWaitEvent(ctx->stream, join, 0));
    }
    const clk::time_point t_launched = clk::now();
    GSW_HIP_TRY(hipEventRecord(ctx->ev1, ctx->stream));
    GSW_HIP_TRY(hipEventSynchronize(ctx->ev1));
    float ms = 0.f;
    GSW_HIP_TRY(hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
    if (trace) {
        auto d = [](clk::time_point a, clk::time_point b) { return std::chrono::duration<double, std::milli>(b - a).count(); };
        std::fprintf(stderr, "[gsw] cross nq=%d nr=%d launches=%zu: plan %.2f ms, launch +%.2f ms, device done +%.2f ms "
                             "(device %.2f ms)\n", nq, nr, groups.size(), d(t_in, t_planned), d(t_planned, t_launched),
                     d(t_launched, clk::now()), ms);
    }
    // A launch's time runs from its start event to its end event on its stream; when the
    // other stream's launches run beside it, it includes the time it shared the GPU (or
    // waited for CUs).  The launch with the most cells is reported.
    int64_t top_cells = -1;
    for (size_t gi = 0; gi < groups.size(); ++gi) {
        const Group &g = groups[gi];
        const Variant &v = *units[g.u0].v;
        float gms = 0.f;
        GSW_HIP_TRY(hipEventElapsedTime(&gms, ctx->marks[2 * gi], ctx->marks[2 * gi + 1]));
        int64_t qcols = 0;
        double width = 0;
        for (size_t u = g.u0; u < g.u1; ++u) {
            qcols += (int64_t)units[u].a.len + (units[u].has_b ? units[u].b.len : 0);
            width += (double)v.G * v.R * units[u].nstr * (v.packed ? 2 : 1) * (double)rs.total;
        }
        const int64_t cells = qcols * rs.total;
        if (cells > top_cells && gms > 0.f) {
            top_cells = cells;
            res->top_ms = gms;
            res->top_cells = cells;
            res->top_variant = v.id;
            res->top_striped = g.striped ? 1 : 0;
        }
        if (trace)
            std::fprintf(stderr, "[gsw] launch %zu: G=%d R=%d %s %s%s units=%zu items=%zu cells=%.3e ms=%.2f "
                         "GCUPS=%.0f padded_GCUPS=%.0f\n", gi, v.G, v.R, kind_name(v.kind),
                         v.local ? "local" : "nonlocal", g.striped ? " striped" : "", g.u1 - g.u0, g.i1 - g.i0,
                         (double)cells, gms, (double)cells / (gms * 1e6), width / (gms * 1e6));
    }
    res->kernel_ms += ms;
    res->variant = dominant;

    // -- long queries: plain kernel
    if (!plain_q.empty()) {
        int maxq = 1;
        for (int q : plain_q) maxq = std::max(maxq, qs.len[q]);
        bool use16 = dtype_policy != GSW_DTYPE_INT32;
        for (int q : plain_q) use16 = use16 && fits_int16(s, qs.len[q], rs.max_len) && !local;
        PlainJob job{};
        job.qcodes = qs.d_codes; job.qoff = qs.d_off; job.rcodes = rs.d_codes; job.roff = rs.d_off;
        job.dtype = use16 ? GSW_DTYPE_INT16 : GSW_DTYPE_INT32;
        job.threads = 256; job.maxq = maxq;
        job.h_qlist = plain_q.data(); job.h_outrow = plain_rows.data(); job.nq = (int32_t)plain_q.size();
        job.nr = nr; job.ldo = nr; job.d_out = d_out;
        float pms = 0.f;
        GSW_TRY(plain_run(ctx, s, job, &pms));
        res->kernel_ms += pms;
        res->plain_pairs += (int64_t)plain_q.size() * nr;
        res->launches++;
    }

    // -- edges and saturation rescoring
    GSW_TRY(scatter_values(ctx, d_out, e_idx, e_val));
    std::vector<SatEntry> sat;
    bool overflowed = false;
    GSW_TRY(read_saturated(ctx, sb, &sat, &overflowed));
    if (overflowed) {
        // too many listed results: every row of a packed local unit again (the rows'
        // empty-reference cells are written again with the same values), half2 rows on the
        // int16 tiles (whose own saturated results are listed in turn), int16 rows on int32
        for (int kind = KIND_I16; kind <= KIND_F16; ++kind) {
            std::vector<int32_t> rows, ql;
            for (const CrossUnit &u : units) {
                if (u.v->kind != kind) continue;
                rows.push_back((int32_t)(u.a.out / nr)); ql.push_back(u.a.id);
                if (u.has_b) { rows.push_back((int32_t)(u.b.out / nr)); ql.push_back(u.b.id); }
            }
            if (ql.empty()) continue;
            BatchResult r;
            GSW_TRY(batch_cross(ctx, s, kind == KIND_F16 ? GSW_DTYPE_PK16 : GSW_DTYPE_INT32, qs, ql.data(),
                                (int32_t)ql.size(), rs, d_out, &r, rows.data()));
            res->kernel_ms += r.kernel_ms;
            res->launches += r.launches;
            res->plain_pairs += r.plain_pairs;
            res->rescored += (int64_t)ql.size() * nr + r.rescored;
        }
    } else if (!sat.empty()) {
        // the listed results as explicit pairs: the output row names the query
        std::vector<std::pair<int64_t, int32_t>> row_query((size_t)nq);
        for (int a = 0; a < nq; ++a) row_query[a] = {orow(a), h_qlist[a]};
        std::sort(row_query.begin(), row_query.end());
        std::vector<int32_t> qi, rj;
        std::vector<int64_t> oi;
        for (const SatEntry &e : sat) {
            const int64_t row = e.o / nr;
            auto it = std::lower_bound(row_query.begin(), row_query.end(), std::make_pair(row, INT32_MIN));
            if (it == row_query.end() || it->first != row) return fail(GSW_ERR_HIP, "saturation list names no row");
            qi.push_back(it->second);
            rj.push_back(e.r);
            oi.push_back(e.o);
        }
        GSW_TRY(rescore_pairs_int32(ctx, s, qs, rs, qi, rj, oi, d_out, res));
    }
    return GSW_OK;
}

// ---------------------------------------------------------------- long explicit pairs
// Explicit pairs that no pairs tile takes on either side (both longer than the widest
// tile) run on the int32 stripe kernels.  The pairs of one query become a list of
// positions, its references longest first, each position with its own output index
// (KArgs::oidx), and the list is cut into items of whole references as a cross product's
// reference range is; the stripe kernel then sweeps them as it sweeps a reference range.
// int32 throughout: no result is listed for rescoring.  One launch per tile, on the
// library stream.  Pairs of queries that no int32 stripe tile takes (a forced tile
// without stripes, or a pair outside the int32 range) are returned in rq/rr/ro.
static int batch_long_pairs(DeviceCtx *ctx, const SchemeDev &s, const SeqSet &qs, const SeqSet &rs,
                            const std::vector<int32_t> &pq, const std::vector<int32_t> &pr,
                            const std::vector<int64_t> &po, int32_t *d_out, BatchResult *res,
                            std::vector<int32_t> *rq, std::vector<int32_t> *rr, std::vector<int64_t> *ro) {
    const bool local = s.mode == GSW_MODE_LOCAL;
    const bool track_last_col = s.mode == GSW_MODE_SEMIGLOBAL && s.free_q;
    const size_t n = pq.size();
    std::vector<size_t> idx(n);
    std::iota(idx.begin(), idx.end(), (size_t)0);
    std::stable_sort(idx.begin(), idx.end(), [&](size_t x, size_t y) {
        return pq[x] != pq[y] ? pq[x] < pq[y] : rs.len[pr[x]] > rs.len[pr[y]];
    });
    struct QGroup { size_t i0, i1; const Variant *v; int nstr, shift; };
    std::vector<QGroup> qg;
    int force_g, force_r;
    forced_tile(&force_g, &force_r);
    // A pair runs on one lane group: with fewer pairs than the GPU has lanes for (two waves
    // per SIMD), a wider group sweeps each pair with more lanes, so the stripe cost of a
    // G-lane tile is scaled by how much of the GPU its groups leave idle.
    const double lanes = 2.0 * 4 * 64 * std::max(ctx->num_cus, 1);
    auto pick = [&](int lq, double lr, int *nstr) -> const Variant * {
        if (force_g) return pick_cross_tile(lq, lr, KIND_I32, local, s.ns, (size_t)ctx->max_smem, track_last_col, nstr);
        const Variant *best = nullptr;
        double best_cost = 1e30;
        for (const Variant &v : variant_table()) {
            if (!v.stripe_fn || v.kind != KIND_I32 || v.local != local) continue;
            if (cross_lds_bytes(v, s.ns) > (size_t)ctx->max_smem) continue;
            const int ns_ = stripes_for(v, lq, track_last_col);
            const double idle = std::max(1.0, lanes / ((double)n * v.G));
            const double c = stripe_cost(v.G, v.R, lq, lr, ns_, KIND_I32) * tile_occupancy_cost(v, s.ns) * idle;
            if (c < best_cost) {
                best_cost = c;
                best = &v;
                *nstr = ns_;
            }
        }
        return best;
    };
    for (size_t i0 = 0; i0 < n;) {
        const int q = pq[idx[i0]];
        size_t i1 = i0;
        int64_t rows = 0;
        bool ok = true;
        for (; i1 < n && pq[idx[i1]] == q; ++i1) {
            rows += rs.len[pr[idx[i1]]];
            ok = ok && fits_int32(s, qs.len[q], rs.len[pr[idx[i1]]]);
        }
        int nstr = 1;
        const Variant *v = ok ? pick(qs.len[q], (double)rows / (double)(i1 - i0), &nstr) : nullptr;
        if (v && v->stripe_fn) {
            qg.push_back(QGroup{i0, i1, v, nstr, track_last_col ? (v->R - qs.len[q] % v->R) % v->R : 0});
        } else {
            for (size_t i = i0; i < i1; ++i) {
                rq->push_back(pq[idx[i]]);
                rr->push_back(pr[idx[i]]);
                ro->push_back(po[idx[i]]);
            }
        }
        i0 = i1;
    }
    if (qg.empty()) return GSW_OK;
    std::stable_sort(qg.begin(), qg.end(), [](const QGroup &x, const QGroup &y) { return x.v->id < y.v->id; });

    // positions in group order; items per launch, each of whole references: at least a
    // wave's worth (64/G references) of its query's rows, else an equal share of the
    // launch's rows (two items per workgroup), longest first.  Workgroups of one wave where
    // the tile has that launch (a query's references are often few), their boundary
    // regions within 512 MiB.
    std::vector<int32_t> vorder;
    std::vector<int64_t> vbpos{0}, voidx;
    std::vector<CrossItem> items;
    struct Launch { const Variant *v; size_t it0, it1; int grid, threads; int64_t widest; };
    std::vector<Launch> launches;
    int64_t bnd_words = 0;
    for (size_t g0 = 0; g0 < qg.size();) {
        size_t g1 = g0;
        int64_t rows = 0, npairs = 0;
        int max_lr = 0;
        for (; g1 < qg.size() && qg[g1].v == qg[g0].v; ++g1)
            for (size_t i = qg[g1].i0; i < qg[g1].i1; ++i) {
                rows += (int64_t)rs.len[pr[idx[i]]] * qg[g1].nstr;
                max_lr = std::max(max_lr, rs.len[pr[idx[i]]]);
                ++npairs;
            }
        const Variant &v = *qg[g0].v;
        Launch L{&v, items.size(), 0, 0, v.stripe1 ? 64 : CROSS_WAVES * 64, 0};
        GSW_TRY(occupancy_grid(ctx, v.stripe_fn, L.threads, cross_lds_bytes(v, s.ns), &L.grid));
        const int64_t cap = stripe_item_rows(rs.n_nonempty ? (double)rs.total / rs.n_nonempty : 1.0, rs.max_len);
        const int64_t wave_rows = (int64_t)(64 / v.G) * (rows / std::max<int64_t>(npairs, 1));
        const int64_t target = std::min<int64_t>(
            cap, std::max<int64_t>({1, wave_rows, rows / (2 * (int64_t)std::max(L.grid, 1))}));
        for (size_t gi = g0; gi < g1; ++gi) {
            const QGroup &G = qg[gi];
            const int q = pq[idx[G.i0]];
            int64_t item_rows = 0;
            int32_t rb = (int32_t)vorder.size();
            for (size_t i = G.i0; i < G.i1; ++i) {
                const int lr = rs.len[pr[idx[i]]];
                vorder.push_back(pr[idx[i]]);
                voidx.push_back(po[idx[i]]);
                vbpos.push_back(vbpos.back() + lr);
                item_rows += lr;
                if (i + 1 == G.i1 || item_rows * G.nstr >= target || item_rows + max_lr > cap) {
                    const int32_t re = (int32_t)vorder.size();
                    items.push_back(CrossItem{qs.h_off[q], qs.len[q], 0, -1, 0, -1, rb, re, G.shift, G.nstr});
                    L.widest = std::max(L.widest, item_rows);
                    rb = re;
                    item_rows = 0;
                }
            }
        }
        L.it1 = items.size();
        std::stable_sort(items.begin() + L.it0, items.end(), [&](const CrossItem &x, const CrossItem &y) {
            return (vbpos[x.rend] - vbpos[x.rbeg]) * x.nstr > (vbpos[y.rend] - vbpos[y.rbeg]) * y.nstr;
        });
        L.grid = std::min<int>(L.grid, (int)(L.it1 - L.it0));
        L.grid = (int)std::max<int64_t>(1, std::min<int64_t>(L.grid, ((int64_t)1 << 26) / std::max<int64_t>(L.widest, 1)));
        bnd_words = std::max(bnd_words, 2 * L.widest * L.grid);
        launches.push_back(L);
        g0 = g1;
    }
    if (launches.size() > (size_t)MAX_GROUPS) return fail(GSW_ERR_LIMIT, "too many tile groups (%zu)", launches.size());
    GSW_TRY(ensure_buffer(&ctx->bnd, &ctx->bnd_bytes, sizeof(uint32_t) * (size_t)bnd_words));

    // one device block for the items and the position arrays, freed after the launches
    const size_t npos = vorder.size();
    auto up = [](size_t b) { return (b + 255) & ~(size_t)255; };
    const size_t o_items = 0, o_order = up(sizeof(CrossItem) * items.size());
    const size_t o_bpos = o_order + up(sizeof(int32_t) * npos), o_oidx = o_bpos + up(sizeof(int64_t) * (npos + 1));
    const size_t bytes = o_oidx + up(sizeof(int64_t) * npos);
    char *d_mem = nullptr;
    GSW_HIP_TRY(hipMalloc(&d_mem, bytes));
    struct Free { char *p; ~Free() { (void)hipFree(p); } } free_mem{d_mem};
    GSW_HIP_TRY(hipMemcpyAsync(d_mem + o_items, items.data(), sizeof(CrossItem) * items.size(), hipMemcpyHostToDevice,
                               ctx->stream));
    GSW_HIP_TRY(hipMemcpyAsync(d_mem + o_order, vorder.data(), sizeof(int32_t) * npos, hipMemcpyHostToDevice, ctx->stream));
    GSW_HIP_TRY(hipMemcpyAsync(d_mem + o_bpos, vbpos.data(), sizeof(int64_t) * (npos + 1), hipMemcpyHostToDevice,
                               ctx->stream));
    GSW_HIP_TRY(hipMemcpyAsync(d_mem + o_oidx, voidx.data(), sizeof(int64_t) * npos, hipMemcpyHostToDevice, ctx->stream));
    GSW_HIP_TRY(hipMemsetAsync(ctx->counters + 2, 0, sizeof(int32_t) * launches.size(), ctx->stream));
    GSW_HIP_TRY(hipEventRecord(ctx->ev0, ctx->stream));
    for (size_t li = 0; li < launches.size(); ++li) {
        const Launch &L = launches[li];
        KArgs a = base_args(s, qs, rs, d_out);
        a.items = reinterpret_cast<const CrossItem *>(d_mem + o_items) + L.it0;
        a.nitems = (int32_t)(L.it1 - L.it0);
        a.counter = ctx->counters + 2 + li;
        a.rorder = reinterpret_cast<const int32_t *>(d_mem + o_order);
        a.bpos = reinterpret_cast<const int64_t *>(d_mem + o_bpos);
        a.oidx = reinterpret_cast<const int64_t *>(d_mem + o_oidx);
        a.bnd = static_cast<uint32_t *>(ctx->bnd);
        a.bnd_stride = L.widest;
        a.sat_check = 0;
        (L.threads == 64 ? L.v->stripe1 : L.v->stripe)(a, L.grid, cross_lds_bytes(*L.v, s.ns), ctx->stream);
        GSW_HIP_TRY(hipGetLastError());
        res->launches++;
    }
    GSW_HIP_TRY(hipEventRecord(ctx->ev1, ctx->stream));
    GSW_HIP_TRY(hipEventSynchronize(ctx->ev1));
    float ms = 0.f;
    GSW_HIP_TRY(hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
    res->kernel_ms += ms;
    return GSW_OK;
}

// ---------------------------------------------------------------- pairs
// Explicit pairs run in batches of PAIR_BATCH pairs.  Each batch is planned on the host
// in O(n) (counting sorts by reference length rank, then by last-column register) into a
// pinned buffer of its slot, copied to the slot's device buffer and run by replaying the
// slot's HIP graph (the persistent pairs kernel with this call's arguments), both on one
// of two streams that batches alternate between.  The header carries the task count and a
// zeroed work counter, so no memset is needed.  Planning later batches and copying their
// tasks overlap the kernels of earlier ones.
constexpr int64_t PAIR_BATCH = 131072;
constexpr size_t PAIR_HDR = 64;

struct PairGraph {
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    hipGraphNode_t kern = nullptr;
    char *d_buf = nullptr, *h_buf = nullptr;
    size_t cap = 0;
    hipEvent_t copied = nullptr;     // the copy has read h_buf and written d_buf
    hipEvent_t consumed = nullptr;   // the kernel that read d_buf has finished
    int grid = 0, waves = 0;
    size_t lds = 0;
};

static std::mutex g_graph_mu;
static std::map<std::tuple<int, int, int>, std::unique_ptr<PairGraph>> g_pair_graphs;

static int pair_graph(DeviceCtx *ctx, const Variant &v, bool tab, bool perm, int slot, int ns, int rows,
                      PairGraph **out) {
    std::lock_guard<std::mutex> lk(g_graph_mu);
    perm = perm && tab && v.pairs_perm_fn;
    auto key = std::make_tuple(ctx->device, 4 * v.id + (tab ? 1 : 0) + (perm ? 2 : 0), slot);
    auto &pg = g_pair_graphs[key];
    const int waves = tab ? PAIRS_TAB_WAVES : pairs_waves(v, rows, (size_t)ctx->max_smem);
    if (waves == 0) return fail(GSW_ERR_LIMIT, "pairs tile (%d, %d) does not fit in LDS", v.G, v.R);
    const size_t lds = tab ? tab_kernel_lds(v, perm, ns) : (size_t)waves * pairs_wave_bytes(v, rows);
    if (pg && pg->lds == lds && pg->waves == waves) {
        *out = pg.get();
        return GSW_OK;
    }
    if (pg) {   // table size changed: rebuild
        (void)hipGraphExecDestroy(pg->exec);
        (void)hipGraphDestroy(pg->graph);
    } else {
        pg = std::make_unique<PairGraph>();
        const int gpw = 64 / v.G;
        pg->cap = PAIR_HDR + sizeof(PairTask) * (size_t)(PAIR_BATCH + 64 * gpw + gpw);
        GSW_HIP_TRY(hipMalloc(&pg->d_buf, pg->cap));
        GSW_HIP_TRY(hipHostMalloc(&pg->h_buf, pg->cap, hipHostMallocDefault));
        std::memset(pg->h_buf, 0, PAIR_HDR);
        GSW_HIP_TRY(hipEventCreateWithFlags(&pg->copied, hipEventDisableTiming));
        GSW_HIP_TRY(hipEventCreateWithFlags(&pg->consumed, hipEventDisableTiming));
    }
    pg->lds = lds;
    pg->waves = waves;
    const void *fn = tab ? tab_kernel_fn(v, perm) : v.pairs_fn;
    GSW_TRY(occupancy_grid(ctx, fn, waves * 64, lds, &pg->grid));
    GSW_HIP_TRY(hipGraphCreate(&pg->graph, 0));
    KArgs a{};
    void *params[] = {&a};
    hipKernelNodeParams kp{};
    kp.func = const_cast<void *>(fn);
    kp.gridDim = dim3(pg->grid);
    kp.blockDim = dim3(waves * 64);
    kp.sharedMemBytes = (unsigned)lds;
    kp.kernelParams = params;
    kp.extra = nullptr;
    GSW_HIP_TRY(hipGraphAddKernelNode(&pg->kern, pg->graph, nullptr, 0, &kp));
    GSW_HIP_TRY(hipGraphInstantiate(&pg->exec, pg->graph, nullptr, nullptr, 0));
    *out = pg.get();
    return GSW_OK;
}

// Planning threads for explicit pairs (GSW_HOST_THREADS overrides; default up to 8).
static int host_threads() {
    static const int t = [] {
        const char *e = std::getenv("GSW_HOST_THREADS");
        int v = e ? std::atoi(e) : 0;
        if (v <= 0) v = (int)std::min(8u, std::max(1u, std::thread::hardware_concurrency()));
        return std::min(v, 64);
    }();
    return t;
}

// The library's one pool (a function-local static in the template below would give each
// caller's lambda type a pool of its own).
static HostPool &host_pool() {
    static HostPool pool;
    return pool;
}

// fn(tid) on T threads (tid 0 on the calling thread).
template <typename F> static void run_threads(int T, F &&fn) {
    if (T <= 1) {
        fn(0);
        return;
    }
    host_pool().run(T, fn);
}

constexpr int MAX_PAIR_SLOTS = 8;   // pinned buffers + graphs per tile

int batch_pairs(DeviceCtx *ctx, const SchemeDev &s, int dtype_policy, const SeqSet &qs,
                const SeqSet &rs, const int32_t *h_qi, const int32_t *h_rj, int64_t n,
                int32_t *d_out, BatchResult *res, const int64_t *h_outidx) {
    if (n == 0) return GSW_OK;
    auto oidx = [h_outidx](int64_t k) -> int64_t { return h_outidx ? h_outidx[k] : k; };
    // GSW_TRACE_LAUNCHES=1: host phases of the call to stderr
    static const bool trace = [] { const char *e = std::getenv("GSW_TRACE_LAUNCHES"); return e && *e == '1'; }();
    using clk = std::chrono::steady_clock;
    const clk::time_point t_in = clk::now();
    clk::time_point t_first{}, t_head{}, t_rest{};
    const bool local = s.mode == GSW_MODE_LOCAL;
    const bool track_last_col = s.mode == GSW_MODE_SEMIGLOBAL && s.free_q;
    const int wmax = max_width();
    const double mean_lr = rs.n_nonempty ? (double)rs.total / rs.n_nonempty : 1.0;
    const int T = n >= 65536 ? host_threads() : 1;
    const bool half = half_enabled();
    // half2 pairs on byte-select tables where the scheme allows (GSW_PAIRS_PERM=0: LDS tables)
    const bool perm = half && half_perm_ok(s) && [] {
        const char *e = std::getenv("GSW_PAIRS_PERM");
        return !(e && e[0] == '0');
    }();
    // profile rows of the profile kernel: codes past the set's largest one never occur
    const int prow = std::min(s.ns - 1, rs.max_code);
    const size_t nv = 2 * variant_table().size();   // kinds: 2 * variant id + (table kernel)
    // The pair list runs in segments.  A call of at least four batches classifies, plans and
    // launches a head of 21/8 batches first and classifies the rest while the device runs
    // the head (classifying 1,000,000 pairs and listing each tile's members take about
    // 1 ms, which otherwise precedes the first launch); a smaller call is one segment.  The
    // call's first tile ramps up (1/8, 1/4, 1/2, 3/4 of a batch, then full batches), so the
    // device starts after a short plan and has work while the first full batch is planned.
    static constexpr int64_t RAMP[4] = {PAIR_BATCH / 8, PAIR_BATCH / 4, PAIR_BATCH / 2, PAIR_BATCH * 3 / 4};
    const int64_t head = n >= 4 * PAIR_BATCH ? PAIR_BATCH * 21 / 8 : n;

    // -- classify a segment in parallel over contiguous index ranges: empty side (closed
    //    form), no tile (comparator), else a tile
    struct Part {
        std::vector<int64_t> e_idx, p_oi;
        std::vector<int32_t> e_val, p_qi, p_rj;
        std::vector<int64_t> count;
        int64_t cells = 0;
        int64_t bad = -1;
        int bad_code = GSW_OK;
    };
    std::vector<Part> parts(T);
    // per-pair tile and tile members: kept between calls (grow-only), so that a call does
    // not fault in fresh pages for them
    // (local references: the planning threads must use the calling thread's vectors)
    static thread_local std::vector<int16_t> tl_vid;
    static thread_local std::vector<int64_t> tl_members;
    std::vector<int16_t> &vid = tl_vid;
    std::vector<int64_t> &members = tl_members;
    struct Trim {   // calls of more than 2^24 pairs give their buffers back
        std::vector<int16_t> &a;
        std::vector<int64_t> &b;
        ~Trim() {
            if (a.size() > ((size_t)1 << 24)) std::vector<int16_t>().swap(a);
            if (b.size() > ((size_t)1 << 24)) std::vector<int64_t>().swap(b);
        }
    } trim{vid, members};
    if (vid.size() < (size_t)n) vid.resize((size_t)n);
    // The int16 and int32 range checks are monotone in the reference length for a given
    // query length: the largest reference length each allows is found once per query
    // length (binary search) and a pair costs one comparison per check.
    auto lr_limit = [&](bool (*fits)(const SchemeDev &, int, int), int lq) -> int {
        if (!fits(s, lq, 1)) return 0;
        int lo = 1, hi = INT32_MAX;   // fits(lo) holds
        while (lo < hi) {
            const int mid = lo + (hi - lo + 1) / 2;
            if (fits(s, lq, mid)) lo = mid;
            else hi = mid - 1;
        }
        return lo;
    };
    std::vector<int64_t> e_idx, p_oi;
    std::vector<int32_t> e_val, p_qi, p_rj;
    std::vector<int64_t> vbeg(nv + 1, 0);
    std::vector<std::vector<int64_t>> pos(T, std::vector<int64_t>(nv));
    res->cells = 0;
    // one thread's share [a, b) of a segment
    auto classify_part = [&](Part &P, int64_t a, int64_t b) {
        P.e_idx.clear(); P.e_val.clear(); P.p_qi.clear(); P.p_rj.clear(); P.p_oi.clear();
        P.count.assign(nv, 0);
        P.cells = 0;
        P.bad = -1;
        P.bad_code = GSW_OK;
        std::vector<const Variant *> vcache(3 * (wmax + 1), nullptr);
        std::vector<char> vdone(3 * (wmax + 1), 0), vtab(3 * (wmax + 1), 0);
        std::vector<int> limh(wmax + 1, -1), lim16(wmax + 1, -1), lim32(wmax + 1, -1);   // -1: not computed
        for (int64_t k = a; k < b; ++k) {
            const int q = h_qi[k], r = h_rj[k];
            vid[k] = -1;
            if (q < 0 || q >= qs.n || r < 0 || r >= rs.n) {
                P.bad = k;
                P.bad_code = GSW_ERR_INVALID;
                return;
            }
            const int lq = qs.len[q], lr = rs.len[r];
            P.cells += (int64_t)lq * lr;
            if (lq == 0 || lr == 0) {
                P.e_idx.push_back(oidx(k));
                P.e_val.push_back(empty_edge_score(s, lq, lr));
                continue;
            }
            bool ok32, ok16, okh;
            if (lq <= wmax) {
                if (lim32[lq] < 0) {
                    lim32[lq] = lr_limit(fits_int32, lq);
                    lim16[lq] = lr_limit(sweep_fits_int16, lq);
                    limh[lq] = lr_limit(sweep_fits_half, lq);
                }
                ok32 = lr <= lim32[lq];
                ok16 = lr <= lim16[lq];
                okh = lr <= limh[lq];
            } else {
                ok32 = fits_int32(s, lq, lr);
                ok16 = sweep_fits_int16(s, lq, lr);
                okh = sweep_fits_half(s, lq, lr);
            }
            if (!ok32) {
                P.bad = k;
                P.bad_code = GSW_ERR_OVERFLOW;
                return;
            }
            // GSW_DTYPE_INT16 runs every local pair packed (results at the type's top are listed)
            const int kind = pair_kind(s, dtype_policy, ok16, okh, half);
            const Variant *v = nullptr;
            bool tab = false;
            if (lq <= wmax) {
                const size_t ci = (size_t)kind * (wmax + 1) + lq;
                if (!vdone[ci]) {
                    bool tb = false;
                    vcache[ci] = pick_pairs_tile(lq, mean_lr, kind, local, s.ns, prow + 1, (size_t)ctx->max_smem,
                                                 &tb);
                    vtab[ci] = tb;
                    vdone[ci] = 1;
                }
                v = vcache[ci];
                tab = vtab[ci];
            }
            if (!v) {
                P.p_qi.push_back(q);
                P.p_rj.push_back(r);
                P.p_oi.push_back(oidx(k));
                continue;
            }
            vid[k] = (int16_t)(2 * v->id + (tab ? 1 : 0));
            P.count[2 * v->id + (tab ? 1 : 0)]++;
        }
    };
    // Classify pairs [k0, k1) and list each tile's members of the segment, in index order,
    // in members[vbeg[kind], vbeg[kind + 1]).
    auto classify = [&](int64_t k0, int64_t k1) -> int {
        const int64_t m = k1 - k0;
        const int Tc = m >= 65536 ? T : 1;
        auto lo = [&](int t) { return k0 + m * t / Tc; };
        run_threads(Tc, [&](int tid) { classify_part(parts[tid], lo(tid), lo(tid + 1)); });
        for (int t = 0; t < Tc; ++t) {   // ranges are in index order: the first error is the lowest index
            const Part &P = parts[t];
            if (P.bad >= 0)
                return fail(P.bad_code, P.bad_code == GSW_ERR_INVALID ? "pair %lld index out of range"
                                                                      : "score range of pair %lld exceeds the int32 kernels",
                            (long long)P.bad);
        }
        for (int t = 0; t < Tc; ++t) {
            const Part &P = parts[t];
            res->cells += P.cells;
            e_idx.insert(e_idx.end(), P.e_idx.begin(), P.e_idx.end());
            e_val.insert(e_val.end(), P.e_val.begin(), P.e_val.end());
            p_qi.insert(p_qi.end(), P.p_qi.begin(), P.p_qi.end());
            p_rj.insert(p_rj.end(), P.p_rj.begin(), P.p_rj.end());
            p_oi.insert(p_oi.end(), P.p_oi.begin(), P.p_oi.end());
        }
        // members of each tile (parallel scatter)
        for (size_t v = 0; v < nv; ++v) {
            int64_t p = vbeg[v];
            for (int t = 0; t < Tc; ++t) {
                pos[t][v] = p;
                p += parts[t].count[v];
            }
            vbeg[v + 1] = p;
        }
        if (members.size() < (size_t)vbeg[nv]) members.resize((size_t)vbeg[nv]);
        run_threads(Tc, [&](int tid) {
            std::vector<int64_t> &ps = pos[tid];
            for (int64_t k = lo(tid); k < lo(tid + 1); ++k)
                if (vid[k] >= 0) members[ps[vid[k]]++] = k;
        });
        return GSW_OK;
    };

    // -- a segment's batches: PAIR_BATCH members of one tile each (the call's first tile
    //    ramps up); a tile's batches cycle through its MAX_PAIR_SLOTS slots, and a batch
    //    records the one that used its slot before it (-1: none in this call)
    struct Batch { const Variant *v; bool tab; int kind; int64_t b0, b1; int slot; int64_t prev; };
    std::vector<Batch> batches;
    std::vector<int64_t> kind_pairs(nv, 0), kind_batches(nv, 0), last_use(nv * MAX_PAIR_SLOTS, -1);
    auto add_batches = [&]() {
        for (size_t kind = 0; kind < nv; ++kind) {
            if (vbeg[kind + 1] == vbeg[kind]) continue;
            kind_pairs[kind] += vbeg[kind + 1] - vbeg[kind];
            const Variant &v = variant_table()[kind / 2];
            const bool ramp = batches.empty();
            int bi = 0;
            for (int64_t b0 = vbeg[kind]; b0 < vbeg[kind + 1]; ++bi) {
                const int64_t size = ramp && bi < 4 ? RAMP[bi] : PAIR_BATCH;
                const int64_t b1 = std::min(vbeg[kind + 1], b0 + size);
                const int slot = (int)(kind_batches[kind]++ % MAX_PAIR_SLOTS);
                int64_t &last = last_use[kind * MAX_PAIR_SLOTS + slot];
                batches.push_back(Batch{&v, (kind & 1) != 0, (int)kind, b0, b1, slot, last});
                last = (int64_t)batches.size() - 1;
                b0 = b1;
            }
        }
    };

    SatBuffers sb;
    GSW_TRY(prepare_sat(ctx, &sb, 1 << 20));
    // graphs of the slots a segment's batches use, with this call's pointers in their kernel
    // nodes (a slot is set up once per call), and timing events: fork and join in marks[0]
    // and marks[1], batch b's start and end in marks[2 + 2b] and marks[3 + 2b]
    std::map<std::pair<int, int>, PairGraph *> slot_of;
    auto configure = [&](size_t nb0, size_t nb1) -> int {
        while (ctx->marks.size() < 2 * nb1 + 2) {
            hipEvent_t e;
            GSW_HIP_TRY(hipEventCreate(&e));
            ctx->marks.push_back(e);
        }
        for (size_t bb = nb0; bb < nb1; ++bb) {
            const Batch &B = batches[bb];
            if (slot_of.count({B.kind, B.slot})) continue;
            PairGraph *pg = nullptr;
            GSW_TRY(pair_graph(ctx, *B.v, B.tab, perm, B.slot, s.ns, prow + 1, &pg));
            slot_of[{B.kind, B.slot}] = pg;
            KArgs a = base_args(s, qs, rs, d_out);
            a.sat = sb.d_sat;
            a.sat_count = ctx->counters + 1;
            a.sat_cap = sb.cap;
            a.sat_check = (B.v->packed && local) ? 1 : 0;
            a.items = pg->d_buf + PAIR_HDR;
            a.nitems_ptr = reinterpret_cast<const int32_t *>(pg->d_buf);
            a.counter = reinterpret_cast<int32_t *>(pg->d_buf) + 1;
            void *params[] = {&a};
            hipKernelNodeParams kp{};
            a.tab_copies = B.tab ? pairs_tab_copies(*B.v, s.ns) : 1;
            a.prow = prow;
            kp.func = const_cast<void *>(B.tab ? tab_kernel_fn(*B.v, perm) : B.v->pairs_fn);
            kp.gridDim = dim3(pg->grid);
            kp.blockDim = dim3(pg->waves * 64);
            kp.sharedMemBytes = (unsigned)pg->lds;
            kp.kernelParams = params;
            kp.extra = nullptr;
            GSW_HIP_TRY(hipGraphExecKernelNodeSetParams(pg->exec, pg->kern, &kp));
        }
        return GSW_OK;
    };

    // Plan one batch into its slot's pinned buffer (plan_pair_batch: by reference length
    // rank, then front padding; adjacent members of one reference packed into one task).
    using Scratch = PairPlanScratch;
    std::vector<double> batch_cells;
    const PairPlanView view{h_qi, h_rj, nullptr, h_outidx, qs.len.data(), rs.len.data(), rs.rank.data(),
                            rs.n_nonempty, wmax};
    auto plan = [&](const Batch &B, PairGraph &g, Scratch &S, double *cells) -> int32_t {
        PairPlanView in = view;
        in.members = members.data();
        const int32_t nt = plan_pair_batch(in, B.b0, B.b1, B.v->G, B.v->R, B.v->packed, track_last_col, S,
                                           reinterpret_cast<PairTask *>(g.h_buf + PAIR_HDR), cells);
        int32_t *hdr = reinterpret_cast<int32_t *>(g.h_buf);
        hdr[0] = nt;   // task count
        hdr[1] = 0;    // work counter (the graph's copy zeroes it on the device)
        return nt;
    };

    // -- planners (worker threads) and the launching thread of a segment, batches
    //    [seg_begin, seg_end).  A planner reuses a slot's pinned buffer only after the batch
    //    that last used it was launched and its copy has read the buffer.
    std::vector<int> state;   // per batch: 0 pending, 1 planned, 2 launched
    std::vector<int32_t> ntasks;
    size_t seg_begin = 0, seg_end = 0;
    std::mutex m;
    std::condition_variable cv;
    bool abort = false;
    int werr = GSW_OK;
    std::string werr_msg;
    std::atomic<size_t> next{0};
    auto worker = [&]() {
        Scratch S;
        for (;;) {
            const size_t b = next.fetch_add(1);
            if (b >= seg_end) return;
            const Batch &B = batches[b];
            PairGraph &g = *slot_of.at({B.kind, B.slot});
            if (B.prev >= 0) {   // launched already when it belongs to an earlier segment
                std::unique_lock<std::mutex> lk(m);
                cv.wait(lk, [&] { return state[B.prev] == 2 || abort; });
                if (abort) return;
            }
            // the slot's last copy (of this call or an earlier one) has read h_buf
            const hipError_t e = hipEventSynchronize(g.copied);
            if (e != hipSuccess) {
                std::lock_guard<std::mutex> lk2(m);
                abort = true;
                werr = GSW_ERR_HIP;
                werr_msg = hipGetErrorString(e);
                cv.notify_all();
                return;
            }
            const int32_t nt = plan(B, g, S, &batch_cells[b]);
            std::lock_guard<std::mutex> lk(m);
            ntasks[b] = nt;
            state[b] = 1;
            cv.notify_all();
        }
    };
    int rc = GSW_OK;
    // Batches alternate between the library stream and the second one; on its stream a
    // batch's task copy follows the kernel two batches back, so it overlaps the other
    // stream's kernel, and its kernel takes the CUs as that kernel's last waves finish.
    const hipStream_t streams[2] = {ctx->stream, ctx->stream2};
    auto launcher = [&]() {
        for (size_t b = seg_begin; b < seg_end && rc == GSW_OK; ++b) {
            {
                std::unique_lock<std::mutex> lk(m);
                cv.wait(lk, [&] { return state[b] >= 1 || abort; });
                if (abort) {
                    rc = fail(werr != GSW_OK ? werr : GSW_ERR_HIP, "pairs planning failed: %s", werr_msg.c_str());
                    break;
                }
            }
            const Batch &B = batches[b];
            PairGraph &g = *slot_of.at({B.kind, B.slot});
            // copy once the slot's previous kernel (on either stream) is done with d_buf, then
            // run; planners reuse h_buf once `copied` has passed
            const hipStream_t st = streams[b & 1];
            hipError_t e = hipStreamWaitEvent(st, g.consumed, 0);
            if (e == hipSuccess)
                e = hipMemcpyAsync(g.d_buf, g.h_buf, PAIR_HDR + sizeof(PairTask) * (size_t)ntasks[b],
                                   hipMemcpyHostToDevice, st);
            if (e == hipSuccess) e = hipEventRecord(g.copied, st);
            if (e == hipSuccess) e = hipEventRecord(ctx->marks[2 + 2 * b], st);
            if (e == hipSuccess) e = hipGraphLaunch(g.exec, st);
            if (e == hipSuccess) e = hipEventRecord(ctx->marks[3 + 2 * b], st);
            if (e == hipSuccess) e = hipEventRecord(g.consumed, st);
            if (e != hipSuccess) {
                rc = fail(GSW_ERR_HIP, "pairs batch launch: %s", hipGetErrorString(e));
                break;
            }
            res->launches++;
            if (b == 0) t_first = clk::now();
            {
                std::lock_guard<std::mutex> lk(m);
                state[b] = 2;
            }
            cv.notify_all();
        }
        if (rc != GSW_OK) {   // release planners waiting for a launch
            {
                std::lock_guard<std::mutex> lk(m);
                abort = true;
            }
            cv.notify_all();
        }
    };
    // a segment: its graphs and events, then the launcher on the calling thread and up to T
    // planners beside it
    auto run_segment = [&]() -> int {
        state.resize(seg_end, 0);
        ntasks.resize(seg_end, 0);
        batch_cells.resize(seg_end, 0.0);
        GSW_TRY(configure(seg_begin, seg_end));
        if (seg_end == seg_begin) return GSW_OK;
        next.store(seg_begin);
        const int W = (int)std::min<size_t>((size_t)T, seg_end - seg_begin);
        run_threads(W + 1, [&](int tid) {
            if (tid == 0)
                launcher();
            else
                worker();
        });
        return rc;
    };
    auto stop = [&](int code) {   // nothing may still run when the caller sees the error
        (void)hipStreamSynchronize(ctx->stream2);
        (void)hipStreamSynchronize(ctx->stream);
        return code;
    };

    GSW_TRY(classify(0, head));
    const clk::time_point t_classified = clk::now();
    add_batches();
    seg_end = batches.size();
    GSW_TRY(configure(0, 0));   // fork and join events
    const hipEvent_t fork = ctx->marks[0], join = ctx->marks[1];
    // device time from here; the second stream starts after the saturation counter's reset
    GSW_HIP_TRY(hipEventRecord(ctx->ev0, ctx->stream));
    GSW_HIP_TRY(hipEventRecord(fork, ctx->stream));
    GSW_HIP_TRY(hipStreamWaitEvent(ctx->stream2, fork, 0));
    if (int e = run_segment(); e != GSW_OK) return stop(e);
    t_head = t_rest = clk::now();
    if (head < n) {   // the rest, while the device runs the head
        if (int e = classify(head, n); e != GSW_OK) return stop(e);
        t_rest = clk::now();
        seg_begin = seg_end;
        add_batches();
        seg_end = batches.size();
        if (int e = run_segment(); e != GSW_OK) return stop(e);
    }
    const size_t nb = batches.size();
    if (nb == 0) t_first = t_rest;
    const clk::time_point t_launched = clk::now();
    GSW_HIP_TRY(hipEventRecord(join, ctx->stream2));
    GSW_HIP_TRY(hipStreamWaitEvent(ctx->stream, join, 0));
    GSW_HIP_TRY(hipEventRecord(ctx->ev1, ctx->stream));
    GSW_HIP_TRY(hipEventSynchronize(ctx->ev1));
    float ms = 0.f;
    GSW_HIP_TRY(hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
    res->kernel_ms += ms;
    if (trace) {
        auto d = [](clk::time_point a, clk::time_point b) { return std::chrono::duration<double, std::milli>(b - a).count(); };
        std::fprintf(stderr, "[gsw] pairs n=%lld head=%lld batches=%zu: head classified %.2f ms, first launch +%.2f ms, "
                             "head launched +%.2f ms, rest classified +%.2f ms, all launched +%.2f ms, device done "
                             "+%.2f ms (device %.2f ms)\n",
                     (long long)n, (long long)head, nb, d(t_in, t_classified), d(t_classified, t_first),
                     d(t_first, t_head), d(t_head, t_rest), d(t_rest, t_launched), d(t_launched, clk::now()), ms);
    }
    int dominant = -1;   // kind with the most pairs
    for (size_t kind = 0; kind < nv; ++kind)
        if (kind_pairs[kind] > 0 && (dominant < 0 || kind_pairs[kind] > kind_pairs[dominant])) dominant = (int)kind;
    res->variant = dominant >= 0 ? dominant / 2 : -1;
    // the dominant tile's full batches (all of its batches if none is full): device time
    // from the first one's start event to the last one's end event, per batch (consecutive
    // batches overlap on the two streams, so one batch's own interval would count shared
    // time), and mean cells
    {
        bool any_full = false;
        for (size_t b = 0; b < nb; ++b)
            any_full |= batches[b].kind == dominant && batches[b].b1 - batches[b].b0 == PAIR_BATCH;
        double sum_cells = 0, t_lo = 1e30, t_hi = -1e30;
        int cnt = 0;
        for (size_t b = 0; b < nb; ++b) {
            if (batches[b].kind != dominant) continue;
            if (any_full && batches[b].b1 - batches[b].b0 != PAIR_BATCH) continue;
            float t0 = 0.f, t1 = 0.f;
            GSW_HIP_TRY(hipEventElapsedTime(&t0, ctx->ev0, ctx->marks[2 + 2 * b]));
            GSW_HIP_TRY(hipEventElapsedTime(&t1, ctx->ev0, ctx->marks[3 + 2 * b]));
            t_lo = std::min(t_lo, (double)t0);
            t_hi = std::max(t_hi, (double)t1);
            sum_cells += batch_cells[b];
            ++cnt;
        }
        if (cnt && t_hi > t_lo) {
            res->top_ms = (t_hi - t_lo) / cnt;
            res->top_cells = (int64_t)(sum_cells / cnt);
            res->top_variant = dominant / 2;
            // pairs kernel (3: table kernel, 4: its byte-select form)
            res->top_striped = !(dominant & 1) ? 2 : perm && variant_table()[dominant / 2].pairs_perm_fn ? 4 : 3;
        }
    }

    // local results at the top of a packed type (int16 32767, half2 2048) are listed before
    // any other call resets the list; they run again below: the listed ones on the int32
    // tiles or, when more were listed than the list holds, every pair that ran packed (half2
    // ones on the int16 tiles, whose own saturated results are listed in turn)
    std::vector<int32_t> rqi, rrj, hqi, hrj;
    std::vector<int64_t> roi, hoi;
    {
        std::vector<SatEntry> sat;
        bool overflowed = false;
        GSW_TRY(read_saturated(ctx, sb, &sat, &overflowed));
        if (overflowed) {
            const std::vector<Variant> &vt = variant_table();
            for (int64_t k = 0; k < n; ++k) {
                if (vid[k] < 0) continue;
                const Variant &v = vt[vid[k] / 2];
                if (!v.packed || !v.local) continue;
                const bool h = v.kind == KIND_F16;
                (h ? hqi : rqi).push_back(h_qi[k]);
                (h ? hrj : rrj).push_back(h_rj[k]);
                (h ? hoi : roi).push_back(oidx(k));
            }
        } else {
            for (const SatEntry &e : sat) {
                rqi.push_back(e.q);
                rrj.push_back(e.r);
                roi.push_back(e.o);
            }
        }
    }
    // (the calls below reuse this thread's planning buffers vid and members: nothing reads
    // them after this point)

    // Pairs without a tile have a query wider than every pairs tile.  Where the reference
    // fits one, the pair is scored with query and reference exchanged under the transposed
    // scheme (same score in every mode); pairs longer than every tile on both sides run on
    // the int32 stripe kernels (batch_long_pairs), and only what no stripe tile takes on the
    // one-thread-per-pair comparator.
    if (!p_qi.empty()) {
        std::vector<int32_t> tq, tr, cq, cr;
        std::vector<int64_t> to, co;
        for (size_t i = 0; i < p_qi.size(); ++i) {
            const bool swap = qs.len[p_qi[i]] > wmax && rs.len[p_rj[i]] <= wmax;
            (swap ? tq : cq).push_back(p_qi[i]);
            (swap ? tr : cr).push_back(p_rj[i]);
            (swap ? to : co).push_back(p_oi[i]);
        }
        if (!tq.empty()) {
            SchemeDev st;
            GSW_TRY(scheme_transpose(ctx, s, &st));
            BatchResult r;
            GSW_TRY(batch_pairs(ctx, st, dtype_policy, rs, qs, tr.data(), tq.data(), (int64_t)tq.size(), d_out, &r,
                                to.data()));
            res->kernel_ms += r.kernel_ms;
            res->launches += r.launches;
            res->plain_pairs += r.plain_pairs;
            res->rescored += r.rescored;
        }
        if (!cq.empty()) {   // the int32 stripe kernels, as lists of references per query
            std::vector<int32_t> lq, lr;
            std::vector<int64_t> lo;
            GSW_TRY(batch_long_pairs(ctx, s, qs, rs, cq, cr, co, d_out, res, &lq, &lr, &lo));
            cq.swap(lq);
            cr.swap(lr);
            co.swap(lo);
        }
        if (!cq.empty()) {
            int maxq = 1;
            bool use16 = dtype_policy != GSW_DTYPE_INT32 && !local;
            for (size_t i = 0; i < cq.size(); ++i) {
                maxq = std::max(maxq, qs.len[cq[i]]);
                use16 = use16 && fits_int16(s, qs.len[cq[i]], rs.len[cr[i]]);
            }
            PlainJob job{};
            job.qcodes = qs.d_codes; job.qoff = qs.d_off; job.rcodes = rs.d_codes; job.roff = rs.d_off;
            job.dtype = use16 ? GSW_DTYPE_INT16 : GSW_DTYPE_INT32;
            job.threads = 256; job.maxq = maxq;
            job.h_qi = cq.data(); job.h_rj = cr.data(); job.h_outidx = co.data();
            job.npairs = (int64_t)cq.size(); job.d_out = d_out;
            float pms = 0.f;
            GSW_TRY(plain_run(ctx, s, job, &pms));
            res->kernel_ms += pms;
            res->plain_pairs += (int64_t)cq.size();
            res->launches++;
        }
    }
    GSW_TRY(scatter_values(ctx, d_out, e_idx, e_val));
    if (!hqi.empty()) {
        BatchResult r;
        GSW_TRY(batch_pairs(ctx, s, GSW_DTYPE_PK16, qs, rs, hqi.data(), hrj.data(), (int64_t)hqi.size(), d_out, &r,
                            hoi.data()));
        res->kernel_ms += r.kernel_ms;
        res->launches += r.launches;
        res->plain_pairs += r.plain_pairs;
        res->rescored += (int64_t)hqi.size() + r.rescored;
    }
    return rescore_pairs_int32(ctx, s, qs, rs, rqi, rrj, roi, d_out, res);
}

}  // namespace gsw

using namespace gsw;

static int batch_common(const gsw_scheme *s, int32_t dtype_policy, DeviceCtx **ctx, SchemeDev *sd) {
    if (dtype_policy < GSW_DTYPE_AUTO || dtype_policy > GSW_DTYPE_INT32)
        return fail(GSW_ERR_INVALID, "dtype_policy %d out of range", dtype_policy);
    GSW_TRY(ctx_get(ctx));
    GSW_TRY(scheme_prepare(s, *ctx, sd));
    return GSW_OK;
}

static void fill_stats(gsw_stats *st, const BatchResult &r, int64_t cells, int64_t pairs,
                       std::chrono::steady_clock::time_point t0) {
    if (!st) return;
    *st = gsw_stats{};
    st->kernel_ms = r.kernel_ms;
    st->cells = cells;
    st->pairs = pairs;
    st->rescored = r.rescored;
    st->plain_pairs = r.plain_pairs;
    st->launches = r.launches;
    st->variant = r.variant;
    st->top_ms = r.top_ms;
    st->top_cells = r.top_cells;
    st->top_variant = r.top_variant;
    st->top_striped = r.top_striped;
    st->wall_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
}

extern "C" int gsw_score_cross(const gsw_scheme *s, int32_t dtype_policy, gsw_seqset queries,
                               const int32_t *h_qlist, int32_t nq, gsw_seqset refs, int32_t *d_out,
                               gsw_stats *stats) {
    auto t0 = std::chrono::steady_clock::now();
    const SeqSet *qs = reinterpret_cast<const SeqSet *>(queries);
    const SeqSet *rs = reinterpret_cast<const SeqSet *>(refs);
    if (!qs || !rs) return fail(GSW_ERR_INVALID, "null sequence set");
    if (nq < 0) return fail(GSW_ERR_INVALID, "negative query count");
    DeviceCtx *ctx = nullptr;
    SchemeDev sd;
    GSW_TRY(batch_common(s, dtype_policy, &ctx, &sd));
    if (qs->device != ctx->device || rs->device != ctx->device)
        return fail(GSW_ERR_INVALID, "sequence set belongs to another device");
    std::lock_guard<std::mutex> lk(ctx->mu);
    BatchResult r;
    GSW_TRY(batch_cross(ctx, sd, dtype_policy, *qs, h_qlist, nq, *rs, d_out, &r));
    int64_t qsum = 0;
    for (int a = 0; a < nq; ++a) qsum += qs->len[h_qlist[a]];
    fill_stats(stats, r, qsum * rs->total, (int64_t)nq * rs->n, t0);
    return GSW_OK;
}

extern "C" int gsw_score_pairs(const gsw_scheme *s, int32_t dtype_policy, gsw_seqset queries,
                               gsw_seqset refs, const int32_t *h_qi, const int32_t *h_rj, int64_t n,
                               int32_t *d_out, gsw_stats *stats) {
    auto t0 = std::chrono::steady_clock::now();
    const SeqSet *qs = reinterpret_cast<const SeqSet *>(queries);
    const SeqSet *rs = reinterpret_cast<const SeqSet *>(refs);
    if (!qs || !rs) return fail(GSW_ERR_INVALID, "null sequence set");
    if (n < 0) return fail(GSW_ERR_INVALID, "negative pair count");
    DeviceCtx *ctx = nullptr;
    SchemeDev sd;
    GSW_TRY(batch_common(s, dtype_policy, &ctx, &sd));
    if (qs->device != ctx->device || rs->device != ctx->device)
        return fail(GSW_ERR_INVALID, "sequence set belongs to another device");
    std::lock_guard<std::mutex> lk(ctx->mu);
    BatchResult r;
    GSW_TRY(batch_pairs(ctx, sd, dtype_policy, *qs, *rs, h_qi, h_rj, n, d_out, &r));
    int64_t cells = r.cells;
    if (cells < 0) {
        cells = 0;
        for (int64_t k = 0; k < n; ++k) cells += (int64_t)qs->len[h_qi[k]] * rs->len[h_rj[k]];
    }
    fill_stats(stats, r, cells, n, t0);
    return GSW_OK;
}

// Tuning hook: force the (G, R) tile of the batched kernels (0, 0 = automatic).
extern "C" int gsw_force_variant(int32_t G, int32_t R) {
    if (G == 0 && R == 0) {
        g_force.store(0);
        return GSW_OK;
    }
    for (const Variant &v : variant_table())
        if (v.G == G && v.R == R) {
            g_force.store((G << 16) | R);
            return GSW_OK;
        }
    return fail(GSW_ERR_INVALID, "no compiled variant G=%d R=%d", G, R);
}

// Install a measured tile table: n rules of 6 int32 (lo, hi, packed, local, G, R); n = 0
// clears it.  Rules are consulted before the cost model.
extern "C" int gsw_set_tile_table(const int32_t *rules, int32_t n) {
    if (n < 0 || (n > 0 && !rules)) return fail(GSW_ERR_INVALID, "bad tile table");
    std::vector<TileRule> t;
    for (int i = 0; i < n; ++i) {
        const int32_t *r = rules + 6 * i;
        t.push_back(TileRule{r[0], r[1], r[2] ? 1 : 0, r[3] ? 1 : 0, r[4], r[5]});
    }
    std::lock_guard<std::mutex> lk(g_rules_mu);
    g_rules.swap(t);
    return GSW_OK;
}

// List compiled (G, R) tiles: writes up to cap (G, R) pairs, returns the count in *n.
// (G, R) of the tiles that also have a stripe kernel (queries wider than the tile).
extern "C" int gsw_list_stripe_variants(int32_t *gr, int32_t cap, int32_t *n) {
    if (!n || (cap > 0 && !gr)) return fail(GSW_ERR_INVALID, "null argument");
    int cnt = 0, last_g = -1, last_r = -1;
    for (const Variant &v : variant_table()) {
        if (!v.stripe_fn || (v.G == last_g && v.R == last_r)) continue;
        last_g = v.G;
        last_r = v.R;
        if (cnt < cap) {
            gr[2 * cnt] = v.G;
            gr[2 * cnt + 1] = v.R;
        }
        ++cnt;
    }
    *n = cnt;
    return GSW_OK;
}

extern "C" int gsw_variant_info(int32_t id, int32_t *info) {
    const std::vector<Variant> &vt = variant_table();
    if (!info || id < 0 || id >= (int32_t)vt.size()) return fail(GSW_ERR_INVALID, "no variant %d", id);
    info[0] = vt[id].G;
    info[1] = vt[id].R;
    info[2] = vt[id].kind;
    info[3] = vt[id].local ? 1 : 0;
    return GSW_OK;
}

extern "C" int gsw_list_variants(int32_t *gr, int32_t cap, int32_t *n) {
    int cnt = 0;
    int last_g = -1, last_r = -1;
    for (const Variant &v : variant_table()) {
        if (v.G == last_g && v.R == last_r) continue;
        last_g = v.G;
        last_r = v.R;
        if (cnt < cap) {
            gr[2 * cnt] = v.G;
            gr[2 * cnt + 1] = v.R;
        }
        ++cnt;
    }
    *n = cnt;
    return GSW_OK;
}
