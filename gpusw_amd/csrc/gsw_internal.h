// Internal declarations shared by the libgsw_hip.so translation units.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <mutex>
#include <string>
#include <vector>

#include "gsw_hip.h"

namespace gsw {

// ---------------------------------------------------------------- errors
void set_error(const char *fmt, ...);
int fail(int code, const char *fmt, ...);

#define GSW_HIP_TRY(expr)                                                              \
    do {                                                                               \
        hipError_t _e = (expr);                                                        \
        if (_e != hipSuccess)                                                          \
            return ::gsw::fail(GSW_ERR_HIP, "%s failed: %s (%s:%d)", #expr,           \
                               hipGetErrorString(_e), __FILE__, __LINE__);             \
    } while (0)

#define GSW_TRY(expr)                                                                  \
    do {                                                                               \
        int _rc = (expr);                                                              \
        if (_rc != GSW_OK) return _rc;                                                 \
    } while (0)

// ---------------------------------------------------------------- device context
// Device counters of a context: [1] the saturation count, [2, 2 + MAX_GROUPS) one work
// counter per launch of a call, [NCOUNTERS - 1] the max-code reduction.  A call forms at
// most one launch per (variant, striped), and the variant table holds fewer than
// MAX_GROUPS / 2 rows of one mode, so the group limit is never what stops a call.
constexpr int NCOUNTERS = 512;
constexpr int MAX_GROUPS = NCOUNTERS - 3;

// One library stream and a set of grow-only scratch buffers per device.
struct DeviceCtx {
    int device = -1;
    int num_cus = 0;
    int max_smem = 0;
    hipStream_t stream = nullptr;
    // second stream (default priority): every other cross launch and every other pair batch
    // (its task copy and kernel), so that consecutive launches overlap at their ends
    hipStream_t stream2 = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    // grow-only scratch buffers owned by the library
    void *plain_ws = nullptr;  size_t plain_ws_bytes = 0;
    void *meta = nullptr;      size_t meta_bytes = 0;      // task lists, lists of ids
    void *meta2 = nullptr;     size_t meta2_bytes = 0;
    void *bnd = nullptr;       size_t bnd_bytes = 0;       // stripe boundaries
    int32_t *counters = nullptr;                            // NCOUNTERS int32 counters
    std::vector<hipEvent_t> marks;                          // per-launch timing events
    std::mutex mu;                                          // serialises calls on this device
};

int ctx_get(DeviceCtx **out);                    // context of the current device
int ensure_buffer(void **buf, size_t *cap, size_t need);

// Scheme validated and translated for the kernels.
struct SchemeDev {
    int mode, go, ge, free_q, free_r, ns;
    int sub_min, sub_max;                       // table extremes
    const int32_t *d_sub;                       // device copy (cached by content)
    std::vector<int32_t> h_sub;
};
int scheme_validate(const gsw_scheme *s, SchemeDev *out);   // host only; d_sub left null
int scheme_prepare(const gsw_scheme *s, DeviceCtx *ctx, SchemeDev *out);
// The scheme with query and reference exchanged: transposed table, free-end flags swapped.
// score(q, r) under s equals score(r, q) under the transposed scheme, in every mode.
int scheme_transpose(DeviceCtx *ctx, const SchemeDev &s, SchemeDev *out);

// Element type of a tile: one int32 alignment per lane, or two per lane as int16 or as
// binary16 (exact integers in [-2048, 2048], three-input max; see gsw_sweep.h).
enum : int { KIND_I32 = 0, KIND_I16 = 1, KIND_F16 = 2 };

// Static bound check used to choose int16 per pair (sound for every DP cell).
bool int16_entries(const SchemeDev &s);   // table and gap scores fit int16
bool fits_int16(const SchemeDev &s, int lq, int lr);
bool fits_int32(const SchemeDev &s, int lq, int lr);
// Element type of a pair under a dtype policy: GSW_DTYPE_AUTO applies the envelope per pair
// (int16 only where no value can saturate); GSW_DTYPE_INT16 runs every local pair in int16
// and relies on the kernels' saturation list; GSW_DTYPE_INT32 never uses int16.
bool use_int16(const SchemeDev &s, int dtype_policy, int lq, int lr);

// Binary16 tiles (exact integers in [-2048, 2048]): the table and gap scores fit, and for a
// pair every DP value does (global/semiglobal; local pairs also run in half2 outside this
// bound, their results of 2048 or more listed and rescored).  GSW_HALF=0 in the environment
// turns the half2 tiles off (read per call).
bool half_entries(const SchemeDev &s);
bool fits_half(const SchemeDev &s, int lq, int lr);
bool half_enabled();
// The envelopes of the packed batched sweeps, which carry global/semiglobal values in the
// diagonal-offset form (gsw_sweep.h): local mode as fits_int16/fits_half; otherwise every
// offset value of the pair's real cells fits the type.  (fits_int16 keeps the plain values'
// envelope for the one-pair-per-thread comparator's int16 rows.)
bool sweep_fits_int16(const SchemeDev &s, int lq, int lr);
bool sweep_fits_half(const SchemeDev &s, int lq, int lr);
// Internal dtype policy: packed int16 where GSW_DTYPE_INT16 would use it, never half2 (the
// rerun of local work whose half2 results overflowed the saturation list).
constexpr int GSW_DTYPE_PK16 = 16;
// Internal dtype policy: each pair on the narrowest type its envelope proves exact, so no
// result is ever listed for rescoring (the serving loop, which has no rescoring pass).
constexpr int GSW_DTYPE_SAFE = 17;
// As GSW_DTYPE_SAFE without half2 tiles (the serving loop's rerun of listed half2 results).
constexpr int GSW_DTYPE_SAFE16 = 18;
// Tile element kind (KIND_I32 / KIND_I16 / KIND_F16, gsw_batch.h) of a pair under a policy,
// from its int16 and half2 envelope checks; half: half2 tiles allowed for this call.
int pair_kind(const SchemeDev &s, int dtype_policy, bool ok16, bool okh, bool half);

// Closed-form score when one side is empty (src/gpusw/kernel.py:63-80).
int32_t empty_edge_score(const SchemeDev &s, int lq, int lr);

// ---------------------------------------------------------------- seqset
struct SeqSet {
    const uint8_t *d_codes = nullptr;
    const int32_t *d_off = nullptr;
    std::vector<int32_t> h_off;
    std::vector<int32_t> len;
    int32_t n = 0;
    int32_t max_len = 0;
    int64_t total = 0;
    int32_t max_code = 255;            // largest code of the set (read once at creation)
    std::vector<int32_t> order_desc;   // indices sorted by length, longest first
    std::vector<int32_t> rank;         // rank[i] = position of sequence i in order_desc
    int32_t *d_order_desc = nullptr;   // device copy of order_desc (non-empty only)
    std::vector<int64_t> bpos;         // bpos[p] = total length of order_desc[0..p)
    int64_t *d_bpos = nullptr;         // device copy (n + 1)
    int32_t n_nonempty = 0;
    int device = -1;
    SeqSet() = default;
    SeqSet(const SeqSet &) = delete;
    SeqSet &operator=(const SeqSet &) = delete;
    ~SeqSet() {
        if (d_order_desc) (void)hipFree(d_order_desc);
        if (d_bpos) (void)hipFree(d_bpos);
    }
};

// ---------------------------------------------------------------- plain kernel
// Score pairs with the one-thread-per-pair kernel.  Either a cross product of
// qrows (query ids) x all nr refs, writing out[out_row[a]*ldo + b], or explicit
// pairs writing out[out_idx[k]] (out_idx may be null for identity).
struct PlainJob {
    const uint8_t *qcodes; const int32_t *qoff;
    const uint8_t *rcodes; const int32_t *roff;
    int dtype;            // GSW_DTYPE_INT16 / INT32 (store type)
    int threads;
    int maxq;
    // cross mode
    const int32_t *h_qlist = nullptr; const int32_t *h_outrow = nullptr; int32_t nq = 0;
    int32_t nr = 0; int64_t ldo = 0;
    // pairs mode
    const int32_t *h_qi = nullptr; const int32_t *h_rj = nullptr; const int64_t *h_outidx = nullptr;
    const int32_t *h_qstart = nullptr; const int32_t *h_qlen = nullptr;   // instead of h_qi
    int64_t npairs = 0;
    int32_t *d_out;
};
int plain_run(DeviceCtx *ctx, const SchemeDev &s, const PlainJob &job, float *ms);

// ---------------------------------------------------------------- batched kernels
struct BatchResult { double kernel_ms = 0; int64_t rescored = 0; int64_t plain_pairs = 0;
                     int launches = 0; int variant = -1; int64_t cells = -1; /* -1: not counted */
                     double top_ms = 0; int64_t top_cells = 0; int top_variant = -1, top_striped = 0; };

// Cross product of queries h_qlist[a] x every reference into d_out[outrow(a) * nr + b],
// outrow(a) = h_outrow ? h_outrow[a] : a.
int batch_cross(DeviceCtx *ctx, const SchemeDev &s, int dtype_policy, const SeqSet &qs,
                const int32_t *h_qlist, int32_t nq, const SeqSet &rs, int32_t *d_out,
                BatchResult *res, const int32_t *h_outrow = nullptr);
// Explicit pairs (h_qi[k], h_rj[k]) into d_out[h_outidx ? h_outidx[k] : k].
int batch_pairs(DeviceCtx *ctx, const SchemeDev &s, int dtype_policy, const SeqSet &qs,
                const SeqSet &rs, const int32_t *h_qi, const int32_t *h_rj, int64_t n,
                int32_t *d_out, BatchResult *res, const int64_t *h_outidx = nullptr);

}  // namespace gsw
