// Runtime part of libgsw_hip.so: error state, device contexts, memory, scheme
// preparation and indexed sequence sets.
#include <algorithm>
#include <cstdarg>
#include <cstdlib>
#include <cstring>
#include <map>
#include <memory>
#include <numeric>

#include "gsw_internal.h"

namespace gsw {

static thread_local std::string g_error;

__device__ __forceinline__ int wave_max(int v) {
    for (int off = 32; off >= 1; off >>= 1) v = max(v, __shfl_xor(v, off, 64));
    return v;
}

void set_error(const char *fmt, ...) {
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_error = buf;
}

int fail(int code, const char *fmt, ...) {
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_error = buf;
    return code;
}

// ---------------------------------------------------------------- contexts
static std::mutex g_ctx_mu;
static std::map<int, std::unique_ptr<DeviceCtx>> g_ctx;

static int init_runtime(int *count) {
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess || n <= 0) {
        return fail(GSW_ERR_NO_DEVICE, "no AMD GPU visible to the HIP runtime (%s)",
                    e == hipSuccess ? "0 devices" : hipGetErrorString(e));
    }
    *count = n;
    return GSW_OK;
}

int ctx_get(DeviceCtx **out) {
    int count = 0;
    GSW_TRY(init_runtime(&count));
    int dev = 0;
    GSW_HIP_TRY(hipGetDevice(&dev));
    std::lock_guard<std::mutex> lk(g_ctx_mu);
    auto it = g_ctx.find(dev);
    if (it != g_ctx.end()) {
        *out = it->second.get();
        return GSW_OK;
    }
    auto ctx = std::make_unique<DeviceCtx>();
    ctx->device = dev;
    hipDeviceProp_t prop;
    GSW_HIP_TRY(hipGetDeviceProperties(&prop, dev));
    ctx->num_cus = prop.multiProcessorCount;
    ctx->max_smem = (int)prop.sharedMemPerBlock;
    // Default-priority stream (non-default priorities are avoided on purpose).
    GSW_HIP_TRY(hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking));
    GSW_HIP_TRY(hipStreamCreateWithFlags(&ctx->stream2, hipStreamNonBlocking));
    GSW_HIP_TRY(hipEventCreate(&ctx->ev0));
    GSW_HIP_TRY(hipEventCreate(&ctx->ev1));
    GSW_HIP_TRY(hipMalloc(&ctx->counters, NCOUNTERS * sizeof(int32_t)));
    *out = ctx.get();
    g_ctx[dev] = std::move(ctx);
    return GSW_OK;
}

int ensure_buffer(void **buf, size_t *cap, size_t need) {
    if (need <= *cap && *buf) return GSW_OK;
    size_t want = std::max(need, *cap + *cap / 2);
    want = (want + 255) & ~size_t(255);
    if (*buf) GSW_HIP_TRY(hipFree(*buf));
    *buf = nullptr;
    *cap = 0;
    GSW_HIP_TRY(hipMalloc(buf, want));
    *cap = want;
    return GSW_OK;
}

// ---------------------------------------------------------------- scheme
struct SubCacheKey {
    int device;
    std::vector<int32_t> table;
    bool operator<(const SubCacheKey &o) const {
        if (device != o.device) return device < o.device;
        return table < o.table;
    }
};
static std::mutex g_sub_mu;
static std::map<SubCacheKey, int32_t *> g_sub_cache;   // src/gpusw/_compile.py:83-93

int scheme_validate(const gsw_scheme *s, SchemeDev *out) {
    if (!s) return fail(GSW_ERR_INVALID, "scheme is null");
    if (s->mode < 0 || s->mode > 2) return fail(GSW_ERR_INVALID, "mode %d out of range", s->mode);
    if (s->gap_open > 0 || s->gap_extend > 0)
        return fail(GSW_ERR_INVALID, "gap_open and gap_extend must be <= 0");
    if (s->ns < 1 || s->ns > 64) return fail(GSW_ERR_INVALID, "ns=%d out of range 1..64", s->ns);
    if (!s->sub) return fail(GSW_ERR_INVALID, "substitution table is null");
    if (s->gap_open < -(1 << 20) || s->gap_extend < -(1 << 20))
        return fail(GSW_ERR_OVERFLOW, "gap scores below -2^20 are not supported");
    out->mode = s->mode;
    out->go = s->gap_open;
    out->ge = s->gap_extend;
    out->free_q = s->free_q ? 1 : 0;
    out->free_r = s->free_r ? 1 : 0;
    out->ns = s->ns;
    out->h_sub.assign(s->sub, s->sub + (size_t)s->ns * s->ns);
    out->sub_min = *std::min_element(out->h_sub.begin(), out->h_sub.end());
    out->sub_max = *std::max_element(out->h_sub.begin(), out->h_sub.end());
    if (out->sub_min < -(1 << 20) || out->sub_max > (1 << 20))
        return fail(GSW_ERR_OVERFLOW, "substitution scores beyond +/-2^20 are not supported");
    out->d_sub = nullptr;
    return GSW_OK;
}

int scheme_prepare(const gsw_scheme *s, DeviceCtx *ctx, SchemeDev *out) {
    GSW_TRY(scheme_validate(s, out));
    std::lock_guard<std::mutex> lk(g_sub_mu);
    SubCacheKey key{ctx->device, out->h_sub};
    auto it = g_sub_cache.find(key);
    if (it == g_sub_cache.end()) {
        int32_t *d = nullptr;
        GSW_HIP_TRY(hipMalloc(&d, out->h_sub.size() * sizeof(int32_t)));
        GSW_HIP_TRY(hipMemcpy(d, out->h_sub.data(), out->h_sub.size() * sizeof(int32_t),
                              hipMemcpyHostToDevice));
        it = g_sub_cache.emplace(key, d).first;
    }
    out->d_sub = it->second;
    return GSW_OK;
}

int scheme_transpose(DeviceCtx *ctx, const SchemeDev &s, SchemeDev *out) {
    std::vector<int32_t> t((size_t)s.ns * s.ns);
    for (int a = 0; a < s.ns; ++a)
        for (int b = 0; b < s.ns; ++b) t[(size_t)b * s.ns + a] = s.h_sub[(size_t)a * s.ns + b];
    const gsw_scheme g{s.mode, s.go, s.ge, s.free_r, s.free_q, s.ns, t.data()};
    return scheme_prepare(&g, ctx, out);
}

// Bounds of any DP value for a (lq, lr) pair: every H lies in [lower, upper]
// (upper: at most min(lq,lr) positive substitutions; lower: a diagonal run plus
// one gap run, each step costing at least min(go, ge, min_sub, 0)).
static void dp_bounds(const SchemeDev &s, int64_t lq, int64_t lr, int64_t *lower, int64_t *upper) {
    *upper = std::min(lq, lr) * (int64_t)std::max(s.sub_max, 0);
    if (s.mode == GSW_MODE_LOCAL) {
        *lower = 0;
    } else {
        int64_t worst = std::min({(int64_t)s.go, (int64_t)s.ge, (int64_t)s.sub_min, (int64_t)0});
        *lower = std::max(lq, lr) * worst;
    }
}

bool int16_entries(const SchemeDev &s) {
    return s.sub_min >= -32767 && s.sub_max <= 32767 && s.go >= -32767 && s.ge >= -32767;
}

// The envelope of src/gpusw/scheme.py:116-141 for one pair, with the margins of the packed
// kernels.  Local: H, E and F are alignment scores of prefixes (>= 0 in the clamped form),
// at most min(lq, lr) * max_sub, so below 32767 no value saturates and no result can be
// mistaken for a saturated one.  Global/semiglobal: every value lies in [lo, hi].
bool fits_int16(const SchemeDev &s, int lq, int lr) {
    if (!int16_entries(s)) return false;
    int64_t lo, hi;
    dp_bounds(s, lq, lr, &lo, &hi);
    if (s.mode == GSW_MODE_LOCAL) return hi + std::max(s.sub_max, 0) + 16 <= 32767;
    int64_t margin = -(int64_t)s.go - (int64_t)s.ge - (int64_t)std::min(s.sub_min, 0) + 16;
    return lo - margin >= -32768 && hi + std::max(s.sub_max, 0) + 16 <= 32767;
}

bool use_int16(const SchemeDev &s, int dtype_policy, int lq, int lr) {
    if (dtype_policy == GSW_DTYPE_INT32) return false;
    // a caller-checked int16 local pair runs packed whatever its bound: saturated results
    // (32767) are listed by the kernel and rescored in int32
    if (dtype_policy == GSW_DTYPE_INT16 && s.mode == GSW_MODE_LOCAL) return int16_entries(s);
    return fits_int16(s, lq, lr);
}

bool half_entries(const SchemeDev &s) {
    return s.sub_min >= -2048 && s.sub_max <= 2048 && s.go >= -2048 && s.ge >= -2048;
}

// As fits_int16 with the exact integer range of binary16, [-2048, 2048].
bool fits_half(const SchemeDev &s, int lq, int lr) {
    if (!half_entries(s)) return false;
    int64_t lo, hi;
    dp_bounds(s, lq, lr, &lo, &hi);
    if (s.mode == GSW_MODE_LOCAL) return hi + std::max(s.sub_max, 0) + 16 <= 2047;
    int64_t margin = -(int64_t)s.go - (int64_t)s.ge - (int64_t)std::min(s.sub_min, 0) + 16;
    return lo - margin >= -2048 && hi + std::max(s.sub_max, 0) + 16 <= 2047;
}

// The packed sweeps carry global/semiglobal values in the diagonal-offset form X + (i+j)*D,
// D = -gap_extend (gsw_sweep.h).  Every value they compute for a real cell lies in [lo, hi]:
//   H >= min(0, go + D, 2go + 2D, -64D): the two-gap path from the corner gives
//        H(i,j) >= 2go + (i+j-2)*ge, the boundaries are go + D, i*D or j*D, and front
//        padding (fewer than 64 columns) carries (i+j)*D with j > -64;
//   H <= max(sub,0)*min(lq,lr) + (lq+lr+64)*D: at most min(lq,lr) substitutions score;
//   the sums diag + s + 2D and H + go + D, and E and F (maxima of the latter), lie within
//   H's range widened by min/max(0, sub + 2D, go + D).
// The last-column maximum of semiglobal free query ends is taken over H(i, lq) + lq*D, which
// is >= go + D (a gap from the free column-0 boundary) and <= hi, so it is exact too.
static bool offset_fits(const SchemeDev &s, int lq, int lr, int64_t vmin, int64_t vmax) {
    const int64_t D = -(int64_t)s.ge, go = s.go, smin = s.sub_min, smax = s.sub_max;
    const int64_t hmin = std::min<int64_t>({0, go + D, 2 * go + 2 * D, -64 * D});
    const int64_t hmax = (int64_t)std::min(lq, lr) * std::max<int64_t>(smax, 0) + ((int64_t)lq + lr + 64) * D;
    const int64_t lo = hmin + std::min<int64_t>({0, smin + 2 * D, go + D}) - 16;
    const int64_t hi = hmax + std::max<int64_t>({0, smax + 2 * D, go + D}) + 16;
    return lo >= vmin && hi <= vmax;
}

bool sweep_fits_int16(const SchemeDev &s, int lq, int lr) {
    if (s.mode == GSW_MODE_LOCAL) return fits_int16(s, lq, lr);
    return int16_entries(s) && offset_fits(s, lq, lr, -32767, 32766);
}

bool sweep_fits_half(const SchemeDev &s, int lq, int lr) {
    if (s.mode == GSW_MODE_LOCAL) return fits_half(s, lq, lr);
    return half_entries(s) && offset_fits(s, lq, lr, -2047, 2047);
}

bool half_enabled() {
    const char *e = std::getenv("GSW_HALF");
    return !(e && e[0] == '0');
}

int pair_kind(const SchemeDev &s, int dtype_policy, bool ok16, bool okh, bool half) {
    if (dtype_policy == GSW_DTYPE_INT32) return KIND_I32;
    const bool pk16 = dtype_policy == GSW_DTYPE_PK16;
    if (dtype_policy == GSW_DTYPE_SAFE16) half = false;
    if (s.mode == GSW_MODE_LOCAL && dtype_policy != GSW_DTYPE_SAFE && dtype_policy != GSW_DTYPE_SAFE16) {
        // int16 allowed: by the envelope, or everywhere when the caller vouched for it (its
        // saturated results are listed); half2 wherever int16 is, results >= 2048 listed
        const bool any16 = (dtype_policy == GSW_DTYPE_INT16 || pk16) ? int16_entries(s) : ok16;
        if (!pk16 && half && half_entries(s) && (any16 || okh)) return KIND_F16;
        return any16 ? KIND_I16 : KIND_I32;
    }
    if (!pk16 && half && okh) return KIND_F16;
    return ok16 ? KIND_I16 : KIND_I32;
}

bool fits_int32(const SchemeDev &s, int lq, int lr) {
    int64_t lo, hi;
    dp_bounds(s, lq, lr, &lo, &hi);
    return lo >= -(INT64_C(1) << 29) && hi <= (INT64_C(1) << 29);
}

int32_t empty_edge_score(const SchemeDev &s, int lq, int lr) {
    // src/gpusw/kernel.py:63-80
    if (s.mode == GSW_MODE_LOCAL || (lq == 0 && lr == 0)) return 0;
    bool semi = s.mode == GSW_MODE_SEMIGLOBAL;
    if (lq == 0) return (semi && s.free_q) ? 0 : (int32_t)(s.go + (int64_t)(lr - 1) * s.ge);
    return (semi && s.free_r) ? 0 : (int32_t)(s.go + (int64_t)(lq - 1) * s.ge);
}

// Largest byte of codes[0, n) into *out (which starts at zero).
__global__ void max_code_kernel(const uint8_t *codes, int64_t n, int32_t *out) {
    int m = 0;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        m = max(m, (int)codes[i]);
    m = wave_max(m);
    if ((threadIdx.x & 63) == 0 && m > 0) atomicMax(out, m);
}

}  // namespace gsw

using namespace gsw;

// ====================================================================== C ABI
extern "C" {

const char *gsw_last_error(void) { return g_error.c_str(); }

const char *gsw_version(void) { return "gsw_hip 0.2.0+hip.1 (gfx950)"; }

int gsw_device_count(int32_t *count) {
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess) n = 0;
    if (count) *count = n;
    return GSW_OK;
}

int gsw_set_device(int32_t device) {
    int count = 0;
    GSW_TRY(init_runtime(&count));
    if (device < 0 || device >= count)
        return fail(GSW_ERR_INVALID, "device %d out of range (have %d)", device, count);
    GSW_HIP_TRY(hipSetDevice(device));
    return GSW_OK;
}

int gsw_get_device(int32_t *device) {
    int count = 0;
    GSW_TRY(init_runtime(&count));
    int d = 0;
    GSW_HIP_TRY(hipGetDevice(&d));
    *device = d;
    return GSW_OK;
}

int gsw_device_name(int32_t device, char *buf, int32_t buflen) {
    int count = 0;
    GSW_TRY(init_runtime(&count));
    if (device < 0 || device >= count) return fail(GSW_ERR_INVALID, "device %d out of range", device);
    if (!buf || buflen <= 0) return fail(GSW_ERR_INVALID, "null name buffer");
    hipDeviceProp_t prop;
    GSW_HIP_TRY(hipGetDeviceProperties(&prop, device));
    snprintf(buf, buflen, "%s (%s, %d CUs, %d MHz, LDS/block %zu KiB, %zu GiB)",
             prop.name[0] ? prop.name : "AMD GPU", prop.gcnArchName,
             prop.multiProcessorCount, prop.clockRate / 1000, prop.sharedMemPerBlock / 1024,
             prop.totalGlobalMem >> 30);
    return GSW_OK;
}

int gsw_malloc(void **dptr, size_t bytes) {
    int count = 0;
    GSW_TRY(init_runtime(&count));
    *dptr = nullptr;
    if (bytes == 0) bytes = 16;
    GSW_HIP_TRY(hipMalloc(dptr, bytes));
    return GSW_OK;
}

int gsw_free(void *dptr) {
    if (dptr) GSW_HIP_TRY(hipFree(dptr));
    return GSW_OK;
}

int gsw_memcpy_h2d(void *dst, const void *src, size_t bytes) {
    if (bytes) GSW_HIP_TRY(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice));
    return GSW_OK;
}

int gsw_memcpy_d2h(void *dst, const void *src, size_t bytes) {
    if (bytes) GSW_HIP_TRY(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost));
    return GSW_OK;
}

int gsw_memset(void *dptr, int32_t value, size_t bytes) {
    if (bytes) GSW_HIP_TRY(hipMemset(dptr, value, bytes));
    return GSW_OK;
}

int gsw_memcpy_d2d(void *dst, const void *src, size_t bytes) {
    if (bytes) GSW_HIP_TRY(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToDevice));
    return GSW_OK;
}

int gsw_synchronize(void) {
    GSW_HIP_TRY(hipDeviceSynchronize());
    return GSW_OK;
}

int gsw_seqset_create(const uint8_t *d_codes, const int32_t *d_off, const int32_t *h_off,
                      int32_t n, gsw_seqset *out) {
    if (!out) return fail(GSW_ERR_INVALID, "out is null");
    *out = nullptr;
    if (n < 0) return fail(GSW_ERR_INVALID, "negative sequence count");
    if (n > 0 && (!h_off || !d_off)) return fail(GSW_ERR_INVALID, "offsets are null");
    DeviceCtx *ctx = nullptr;
    GSW_TRY(ctx_get(&ctx));
    auto set = std::make_unique<SeqSet>();
    set->d_codes = d_codes;
    set->d_off = d_off;
    set->n = n;
    set->device = ctx->device;
    set->h_off.assign(h_off, h_off + (n + 1));
    set->len.resize(n);
    for (int i = 0; i < n; ++i) {
        int l = h_off[i + 1] - h_off[i];
        if (l < 0 || h_off[i] < 0) return fail(GSW_ERR_INVALID, "offsets must be non-decreasing");
        set->len[i] = l;
        set->max_len = std::max(set->max_len, l);
        set->total += l;
    }
    set->order_desc.resize(n);
    std::iota(set->order_desc.begin(), set->order_desc.end(), 0);
    std::stable_sort(set->order_desc.begin(), set->order_desc.end(),
                     [&](int a, int b) { return set->len[a] > set->len[b]; });
    set->rank.resize(n);
    for (int i = 0; i < n; ++i) set->rank[set->order_desc[i]] = i;
    set->n_nonempty = 0;
    while (set->n_nonempty < n && set->len[set->order_desc[set->n_nonempty]] > 0) set->n_nonempty++;
    // the largest code: explicit-pairs profiles need rows only up to it (codes at or past the
    // table side read its last row)
    if (set->total > 0) {
        GSW_HIP_TRY(hipMemsetAsync(ctx->counters + (NCOUNTERS - 1), 0, sizeof(int32_t), ctx->stream));
        const int64_t first = h_off[0];
        hipLaunchKernelGGL(max_code_kernel, dim3((unsigned)std::min<int64_t>(1024, (set->total + 255) / 256)),
                           dim3(256), 0, ctx->stream, d_codes + first, set->total, ctx->counters + (NCOUNTERS - 1));
        GSW_HIP_TRY(hipGetLastError());
        int32_t mc = 0;
        GSW_HIP_TRY(hipMemcpyAsync(&mc, ctx->counters + (NCOUNTERS - 1), sizeof(int32_t), hipMemcpyDeviceToHost, ctx->stream));
        GSW_HIP_TRY(hipStreamSynchronize(ctx->stream));
        set->max_code = mc;
    } else {
        set->max_code = 0;
    }
    set->bpos.assign((size_t)n + 1, 0);
    for (int p = 0; p < n; ++p) set->bpos[p + 1] = set->bpos[p] + set->len[set->order_desc[p]];
    if (n > 0) {
        GSW_HIP_TRY(hipMalloc(&set->d_order_desc, sizeof(int32_t) * n));
        GSW_HIP_TRY(hipMemcpy(set->d_order_desc, set->order_desc.data(), sizeof(int32_t) * n,
                              hipMemcpyHostToDevice));
    }
    GSW_HIP_TRY(hipMalloc(&set->d_bpos, sizeof(int64_t) * (n + 1)));
    GSW_HIP_TRY(hipMemcpy(set->d_bpos, set->bpos.data(), sizeof(int64_t) * (n + 1), hipMemcpyHostToDevice));
    *out = reinterpret_cast<gsw_seqset>(set.release());
    return GSW_OK;
}

int gsw_seqset_destroy(gsw_seqset set) {
    SeqSet *s = reinterpret_cast<SeqSet *>(set);
    if (!s) return GSW_OK;
    delete s;
    return GSW_OK;
}

}  // extern "C"
