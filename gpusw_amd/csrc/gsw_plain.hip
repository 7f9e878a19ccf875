// The plain comparator: one thread scores one (query, reference) pair.
//
// This is the MI355X form of the reference's one-thread-per-pair kernel
// (src/gpusw/kernel.py:56-190): same recurrence, boundaries, empty-sequence edges and
// mode results, DP values in 32-bit registers, DP rows stored as int16 or int32.
// Instead of per-thread local arrays capped at MAXQ, each thread keeps its previous
// DP row (H and F) in a global workspace laid out column-major across threads
// (element j of thread t at ws[j*stride + t]) so that the 64 lanes of a wave touch
// consecutive addresses; a grid-stride loop bounds the workspace.  The substitution
// table is staged in LDS.
#include <algorithm>
#include <chrono>

#include "gsw_internal.h"

namespace gsw {

struct PlainArgs {
    const uint8_t *qcodes;
    const int32_t *qoff;
    const uint8_t *rcodes;
    const int32_t *roff;
    // cross mode (nr > 0): pair k -> (qlist[k / nr], k % nr), out[outrow[k/nr]*ldo + k%nr]
    const int32_t *qlist;
    const int32_t *outrow;
    int32_t nr;
    int64_t ldo;
    // pairs mode (nr == 0): pair k -> (qi[k], rj[k]), out[outidx ? outidx[k] : k]
    const int32_t *qi;
    const int32_t *rj;
    const int64_t *outidx;
    const int32_t *qstart;   // descriptor mode: query start/length per pair instead of qi
    const int32_t *qlen;
    int64_t npairs;
    const int32_t *sub;
    int32_t ns, go, ge, free_q, free_r;
    void *ws;
    int64_t stride;  // total threads in the grid
    int32_t *out;
};

template <typename ST> struct StoreT;
template <> struct StoreT<int16_t> {
    static constexpr int NEG = -32768;
    __device__ static int16_t put(int v) { return (int16_t)min(max(v, -32768), 32767); }
};
template <> struct StoreT<int32_t> {
    static constexpr int NEG = -(1 << 30);   // the oracle's sentinel (reference.py:27)
    __device__ static int32_t put(int v) { return v; }
};

template <int MODE, typename ST>
__device__ int plain_score(const uint8_t *__restrict__ q, int lq, const uint8_t *__restrict__ r,
                           int lr, const int32_t *__restrict__ sub, const PlainArgs &a,
                           ST *__restrict__ hp, ST *__restrict__ fp, int64_t stride) {
    using S = StoreT<ST>;
    const int go = a.go, ge = a.ge, ns = a.ns;
    const bool free_q = a.free_q, free_r = a.free_r;
    // empty-sequence edges (kernel.py:63-80)
    if (lq == 0 || lr == 0) {
        if (MODE == GSW_MODE_LOCAL || (lq == 0 && lr == 0)) return 0;
        if (lq == 0) return (MODE == GSW_MODE_SEMIGLOBAL && free_q) ? 0 : go + (lr - 1) * ge;
        return (MODE == GSW_MODE_SEMIGLOBAL && free_r) ? 0 : go + (lq - 1) * ge;
    }
    hp[0] = 0;
    for (int j = 1; j <= lq; ++j) {          // row 0 (kernel.py:85-97)
        int v = 0;
        if (MODE == GSW_MODE_GLOBAL || (MODE == GSW_MODE_SEMIGLOBAL && !free_r)) v = go + (j - 1) * ge;
        hp[(int64_t)j * stride] = S::put(v);
        fp[(int64_t)j * stride] = S::put(S::NEG);
    }
    int best = 0, corner = S::NEG, endcol = S::NEG, endrow = S::NEG;
    for (int i = 1; i <= lr; ++i) {          // rows = reference (kernel.py:102-139)
        const int32_t *srow = sub + min((int)r[i - 1], ns - 1) * ns;   // codes >= ns: sentinel
        int hleft = 0;
        if (MODE == GSW_MODE_GLOBAL || (MODE == GSW_MODE_SEMIGLOBAL && !free_q)) hleft = go + (i - 1) * ge;
        int diag = hp[0];
        hp[0] = S::put(hleft);
        int eleft = S::NEG;
        for (int j = 1; j <= lq; ++j) {
            const int64_t o = (int64_t)j * stride;
            int hup = hp[o];
            int fj = max(hup + go, (int)fp[o] + ge);
            int ej = max(hleft + go, eleft + ge);
            int h = max(max(diag + srow[min((int)q[j - 1], ns - 1)], ej), fj);
            if (MODE == GSW_MODE_LOCAL) {
                h = max(h, 0);
                best = max(best, h);
            } else {
                if (i == lr && j == lq) corner = h;
                if (MODE == GSW_MODE_SEMIGLOBAL) {
                    if (free_q && j == lq) endcol = max(endcol, h);
                    if (free_r && i == lr) endrow = max(endrow, h);
                }
            }
            diag = hup;
            hp[o] = S::put(h);
            fp[o] = S::put(fj);
            hleft = h;
            eleft = ej;
        }
    }
    if (MODE == GSW_MODE_LOCAL) return best;     // kernel.py:141-157
    if (MODE == GSW_MODE_GLOBAL) return corner;
    int res = corner;
    if (free_q) res = max(res, endcol);
    if (free_r) res = max(res, endrow);
    if (free_q && free_r) res = max(res, 0);
    return res;
}

template <int MODE, typename ST>
__global__ void __launch_bounds__(256) plain_kernel_body(PlainArgs a, int64_t total, int32_t maxq) {
    __shared__ int32_t s_sub[64 * 64];
    for (int i = threadIdx.x; i < a.ns * a.ns; i += blockDim.x) s_sub[i] = a.sub[i];
    __syncthreads();
    const int64_t tid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (tid >= a.stride) return;
    ST *hp = reinterpret_cast<ST *>(a.ws) + tid;
    ST *fp = hp + (int64_t)(maxq + 1) * a.stride;
    for (int64_t k = tid; k < total; k += a.stride) {
        int qi, rj;
        int64_t o;
        if (a.nr > 0) {
            const int64_t row = k / a.nr;
            rj = (int)(k - row * a.nr);
            qi = a.qlist[row];
            o = (int64_t)a.outrow[row] * a.ldo + rj;
        } else {
            qi = a.qstart ? 0 : a.qi[k];
            rj = a.rj[k];
            o = a.outidx ? a.outidx[k] : k;
        }
        const int qs = a.qstart ? a.qstart[k] : a.qoff[qi];
        const int rs = a.roff[rj];
        // the workspace holds maxq+1 columns; longer queries are truncated as the
        // reference truncates at MAXQ (kernel.py:61) -- callers pass the true maximum
        const int lq = min(a.qstart ? a.qlen[k] : a.qoff[qi + 1] - qs, maxq);
        a.out[o] = plain_score<MODE, ST>(a.qcodes + qs, lq, a.rcodes + rs,
                                         a.roff[rj + 1] - rs, s_sub, a, hp, fp, a.stride);
    }
}

template <int MODE, typename ST>
static void launch_plain(const PlainArgs &a, int64_t total, int maxq, int blocks, int threads,
                         hipStream_t st) {
    hipLaunchKernelGGL((plain_kernel_body<MODE, ST>), dim3(blocks), dim3(threads), 0, st, a, total,
                       maxq);
}

int plain_run(DeviceCtx *ctx, const SchemeDev &s, const PlainJob &job, float *ms) {
    if (ms) *ms = 0.f;
    const bool cross = job.nr > 0 && job.h_qlist;
    const int64_t total = cross ? (int64_t)job.nq * job.nr : job.npairs;
    if (total <= 0) return GSW_OK;
    int threads = job.threads > 0 ? job.threads : 128;
    threads = std::min(256, std::max(64, (threads + 63) / 64 * 64));
    const int maxq = std::max(job.maxq, 1);
    const size_t st_bytes = job.dtype == GSW_DTYPE_INT16 ? 2 : 4;
    const size_t per_thread = 2 * (size_t)(maxq + 1) * st_bytes;
    // grid: at most ~1024 resident threads per CU, and a workspace of <= 4 GiB
    int64_t cap = (int64_t)ctx->num_cus * 1024;
    cap = std::min<int64_t>(cap, (int64_t)((size_t(4) << 30) / per_thread));
    cap = std::max<int64_t>(cap, threads);
    int64_t nthr = std::min<int64_t>(total, cap);
    nthr = (nthr + threads - 1) / threads * threads;
    const int blocks = (int)(nthr / threads);
    GSW_TRY(ensure_buffer(&ctx->plain_ws, &ctx->plain_ws_bytes, per_thread * nthr));

    // stage the index lists in device scratch
    size_t need = 0;
    if (cross) need = sizeof(int32_t) * 2 * (size_t)job.nq;
    else need = sizeof(int32_t) * 4 * (size_t)job.npairs + (job.h_outidx ? sizeof(int64_t) * job.npairs : 0);
    GSW_TRY(ensure_buffer(&ctx->meta, &ctx->meta_bytes, need + 256));
    PlainArgs a{};
    a.qcodes = job.qcodes;
    a.qoff = job.qoff;
    a.rcodes = job.rcodes;
    a.roff = job.roff;
    a.sub = s.d_sub;
    a.ns = s.ns;
    a.go = s.go;
    a.ge = s.ge;
    a.free_q = s.free_q;
    a.free_r = s.free_r;
    a.ws = ctx->plain_ws;
    a.stride = nthr;
    a.out = job.d_out;
    char *m = static_cast<char *>(ctx->meta);
    if (cross) {
        int32_t *d_ql = reinterpret_cast<int32_t *>(m);
        int32_t *d_or = d_ql + job.nq;
        GSW_HIP_TRY(hipMemcpyAsync(d_ql, job.h_qlist, sizeof(int32_t) * job.nq, hipMemcpyHostToDevice, ctx->stream));
        if (job.h_outrow) {
            GSW_HIP_TRY(hipMemcpyAsync(d_or, job.h_outrow, sizeof(int32_t) * job.nq, hipMemcpyHostToDevice, ctx->stream));
        } else {
            std::vector<int32_t> ident(job.nq);
            for (int i = 0; i < job.nq; ++i) ident[i] = i;
            GSW_HIP_TRY(hipMemcpyAsync(d_or, ident.data(), sizeof(int32_t) * job.nq, hipMemcpyHostToDevice, ctx->stream));
            GSW_HIP_TRY(hipStreamSynchronize(ctx->stream));
        }
        a.qlist = d_ql;
        a.outrow = d_or;
        a.nr = job.nr;
        a.ldo = job.ldo > 0 ? job.ldo : job.nr;
    } else {
        int32_t *d_qi = reinterpret_cast<int32_t *>(m);
        int32_t *d_rj = d_qi + job.npairs;
        int32_t *d_qs = d_rj + job.npairs;
        int32_t *d_ql = d_qs + job.npairs;
        int64_t *d_oi = reinterpret_cast<int64_t *>(
            (reinterpret_cast<uintptr_t>(d_ql + job.npairs) + 15) & ~uintptr_t(15));
        if (job.h_qstart) {
            GSW_HIP_TRY(hipMemcpyAsync(d_qs, job.h_qstart, sizeof(int32_t) * job.npairs, hipMemcpyHostToDevice, ctx->stream));
            GSW_HIP_TRY(hipMemcpyAsync(d_ql, job.h_qlen, sizeof(int32_t) * job.npairs, hipMemcpyHostToDevice, ctx->stream));
            a.qstart = d_qs;
            a.qlen = d_ql;
        } else {
            GSW_HIP_TRY(hipMemcpyAsync(d_qi, job.h_qi, sizeof(int32_t) * job.npairs, hipMemcpyHostToDevice, ctx->stream));
        }
        GSW_HIP_TRY(hipMemcpyAsync(d_rj, job.h_rj, sizeof(int32_t) * job.npairs, hipMemcpyHostToDevice, ctx->stream));
        if (job.h_outidx)
            GSW_HIP_TRY(hipMemcpyAsync(d_oi, job.h_outidx, sizeof(int64_t) * job.npairs, hipMemcpyHostToDevice, ctx->stream));
        a.qi = d_qi;
        a.rj = d_rj;
        a.outidx = job.h_outidx ? d_oi : nullptr;
        a.npairs = job.npairs;
        a.nr = 0;
    }
    GSW_HIP_TRY(hipEventRecord(ctx->ev0, ctx->stream));
    const bool i16 = job.dtype == GSW_DTYPE_INT16;
    switch (s.mode) {
    case GSW_MODE_LOCAL:
        i16 ? launch_plain<GSW_MODE_LOCAL, int16_t>(a, total, maxq, blocks, threads, ctx->stream)
            : launch_plain<GSW_MODE_LOCAL, int32_t>(a, total, maxq, blocks, threads, ctx->stream);
        break;
    case GSW_MODE_GLOBAL:
        i16 ? launch_plain<GSW_MODE_GLOBAL, int16_t>(a, total, maxq, blocks, threads, ctx->stream)
            : launch_plain<GSW_MODE_GLOBAL, int32_t>(a, total, maxq, blocks, threads, ctx->stream);
        break;
    default:
        i16 ? launch_plain<GSW_MODE_SEMIGLOBAL, int16_t>(a, total, maxq, blocks, threads, ctx->stream)
            : launch_plain<GSW_MODE_SEMIGLOBAL, int32_t>(a, total, maxq, blocks, threads, ctx->stream);
        break;
    }
    GSW_HIP_TRY(hipGetLastError());
    GSW_HIP_TRY(hipEventRecord(ctx->ev1, ctx->stream));
    GSW_HIP_TRY(hipEventSynchronize(ctx->ev1));
    float t = 0.f;
    GSW_HIP_TRY(hipEventElapsedTime(&t, ctx->ev0, ctx->ev1));
    if (ms) *ms = t;
    return GSW_OK;
}

}  // namespace gsw

using namespace gsw;

static int plain_common(const gsw_scheme *s, int32_t dtype, DeviceCtx **ctx, SchemeDev *sd) {
    if (dtype != GSW_DTYPE_INT16 && dtype != GSW_DTYPE_INT32)
        return fail(GSW_ERR_INVALID, "plain kernels take dtype int16 or int32, got %d", dtype);
    GSW_TRY(ctx_get(ctx));
    GSW_TRY(scheme_prepare(s, *ctx, sd));
    return GSW_OK;
}

extern "C" int gsw_score_cross_plain(const gsw_scheme *s, int32_t dtype, int32_t threads,
                                     const uint8_t *d_qcodes, const int32_t *d_qoff,
                                     const uint8_t *d_rcodes, const int32_t *d_roff,
                                     const int32_t *h_qlist, int32_t nq, int32_t nr, int32_t maxq,
                                     int32_t *d_out, gsw_stats *stats) {
    auto t0 = std::chrono::steady_clock::now();
    DeviceCtx *ctx = nullptr;
    SchemeDev sd;
    GSW_TRY(plain_common(s, dtype, &ctx, &sd));
    if (nq < 0 || nr < 0) return fail(GSW_ERR_INVALID, "negative sizes");
    std::lock_guard<std::mutex> lk(ctx->mu);
    PlainJob job{};
    job.qcodes = d_qcodes; job.qoff = d_qoff; job.rcodes = d_rcodes; job.roff = d_roff;
    job.dtype = dtype; job.threads = threads; job.maxq = maxq;
    job.h_qlist = h_qlist; job.nq = nq; job.nr = nr; job.ldo = nr; job.d_out = d_out;
    float ms = 0.f;
    if (nq > 0 && nr > 0) GSW_TRY(plain_run(ctx, sd, job, &ms));
    if (stats) {
        *stats = gsw_stats{};
        stats->kernel_ms = ms;
        stats->pairs = (int64_t)nq * nr;
        stats->plain_pairs = stats->pairs;
        stats->launches = (nq > 0 && nr > 0) ? 1 : 0;
        stats->variant = -1;
        stats->wall_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    }
    return GSW_OK;
}

extern "C" int gsw_score_pairs_plain(const gsw_scheme *s, int32_t dtype, int32_t threads,
                                     const uint8_t *d_qcodes, const int32_t *d_qoff,
                                     const uint8_t *d_rcodes, const int32_t *d_roff,
                                     const int32_t *h_qi, const int32_t *h_rj, int64_t n, int32_t maxq,
                                     int32_t *d_out, gsw_stats *stats) {
    auto t0 = std::chrono::steady_clock::now();
    DeviceCtx *ctx = nullptr;
    SchemeDev sd;
    GSW_TRY(plain_common(s, dtype, &ctx, &sd));
    if (n < 0) return fail(GSW_ERR_INVALID, "negative pair count");
    std::lock_guard<std::mutex> lk(ctx->mu);
    PlainJob job{};
    job.qcodes = d_qcodes; job.qoff = d_qoff; job.rcodes = d_rcodes; job.roff = d_roff;
    job.dtype = dtype; job.threads = threads; job.maxq = maxq;
    job.h_qi = h_qi; job.h_rj = h_rj; job.npairs = n; job.d_out = d_out;
    float ms = 0.f;
    if (n > 0) GSW_TRY(plain_run(ctx, sd, job, &ms));
    if (stats) {
        *stats = gsw_stats{};
        stats->kernel_ms = ms;
        stats->pairs = n;
        stats->plain_pairs = n;
        stats->launches = n > 0 ? 1 : 0;
        stats->variant = -1;
        stats->wall_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    }
    return GSW_OK;
}
