// Explicit pairs without device allocation: gsw_score_pairs_noalloc.
//
// The same batched table kernels as gsw_score_pairs (pairs_tab_kernel), for callers that
// cannot allocate device memory.  Everything a call needs beyond its inputs and outputs
// -- the task list, the substitution table, the work counters, the saturation list and the
// closed-form scores of empty pairs -- lives in `scratch`, a file-scope device array of
// SCRATCH_BYTES.  The array's device address is looked up once per device with
// hipGetSymbolAddress (the name is only a host-side stand-in for it) and kept per device.
// A call plans on the host, copies its tasks into the array, launches one persistent
// kernel per tile on the null stream and waits for them; int16 results that saturate
// (local mode under GSW_DTYPE_INT16) are rescored on the int32 tiles by a second pass
// through the same array.  No hipMalloc, hipFree, stream or event is created by a call.
#include <algorithm>
#include <chrono>
#include <map>
#include <memory>
#include <mutex>
#include <vector>

#include "gsw_batch.h"
#include "gsw_internal.h"

namespace gsw {
namespace {

constexpr int64_t MAX_PAIRS = int64_t(1) << 18;          // pairs per call
constexpr int64_t MAX_TASKS = MAX_PAIRS + 65536;          // + block padding of every tile
constexpr int SAT_SLOT = 255;                             // header word of the saturation count
// layout of the array (byte offsets, each 256-aligned)
constexpr size_t HDR_AT = 0, HDR_BYTES = 256 * sizeof(int32_t);   // per-tile work counters, sat count
constexpr size_t SUB_AT = HDR_AT + HDR_BYTES, SUB_BYTES = 64 * 64 * sizeof(int32_t);
constexpr size_t SAT_AT = SUB_AT + SUB_BYTES, SAT_BYTES = sizeof(SatEntry) * MAX_PAIRS;
constexpr size_t EIDX_AT = SAT_AT + SAT_BYTES, EIDX_BYTES = sizeof(int64_t) * MAX_PAIRS;
constexpr size_t EVAL_AT = EIDX_AT + EIDX_BYTES, EVAL_BYTES = sizeof(int32_t) * MAX_PAIRS;
constexpr size_t TASK_AT = EVAL_AT + EVAL_BYTES, TASK_BYTES = sizeof(PairTask) * MAX_TASKS;
constexpr size_t SCRATCH_BYTES = TASK_AT + TASK_BYTES;   // about 18 MiB
static_assert(SCRATCH_BYTES % 256 == 0 && TASK_AT % 256 == 0 && EIDX_AT % 256 == 0, "layout alignment");

}  // namespace

static __device__ __align__(256) char scratch[SCRATCH_BYTES];

namespace {

// Closed-form scores of pairs with an empty side: out[idx[i]] = val[i].
__global__ void edge_kernel(int32_t *out, const int64_t *idx, const int32_t *val, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[idx[i]] = val[i];
}

// Per device: the array's address (hipGetSymbolAddress with the device current) and the
// device's launch limits.  Calls on one device are serialised by its lock: they share the
// array, which keeps the last call's contents.
struct NoAllocDevice {
    char *ptr = nullptr;
    DeviceCtx info;   // device, CU count and LDS limit only (no streams, nothing allocated)
    std::mutex mu;
};

std::mutex g_dev_mu;
std::map<int, std::unique_ptr<NoAllocDevice>> g_dev;

int noalloc_device(NoAllocDevice **out) {
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count <= 0)
        return fail(GSW_ERR_NO_DEVICE, "no AMD GPU visible to the HIP runtime");
    int dev = 0;
    GSW_HIP_TRY(hipGetDevice(&dev));
    std::lock_guard<std::mutex> lk(g_dev_mu);
    auto it = g_dev.find(dev);
    if (it != g_dev.end()) {
        *out = it->second.get();
        return GSW_OK;
    }
    auto d = std::make_unique<NoAllocDevice>();
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
        return fail(GSW_ERR_HIP, "hipGetSymbolAddress(scratch) on device %d: %s", dev, hipGetErrorString(e));
    d->ptr = static_cast<char *>(p);
    d->info.device = dev;
    GSW_HIP_TRY(hipDeviceGetAttribute(&d->info.num_cus, hipDeviceAttributeMultiprocessorCount, dev));
    GSW_HIP_TRY(hipDeviceGetAttribute(&d->info.max_smem, hipDeviceAttributeMaxSharedMemoryPerBlock, dev));
    *out = d.get();
    g_dev[dev] = std::move(d);
    return GSW_OK;
}

struct Pass {
    int64_t launches = 0;
    int64_t cells = 0;
};

// Score pairs (qi[i], rj[i]) into d_out[oi[i]] under a dtype policy, through the array.
int run_pass(NoAllocDevice &dev, const SchemeDev &s, int dtype_policy, const SeqSet &qs, const SeqSet &rs,
             const int32_t *qi, const int32_t *rj, const int64_t *oi, int64_t n, int32_t *d_out, Pass *pass) {
    const bool local = s.mode == GSW_MODE_LOCAL;
    const bool track_last_col = s.mode == GSW_MODE_SEMIGLOBAL && s.free_q;
    const bool local16 = dtype_policy == GSW_DTYPE_INT16 && local && int16_entries(s);
    const double mean_lr = rs.n_nonempty ? (double)rs.total / rs.n_nonempty : 1.0;
    const int wmax = max_width();
    const std::vector<Variant> &vt = variant_table();
    // -- classify: closed form, or a table-kernel tile (by variant id)
    static thread_local std::vector<int64_t> e_idx;
    static thread_local std::vector<int32_t> e_val;
    static thread_local std::vector<std::vector<int64_t>> members;
    e_idx.clear();
    e_val.clear();
    members.resize(vt.size());
    for (auto &m : members) m.clear();
    for (int64_t i = 0; i < n; ++i) {
        const int q = qi[i], r = rj[i];
        if (q < 0 || q >= qs.n || r < 0 || r >= rs.n) return fail(GSW_ERR_INVALID, "pair %lld index out of range", (long long)i);
        const int lq = qs.len[q], lr = rs.len[r];
        pass->cells += (int64_t)lq * lr;
        if (lq == 0 || lr == 0) {
            e_idx.push_back(oi[i]);
            e_val.push_back(empty_edge_score(s, lq, lr));
            continue;
        }
        if (!fits_int32(s, lq, lr))
            return fail(GSW_ERR_OVERFLOW, "score range of pair %lld exceeds the int32 kernels", (long long)i);
        if (lq > wmax)
            return fail(GSW_ERR_LIMIT, "pair %lld: query of %d columns; this entry point scores queries of up to %d",
                        (long long)i, lq, wmax);
        const bool use16 = dtype_policy == GSW_DTYPE_INT32 ? false : local16 ? true : sweep_fits_int16(s, lq, lr);
        bool tab = false;
        const Variant *v = pick_pairs_tile(lq, mean_lr, use16 ? KIND_I16 : KIND_I32, local, s.ns, 0,
                                           (size_t)dev.info.max_smem, &tab);
        if (!v || !tab) return fail(GSW_ERR_LIMIT, "pair %lld has no table-kernel tile", (long long)i);
        members[v->id].push_back(i);
    }
    // -- tasks per tile: sorted by reference length rank (then front padding), adjacent
    //    pairs of one reference packed into one task, blocks of 64/G tasks padded where the
    //    front padding changes
    static thread_local std::vector<PairTask> tasks;
    tasks.clear();
    struct Launch { const Variant *v; int64_t t0, t1; };
    std::vector<Launch> launches;
    for (const Variant &v : vt) {
        std::vector<int64_t> &mem = members[v.id];
        if (mem.empty()) continue;
        const int gpw = 64 / v.G;
        auto shift_of = [&](int64_t i) { return track_last_col ? (v.R - qs.len[qi[i]] % v.R) % v.R : 0; };
        std::stable_sort(mem.begin(), mem.end(), [&](int64_t a, int64_t b) {
            const int sa = shift_of(a), sb = shift_of(b);
            if (sa != sb) return sa < sb;
            return rs.rank[rj[a]] < rs.rank[rj[b]];
        });
        const int64_t t0 = (int64_t)tasks.size();
        int cur = -1;
        for (size_t i = 0; i < mem.size();) {
            const int64_t x = mem[i];
            const int sx = shift_of(x);
            if (sx != cur) {
                while (((int64_t)tasks.size() - t0) % gpw) tasks.push_back(PairTask{-1, -1, -1, cur, -1, -1});
                cur = sx;
            }
            if (v.packed && i + 1 < mem.size() && rj[mem[i + 1]] == rj[x] && shift_of(mem[i + 1]) == sx) {
                const int64_t y = mem[i + 1];
                tasks.push_back(PairTask{qi[x], qi[y], rj[x], sx, oi[x], oi[y]});
                i += 2;
            } else {
                tasks.push_back(PairTask{qi[x], -1, rj[x], sx, oi[x], -1});
                i += 1;
            }
        }
        while (((int64_t)tasks.size() - t0) % gpw) tasks.push_back(PairTask{-1, -1, -1, cur, -1, -1});
        launches.push_back(Launch{&v, t0, (int64_t)tasks.size()});
    }
    if ((int64_t)tasks.size() > MAX_TASKS) return fail(GSW_ERR_LIMIT, "task list exceeds the scratch array");
    if ((int)launches.size() > SAT_SLOT)   // header words [0, SAT_SLOT) are the launches' work counters
        return fail(GSW_ERR_LIMIT, "%zu tiles exceed the scratch array's %d work counters", launches.size(), SAT_SLOT);
    // -- inputs into the array (its counters start at zero), then the launches
    static const int32_t zeros[HDR_BYTES / sizeof(int32_t)] = {};
    char *base = dev.ptr;
    const hipStream_t st = nullptr;   // the null stream: nothing to create
    GSW_HIP_TRY(hipMemcpyAsync(base + HDR_AT, zeros, HDR_BYTES, hipMemcpyHostToDevice, st));
    GSW_HIP_TRY(hipMemcpyAsync(base + SUB_AT, s.h_sub.data(), sizeof(int32_t) * s.h_sub.size(),
                               hipMemcpyHostToDevice, st));
    if (!tasks.empty())
        GSW_HIP_TRY(hipMemcpyAsync(base + TASK_AT, tasks.data(), sizeof(PairTask) * tasks.size(),
                                   hipMemcpyHostToDevice, st));
    if (!e_idx.empty()) {
        GSW_HIP_TRY(hipMemcpyAsync(base + EIDX_AT, e_idx.data(), sizeof(int64_t) * e_idx.size(),
                                   hipMemcpyHostToDevice, st));
        GSW_HIP_TRY(hipMemcpyAsync(base + EVAL_AT, e_val.data(), sizeof(int32_t) * e_val.size(),
                                   hipMemcpyHostToDevice, st));
    }
    int32_t *hdr = reinterpret_cast<int32_t *>(base + HDR_AT);
    for (size_t li = 0; li < launches.size(); ++li) {
        const Launch &L = launches[li];
        const Variant &v = *L.v;
        KArgs a{};
        a.qcodes = qs.d_codes;
        a.qoff = qs.d_off;
        a.rcodes = rs.d_codes;
        a.roff = rs.d_off;
        a.rorder = rs.d_order_desc;
        a.sub = reinterpret_cast<const int32_t *>(base + SUB_AT);
        a.ns = s.ns;
        a.k = SweepConst{s.mode, s.go, s.ge, s.free_q, s.free_r, s.ns};
        a.out = d_out;
        a.sat = reinterpret_cast<SatEntry *>(base + SAT_AT);
        a.sat_count = hdr + SAT_SLOT;
        a.sat_cap = (int32_t)MAX_PAIRS;
        a.sat_check = (v.packed && local) ? 1 : 0;
        a.items = reinterpret_cast<const PairTask *>(base + TASK_AT) + L.t0;
        a.nitems = (int32_t)(L.t1 - L.t0);
        a.nitems_ptr = nullptr;
        a.counter = hdr + li;   // one zeroed work counter per launch
        a.tab_copies = pairs_tab_copies(v, s.ns);
        const size_t lds = pairs_tab_lds_bytes(v, s.ns);
        int grid = 0;
        GSW_TRY(occupancy_grid(&dev.info, v.pairs_tab_fn, PAIRS_TAB_WAVES * 64, lds, &grid));
        void *args[] = {&a};
        GSW_HIP_TRY(hipLaunchKernel(v.pairs_tab_fn, dim3(grid), dim3(PAIRS_TAB_WAVES * 64), args, lds, st));
        pass->launches++;
    }
    if (!e_idx.empty()) {
        const int threads = 256;
        hipLaunchKernelGGL(edge_kernel, dim3((unsigned)((e_idx.size() + threads - 1) / threads)), dim3(threads), 0,
                           st, d_out, reinterpret_cast<const int64_t *>(base + EIDX_AT),
                           reinterpret_cast<const int32_t *>(base + EVAL_AT), (int64_t)e_idx.size());
        GSW_HIP_TRY(hipGetLastError());
        pass->launches++;
    }
    GSW_HIP_TRY(hipStreamSynchronize(st));
    return GSW_OK;
}

}  // namespace
}  // namespace gsw

using namespace gsw;

extern "C" int gsw_score_pairs_noalloc(const gsw_scheme *s, int32_t dtype_policy, gsw_seqset queries,
                                       gsw_seqset refs, const int32_t *h_qi, const int32_t *h_rj, int64_t n,
                                       int32_t *d_out, gsw_stats *stats) {
    const auto t0 = std::chrono::steady_clock::now();
    if (n < 0) return fail(GSW_ERR_INVALID, "negative pair count");
    if (n > MAX_PAIRS)
        return fail(GSW_ERR_LIMIT, "gsw_score_pairs_noalloc scores at most %lld pairs per call (got %lld)",
                    (long long)MAX_PAIRS, (long long)n);
    if (dtype_policy < GSW_DTYPE_AUTO || dtype_policy > GSW_DTYPE_INT32)
        return fail(GSW_ERR_INVALID, "dtype_policy %d out of range", dtype_policy);
    const SeqSet *qs = reinterpret_cast<const SeqSet *>(queries);
    const SeqSet *rs = reinterpret_cast<const SeqSet *>(refs);
    if (!qs || !rs) return fail(GSW_ERR_INVALID, "null sequence set");
    if (n > 0 && (!h_qi || !h_rj || !d_out)) return fail(GSW_ERR_INVALID, "null pair list or output");
    SchemeDev sd;
    GSW_TRY(scheme_validate(s, &sd));
    NoAllocDevice *dev = nullptr;
    GSW_TRY(noalloc_device(&dev));
    if (qs->device != dev->info.device || rs->device != dev->info.device)
        return fail(GSW_ERR_INVALID, "sequence set belongs to another device");
    std::lock_guard<std::mutex> lk(dev->mu);
    static thread_local std::vector<int64_t> ident;
    if ((int64_t)ident.size() < n) {
        const int64_t k0 = (int64_t)ident.size();
        ident.resize((size_t)n);
        for (int64_t k = k0; k < n; ++k) ident[k] = k;
    }
    Pass pass;
    GSW_TRY(run_pass(*dev, sd, dtype_policy, *qs, *rs, h_qi, h_rj, ident.data(), n, d_out, &pass));
    const int64_t cells = pass.cells;
    // saturated int16 results (GSW_DTYPE_INT16, local mode): the list holds every pair of the
    // call, so it cannot overflow; its pairs run again on the int32 tiles
    int32_t nsat = 0;
    GSW_HIP_TRY(hipMemcpy(&nsat, dev->ptr + HDR_AT + SAT_SLOT * sizeof(int32_t), sizeof(int32_t),
                          hipMemcpyDeviceToHost));
    if (nsat > 0) {
        std::vector<SatEntry> sat((size_t)nsat);
        GSW_HIP_TRY(hipMemcpy(sat.data(), dev->ptr + SAT_AT, sizeof(SatEntry) * (size_t)nsat, hipMemcpyDeviceToHost));
        std::vector<int32_t> qi((size_t)nsat), rj((size_t)nsat);
        std::vector<int64_t> oi((size_t)nsat);
        for (int32_t i = 0; i < nsat; ++i) {
            qi[i] = sat[i].q;
            rj[i] = sat[i].r;
            oi[i] = sat[i].o;
        }
        Pass again;
        GSW_TRY(run_pass(*dev, sd, GSW_DTYPE_INT32, *qs, *rs, qi.data(), rj.data(), oi.data(), nsat, d_out, &again));
        pass.launches += again.launches;
    }
    if (stats) {
        *stats = gsw_stats{};
        stats->wall_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
        stats->kernel_ms = stats->wall_ms;   // no events: the call's wall time
        stats->cells = cells;
        stats->pairs = n;
        stats->rescored = nsat;
        stats->launches = (int32_t)pass.launches;
        stats->variant = -1;
        stats->top_variant = -1;
    }
    return GSW_OK;
}
