// Serving loop: resident reference databases, memory-budget admission and continuous
// batching of top-k requests into the persistent cross kernel.
//
// A worker thread owns a private stream and one stream per launch slot.  Submitted
// requests wait in FIFO order until
// their device memory (query codes + score rows + top-k output) fits in the server's
// arenas; admitted requests are planned into cross work items (query pair x chunk of
// length-sorted references).  Every launch takes items round-robin from all admitted
// requests of one (database, tile), up to a cell budget, so requests that arrive while
// others run join the next launch and share the device with them.  Two launches are kept
// in flight on the two slots' streams, so the second takes the CUs as the first one's last
// workgroups finish.  When a request's last item completes, a device top-k of its score rows
// (score descending, ties to the lower reference index) is copied back and its memory is
// released.  Local queries whose values fit int16 (src/gpusw/scheme.py:116-141, applied per
// query) run on the half2 tiles; a launch lists the results of 2048 or more (binary16 may have
// rounded them) in its slot, and when it retires the listed query rows are planned again on
// the int16 tiles (or int32, by the envelope), whose results the envelope proves exact, so a
// request's top-k waits for them.  Global/semiglobal queries run on half2 or int16 only where
// the envelope proves every value exact, otherwise on the int32 tiles.
#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <thread>

#include "gsw_batch.h"
#include "gsw_internal.h"

namespace gsw {

namespace {



struct ServerDb {
    SchemeDev s;
    const SeqSet *rs;
    double mean_lr;
    std::vector<int64_t> prefix;   // prefix sums of reference lengths in sorted order
};

struct Request {
    int64_t id = 0;
    int db = 0, nq = 0, k = 0, kk = 0;
    std::vector<uint8_t> codes;
    std::vector<int32_t> off;
    size_t codes_at = 0, codes_bytes = 0;     // in the code arena
    size_t scores_at = 0, scores_bytes = 0;   // in the score arena (scores + top-k output)
    std::vector<CrossItem> items;             // grouped by tile
    std::vector<int> item_tile;               // per item: 2 * variant id + 1 if striped
    size_t next = 0;
    int inflight = 0;
    std::vector<int32_t> idx, sc;
    int status = GSW_OK;
    std::string error;
    bool done = false;
    bool finishing = false;                   // its top-k is queued: the completer thread finishes it
    hipEvent_t done_ev = nullptr;             // after its top-k results reach the host
    bool claimed = false;                     // a gsw_server_wait is waiting for it
    double t_submit = 0, t_admit = 0, t_done = 0;
    int64_t cells = 0;
    int launches = 0;
};

struct Slot {
    hipStream_t st = nullptr;      // the slot's own stream (default priority): the two slots'
                                   // launches overlap, each one's workgroups taking CUs as the
                                   // other's finish
    hipEvent_t ev = nullptr;
    void *items = nullptr;         // SLOT_HDR bytes (the work counter first), then the items
    size_t items_cap = 0;
    char *h_items = nullptr;       // pinned staging of the same bytes
    size_t h_cap = 0;
    int db = -1;
    std::vector<std::pair<Request *, int>> reqs;
    std::vector<CrossItem> host_items;
    void *bnd = nullptr;           // stripe boundaries of this slot's launch
    size_t bnd_bytes = 0;
    SatEntry *sat = nullptr;       // results listed by this slot's launch (half2 >= 2048)
    SatEntry *h_sat = nullptr;     // pinned copy of the listed entries
    int32_t *h_hdr = nullptr;      // pinned copy of the launch's header (counters) once it ran
};

constexpr int NSLOTS = 2;
// A launch's work counter (word 0) and listed-result counter (word 2) travel in the header of
// its item copy: a memset would be a fill kernel, which cannot start before the other slot's
// persistent launch frees a CU.
constexpr size_t SLOT_HDR = 64;
constexpr int32_t SLOT_SAT_CAP = 1 << 16;   // listed results per launch (more: rerun its rows)

}  // namespace

struct Server {
    DeviceCtx ctx;                       // private stream (query-code uploads), events and scratch
    hipEvent_t codes_ready = nullptr;    // after the latest query-code upload on ctx.stream
    hipStream_t finish = nullptr;        // finished requests' top-k and result copies
    uint8_t *code_base = nullptr;
    uint8_t *score_base = nullptr;
    Arena code_arena, score_arena;
    double launch_cells = 2e11;
    std::vector<std::unique_ptr<ServerDb>> dbs;
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::deque<Request *> queue;
    std::vector<Request *> active;
    std::map<int64_t, std::unique_ptr<Request>> reqs;
    Slot slots[NSLOTS];
    std::deque<int> inflight;
    std::thread worker;
    // The completer thread waits for queued top-k results, so the worker never waits for a
    // kernel that needs CUs while the slots' persistent launches hold them all.
    std::deque<Request *> finishing;
    std::condition_variable cv_finish;
    std::thread completer;
    bool stop_completer = false;
    // buffers replaced by larger ones: kept to the end (hipFree and hipHostFree wait for the
    // whole device, which would stall the other slot's launch)
    std::vector<void *> graveyard, host_graveyard;
    bool stop = false;
    int64_t next_id = 1;
    std::chrono::steady_clock::time_point t0 = std::chrono::steady_clock::now();
    // statistics
    int64_t submitted = 0, completed = 0, rejected = 0, launches = 0, cells_done = 0, rescored = 0;
    double gpu_ms = 0;
    std::string fatal;

    double now_ms() const {
        return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    }
    void run();
    void complete_loop();
    int grow(void **buf, size_t *cap, size_t need, bool pinned);
    int admit_locked();
    int plan_locked(Request *r);
    int append_units_locked(Request *r, const std::vector<CrossUnit> &units, const std::vector<CrossQuery> &plain);
    int rerun_rows_locked(Request *r, std::vector<int> rows);
    Request *next_issuable_locked();
    int launch_locked(int slot, Request *head);
    int retire(int slot);
    int finalize(Request *r);
};

// Grow a device (or pinned host) buffer to `need` bytes without freeing the old one.
int Server::grow(void **buf, size_t *cap, size_t need, bool pinned) {
    if (need <= *cap && *buf) return GSW_OK;
    size_t want = std::max(need, *cap + *cap / 2);
    want = (want + 255) & ~size_t(255);
    void *p = nullptr;
    if (pinned) GSW_HIP_TRY(hipHostMalloc(&p, want, hipHostMallocDefault));
    else GSW_HIP_TRY(hipMalloc(&p, want));
    if (*buf) (pinned ? host_graveyard : graveyard).push_back(*buf);
    *buf = p;
    *cap = want;
    return GSW_OK;
}

// ------------------------------------------------------------------ planning
int Server::plan_locked(Request *r) {
    const ServerDb &db = *dbs[r->db];
    const SeqSet &rs = *db.rs;
    const int32_t nr = rs.n;
    std::vector<CrossQuery> cq;
    std::vector<int64_t> e_idx;
    std::vector<int32_t> e_val;
    for (int a = 0; a < r->nq; ++a) {
        const int32_t len = r->off[a + 1] - r->off[a];
        const int64_t row = (int64_t)(r->scores_at / 4) + (int64_t)a * nr;
        if (len == 0 || rs.n_nonempty == 0) {
            for (int b = 0; b < nr; ++b) {
                e_idx.push_back(row + b);
                e_val.push_back(empty_edge_score(db.s, len, rs.len[b]));
            }
            continue;
        }
        for (int p = rs.n_nonempty; p < nr; ++p) {
            e_idx.push_back(row + rs.order_desc[p]);
            e_val.push_back(empty_edge_score(db.s, len, 0));
        }
        cq.push_back(CrossQuery{(int32_t)r->codes_at + r->off[a], len, row, a});
        r->cells += (int64_t)len * rs.total;
    }
    std::vector<CrossUnit> units;
    std::vector<CrossQuery> plain;
    GSW_TRY(plan_cross_units(db.s, GSW_DTYPE_AUTO, cq, rs.max_len, db.mean_lr, (size_t)ctx.max_smem, &units,
                             &plain));
    GSW_TRY(append_units_locked(r, units, plain));
    if (!e_idx.empty()) GSW_TRY(scatter_values(&ctx, reinterpret_cast<int32_t *>(score_base), e_idx, e_val));
    return GSW_OK;
}

// Cross items of planned units appended to the request (grouped by tile); queries no tile
// covers are scored now with the comparator.
int Server::append_units_locked(Request *r, const std::vector<CrossUnit> &units, const std::vector<CrossQuery> &plain) {
    const ServerDb &db = *dbs[r->db];
    const SeqSet &rs = *db.rs;
    const int32_t nr = rs.n;
    // chunk references into small items: a launch gathers items of one tile from the
    // admitted requests, and small requests or rare tiles must still give it more items than
    // workgroups (with 1e8-cell items, 29 of 120 launches of the serving mix had fewer)
    for (size_t u0 = 0; u0 < units.size();) {
        size_t u1 = u0;
        const bool striped = units[u0].nstr > 1;
        while (u1 < units.size() && units[u1].v == units[u0].v && (units[u1].nstr > 1) == striped) ++u1;
        const Variant &v = *units[u0].v;
        const int gpw = 64 / v.G;
        if (striped) {
            // a quarter of the batch path's item, at least four blocks per wave per stripe
            const int64_t rows = std::max<int64_t>(stripe_item_rows(db.mean_lr, rs.max_len) / 4,
                                                   (int64_t)(4.0 * gpw * CROSS_WAVES * db.mean_lr));
            expand_stripe_items(units.data() + u0, units.data() + u1, rs, rows, &r->items);
        } else {
            // items of ~5e7 cells, at least two blocks of references per wave (the waves of a
            // workgroup then finish an item within about a block of each other)
            const double qcols = (double)(units[u0].a.len + std::max(units[u0].b.len, 0));
            int64_t rows = (int64_t)(5e7 / std::max(qcols, 1.0));
            rows = std::max<int64_t>(rows, (int64_t)(2.0 * gpw * CROSS_WAVES * db.mean_lr));
            expand_cross_items(units.data() + u0, units.data() + u1, rs, rows, gpw, &r->items);
        }
        r->item_tile.resize(r->items.size(), 2 * v.id + (striped ? 1 : 0));
        u0 = u1;
    }
    // queries no tile covers are scored now with the comparator
    if (!plain.empty()) {
        std::vector<int32_t> qs, ql, rj;
        std::vector<int64_t> oi;
        int maxq = 1;
        for (const CrossQuery &q : plain) {
            maxq = std::max(maxq, q.len);
            for (int b = 0; b < nr; ++b) {
                if (rs.len[b] == 0) continue;
                qs.push_back(q.start);
                ql.push_back(q.len);
                rj.push_back(b);
                oi.push_back(q.out + b);
            }
        }
        PlainJob job{};
        job.qcodes = code_base; job.qoff = nullptr; job.rcodes = rs.d_codes; job.roff = rs.d_off;
        job.dtype = GSW_DTYPE_INT32; job.threads = 256; job.maxq = maxq;
        job.h_qstart = qs.data(); job.h_qlen = ql.data(); job.h_rj = rj.data(); job.h_outidx = oi.data();
        job.npairs = (int64_t)qs.size();
        job.d_out = reinterpret_cast<int32_t *>(score_base);
        float ms = 0.f;
        GSW_TRY(plain_run(&ctx, db.s, job, &ms));
        gpu_ms += ms;
    }
    return GSW_OK;
}

// Query rows of a request whose half2 results were listed: planned again on tiles whose
// results the envelope proves exact (int16, else int32); the new items follow the request's
// others, so its top-k waits for them.  A half2 launch still running other items of a row
// can only write exact values below 2048, or list its own.
int Server::rerun_rows_locked(Request *r, std::vector<int> rows) {
    std::sort(rows.begin(), rows.end());
    rows.erase(std::unique(rows.begin(), rows.end()), rows.end());
    const ServerDb &db = *dbs[r->db];
    const SeqSet &rs = *db.rs;
    std::vector<CrossQuery> cq;
    for (int a : rows) {
        const int32_t len = r->off[a + 1] - r->off[a];
        if (len == 0 || rs.n_nonempty == 0) continue;
        cq.push_back(CrossQuery{(int32_t)r->codes_at + r->off[a], len,
                                (int64_t)(r->scores_at / 4) + (int64_t)a * rs.n, a});
    }
    std::vector<CrossUnit> units;
    std::vector<CrossQuery> plain;
    GSW_TRY(plan_cross_units(db.s, GSW_DTYPE_SAFE16, cq, rs.max_len, db.mean_lr, (size_t)ctx.max_smem, &units,
                             &plain));
    GSW_TRY(append_units_locked(r, units, plain));
    rescored += (int64_t)cq.size() * rs.n;
    return GSW_OK;
}

int Server::admit_locked() {
    while (!queue.empty()) {
        Request *r = queue.front();
        const SeqSet &rs = *dbs[r->db]->rs;
        r->codes_bytes = std::max<size_t>(r->codes.size(), 1);
        r->scores_bytes = 4 * ((size_t)r->nq * rs.n + 2 * (size_t)r->nq * std::max(r->kk, 1));
        size_t ca = 0, sa = 0;
        const Admission adm = admit_head(code_arena, score_arena, r->codes_bytes, r->scores_bytes, &ca, &sa);
        if (adm == Admission::WAIT) break;
        if (adm == Admission::REJECT) {
            r->status = GSW_ERR_BUSY;
            r->error = "request needs more device memory than the server budget";
            r->done = true;
            r->t_done = now_ms();
            rejected++;
            queue.pop_front();
            cv_done.notify_all();
            continue;
        }
        queue.pop_front();
        r->codes_at = ca;
        r->scores_at = sa;
        r->t_admit = now_ms();
        if (!r->codes.empty()) {
            GSW_HIP_TRY(hipMemcpyAsync(code_base + ca, r->codes.data(), r->codes.size(), hipMemcpyHostToDevice,
                                       ctx.stream));
            // the slots' streams wait for it before their next launch
            GSW_HIP_TRY(hipEventRecord(codes_ready, ctx.stream));
        }
        int rc = plan_locked(r);
        if (rc != GSW_OK) {
            r->status = rc;
            r->error = gsw_last_error();
            code_arena.release(r->codes_at, r->codes_bytes);
            score_arena.release(r->scores_at, r->scores_bytes);
            r->done = true;
            r->t_done = now_ms();
            cv_done.notify_all();
            continue;
        }
        active.push_back(r);
        if (r->items.empty() && r->inflight == 0) {
            GSW_TRY(finalize(r));
        }
    }
    return GSW_OK;
}

Request *Server::next_issuable_locked() {
    for (Request *r : active)
        if (r->next < r->items.size()) return r;
    return nullptr;
}

// ------------------------------------------------------------------ launching
int Server::launch_locked(int si, Request *head) {
    Slot &sl = slots[si];
    const int db = head->db;
    const int tile = head->item_tile[head->next];
    const ServerDb &d = *dbs[db];
    const Variant &v = variant_table()[tile / 2];
    const bool striped = tile & 1;
    sl.host_items.clear();
    sl.reqs.clear();
    sl.db = db;
    // round-robin over admitted requests of this (database, tile), oldest first
    double cells = 0;
    bool progress = true;
    while (progress && cells < launch_cells) {
        progress = false;
        for (Request *r : active) {
            if (r->db != db || r->next >= r->items.size() || r->item_tile[r->next] != tile) continue;
            const CrossItem &it = r->items[r->next++];
            sl.host_items.push_back(it);
            if (sl.reqs.empty() || sl.reqs.back().first != r) {
                bool seen = false;
                for (auto &pr : sl.reqs)
                    if (pr.first == r) { pr.second++; seen = true; break; }
                if (!seen) sl.reqs.emplace_back(r, 1);
            } else {
                sl.reqs.back().second++;
            }
            cells += (double)(it.lqa + std::max(it.lqb, 0)) * (double)(d.rs->bpos[it.rend] - d.rs->bpos[it.rbeg]);
            progress = true;
            if (cells >= launch_cells) break;
        }
    }
    for (auto &pr : sl.reqs) {
        pr.first->inflight++;
        pr.first->launches++;
    }
    // longest first, so the launch's work queue ends on the smallest items
    const int64_t *bp = d.rs->bpos.data();
    std::stable_sort(sl.host_items.begin(), sl.host_items.end(), [&](const CrossItem &x, const CrossItem &y) {
        return (bp[x.rend] - bp[x.rbeg]) * (int64_t)(x.lqa + std::max(x.lqb, 0)) * x.nstr >
               (bp[y.rend] - bp[y.rbeg]) * (int64_t)(y.lqa + std::max(y.lqb, 0)) * y.nstr;
    });
    // header (zeroed work counter) and items in one copy from pinned memory: the slot's previous
    // copy was read before its launch, which has retired
    const size_t bytes = SLOT_HDR + sizeof(CrossItem) * sl.host_items.size();
    GSW_TRY(grow(&sl.items, &sl.items_cap, bytes, false));
    GSW_TRY(grow(reinterpret_cast<void **>(&sl.h_items), &sl.h_cap, bytes, true));
    std::memset(sl.h_items, 0, SLOT_HDR);
    std::memcpy(sl.h_items + SLOT_HDR, sl.host_items.data(), bytes - SLOT_HDR);
    GSW_HIP_TRY(hipStreamWaitEvent(sl.st, codes_ready, 0));   // this launch's query codes are uploaded
    GSW_HIP_TRY(hipMemcpyAsync(sl.items, sl.h_items, bytes, hipMemcpyHostToDevice, sl.st));
    const size_t lds = cross_lds_bytes(v, d.s.ns);
    int grid = 0;
    GSW_TRY(occupancy_grid(&ctx, striped ? v.stripe_fn : v.cross_fn, CROSS_WAVES * 64, lds, &grid));
    // GSW_TRACE_LAUNCHES=1: one line per launch to stderr
    static const bool trace = [] { const char *e = std::getenv("GSW_TRACE_LAUNCHES"); return e && *e == '1'; }();
    if (trace)
        std::fprintf(stderr, "[gsw] serve launch %lld: db=%d G=%d R=%d%s requests=%zu items=%zu grid=%d cells=%.3e\n",
                     (long long)launches, db, v.G, v.R, striped ? " striped" : "", sl.reqs.size(),
                     sl.host_items.size(), grid, cells);
    grid = std::min<int>(grid, (int)sl.host_items.size());
    int64_t rows = 0;   // boundary entries per workgroup (striped)
    if (striped) {
        for (const CrossItem &it : sl.host_items) rows = std::max(rows, d.rs->bpos[it.rend] - d.rs->bpos[it.rbeg]);
        // launches of both slots may run at once: each slot has its own half of the scratch
        GSW_TRY(grow(&sl.bnd, &sl.bnd_bytes, 2 * sizeof(uint32_t) * (size_t)rows * grid, false));
    }
    KArgs a{};
    a.qcodes = code_base;
    a.rcodes = d.rs->d_codes;
    a.roff = d.rs->d_off;
    a.rorder = d.rs->d_order_desc;
    a.sub = d.s.d_sub;
    a.ns = d.s.ns;
    a.k = SweepConst{d.s.mode, d.s.go, d.s.ge, d.s.free_q, d.s.free_r, d.s.ns};
    a.out = reinterpret_cast<int32_t *>(score_base);
    a.sat = sl.sat;
    a.sat_count = reinterpret_cast<int32_t *>(static_cast<char *>(sl.items) + 8);
    a.sat_cap = SLOT_SAT_CAP;
    a.sat_check = (v.kind == KIND_F16 && d.s.mode == GSW_MODE_LOCAL) ? 1 : 0;
    a.items = static_cast<char *>(sl.items) + SLOT_HDR;
    a.nitems = (int32_t)sl.host_items.size();
    a.counter = static_cast<int32_t *>(sl.items);
    if (striped) {
        a.bnd = static_cast<uint32_t *>(sl.bnd);
        a.bnd_stride = rows;
        a.bpos = d.rs->d_bpos;
        v.stripe(a, grid, lds, sl.st);
    } else {
        v.cross(a, grid, lds, sl.st);
    }
    GSW_HIP_TRY(hipGetLastError());
    // the header's counters come back with the launch (DMA: no CU needed)
    GSW_HIP_TRY(hipMemcpyAsync(sl.h_hdr, sl.items, 16, hipMemcpyDeviceToHost, sl.st));
    GSW_HIP_TRY(hipEventRecord(sl.ev, sl.st));
    launches++;
    inflight.push_back(si);
    return GSW_OK;
}

// A request whose launches have all retired: its top-k and result copies are queued on the
// finish stream, and the completer thread marks it done once they land.
int Server::finalize(Request *r) {
    const SeqSet &rs = *dbs[r->db]->rs;
    const int32_t nr = rs.n;
    r->idx.assign((size_t)r->nq * r->kk, 0);
    r->sc.assign((size_t)r->nq * r->kk, 0);
    if (r->nq > 0 && r->kk > 0) {
        int32_t *scores = reinterpret_cast<int32_t *>(score_base + r->scores_at);
        int32_t *d_idx = scores + (size_t)r->nq * nr;
        int32_t *d_sc = d_idx + (size_t)r->nq * r->kk;
        topk_launch(scores, r->nq, nr, r->kk, d_idx, d_sc, finish);
        GSW_HIP_TRY(hipGetLastError());
        GSW_HIP_TRY(hipMemcpyAsync(r->idx.data(), d_idx, 4 * r->idx.size(), hipMemcpyDeviceToHost, finish));
        GSW_HIP_TRY(hipMemcpyAsync(r->sc.data(), d_sc, 4 * r->sc.size(), hipMemcpyDeviceToHost, finish));
        GSW_HIP_TRY(hipEventCreateWithFlags(&r->done_ev, hipEventDisableTiming));
        GSW_HIP_TRY(hipEventRecord(r->done_ev, finish));
    }
    r->codes.clear();
    r->codes.shrink_to_fit();
    r->items.clear();
    r->items.shrink_to_fit();
    r->finishing = true;
    active.erase(std::find(active.begin(), active.end(), r));
    finishing.push_back(r);
    cv_finish.notify_one();
    return GSW_OK;
}

void Server::complete_loop() {
    (void)hipSetDevice(ctx.device);
    std::unique_lock<std::mutex> lk(mu);
    for (;;) {
        cv_finish.wait(lk, [&] { return !finishing.empty() || stop_completer; });
        if (finishing.empty()) return;   // stopped, nothing left
        Request *r = finishing.front();
        finishing.pop_front();
        const hipEvent_t ev = r->done_ev;
        lk.unlock();
        const hipError_t e = ev ? hipEventSynchronize(ev) : hipSuccess;
        lk.lock();
        if (ev) (void)hipEventDestroy(ev);
        r->done_ev = nullptr;
        if (e != hipSuccess && r->status == GSW_OK) {
            r->status = GSW_ERR_HIP;
            r->error = hipGetErrorString(e);
        }
        code_arena.release(r->codes_at, r->codes_bytes);
        score_arena.release(r->scores_at, r->scores_bytes);
        r->done = true;
        r->t_done = now_ms();
        completed++;
        cells_done += r->cells;
        cv_done.notify_all();
        cv_work.notify_all();   // memory came back: the worker may admit more
    }
}

int Server::retire(int si) {
    Slot &sl = slots[si];
    GSW_HIP_TRY(hipEventSynchronize(sl.ev));
    const int32_t listed = sl.h_hdr[2];
    if (listed > 0 && listed <= SLOT_SAT_CAP) {
        GSW_HIP_TRY(hipMemcpyAsync(sl.h_sat, sl.sat, sizeof(SatEntry) * (size_t)listed, hipMemcpyDeviceToHost, sl.st));
        GSW_HIP_TRY(hipStreamSynchronize(sl.st));
    }
    std::unique_lock<std::mutex> lk(mu);
    if (listed > 0) {
        // the listed output cells name their request and query row; more than the list
        // holds: every query row of the launch's items
        std::vector<int64_t> outs;
        if (listed <= SLOT_SAT_CAP) {
            for (int32_t i = 0; i < listed; ++i) outs.push_back(sl.h_sat[i].o);
        } else {
            for (const CrossItem &it : sl.host_items) {
                outs.push_back(it.outa);
                if (it.outb >= 0) outs.push_back(it.outb);
            }
        }
        for (auto &pr : sl.reqs) {
            Request *r = pr.first;
            const int32_t nr = dbs[r->db]->rs->n;
            const int64_t base = (int64_t)(r->scores_at / 4), end = base + (int64_t)r->nq * nr;
            std::vector<int> rows;
            for (int64_t o : outs)
                if (o >= base && o < end) rows.push_back((int)((o - base) / nr));
            if (!rows.empty()) GSW_TRY(rerun_rows_locked(r, std::move(rows)));
        }
    }
    for (auto &pr : sl.reqs) {
        Request *r = pr.first;
        r->inflight--;
        if (r->inflight == 0 && r->next >= r->items.size() && !r->done) GSW_TRY(finalize(r));
    }
    sl.reqs.clear();
    return GSW_OK;
}

void Server::run() {
    (void)hipSetDevice(ctx.device);
    std::unique_lock<std::mutex> lk(mu);
    for (;;) {
        int rc = admit_locked();
        Request *head = rc == GSW_OK ? next_issuable_locked() : nullptr;
        if (rc == GSW_OK && head && (int)inflight.size() < NSLOTS) {
            int free_slot = -1;
            for (int i = 0; i < NSLOTS; ++i)
                if (std::find(inflight.begin(), inflight.end(), i) == inflight.end()) free_slot = i;
            rc = launch_locked(free_slot, head);
            if (rc == GSW_OK) continue;
        }
        if (rc == GSW_OK && !inflight.empty()) {
            const int si = inflight.front();
            inflight.pop_front();
            lk.unlock();
            const double t = now_ms();
            rc = retire(si);
            lk.lock();
            gpu_ms += now_ms() - t;
            if (rc == GSW_OK) continue;
        }
        if (rc != GSW_OK) {
            // a device failure ends the server: fail everything that is pending
            fatal = gsw_last_error();
            for (auto &kv : reqs) {
                Request *r = kv.second.get();
                if (!r->done && !r->finishing) {   // the completer finishes the others
                    r->status = rc;
                    r->error = fatal;
                    r->done = true;
                }
            }
            queue.clear();
            active.clear();
            inflight.clear();
            cv_done.notify_all();
            stop = true;
            return;
        }
        if (stop && queue.empty() && active.empty()) return;
        cv_work.wait(lk);
    }
}

}  // namespace gsw

using namespace gsw;

extern "C" {

int gsw_server_create(size_t mem_budget_bytes, double launch_cells, gsw_server *out) {
    if (!out) return fail(GSW_ERR_INVALID, "out is null");
    *out = nullptr;
    DeviceCtx *shared = nullptr;
    GSW_TRY(ctx_get(&shared));
    auto srv = std::make_unique<Server>();
    srv->ctx.device = shared->device;
    srv->ctx.num_cus = shared->num_cus;
    srv->ctx.max_smem = shared->max_smem;
    GSW_HIP_TRY(hipStreamCreateWithFlags(&srv->ctx.stream, hipStreamNonBlocking));
    GSW_HIP_TRY(hipStreamCreateWithFlags(&srv->finish, hipStreamNonBlocking));
    GSW_HIP_TRY(hipEventCreate(&srv->ctx.ev0));
    GSW_HIP_TRY(hipEventCreate(&srv->ctx.ev1));
    GSW_HIP_TRY(hipEventCreateWithFlags(&srv->codes_ready, hipEventDisableTiming));
    GSW_HIP_TRY(hipMalloc(&srv->ctx.counters, NCOUNTERS * sizeof(int32_t)));
    if (launch_cells > 0) srv->launch_cells = launch_cells;
    // query codes are addressed with 32-bit offsets: at most 1 GiB of them
    const size_t budget = std::max<size_t>(mem_budget_bytes, size_t(64) << 20);
    const size_t code_cap = std::min<size_t>(budget / 16, size_t(1) << 30);
    const size_t score_cap = budget - code_cap;
    GSW_HIP_TRY(hipMalloc(&srv->code_base, code_cap));
    GSW_HIP_TRY(hipMalloc(&srv->score_base, score_cap));
    srv->code_arena.init(code_cap);
    srv->score_arena.init(score_cap);
    for (int i = 0; i < NSLOTS; ++i) {
        Slot &sl = srv->slots[i];
        GSW_HIP_TRY(hipStreamCreateWithFlags(&sl.st, hipStreamNonBlocking));
        GSW_HIP_TRY(hipEventCreateWithFlags(&sl.ev, hipEventDisableTiming));
        GSW_HIP_TRY(hipMalloc(&sl.sat, sizeof(SatEntry) * SLOT_SAT_CAP));
        GSW_HIP_TRY(hipHostMalloc(&sl.h_sat, sizeof(SatEntry) * SLOT_SAT_CAP, hipHostMallocDefault));
        GSW_HIP_TRY(hipHostMalloc(&sl.h_hdr, SLOT_HDR, hipHostMallocDefault));
    }
    Server *p = srv.get();
    p->worker = std::thread([p] { p->run(); });
    p->completer = std::thread([p] { p->complete_loop(); });
    *out = reinterpret_cast<gsw_server>(srv.release());
    return GSW_OK;
}

int gsw_server_add_db(gsw_server server, const gsw_scheme *s, gsw_seqset refs, int32_t *db_id) {
    Server *srv = reinterpret_cast<Server *>(server);
    const SeqSet *rs = reinterpret_cast<const SeqSet *>(refs);
    if (!srv || !rs || !db_id) return fail(GSW_ERR_INVALID, "null argument");
    if (rs->device != srv->ctx.device) return fail(GSW_ERR_INVALID, "database lives on another device");
    auto db = std::make_unique<ServerDb>();
    GSW_TRY(scheme_prepare(s, &srv->ctx, &db->s));
    db->rs = rs;
    db->mean_lr = rs->n_nonempty ? (double)rs->total / rs->n_nonempty : 1.0;
    std::lock_guard<std::mutex> lk(srv->mu);
    if (srv->stop) return fail(GSW_ERR_HIP, "server stopped: %s", srv->fatal.c_str());
    srv->dbs.push_back(std::move(db));
    *db_id = (int32_t)srv->dbs.size() - 1;
    return GSW_OK;
}

int gsw_server_submit(gsw_server server, int32_t db_id, const uint8_t *h_codes, const int32_t *h_off,
                      int32_t nq, int32_t k, int64_t *req_id) {
    Server *srv = reinterpret_cast<Server *>(server);
    if (!srv || !req_id || (nq > 0 && (!h_off || !h_codes))) return fail(GSW_ERR_INVALID, "null argument");
    if (nq < 0 || k < 0) return fail(GSW_ERR_INVALID, "negative query count or k");
    auto r = std::make_unique<Request>();
    {
        std::lock_guard<std::mutex> lk(srv->mu);
        if (db_id < 0 || db_id >= (int)srv->dbs.size()) return fail(GSW_ERR_INVALID, "unknown database %d", db_id);
        if (srv->stop) return fail(GSW_ERR_HIP, "server stopped: %s", srv->fatal.c_str());
    }
    r->db = db_id;
    r->nq = nq;
    r->k = k;
    r->off.assign(h_off, h_off + nq + 1);
    const int32_t base = nq > 0 ? h_off[0] : 0;
    for (int i = 0; i <= nq; ++i) {
        r->off[i] -= base;
        if (i && r->off[i] < r->off[i - 1]) return fail(GSW_ERR_INVALID, "offsets must be non-decreasing");
    }
    if (nq > 0) r->codes.assign(h_codes + base, h_codes + base + r->off[nq]);
    std::lock_guard<std::mutex> lk(srv->mu);
    r->kk = std::min<int>(k, srv->dbs[db_id]->rs->n);
    r->id = srv->next_id++;
    r->t_submit = srv->now_ms();
    *req_id = r->id;
    srv->queue.push_back(r.get());
    srv->reqs[r->id] = std::move(r);
    srv->submitted++;
    srv->cv_work.notify_all();
    return GSW_OK;
}

int gsw_server_wait(gsw_server server, int64_t req_id, int32_t timeout_ms, int32_t *h_idx, int32_t *h_sc,
                    gsw_request_info *info) {
    Server *srv = reinterpret_cast<Server *>(server);
    if (!srv) return fail(GSW_ERR_INVALID, "null server");
    std::unique_lock<std::mutex> lk(srv->mu);
    auto it = srv->reqs.find(req_id);
    if (it == srv->reqs.end()) return fail(GSW_ERR_INVALID, "unknown request %lld", (long long)req_id);
    Request *r = it->second.get();
    // one waiter at a time: the waiter that returns the result frees the request
    if (r->claimed) return fail(GSW_ERR_INVALID, "request %lld is already being waited for", (long long)req_id);
    r->claimed = true;
    auto ready = [&] { return r->done; };
    if (timeout_ms < 0) {
        srv->cv_done.wait(lk, ready);
    } else if (!srv->cv_done.wait_for(lk, std::chrono::milliseconds(timeout_ms), ready)) {
        r->claimed = false;
        return fail(GSW_ERR_TIMEOUT, "request %lld not finished", (long long)req_id);
    }
    if (info) {
        *info = gsw_request_info{};
        info->submit_ms = r->t_submit;
        info->admit_ms = r->t_admit;
        info->done_ms = r->t_done;
        info->latency_ms = r->t_done - r->t_submit;
        info->cells = r->cells;
        info->nq = r->nq;
        info->k = r->kk;
        info->launches = r->launches;
        info->status = r->status;
    }
    const int status = r->status;
    std::string err = r->error;
    if (status == GSW_OK) {
        if (h_idx && !r->idx.empty()) std::copy(r->idx.begin(), r->idx.end(), h_idx);
        if (h_sc && !r->sc.empty()) std::copy(r->sc.begin(), r->sc.end(), h_sc);
    }
    srv->reqs.erase(it);
    if (status != GSW_OK) return fail(status, "%s", err.c_str());
    return GSW_OK;
}

int gsw_server_stats(gsw_server server, gsw_server_status *st) {
    Server *srv = reinterpret_cast<Server *>(server);
    if (!srv || !st) return fail(GSW_ERR_INVALID, "null argument");
    std::lock_guard<std::mutex> lk(srv->mu);
    *st = gsw_server_status{};
    st->submitted = srv->submitted;
    st->completed = srv->completed;
    st->rejected = srv->rejected;
    st->launches = srv->launches;
    st->cells = srv->cells_done;
    st->rescored = srv->rescored;
    st->queued = (int32_t)srv->queue.size();
    st->active = (int32_t)srv->active.size();
    st->memory_budget = srv->code_arena.cap() + srv->score_arena.cap();
    st->memory_used = srv->code_arena.used() + srv->score_arena.used();
    st->uptime_ms = srv->now_ms();
    return GSW_OK;
}

int gsw_server_destroy(gsw_server server) {
    Server *srv = reinterpret_cast<Server *>(server);
    if (!srv) return GSW_OK;
    {
        std::lock_guard<std::mutex> lk(srv->mu);
        srv->stop = true;
        srv->cv_work.notify_all();
    }
    if (srv->worker.joinable()) srv->worker.join();
    {
        std::lock_guard<std::mutex> lk(srv->mu);
        srv->stop_completer = true;
        srv->cv_finish.notify_all();
    }
    if (srv->completer.joinable()) srv->completer.join();   // it drains the finishing queue first
    (void)hipStreamSynchronize(srv->finish);
    (void)hipStreamSynchronize(srv->ctx.stream);
    for (int i = 0; i < NSLOTS; ++i) {
        if (srv->slots[i].st) {
            (void)hipStreamSynchronize(srv->slots[i].st);
            (void)hipStreamDestroy(srv->slots[i].st);
        }
        (void)hipEventDestroy(srv->slots[i].ev);
        if (srv->slots[i].items) (void)hipFree(srv->slots[i].items);
        if (srv->slots[i].h_items) (void)hipHostFree(srv->slots[i].h_items);
        if (srv->slots[i].bnd) (void)hipFree(srv->slots[i].bnd);
        if (srv->slots[i].sat) (void)hipFree(srv->slots[i].sat);
        if (srv->slots[i].h_sat) (void)hipHostFree(srv->slots[i].h_sat);
        if (srv->slots[i].h_hdr) (void)hipHostFree(srv->slots[i].h_hdr);
    }
    for (void *b : srv->graveyard) (void)hipFree(b);
    for (void *b : srv->host_graveyard) (void)hipHostFree(b);
    (void)hipFree(srv->code_base);
    (void)hipFree(srv->score_base);
    (void)hipFree(srv->ctx.counters);
    if (srv->ctx.plain_ws) (void)hipFree(srv->ctx.plain_ws);
    if (srv->ctx.meta) (void)hipFree(srv->ctx.meta);
    if (srv->ctx.meta2) (void)hipFree(srv->ctx.meta2);
    (void)hipEventDestroy(srv->ctx.ev0);
    (void)hipEventDestroy(srv->ctx.ev1);
    if (srv->codes_ready) (void)hipEventDestroy(srv->codes_ready);
    (void)hipStreamDestroy(srv->ctx.stream);
    (void)hipStreamDestroy(srv->finish);
    delete srv;
    return GSW_OK;
}

}  // extern "C"
