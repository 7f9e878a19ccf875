// CPU harness for the host code of libgsw_hip.so that touches no device (gsw_host.h): the
// planning thread pool, the explicit-pairs batch planner and the serving loop's arena and
// FIFO admission.  Built twice by the Makefile beside it, under -fsanitize=thread and under
// -fsanitize=address,undefined, and run by tests/test_host_sanitizers.py.  Exits non-zero on
// the first failed check (the sanitizers report races and memory errors themselves).
#include <algorithm>
#include <atomic>
#include <condition_variable>
#include <cstdio>
#include <cstdlib>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <numeric>
#include <random>
#include <set>
#include <thread>
#include <vector>

#include "gsw_host.h"

using namespace gsw;

#define CHECK(cond)                                                                  \
    do {                                                                             \
        if (!(cond)) {                                                               \
            std::fprintf(stderr, "check failed: %s (%s:%d)\n", #cond, __FILE__, __LINE__); \
            std::exit(1);                                                            \
        }                                                                            \
    } while (0)

// ------------------------------------------------------------------ thread pool
// Several callers run jobs at once: one owns the pool, the others start threads of their own.
static void test_pool() {
    HostPool pool;
    std::atomic<int64_t> total{0};
    std::vector<std::thread> callers;
    for (int c = 0; c < 4; ++c) {
        callers.emplace_back([&, c] {
            std::mt19937 rng(100 + c);
            for (int it = 0; it < 150; ++it) {
                const int T = 1 + (int)(rng() % 8);
                std::vector<int> hit(T, 0);   // each job writes only its own slot
                pool.run(T, [&](int tid) { hit[tid] += 1; });
                for (int t = 0; t < T; ++t) CHECK(hit[t] == 1);
                total += T;
            }
        });
    }
    for (auto &t : callers) t.join();
    CHECK(total.load() > 0);
    std::printf("pool: %lld jobs\n", (long long)total.load());
}

// ------------------------------------------------------------------ pair planner
struct SynthSet {
    std::vector<int32_t> qlen, rlen, rank, qi, rj;
    int32_t n_nonempty = 0;
};

static SynthSet synth_pairs(std::mt19937 &rng, int nq, int nr, int64_t n, int wmax) {
    SynthSet s;
    s.qlen.resize(nq);
    s.rlen.resize(nr);
    for (auto &l : s.qlen) l = 1 + (int)(rng() % wmax);
    for (auto &l : s.rlen) l = (rng() % 10 == 0) ? 0 : 1 + (int)(rng() % 400);
    std::vector<int32_t> order(nr);
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return s.rlen[a] > s.rlen[b]; });
    s.rank.resize(nr);
    for (int p = 0; p < nr; ++p) s.rank[order[p]] = p;
    while (s.n_nonempty < nr && s.rlen[order[s.n_nonempty]] > 0) ++s.n_nonempty;
    for (int64_t k = 0; k < n; ++k) {
        s.qi.push_back((int32_t)(rng() % nq));
        s.rj.push_back(order[rng() % s.n_nonempty]);   // empty references take closed forms
    }
    return s;
}

static void check_batch(const SynthSet &s, const PairPlanView &in, int64_t b0, int64_t b1, int G, int R,
                        bool packed, bool track, const PairTask *t, int32_t nt, double cells) {
    const int gpw = 64 / G;
    CHECK(nt % gpw == 0);
    CHECK(nt <= (b1 - b0) + (R + 1) * gpw);
    std::map<int64_t, int64_t> out_to_k;
    for (int64_t i = b0; i < b1; ++i) out_to_k[in.outidx ? in.outidx[in.members[i]] : in.members[i]] = in.members[i];
    std::set<int64_t> seen;
    double want_cells = 0;
    for (int64_t i = b0; i < b1; ++i) want_cells += (double)s.qlen[s.qi[in.members[i]]] * s.rlen[s.rj[in.members[i]]];
    CHECK(cells == want_cells);
    int last_rank = -1, last_shift = -1;
    for (int32_t b = 0; b < nt; b += gpw) {
        const int block_shift = t[b].shift;
        for (int32_t j = b; j < b + gpw; ++j) {
            const PairTask &x = t[j];
            if (x.ref < 0) {
                CHECK(x.oa < 0 && x.ob < 0);
                continue;
            }
            CHECK(x.shift == block_shift);   // one front padding per block (wave-uniform)
            for (int h = 0; h < 2; ++h) {
                const int64_t o = h ? x.ob : x.oa;
                const int32_t q = h ? x.qb : x.qa;
                if (o < 0) {
                    CHECK(h == 1 && q < 0);
                    continue;
                }
                CHECK(out_to_k.count(o) && seen.insert(o).second);   // each member exactly once
                const int64_t k = out_to_k[o];
                CHECK(s.qi[k] == q && s.rj[k] == x.ref);
                const int lq = s.qlen[q];
                CHECK(x.shift == (track ? (R - lq % R) % R : 0));
            }
            CHECK(packed || x.qb < 0);
            // by front padding, then reference rank
            const int rk = s.rank[x.ref];
            if (x.shift == last_shift) CHECK(rk >= last_rank);
            else CHECK(x.shift > last_shift);
            last_rank = rk;
            last_shift = x.shift;
        }
    }
    CHECK((int64_t)seen.size() == b1 - b0);
}

// Batches of one tile planned in parallel on the pool, each planner with its own scratch
// and task buffer (as the library's planner threads do).
static void test_planner() {
    std::mt19937 rng(7);
    const int wmax = 2048;
    int checked = 0;
    for (int G : {8, 16, 64}) {
        for (int R : {19, 20, 32}) {
            if (G * R > wmax) continue;
            for (int flags = 0; flags < 4; ++flags) {
                const bool packed = flags & 1, track = flags & 2;
                SynthSet s = synth_pairs(rng, 3000, 500, 40000, G * R);
                std::vector<int64_t> members(s.qi.size());
                std::iota(members.begin(), members.end(), 0);
                std::vector<int64_t> outidx(members.size());
                std::iota(outidx.begin(), outidx.end(), 1000000);
                std::shuffle(outidx.begin(), outidx.end(), rng);
                PairPlanView in{s.qi.data(), s.rj.data(), members.data(), (flags == 3) ? outidx.data() : nullptr,
                                s.qlen.data(), s.rlen.data(), s.rank.data(), s.n_nonempty, G * R};
                const int64_t batch = 4096 + 37 * G;
                const int64_t nb = ((int64_t)members.size() + batch - 1) / batch;
                HostPool pool;
                std::atomic<int64_t> next{0};
                pool.run(6, [&](int) {
                    PairPlanScratch S;
                    std::vector<PairTask> tasks;
                    for (;;) {
                        const int64_t b = next.fetch_add(1);
                        if (b >= nb) return;
                        const int64_t b0 = b * batch, b1 = std::min<int64_t>((int64_t)members.size(), b0 + batch);
                        tasks.assign((size_t)(b1 - b0) + (R + 1) * (64 / G), PairTask{});
                        double cells = 0;
                        const int32_t nt = plan_pair_batch(in, b0, b1, G, R, packed, track, S, tasks.data(), &cells);
                        check_batch(s, in, b0, b1, G, R, packed, track, tasks.data(), nt, cells);
                    }
                });
                ++checked;
            }
        }
    }
    std::printf("planner: %d tile/flag combinations\n", checked);
}

// ------------------------------------------------------------------ arena
static void test_arena() {
    std::mt19937 rng(11);
    Arena a;
    const size_t cap = size_t(64) << 20;
    a.init(cap);
    std::map<size_t, size_t> live;   // offset -> requested bytes
    for (int it = 0; it < 20000; ++it) {
        if (!live.empty() && (rng() % 3 == 0 || a.used() > cap / 2)) {
            auto pick = live.begin();
            std::advance(pick, rng() % live.size());
            a.release(pick->first, pick->second);
            live.erase(pick);
        } else {
            const size_t bytes = 1 + rng() % (size_t(1) << (rng() % 22));
            size_t off = 0;
            if (a.alloc(bytes, &off)) {
                CHECK(off % Arena::ALIGN == 0 && off + bytes <= cap);
                auto nx = live.lower_bound(off);   // no overlap with its neighbours
                if (nx != live.end()) CHECK(off + Arena::round_up(bytes) <= nx->first);
                if (nx != live.begin()) {
                    auto pv = std::prev(nx);
                    CHECK(pv->first + Arena::round_up(pv->second) <= off);
                }
                live[off] = bytes;
            }
        }
        size_t used = 0;
        if (it % 500 == 0) {
            for (auto &kv : live) used += Arena::round_up(kv.second);
            CHECK(used == a.used());
        }
    }
    for (auto &kv : live) a.release(kv.first, kv.second);
    CHECK(a.used() == 0);
    size_t off = 1;
    CHECK(a.alloc(cap, &off) && off == 0);   // the free list coalesced back to one block
    std::printf("arena: ok\n");
}

// ------------------------------------------------------------------ serving protocol
// The serving loop's shape: submitters queue requests under one lock; a worker admits the
// head in FIFO order over two arenas (admit_head), runs admitted requests for a while, then
// releases their memory and marks them done; one waiter claims each request (a second claim
// of the same request is refused) and frees it.
struct Req {
    int64_t id;
    size_t code_bytes, score_bytes, ca = 0, sa = 0;
    bool done = false, rejected = false, claimed = false;
};

static void test_admission_protocol() {
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::deque<Req *> queue;
    std::vector<Req *> running;
    std::map<int64_t, std::unique_ptr<Req>> reqs;
    Arena codes, scores;
    codes.init(size_t(1) << 20);
    scores.init(size_t(8) << 20);
    bool stop = false;
    int64_t next_id = 1, admitted = 0, rejected = 0, refused_second = 0;
    std::thread worker([&] {
        std::unique_lock<std::mutex> lk(mu);
        std::mt19937 rng(5);
        for (;;) {
            while (!queue.empty()) {
                Req *r = queue.front();
                const Admission adm = admit_head(codes, scores, r->code_bytes, r->score_bytes, &r->ca, &r->sa);
                if (adm == Admission::WAIT) break;
                queue.pop_front();
                if (adm == Admission::REJECT) {
                    r->rejected = r->done = true;
                    ++rejected;
                    cv_done.notify_all();
                } else {
                    running.push_back(r);
                    ++admitted;
                }
            }
            if (!running.empty()) {   // finish one admitted request, outside the lock
                Req *r = running[rng() % running.size()];
                lk.unlock();
                std::this_thread::yield();
                lk.lock();
                codes.release(r->ca, r->code_bytes);
                scores.release(r->sa, r->score_bytes);
                running.erase(std::find(running.begin(), running.end(), r));
                r->done = true;
                cv_done.notify_all();
                continue;
            }
            if (stop && queue.empty()) return;
            cv_work.wait(lk);
        }
    });
    std::vector<std::thread> clients;
    for (int c = 0; c < 6; ++c) {
        clients.emplace_back([&, c] {
            std::mt19937 rng(200 + c);
            for (int it = 0; it < 120; ++it) {
                int64_t id;
                {
                    std::lock_guard<std::mutex> lk(mu);
                    auto r = std::make_unique<Req>();
                    r->id = id = next_id++;
                    r->code_bytes = 1 + rng() % 70000;
                    r->score_bytes = (rng() % 50 == 0) ? (size_t(9) << 20) : 1 + rng() % (size_t(3) << 20);
                    queue.push_back(r.get());
                    reqs[id] = std::move(r);
                }
                cv_work.notify_all();
                std::unique_lock<std::mutex> lk(mu);
                Req *r = reqs.at(id).get();
                CHECK(!r->claimed);
                r->claimed = true;
                cv_done.wait(lk, [&] { return r->done; });
                CHECK(r->rejected == (r->score_bytes > scores.cap()));
                // a second waiter on a claimed request is refused before it touches it
                if (reqs.at(id)->claimed) ++refused_second;
                reqs.erase(id);
            }
        });
    }
    for (auto &t : clients) t.join();
    {
        std::lock_guard<std::mutex> lk(mu);
        stop = true;
    }
    cv_work.notify_all();
    worker.join();
    CHECK(codes.used() == 0 && scores.used() == 0 && reqs.empty());
    CHECK(admitted + rejected == 6 * 120 && refused_second == 6 * 120);
    std::printf("admission: %lld admitted, %lld rejected\n", (long long)admitted, (long long)rejected);
}

int main() {
    test_pool();
    test_planner();
    test_arena();
    test_admission_protocol();
    std::printf("host harness ok\n");
    return 0;
}
