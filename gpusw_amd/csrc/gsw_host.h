// Host-side parts of libgsw_hip.so that touch no device: the planning thread pool, the
// explicit-pairs batch planner and the first-fit arena of the serving loop.  Plain C++ with
// no HIP header, so that tests/native builds exactly this code for the CPU and runs it
// under the thread and address sanitizers.
#pragma once

#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstddef>
#include <cstdint>
#include <functional>
#include <map>
#include <mutex>
#include <thread>
#include <vector>

namespace gsw {

// Task of the pairs kernel: queries qa and qb (qb < 0: single) against reference ref.
struct PairTask {
    int32_t qa, qb, ref, shift;
    int64_t oa, ob;          // output indices (ob < 0: none)
};

// Host threads for planning, started on first use and kept: starting and joining 7
// threads costs 0.2-0.4 ms, several times per call.  run(T, fn) runs fn(1..T-1) on pool
// threads and fn(0) on the caller, and returns when all have returned.  One run at a time;
// a caller that finds the pool busy starts threads of its own instead.  A pairs call runs
// several jobs back to back and a sleeping thread takes tens of microseconds to wake, so
// the workers (for their next job) and the caller (for the workers to finish) spin up to
// 200 microseconds before they sleep: 70-90 us per short job become about 15.
class HostPool {
  public:
    template <typename F> void run(int T, F &&fn) {
        std::unique_lock<std::mutex> call(call_mu_, std::try_to_lock);
        if (!call.owns_lock()) {
            std::vector<std::thread> th;
            for (int i = 1; i < T; ++i) th.emplace_back([&fn, i] { fn(i); });
            fn(0);
            for (auto &x : th) x.join();
            return;
        }
        {
            std::lock_guard<std::mutex> lk(mu_);
            while ((int)threads_.size() < T - 1) {
                const int w = (int)threads_.size() + 1;
                threads_.emplace_back([this, w] { loop(w); });
            }
            job_ = [&fn](int i) { fn(i); };
            njobs_ = T;
            remaining_.store(T - 1, std::memory_order_relaxed);
            gen_.fetch_add(1, std::memory_order_release);
        }
        cv_.notify_all();
        fn(0);
        spin([&] { return remaining_.load(std::memory_order_acquire) == 0; });
        std::unique_lock<std::mutex> lk(mu_);
        done_.wait(lk, [&] { return remaining_.load(std::memory_order_acquire) == 0; });
        job_ = nullptr;
    }
    ~HostPool() {
        {
            std::lock_guard<std::mutex> lk(mu_);
            quit_.store(true);
        }
        cv_.notify_all();
        for (auto &t : threads_) t.join();
    }

  private:
    // pred() or 200 microseconds, whichever comes first
    template <typename P> static void spin(P &&pred) {
        using clk = std::chrono::steady_clock;
        const clk::time_point end = clk::now() + std::chrono::microseconds(200);
        for (int i = 0; !pred(); ++i)
            if ((i & 255) == 255 && clk::now() > end) return;
    }
    void loop(int w) {
        uint64_t seen = 0;
        for (;;) {
            spin([&] { return gen_.load(std::memory_order_acquire) != seen || quit_.load(std::memory_order_relaxed); });
            std::function<void(int)> f;
            {
                std::unique_lock<std::mutex> lk(mu_);
                cv_.wait(lk, [&] { return quit_.load() || gen_.load() != seen; });
                if (quit_.load()) return;
                seen = gen_.load();
                if (w >= njobs_) continue;
                f = job_;
            }
            f(w);
            // under the lock, so the caller cannot miss the notification
            std::lock_guard<std::mutex> lk(mu_);
            if (remaining_.fetch_sub(1, std::memory_order_acq_rel) == 1) done_.notify_all();
        }
    }
    std::mutex call_mu_, mu_;
    std::condition_variable cv_, done_;
    std::vector<std::thread> threads_;
    std::function<void(int)> job_;
    int njobs_ = 0;
    std::atomic<int> remaining_{0};
    std::atomic<uint64_t> gen_{0};
    std::atomic<bool> quit_{false};
};

// First-fit allocator over [0, cap) offsets (host bookkeeping for a device arena).
class Arena {
  public:
    static constexpr size_t ALIGN = 256;   // every block starts on a 256-byte boundary
    static size_t round_up(size_t v) { return (v + ALIGN - 1) & ~(ALIGN - 1); }
    void init(size_t cap) {
        cap_ = cap;
        used_ = 0;
        free_.clear();
        if (cap) free_[0] = cap;
    }
    bool alloc(size_t bytes, size_t *off) {
        bytes = round_up(std::max<size_t>(bytes, 1));
        for (auto it = free_.begin(); it != free_.end(); ++it) {
            if (it->second < bytes) continue;
            *off = it->first;
            const size_t rest = it->second - bytes, start = it->first + bytes;
            free_.erase(it);
            if (rest) free_[start] = rest;
            used_ += bytes;
            return true;
        }
        return false;
    }
    void release(size_t off, size_t bytes) {
        bytes = round_up(std::max<size_t>(bytes, 1));
        used_ -= bytes;
        auto it = free_.emplace(off, bytes).first;
        auto nx = std::next(it);
        if (nx != free_.end() && it->first + it->second == nx->first) {
            it->second += nx->second;
            free_.erase(nx);
        }
        if (it != free_.begin()) {
            auto pv = std::prev(it);
            if (pv->first + pv->second == it->first) {
                pv->second += it->second;
                free_.erase(it);
            }
        }
    }
    size_t cap() const { return cap_; }
    size_t used() const { return used_; }

  private:
    std::map<size_t, size_t> free_;
    size_t cap_ = 0, used_ = 0;
};

// FIFO admission of the serving loop's head request over its two arenas (query codes, score
// rows): ADMIT when both blocks fit (their offsets in *ca, *sa), REJECT when a block exceeds
// its arena's whole capacity (it could never fit), else WAIT for memory (nothing taken).
enum class Admission { ADMIT, WAIT, REJECT };
inline Admission admit_head(Arena &codes, Arena &scores, size_t code_bytes, size_t score_bytes, size_t *ca,
                            size_t *sa) {
    if (Arena::round_up(code_bytes) > codes.cap() || Arena::round_up(score_bytes) > scores.cap())
        return Admission::REJECT;
    if (!codes.alloc(code_bytes, ca)) return Admission::WAIT;
    if (!scores.alloc(score_bytes, sa)) {
        codes.release(*ca, code_bytes);
        return Admission::WAIT;
    }
    return Admission::ADMIT;
}

// The inputs of one explicit-pairs call that its batch planner reads.
struct PairPlanView {
    const int32_t *qi, *rj;       // the call's pair list
    const int64_t *members;       // pair indices of one tile (a batch is a range of them)
    const int64_t *outidx;        // output index of pair k (null: k)
    const int32_t *qlen, *rlen;   // sequence lengths
    const int32_t *rank;          // reference position in length order (longest first)
    int32_t n_nonempty;           // references of nonzero length (ranks 0 .. n_nonempty-1)
    int wmax;                     // widest query a tile takes
};

// A planner thread's buffers, kept across batches.
struct PairPlanScratch {
    struct Member { int64_t k; int32_t q, r, shift; };
    std::vector<Member> a, b;
    std::vector<int32_t> cnt;
    std::vector<int32_t> shift_of_len;   // front padding per query length (0..wmax)
    int shift_r = -1;                    // the R it holds (0: last column not tracked)
};

// Plan members[b0, b1) of one tile (G lanes, R columns per lane; packed: two queries per
// task) into tasks: a counting sort by reference rank, then (stable) by front padding when
// the last query column is tracked; adjacent members of one reference and padding share a
// packed task; blocks of 64/G tasks are padded with empty tasks where the padding changes
// and at the end.  Each member's query, reference and padding are gathered once, reading the
// pair list in index order.  Returns the task count (a multiple of 64/G); tasks must hold
// (b1 - b0) + (R + 1) * 64/G entries.  *cells = sum of lq * lr over the members.
inline int32_t plan_pair_batch(const PairPlanView &in, int64_t b0, int64_t b1, int G, int R, bool packed,
                               bool track_last_col, PairPlanScratch &S, PairTask *tasks, double *cells) {
    using Member = PairPlanScratch::Member;
    const int gpw = 64 / G;
    const size_t nm = (size_t)(b1 - b0);
    auto oidx = [&](int64_t k) -> int64_t { return in.outidx ? in.outidx[k] : k; };
    if (S.shift_r != (track_last_col ? R : 0)) {   // a tile's members are at most wmax long
        S.shift_r = track_last_col ? R : 0;
        S.shift_of_len.assign((size_t)in.wmax + 1, 0);
        if (track_last_col)
            for (int lq = 0; lq <= in.wmax; ++lq) S.shift_of_len[lq] = (R - lq % R) % R;
    }
    S.cnt.assign((size_t)in.n_nonempty + 1, 0);
    for (int64_t i = b0; i < b1; ++i) S.cnt[in.rank[in.rj[in.members[i]]] + 1]++;
    for (int32_t r = 0; r < in.n_nonempty; ++r) S.cnt[r + 1] += S.cnt[r];
    S.a.resize(nm);
    double c = 0;
    for (int64_t i = b0; i < b1; ++i) {
        const int64_t k = in.members[i];
        const int32_t q = in.qi[k], r = in.rj[k];
        const int lq = in.qlen[q];
        c += (double)lq * in.rlen[r];
        S.a[S.cnt[in.rank[r]]++] = Member{k, q, r, S.shift_of_len[lq]};
    }
    *cells = c;
    if (track_last_col) {
        S.cnt.assign(R + 1, 0);
        for (const Member &x : S.a) S.cnt[x.shift + 1]++;
        int distinct = 0;
        for (int r = 1; r <= R; ++r) distinct += S.cnt[r] > 0;
        // a second (stable) pass by front padding, unless the whole batch shares one
        if (distinct > 1) {
            for (int r = 0; r < R; ++r) S.cnt[r + 1] += S.cnt[r];
            S.b.resize(nm);
            for (const Member &x : S.a) S.b[S.cnt[x.shift]++] = x;
            S.a.swap(S.b);
        }
    }
    int32_t nt = 0;
    int cur_shift = -1;
    for (size_t i = 0; i < nm;) {
        const Member &x = S.a[i];
        if (x.shift != cur_shift) {
            while (nt % gpw) tasks[nt++] = PairTask{-1, -1, -1, cur_shift, -1, -1};
            cur_shift = x.shift;
        }
        if (packed && i + 1 < nm && S.a[i + 1].r == x.r && S.a[i + 1].shift == x.shift) {
            const Member &y = S.a[i + 1];
            tasks[nt++] = PairTask{x.q, y.q, x.r, x.shift, oidx(x.k), oidx(y.k)};
            i += 2;
        } else {
            tasks[nt++] = PairTask{x.q, -1, x.r, x.shift, oidx(x.k), -1};
            i += 1;
        }
    }
    while (nt % gpw) tasks[nt++] = PairTask{-1, -1, -1, cur_shift, -1, -1};
    return nt;
}

}  // namespace gsw
