// Batched kernels of gfx950: the cross, stripe, pairs and pairs-table kernels, templated
// on the element type (int16x2, half2 or int32) and tile (G lanes x R columns), plus the
// launch stubs and Variant rows that the host schedules them by.  Included by
// gsw_sweep.hip (int16x2 and int32 instantiations) and gsw_sweep_f16.hip (half2), so the
// instantiations compile as two translation units in parallel.
#pragma once

#include "gsw_batch.h"
#include "gsw_sweep.h"

namespace gsw {

// Store a result; with the saturation check on (packed local launches), a result at or
// above the element type's exact range (int16: 32767, half2: 2048) is also listed for
// rescoring in a wider type.
template <typename V>
__device__ __forceinline__ void emit(const KArgs &a, int64_t o, int q, int r, int score) {
    a.out[o] = score;
    if (a.sat_check && score >= VOps<V>::TOP) {
        const int i = atomicAdd(a.sat_count, 1);
        if (i < a.sat_cap) a.sat[i] = SatEntry{q, r, o};
    }
}

// ---------------------------------------------------------------- cross kernel
// A workgroup takes (query pair, reference range) items from a global counter, builds
// the pair's packed profile in LDS once, then its waves take blocks of 64/G references
// from an LDS counter; each group of G lanes sweeps one reference.
// Occupancy target: 4 waves per SIMD (<= 128 VGPRs) for the narrow local tiles (G <= 8,
// R <= 34), whose loop fits in 128 registers (R 33 and 34 spill two registers outside the
// loop); the other tiles are left free.
template <bool LOCAL, int G, int R> struct WavesPerEU { static constexpr int value = (LOCAL && G <= 8 && R <= 34) ? 4 : 1; };

template <typename V, int G, int R, bool LOCAL, int WAVES>
__global__ void __launch_bounds__(WAVES * 64)
__attribute__((amdgpu_waves_per_eu(WavesPerEU<LOCAL, G, R>::value))) cross_kernel(KArgs a) {
    using O = VOps<V>;
    using L = ProfLayout<G, R>;
    constexpr int LW = L::KC * 4;   // padded columns per lane
    constexpr int GPW = 64 / G;
    extern __shared__ __align__(16) char lds[];
    __shared__ int s_item, s_blk;
    const int lane = threadIdx.x & 63;
    const int t = GroupLanes<G>::t_of(lane), g = GroupLanes<G>::g_of(lane);
    const CrossItem *items = static_cast<const CrossItem *>(a.items);
    for (;;) {
        if (threadIdx.x == 0) {
            s_item = atomicAdd(a.counter, 1);
            s_blk = 0;
        }
        __syncthreads();
        const int it = __builtin_amdgcn_readfirstlane(s_item);
        if (it >= a.nitems) break;
        const CrossItem item = items[it];
        const int qsa = item.qsa, lqa = item.lqa;
        const int qsb = item.lqb >= 0 ? item.qsb : qsa, lqb = item.lqb >= 0 ? item.lqb : lqa;
        const int total = a.ns * G * LW;
        for (int idx = threadIdx.x; idx < total; idx += blockDim.x) {
            const int c = idx / (G * LW);
            const int rem = idx - c * (G * LW);
            const int lt = rem / LW, kk = rem - lt * LW;
            const int col = lt * R + kk - item.shift;   // front padding (col < 0) scores 0
            const int va = kk >= R || col >= lqa ? O::NEG
                           : prof_score<V, LOCAL>(a.k, col < 0 ? 0 : a.sub[c * a.ns + min((int)a.qcodes[qsa + col], a.ns - 1)]);
            const int vb = kk >= R || col >= lqb ? O::NEG
                           : prof_score<V, LOCAL>(a.k, col < 0 ? 0 : a.sub[c * a.ns + min((int)a.qcodes[qsb + col], a.ns - 1)]);
            const uint32_t word = O::pack(va, vb);
            char *dst = lds + c * L::ROWB + (kk >> 2) * L::CH + lt * 16 + (kk & 3) * 4;
#pragma unroll
            for (int cp = 0; cp < L::NSLOT; ++cp)   // one copy per group position in an LDS pass
                *reinterpret_cast<uint32_t *>(dst + cp * G * 16) = word;
        }
        __syncthreads();
        const char *prof_lane = lds + L::lane_offset(g, t);
        for (;;) {
            int blk = 0;
            if (lane == 0) blk = atomicAdd(&s_blk, 1);
            blk = __builtin_amdgcn_readfirstlane(blk);
            const int r0 = item.rbeg + blk * GPW;
            if (r0 >= item.rend) break;
            const int rpos = r0 + g;
            const int ref = rpos < item.rend ? a.rorder[rpos] : -1;
            const int rs = ref >= 0 ? a.roff[ref] : 0;
            const int lr = ref >= 0 ? a.roff[ref + 1] - rs : 0;
            const int nsteps = __builtin_amdgcn_readfirstlane(wave_max_i32(lr)) + G - 1;
            const SweepOut res = sweep_group<V, G, R, LOCAL>(a.rcodes + rs, lr, nsteps, prof_lane, t,
                                                             a.k, lqa, lqb, -item.shift);
            if (t == 0 && ref >= 0) {
                emit<V>(a, item.outa + ref, -1, ref, res.a);
                if (item.outb >= 0) emit<V>(a, item.outb + ref, -1, ref, res.b);
            }
        }
        __syncthreads();
    }
}

// ---------------------------------------------------------------- stripe kernel
// Queries wider than a tile: the query is cut into stripes of W = G*R columns, and a
// workgroup takes (query pair, reference range) items as the cross kernel does but runs
// the item stripe by stripe: build the stripe's profile, let the waves sweep all the
// item's references through it, then move right.  Each reference row's H and E at the
// stripe edge are kept in the workgroup's boundary region of global memory (an H plane
// and an E plane of bnd_stride words, one word per reference row of the item in sorted
// order); results of a reference are combined over the stripes with max in the output,
// and the saturation check sees the final value.
template <typename V, int G, int R, bool LOCAL, int WAVES>
__global__ void __launch_bounds__(WAVES * 64)
__attribute__((amdgpu_waves_per_eu(WavesPerEU<LOCAL, G, R>::value))) stripe_kernel(KArgs a) {
    using O = VOps<V>;
    using L = ProfLayout<G, R>;
    constexpr int LW = L::KC * 4;   // padded columns per lane
    constexpr int GPW = 64 / G;
    constexpr int W = G * R;
    extern __shared__ __align__(16) char lds[];
    __shared__ int s_item, s_blk;
    const int lane = threadIdx.x & 63;
    const int t = GroupLanes<G>::t_of(lane), g = GroupLanes<G>::g_of(lane);
    const CrossItem *items = static_cast<const CrossItem *>(a.items);
    uint32_t *bnd_wg = a.bnd + (size_t)blockIdx.x * 2 * (size_t)a.bnd_stride;
    for (;;) {
        if (threadIdx.x == 0) s_item = atomicAdd(a.counter, 1);
        __syncthreads();
        const int it = __builtin_amdgcn_readfirstlane(s_item);
        if (it >= a.nitems) break;
        const CrossItem item = items[it];
        const int qsa = item.qsa, lqa = item.lqa;
        const int qsb = item.lqb >= 0 ? item.qsb : qsa, lqb = item.lqb >= 0 ? item.lqb : lqa;
        const int64_t row_base = a.bpos[item.rbeg];
        for (int st = 0; st < item.nstr; ++st) {
            const int cbase = st * W - item.shift;
            if (threadIdx.x == 0) s_blk = 0;
            const int total = a.ns * G * LW;
            for (int idx = threadIdx.x; idx < total; idx += blockDim.x) {
                const int c = idx / (G * LW);
                const int rem = idx - c * (G * LW);
                const int lt = rem / LW, kk = rem - lt * LW;
                const int col = cbase + lt * R + kk;   // front padding (col < 0) scores 0
                const int va = kk >= R || col >= lqa ? O::NEG
                               : prof_score<V, LOCAL>(a.k, col < 0 ? 0 : a.sub[c * a.ns + min((int)a.qcodes[qsa + col], a.ns - 1)]);
                const int vb = kk >= R || col >= lqb ? O::NEG
                               : prof_score<V, LOCAL>(a.k, col < 0 ? 0 : a.sub[c * a.ns + min((int)a.qcodes[qsb + col], a.ns - 1)]);
                const uint32_t word = O::pack(va, vb);
                char *dst = lds + c * L::ROWB + (kk >> 2) * L::CH + lt * 16 + (kk & 3) * 4;
#pragma unroll
                for (int cp = 0; cp < L::NSLOT; ++cp)
                    *reinterpret_cast<uint32_t *>(dst + cp * G * 16) = word;
            }
            __syncthreads();
            const bool first = st == 0, last = st + 1 == item.nstr;
            const char *prof_lane = lds + L::lane_offset(g, t);
            for (;;) {
                int blk = 0;
                if (lane == 0) blk = atomicAdd(&s_blk, 1);
                blk = __builtin_amdgcn_readfirstlane(blk);
                const int r0 = item.rbeg + blk * GPW;
                if (r0 >= item.rend) break;
                const int rpos = r0 + g;
                const int ref = rpos < item.rend ? a.rorder[rpos] : -1;
                const int rs = ref >= 0 ? a.roff[ref] : 0;
                const int lr = ref >= 0 ? a.roff[ref + 1] - rs : 0;
                const uint32_t boff = ref >= 0 ? (uint32_t)(a.bpos[rpos] - row_base) : 0u;
                const int nsteps = __builtin_amdgcn_readfirstlane(wave_max_i32(lr)) + G - 1;
                const SweepOut res = sweep_group<V, G, R, LOCAL, true>(a.rcodes + rs, lr, nsteps, prof_lane, t, a.k,
                                                                       lqa, lqb, cbase, bnd_wg, bnd_wg + a.bnd_stride,
                                                                       !first, !last, nullptr, 0, boff);
                if (t == 0 && ref >= 0) {
                    // explicit pairs: a position is one (query A, reference) pair with an
                    // output index of its own (and no query B)
                    const int64_t oa = a.oidx ? a.oidx[rpos] : item.outa + ref;
                    int va = res.a, vb = res.b;
                    if (!first) {   // written by this workgroup in an earlier stripe
                        va = max(va, a.out[oa]);
                        if (item.outb >= 0) vb = max(vb, a.out[item.outb + ref]);
                    }
                    if (last) {
                        emit<V>(a, oa, -1, ref, va);
                        if (item.outb >= 0) emit<V>(a, item.outb + ref, -1, ref, vb);
                    } else {
                        a.out[oa] = va;
                        if (item.outb >= 0) a.out[item.outb + ref] = vb;
                    }
                }
            }
            __syncthreads();   // profile, partial results and boundaries complete
        }
    }
}

// ---------------------------------------------------------------- pairs kernel
// Waves take blocks of 64/G tasks from a global counter.  Every lane writes the
// profile entries of its own columns into the wave's LDS region and reads only those,
// so no barrier is needed.  Tasks of one block share their front-padding shift (the host
// pads blocks).  The profile holds rows 0..prow only (prow: the largest reference code of
// the set, or the table's last row), so for DNA references without N a wave's profiles of
// up to 20 columns per lane take 20 KiB and eight waves fit a CU; the sweep is pipelined
// for that occupancy (sweep_pairs_pipe).
template <typename V, int G, int R, bool LOCAL, int WAVES>
__global__ void __launch_bounds__(WAVES * 64) pairs_kernel(KArgs a) {
    using O = VOps<V>;
    using L = ProfLayout<G, R>;
    constexpr int KC = L::KC;
    constexpr int GPW = 64 / G;
    extern __shared__ __align__(16) char lds[];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int t = GroupLanes<G>::t_of(lane), g = GroupLanes<G>::g_of(lane);
    const int nrows = a.prow + 1;
    // blocks of NSLOT groups interleave their private profiles (see ProfLayout)
    const int block_bytes = nrows * L::ROWB;
    char *lane_lds = lds + (size_t)wave * (GPW / L::NSLOT) * block_bytes + (size_t)(g / L::NSLOT) * block_bytes +
                     L::lane_offset(g, t);
    // tables of up to 64 entries (ns <= 8, DNA) are held across the lanes of a register,
    // lane l holding entry l, and read with ds_bpermute: no memory latency in the builds
    const bool lane_table = a.ns * a.ns <= 64;
    const int subv = lane_table && lane < a.ns * a.ns ? a.sub[lane] : 0;
    const PairTask *tasks = static_cast<const PairTask *>(a.items);
    const int nitems = a.nitems_ptr ? __builtin_amdgcn_readfirstlane(*a.nitems_ptr) : a.nitems;
    const int nblocks = (nitems + GPW - 1) / GPW;
    for (;;) {
        int blk = 0;
        if (lane == 0) blk = atomicAdd(a.counter, 1);
        blk = __builtin_amdgcn_readfirstlane(blk);
        if (blk >= nblocks) break;
        const int ti = blk * GPW + g;
        PairTask task{-1, -1, -1, 0, -1, -1};
        if (ti < nitems) task = tasks[ti];
        const int shift = __builtin_amdgcn_readfirstlane(tasks[blk * GPW].shift);
        const bool valid = task.ref >= 0;
        const int qa = valid ? task.qa : 0;
        const int qb = (valid && task.qb >= 0) ? task.qb : qa;
        const int qsa = valid ? a.qoff[qa] : 0, lqa = valid ? a.qoff[qa + 1] - qsa : 0;
        const int qsb = valid ? a.qoff[qb] : 0, lqb = valid ? a.qoff[qb + 1] - qsb : 0;
        const int rs = valid ? a.roff[task.ref] : 0;
        const int lr = valid ? a.roff[task.ref + 1] - rs : 0;
        // this lane's profile columns: all query codes are read first (one round trip), then
        // each row's entries are looked up and written 4 columns at a time
        int ca[KC * 4], cb[KC * 4];   // code, or -1 = back padding (NEG), -2 = front padding (0)
#pragma unroll
        for (int kk = 0; kk < KC * 4; ++kk) {
            const int col = t * R + kk - shift;
            ca[kk] = (kk >= R || col >= lqa) ? -1 : col < 0 ? -2 : min((int)a.qcodes[qsa + col], a.ns - 1);
            cb[kk] = (kk >= R || col >= lqb) ? -1 : col < 0 ? -2 : min((int)a.qcodes[qsb + col], a.ns - 1);
        }
        auto entry = [&](int c, int code) -> int {
            const int v = lane_table ? __builtin_amdgcn_ds_bpermute((c * a.ns + max(code, 0)) * 4, subv)
                                     : a.sub[c * a.ns + max(code, 0)];
            return code == -1 ? O::NEG : prof_score<V, LOCAL>(a.k, code == -2 ? 0 : v);
        };
#pragma unroll 1
        for (int c = 0; c < nrows; ++c) {
#pragma unroll
            for (int kc = 0; kc < KC; ++kc) {
                uint32_t w[4];
#pragma unroll
                for (int e = 0; e < 4; ++e) w[e] = O::pack(entry(c, ca[kc * 4 + e]), entry(c, cb[kc * 4 + e]));
                *reinterpret_cast<int4 *>(lane_lds + c * L::ROWB + kc * L::CH) =
                    int4{(int)w[0], (int)w[1], (int)w[2], (int)w[3]};
            }
        }
        const int nsteps = __builtin_amdgcn_readfirstlane(wave_max_i32(lr)) + G - 1;
        const SweepOut res = sweep_pairs_pipe<V, G, R, LOCAL>(a.rcodes + rs, lr, nsteps, lane_lds, t, a.k, lqa, lqb,
                                                              -shift, a.prow);
        if (t == 0 && valid) {
            emit<V>(a, task.oa, qa, task.ref, res.a);
            if (task.ob >= 0) emit<V>(a, task.ob, qb, task.ref, res.b);
        }
    }
}

// ---------------------------------------------------------------- pairs table kernel
// Explicit pairs with the substitution table in LDS instead of per-lane profiles, for
// alphabets whose profiles do not fit (protein beyond 160 columns) and to free LDS for
// occupancy.  The workgroup builds T[c][a][b] = (sub[c][a] | sub[c][b] << 16) once, where
// a, b run over the query codes plus two padding codes (NEG past a query's end, 0 for
// front padding); int32 tiles use T[c][a].  Each lane keeps its R columns' byte offsets
// into a table row in registers, so a column costs one LDS load and an address add.
// Small tables are stored in pairs_tab_copies() interleaved copies (word w of copy c at
// w*copies + c) and lane l reads copy l mod copies: the lanes of an LDS pass then hit
// distinct banks whatever entries they look up.
//
// PERM (binary16, at most 6 codes, every entry's low byte zero -- half_perm_ok on the host):
// the table is 8 bytes per reference code, byte k the high byte of the entry for query code
// k (k = ns: past the end, k = ns + 1: front padding), and each lane keeps one v_perm_b32
// selector per column instead of a table offset: a row costs one 8-byte LDS load and each
// column one VALU op, with no per-column load to wait for.
template <typename V, int G, int R, bool LOCAL, int WAVES, bool PERM = false>
__global__ void __launch_bounds__(WAVES * 64) pairs_tab_kernel(KArgs a) {
    using O = VOps<V>;
    constexpr int GPW = 64 / G;
    extern __shared__ __align__(16) char lds[];
    const int lane = threadIdx.x & 63;
    const int t = GroupLanes<G>::t_of(lane), g = GroupLanes<G>::g_of(lane);
    const int ns = a.ns, nq = ns + 2, npad = ns, zpad = ns + 1;
    const int row_words = O::PACKED ? nq * nq : nq;
    const int copies = a.tab_copies;
    if constexpr (PERM) {
        static_assert(O::MAX3, "byte-select tables hold binary16 entries");
        for (int idx = threadIdx.x; idx < ns * 8; idx += blockDim.x) {
            const int c = idx >> 3, kq = idx & 7;
            const int v = kq < ns ? prof_score<V, LOCAL>(a.k, a.sub[c * ns + kq])
                          : kq == npad ? PERM_NEG : prof_score<V, LOCAL>(a.k, 0);
            lds[idx] = (char)(O::pack(v, v) >> 8);   // the high byte of v's binary16 bits
        }
    } else {
        for (int idx = threadIdx.x; idx < ns * row_words * copies; idx += blockDim.x) {
            const int w = idx / copies;
            const int c = w / row_words, rem = w - c * row_words;
            const int qa = O::PACKED ? rem / nq : rem, qb = O::PACKED ? rem - (rem / nq) * nq : rem;
            const int va = qa == npad ? O::NEG : prof_score<V, LOCAL>(a.k, qa == zpad ? 0 : a.sub[c * ns + qa]);
            const int vb = qb == npad ? O::NEG : prof_score<V, LOCAL>(a.k, qb == zpad ? 0 : a.sub[c * ns + qb]);
            reinterpret_cast<uint32_t *>(lds)[idx] = O::pack(va, vb);
        }
    }
    __syncthreads();
    const int trow = row_words * copies * 4;
    const char *tab_lane = PERM ? lds : lds + (lane % copies) * 4;
    const PairTask *tasks = static_cast<const PairTask *>(a.items);
    const int nitems = a.nitems_ptr ? __builtin_amdgcn_readfirstlane(*a.nitems_ptr) : a.nitems;
    const int nblocks = (nitems + GPW - 1) / GPW;
    for (;;) {
        int blk = 0;
        if (lane == 0) blk = atomicAdd(a.counter, 1);
        blk = __builtin_amdgcn_readfirstlane(blk);
        if (blk >= nblocks) break;
        const int ti = blk * GPW + g;
        PairTask task{-1, -1, -1, 0, -1, -1};
        if (ti < nitems) task = tasks[ti];
        const int shift = __builtin_amdgcn_readfirstlane(tasks[blk * GPW].shift);
        const bool valid = task.ref >= 0;
        const int qa = valid ? task.qa : 0;
        const int qb = (valid && task.qb >= 0) ? task.qb : qa;
        const int qsa = valid ? a.qoff[qa] : 0, lqa = valid ? a.qoff[qa + 1] - qsa : 0;
        const int qsb = valid ? a.qoff[qb] : 0, lqb = valid ? a.qoff[qb + 1] - qsb : 0;
        const int rs = valid ? a.roff[task.ref] : 0;
        const int lr = valid ? a.roff[task.ref + 1] - rs : 0;
        int toff[R];
#pragma unroll
        for (int kk = 0; kk < R; ++kk) {
            const int col = t * R + kk - shift;
            const int ca = col >= lqa ? npad : col < 0 ? zpad : min((int)a.qcodes[qsa + col], ns - 1);
            const int cb = col >= lqb ? npad : col < 0 ? zpad : min((int)a.qcodes[qsb + col], ns - 1);
            // byte-select: bytes 0 and 2 zero (selector 12), bytes 1 and 3 the table bytes of
            // the two codes (selector k picks byte k of the row's 8)
            toff[kk] = PERM ? 0x0C | ca << 8 | 0x0C << 16 | cb << 24 : (O::PACKED ? ca * nq + cb : ca) * copies * 4;
        }
        const int nsteps = __builtin_amdgcn_readfirstlane(wave_max_i32(lr)) + G - 1;
        const SweepOut res = sweep_group<V, G, R, LOCAL, false, true, true, PERM>(
            a.rcodes + rs, lr, nsteps, tab_lane, t, a.k, lqa, lqb, -shift, nullptr, nullptr, false, false, toff, trow);
        if (t == 0 && valid) {
            emit<V>(a, task.oa, qa, task.ref, res.a);
            if (task.ob >= 0) emit<V>(a, task.ob, qb, task.ref, res.b);
        }
    }
}

// ---------------------------------------------------------------- variants

template <typename V, int G, int R, bool L>
static void launch_cross(const KArgs &a, int grid, size_t lds, hipStream_t st) {
    hipLaunchKernelGGL((cross_kernel<V, G, R, L, CROSS_WAVES>), dim3(grid), dim3(CROSS_WAVES * 64), lds, st, a);
}
template <typename V, int G, int R, bool L>
static void launch_pairs(const KArgs &a, int grid, size_t lds, hipStream_t st) {
    hipLaunchKernelGGL((pairs_kernel<V, G, R, L, PAIRS_WAVES>), dim3(grid), dim3(PAIRS_WAVES * 64), lds, st, a);
}

template <typename V, int G, int R, bool L>
static void launch_stripe(const KArgs &a, int grid, size_t lds, hipStream_t st) {
    hipLaunchKernelGGL((stripe_kernel<V, G, R, L, CROSS_WAVES>), dim3(grid), dim3(CROSS_WAVES * 64), lds, st, a);
}
// One wave per workgroup (and per item): explicit long pairs, whose items hold the few
// references paired with one query, so that a workgroup's other waves do not sit idle.
template <typename V, int G, int R, bool L>
static void launch_stripe1(const KArgs &a, int grid, size_t lds, hipStream_t st) {
    hipLaunchKernelGGL((stripe_kernel<V, G, R, L, CROSS_WAVES>), dim3(grid), dim3(64), lds, st, a);
}

#define GSW_VARIANT(V, KIND, G, R, L)                                                             \
    Variant{0, G, R, KIND != KIND_I32, L,                                                         \
            reinterpret_cast<const void *>(&cross_kernel<V, G, R, L, CROSS_WAVES>),               \
            reinterpret_cast<const void *>(&pairs_kernel<V, G, R, L, PAIRS_WAVES>), nullptr,      \
            &launch_cross<V, G, R, L>, &launch_pairs<V, G, R, L>, nullptr,                        \
            reinterpret_cast<const void *>(&pairs_tab_kernel<V, G, R, L, PAIRS_TAB_WAVES>), KIND}
// local and global/semiglobal rows of a tile for one element type
#define GSW_VARIANT2(V, KIND, G, R) GSW_VARIANT(V, KIND, G, R, true), GSW_VARIANT(V, KIND, G, R, false)
// tiles that also run stripes of longer queries
#define GSW_SVARIANT(V, KIND, G, R, L)                                                            \
    Variant{0, G, R, KIND != KIND_I32, L,                                                         \
            reinterpret_cast<const void *>(&cross_kernel<V, G, R, L, CROSS_WAVES>),               \
            reinterpret_cast<const void *>(&pairs_kernel<V, G, R, L, PAIRS_WAVES>),               \
            reinterpret_cast<const void *>(&stripe_kernel<V, G, R, L, CROSS_WAVES>),              \
            &launch_cross<V, G, R, L>, &launch_pairs<V, G, R, L>, &launch_stripe<V, G, R, L>,     \
            reinterpret_cast<const void *>(&pairs_tab_kernel<V, G, R, L, PAIRS_TAB_WAVES>), KIND,  \
            &launch_stripe1<V, G, R, L>}
#define GSW_SVARIANT2(V, KIND, G, R) GSW_SVARIANT(V, KIND, G, R, true), GSW_SVARIANT(V, KIND, G, R, false)
// a tile for cross products only (no explicit-pairs kernels), and one that also runs stripes
#define GSW_XVARIANT(V, KIND, G, R, L)                                                            \
    Variant{0, G, R, KIND != KIND_I32, L,                                                         \
            reinterpret_cast<const void *>(&cross_kernel<V, G, R, L, CROSS_WAVES>), nullptr,      \
            nullptr, &launch_cross<V, G, R, L>, nullptr, nullptr, nullptr, KIND}
#define GSW_XSVARIANT(V, KIND, G, R, L)                                                           \
    Variant{0, G, R, KIND != KIND_I32, L,                                                         \
            reinterpret_cast<const void *>(&cross_kernel<V, G, R, L, CROSS_WAVES>), nullptr,      \
            reinterpret_cast<const void *>(&stripe_kernel<V, G, R, L, CROSS_WAVES>),              \
            &launch_cross<V, G, R, L>, nullptr, &launch_stripe<V, G, R, L>, nullptr, KIND}

// The rows of one element type over every compiled tile, in the table's tile order (each
// element type's instantiations compile in a translation unit of their own).
#define GSW_TILE_ROWS(V, KIND)                                                                    \
    GSW_VARIANT2(V, KIND, 8, 8), GSW_VARIANT2(V, KIND, 8, 12), GSW_VARIANT2(V, KIND, 8, 16),      \
        GSW_VARIANT2(V, KIND, 8, 19), GSW_SVARIANT2(V, KIND, 8, 20), GSW_SVARIANT2(V, KIND, 8, 24), \
        GSW_VARIANT2(V, KIND, 8, 30), GSW_VARIANT2(V, KIND, 8, 31), GSW_SVARIANT2(V, KIND, 8, 32), \
        GSW_VARIANT2(V, KIND, 8, 33), GSW_VARIANT2(V, KIND, 8, 34), GSW_VARIANT2(V, KIND, 8, 36),  \
        GSW_VARIANT2(V, KIND, 8, 40), GSW_VARIANT2(V, KIND, 8, 44), GSW_VARIANT2(V, KIND, 8, 48),  \
        GSW_VARIANT2(V, KIND, 16, 24), GSW_VARIANT2(V, KIND, 16, 28), GSW_VARIANT2(V, KIND, 16, 32), \
        GSW_VARIANT2(V, KIND, 16, 36), GSW_VARIANT2(V, KIND, 16, 40), GSW_VARIANT2(V, KIND, 16, 44), \
        GSW_VARIANT2(V, KIND, 16, 48), GSW_VARIANT2(V, KIND, 32, 32), GSW_VARIANT2(V, KIND, 32, 48), \
        GSW_VARIANT2(V, KIND, 64, 32)

}  // namespace gsw
