// Traceback of chosen pairs: the alignment behind each score (coordinates, CIGAR runs,
// match/mismatch/gap counts).  The reference scores only and re-aligns its top hits on
// the CPU (reference docs/DESIGN.md:94-96, examples/legacy_16s_gpu_align.py:280-292);
// here the hits are aligned on the GPU.
//
// The DP fill (16 or 32 lanes per pair for queries up to 2,048 columns, one thread per
// pair beyond) is the comparator's int32 recurrence (src/gpusw/kernel.py:82-157), with
// each cell's provenance packed into a direction byte
//   bits 0-1  H source: 0 = zero (local start), 1 = diagonal, 2 = E, 3 = F
//   bit  2    E(i,j) extends E(i,j-1) (else opens from H(i,j-1))
//   bit  3    F(i,j) extends F(i-1,j) (else opens from H(i-1,j))
// stored four cells per 32-bit word in the pair's own region (rows of ceil(lq/4) words).
// Ties: H prefers zero (local), then diagonal, then E, then F; E and F prefer opening.
// End cell: local = first strict maximum in row-major order; global = (lr, lq);
// semiglobal = the corner, else the first best row of the last column (free query ends),
// else the first best column of the last row (free reference ends), in that order on ties.
// The walk then follows the bits back; free end gaps are left out of the CIGAR and show
// in the coordinates, other boundary gaps are runs of I (query residues) or D (reference).
#include <algorithm>
#include <chrono>

#include "gsw_internal.h"

namespace gsw {

enum : uint32_t { OP_EQ = 0, OP_X = 1, OP_I = 2, OP_D = 3 };   // run = (len << 2) | op

struct TbArgs {
    const uint8_t *qcodes;
    const int32_t *qoff;
    const uint8_t *rcodes;
    const int32_t *roff;
    const int32_t *qi, *rj;
    const int64_t *dir_off;    // byte offset of each pair's direction region
    const int64_t *ops_off;    // first CIGAR run of each pair
    int64_t n;
    const int32_t *sub;
    int32_t ns, mode, go, ge, free_q, free_r;
    uint32_t *dir;
    int32_t *ws;               // H and F rows, column-major across threads
    int64_t stride;            // threads in the grid
    int32_t maxq;
    gsw_aln *out;
    uint32_t *ops;
};

constexpr int NEG32 = -(1 << 30);   // the oracle's sentinel

// Appends runs (the walk goes end to start), holding the last one in a register so that
// extending it costs no memory access.  Residue columns always merge with the run before;
// a gap merges only while it continues (merge), so two separate gaps that happen to touch
// stay two runs and the CIGAR's affine cost is the score.
struct RunSink {
    uint32_t *ops;
    int32_t nops = 0;   // runs stored
    uint32_t cur = 0;   // the last run, (length << 2) | op; 0: none yet
    __device__ void push(uint32_t op, int32_t len, bool merge) {
        if (len <= 0) return;
        if (merge && cur && (cur & 3u) == op) {
            cur += (uint32_t)len << 2;
            return;
        }
        if (cur) ops[nops++] = cur;
        cur = ((uint32_t)len << 2) | op;
    }
    __device__ int32_t finish() {
        if (cur) ops[nops++] = cur;
        cur = 0;
        return nops;
    }
};

// What a fill leaves for the walk: the local best (first strict maximum in row-major
// order) and its cell, the corner, the best of the last column (first such row) and of
// the last row (first such column).
struct TbEnd { int32_t best, bi, bj, corner, ec, eci, er, erj; };

// End cell, score and walk of pair k from its fill (direction bytes in dir, trackers e);
// writes a.out[k] and the pair's CIGAR runs.
__device__ void finish_pair(const TbArgs &a, int64_t k, int lq, int lr, const uint8_t *q, const uint8_t *r,
                            const uint32_t *dir, const TbEnd &e) {
    const int go = a.go, ge = a.ge, ns = a.ns;
    const bool local = a.mode == GSW_MODE_LOCAL, semi = a.mode == GSW_MODE_SEMIGLOBAL;
    const bool ladder_r = a.mode == GSW_MODE_GLOBAL || (semi && !a.free_r);
    const bool ladder_q = a.mode == GSW_MODE_GLOBAL || (semi && !a.free_q);
    const int w = (lq + 3) / 4;
    RunSink sink{a.ops + a.ops_off[k]};
    gsw_aln res{};
    // ---- end cell and score (kernel.py:141-157 for the score)
    int ei = 0, ej0 = 0;
    bool empty = false;
    if (lq == 0 || lr == 0) {
        empty = true;   // no aligned columns; the score is the edge formula
        if (local || (lq == 0 && lr == 0)) res.score = 0;
        else if (lq == 0) res.score = (semi && a.free_q) ? 0 : go + (lr - 1) * ge;
        else res.score = (semi && a.free_r) ? 0 : go + (lq - 1) * ge;
        if (!local) {   // a whole-sequence gap unless it is a free end
            if (lq == 0 && lr > 0 && !(semi && a.free_q)) { sink.push(OP_D, lr, false); res.del = lr; res.gap_opens = 1; }
            if (lr == 0 && lq > 0 && !(semi && a.free_r)) { sink.push(OP_I, lq, false); res.ins = lq; res.gap_opens = 1; }
            res.q_end = (semi && a.free_r && lr == 0) ? 0 : lq;
            res.r_end = (semi && a.free_q && lq == 0) ? 0 : lr;
        }
    } else if (local) {
        res.score = e.best;
        ei = e.bi; ej0 = e.bj;
        empty = e.best == 0;
    } else if (!semi) {
        res.score = e.corner;
        ei = lr; ej0 = lq;
    } else {
        int sc = e.corner;
        ei = lr; ej0 = lq;
        if (a.free_q && e.ec > sc) { sc = e.ec; ei = e.eci; ej0 = lq; }
        if (a.free_r && e.er > sc) { sc = e.er; ei = lr; ej0 = e.erj; }
        if (a.free_q && a.free_r && 0 > sc) { sc = 0; empty = true; }
        res.score = sc;
    }
    // ---- walk back from (ei, ej0)
    if (!empty) {
        res.q_end = ej0;
        res.r_end = ei;
        int i = ei, j = ej0, state = 0;   // 0 = H, 1 = E, 2 = F
        bool cont = false;                // the next gap step continues the current gap
        // Diagonal steps are the common ones and each is a dependent read of a direction
        // byte (an HBM miss, rows lie lq bytes apart): the words of the next LA cells on
        // the diagonal are loaded together, la[d] for cell (i - d, j - d), d < la_n.
        constexpr int LA = 8;
        uint32_t la[LA];
        int la_n = 0;
        for (;;) {
            if (state == 0) {
                if (i == 0 || j == 0) {
                    if (i == 0 && j > 0) {
                        if (ladder_r) { sink.push(OP_I, j, false); res.ins += j; res.gap_opens++; j = 0; }
                    } else if (j == 0 && i > 0) {
                        if (ladder_q) { sink.push(OP_D, i, false); res.del += i; res.gap_opens++; i = 0; }
                    }
                    break;
                }
                if (la_n == 0) {
                    la_n = min(LA, min(i, j));
#pragma unroll
                    for (int d = 0; d < LA; ++d)
                        if (d < la_n) la[d] = dir[(int64_t)(i - 1 - d) * w + (j - 1 - d) / 4];
                }
                const uint32_t bits = (la[0] >> (8 * ((j - 1) & 3))) & 0xFFu;
                const uint32_t src = bits & 3u;
                if (src == 0) break;   // local start
                if (src == 1) {
                    const bool eq = min((int)q[j - 1], ns - 1) == min((int)r[i - 1], ns - 1) &&
                                    (int)q[j - 1] < ns - 1;
                    sink.push(eq ? OP_EQ : OP_X, 1, true);
                    if (eq) res.matches++; else res.mismatches++;
                    --i; --j;
#pragma unroll
                    for (int d = 0; d + 1 < LA; ++d) la[d] = la[d + 1];
                    --la_n;
                } else {
                    state = src == 2 ? 1 : 2;
                    res.gap_opens++;
                    cont = false;
                    la_n = 0;   // the path leaves the diagonal
                }
            } else if (state == 1) {   // E: query residue j against a gap
                const uint32_t bits = (dir[(int64_t)(i - 1) * w + (j - 1) / 4] >> (8 * ((j - 1) & 3))) & 0xFFu;
                sink.push(OP_I, 1, cont);
                res.ins++;
                cont = (bits & 4u) != 0;
                if (!cont) state = 0;
                --j;
            } else {                   // F: reference residue i against a gap
                const uint32_t bits = (dir[(int64_t)(i - 1) * w + (j - 1) / 4] >> (8 * ((j - 1) & 3))) & 0xFFu;
                sink.push(OP_D, 1, cont);
                res.del++;
                cont = (bits & 8u) != 0;
                if (!cont) state = 0;
                --i;
            }
        }
        res.q_begin = j;
        res.r_begin = i;
    }
    // runs were pushed end to start
    const int32_t nops = sink.finish();
    uint32_t *ops = sink.ops;
    for (int x = 0, y = nops - 1; x < y; ++x, --y) {
        const uint32_t t = ops[x];
        ops[x] = ops[y];
        ops[y] = t;
    }
    res.n_ops = nops;
    res.ops_offset = a.ops_off[k];
    a.out[k] = res;
}

__global__ void __launch_bounds__(256) traceback_kernel(TbArgs a) {
    __shared__ int32_t s_sub[64 * 64];
    for (int i = threadIdx.x; i < a.ns * a.ns; i += blockDim.x) s_sub[i] = a.sub[i];
    __syncthreads();
    const int64_t tid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (tid >= a.stride) return;
    int32_t *hp = a.ws + tid;
    int32_t *fp = hp + (int64_t)(a.maxq + 1) * a.stride;
    const int go = a.go, ge = a.ge, ns = a.ns;
    const bool local = a.mode == GSW_MODE_LOCAL, semi = a.mode == GSW_MODE_SEMIGLOBAL;
    const bool ladder_r = a.mode == GSW_MODE_GLOBAL || (semi && !a.free_r);   // row 0 is a gap ladder
    const bool ladder_q = a.mode == GSW_MODE_GLOBAL || (semi && !a.free_q);   // column 0 is a gap ladder
    for (int64_t k = tid; k < a.n; k += a.stride) {
        const int qs = a.qoff[a.qi[k]], lq = a.qoff[a.qi[k] + 1] - qs;
        const int rs = a.roff[a.rj[k]], lr = a.roff[a.rj[k] + 1] - rs;
        const uint8_t *q = a.qcodes + qs, *r = a.rcodes + rs;
        uint32_t *dir = a.dir + a.dir_off[k] / 4;
        const int w = (lq + 3) / 4;   // words per direction row
        // ---- fill
        int best = 0, bi = 0, bj = 0;                                   // local
        int corner = NEG32, ec = NEG32, eci = 0, er = NEG32, erj = 0;   // semiglobal / global
        hp[0] = 0;
        for (int j = 1; j <= lq; ++j) {
            hp[(int64_t)j * a.stride] = ladder_r ? go + (j - 1) * ge : 0;
            fp[(int64_t)j * a.stride] = NEG32;
        }
        for (int i = 1; i <= lr; ++i) {
            const int32_t *srow = s_sub + min((int)r[i - 1], ns - 1) * ns;
            int hleft = ladder_q ? go + (i - 1) * ge : 0;
            int diag = hp[0];
            hp[0] = hleft;
            int eleft = NEG32;
            uint32_t word = 0;
            for (int j = 1; j <= lq; ++j) {
                const int64_t o = (int64_t)j * a.stride;
                const int hup = hp[o], fup = fp[o];
                const int fo = hup + go, fx = fup + ge;
                const int eo = hleft + go, ex = eleft + ge;
                const int fj = max(fo, fx), ej = max(eo, ex);
                const int d = diag + srow[min((int)q[j - 1], ns - 1)];
                int h = max(max(d, ej), fj);
                uint32_t src;
                if (local && h <= 0) { h = 0; src = 0; }
                else src = h == d ? 1u : h == ej ? 2u : 3u;
                const uint32_t bits = src | (ex > eo ? 4u : 0u) | (fx > fo ? 8u : 0u);
                word |= bits << (8 * ((j - 1) & 3));
                if (((j - 1) & 3) == 3 || j == lq) {
                    dir[(int64_t)(i - 1) * w + (j - 1) / 4] = word;
                    word = 0;
                }
                if (local) {
                    if (h > best) { best = h; bi = i; bj = j; }
                } else {
                    if (i == lr && j == lq) corner = h;
                    if (j == lq && h > ec) { ec = h; eci = i; }
                    if (i == lr && h > er) { er = h; erj = j; }
                }
                diag = hup;
                hp[o] = h;
                fp[o] = fj;
                hleft = h;
                eleft = ej;
            }
        }
        finish_pair(a, k, lq, lr, q, r, dir, TbEnd{best, bi, bj, corner, ec, eci, er, erj});
    }
}

// Wave-parallel fill for queries of at most TB_G * R columns: TB_G lanes per pair, lane t
// owns query columns [tR, tR + R) and processes row s - t + 1 at step s (a diagonal
// pipeline), taking H and E of the column left of its block, for that row, from lane t - 1
// (one step earlier) and H of the row above from the step before.  Each cell gets the
// values and direction byte of the one-thread fill above, and each lane keeps the trackers
// of its columns in the same scan order; the group reduces them with ties to the first
// cell in row-major order (the one-thread fill's order), so the walk (traceback_walk_kernel)
// starts from the same end cell.
template <int TB_G, int R, bool LOCAL>
__global__ void __launch_bounds__(256) traceback_fill_kernel(TbArgs a, TbEnd *ends) {
    static_assert(R % 4 == 0, "direction words hold four cells");
    static_assert(TB_G == 16 || TB_G == 32, "lane groups of 16 or 32");
    __shared__ int32_t s_sub[64 * 64];
    for (int i = threadIdx.x; i < a.ns * a.ns; i += blockDim.x) s_sub[i] = a.sub[i];
    __syncthreads();
    const int t = (int)(threadIdx.x % TB_G);
    const int64_t k = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / TB_G;
    const int go = a.go, ge = a.ge, ns = a.ns;
    constexpr bool local = LOCAL;   // a compile-time mode keeps the column loop branch-free
    const bool semi = a.mode == GSW_MODE_SEMIGLOBAL;
    const bool ladder_r = a.mode == GSW_MODE_GLOBAL || (semi && !a.free_r);
    const bool ladder_q = a.mode == GSW_MODE_GLOBAL || (semi && !a.free_q);
    int lq = 0, lr = 0;
    const uint8_t *q = a.qcodes, *r = a.rcodes;
    uint32_t *dir = a.dir;
    if (k < a.n) {   // groups past the last pair idle with lq = lr = 0
        const int qs = a.qoff[a.qi[k]], rs = a.roff[a.rj[k]];
        lq = a.qoff[a.qi[k] + 1] - qs;
        lr = a.roff[a.rj[k] + 1] - rs;
        q += qs;
        r += rs;
        dir += a.dir_off[k] / 4;
    }
    const int w = (lq + 3) / 4;
    const int j0 = t * R;   // columns j0 + 1 .. j0 + R (1-based)
    int qc[R], hup[R], fu[R];
#pragma unroll
    for (int kk = 0; kk < R; ++kk) {
        const int j = j0 + kk + 1;
        qc[kk] = j <= lq ? min((int)q[j - 1], ns - 1) : 0;
        hup[kk] = (j <= lq && ladder_r) ? go + (j - 1) * ge : 0;
        fu[kk] = NEG32;
    }
    TbEnd e{0, 0, 0, NEG32, NEG32, 0, NEG32, 0};
    int hin = 0, ein = NEG32, din = 0;   // H(i, j0), E(i, j0), H(i-1, j0) from the left
    int nsteps = lr > 0 ? lr + TB_G - 1 : 0;
    for (int off = 32; off >= 1; off >>= 1) nsteps = max(nsteps, __shfl_xor(nsteps, off, 64));
    for (int s = 0; s < nsteps; ++s) {
        const int i = s - t + 1;
        if (t == 0) {   // column 0
            hin = ladder_q ? go + (i - 1) * ge : 0;
            din = i <= 1 ? 0 : (ladder_q ? go + (i - 2) * ge : 0);
            ein = NEG32;
        } else if (i == 1) {   // row 0 of column j0 (lane t - 1 was idle on it)
            din = ladder_r ? go + (j0 - 1) * ge : 0;
        }
        int hout = 0, eout = NEG32;
        if (i >= 1 && i <= lr && j0 < lq) {
            const int32_t *srow = s_sub + min((int)r[i - 1], ns - 1) * ns;
            int hleft = hin, eleft = ein, diag = din;
            uint32_t word = 0;
            // Every column is computed, those past lq too (their values and direction bits
            // are read by nothing: the walk stays within lq and the lanes to the right are
            // idle), so the loop has no branches and its table reads issue together.
#pragma unroll
            for (int kk = 0; kk < R; ++kk) {
                const int j = j0 + kk + 1;
                const bool in = j <= lq;
                const int fo = hup[kk] + go, fx = fu[kk] + ge;
                const int eo = hleft + go, ex = eleft + ge;
                const int fj = max(fo, fx), ej = max(eo, ex);
                const int d = diag + srow[qc[kk]];
                int h = max(max(d, ej), fj);
                uint32_t src = h == d ? 1u : h == ej ? 2u : 3u;
                if (local && h <= 0) {
                    h = 0;
                    src = 0;
                }
                word |= (src | (ex > eo ? 4u : 0u) | (fx > fo ? 8u : 0u)) << (8 * (kk & 3));
                if (local) {
                    if (in && h > e.best) { e.best = h; e.bi = i; e.bj = j; }
                } else {
                    if (i == lr && j == lq) e.corner = h;
                    if (j == lq && h > e.ec) { e.ec = h; e.eci = i; }
                    if (in && i == lr && h > e.er) { e.er = h; e.erj = j; }
                }
                diag = hup[kk];
                hup[kk] = h;
                fu[kk] = fj;
                hleft = h;
                eleft = ej;
                if ((kk & 3) == 3) {
                    const int widx = (j0 + kk) / 4;
                    if (widx < w) dir[(int64_t)(i - 1) * w + widx] = word;
                    word = 0;
                }
            }
            hout = hleft;
            eout = eleft;
        }
        din = hin;   // H(i, j0) is the next row's diagonal
        hin = __shfl_up(hout, 1, TB_G);
        ein = __shfl_up(eout, 1, TB_G);
    }
    // group reduction; ties go to the first cell in row-major order
#pragma unroll
    for (int off = TB_G / 2; off >= 1; off >>= 1) {
        const int ob = __shfl_xor(e.best, off, TB_G), oi = __shfl_xor(e.bi, off, TB_G), oj = __shfl_xor(e.bj, off, TB_G);
        if (ob > e.best || (ob == e.best && (oi < e.bi || (oi == e.bi && oj < e.bj)))) { e.best = ob; e.bi = oi; e.bj = oj; }
        e.corner = max(e.corner, __shfl_xor(e.corner, off, TB_G));
        const int oec = __shfl_xor(e.ec, off, TB_G), oeci = __shfl_xor(e.eci, off, TB_G);
        if (oec > e.ec) { e.ec = oec; e.eci = oeci; }   // one lane holds column lq
        const int oer = __shfl_xor(e.er, off, TB_G), oerj = __shfl_xor(e.erj, off, TB_G);
        if (oer > e.er || (oer == e.er && oerj < e.erj)) { e.er = oer; e.erj = oerj; }
    }
    if (t == 0 && k < a.n) ends[k] = e;
}

// The walk of each pair after traceback_fill_kernel, one thread per pair.
__global__ void __launch_bounds__(256) traceback_walk_kernel(TbArgs a, const TbEnd *ends) {
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= a.n) return;
    const int qs = a.qoff[a.qi[k]], lq = a.qoff[a.qi[k] + 1] - qs;
    const int rs = a.roff[a.rj[k]], lr = a.roff[a.rj[k] + 1] - rs;
    finish_pair(a, k, lq, lr, a.qcodes + qs, a.rcodes + rs, a.dir + a.dir_off[k] / 4, ends[k]);
}

}  // namespace gsw

using namespace gsw;

extern "C" int gsw_traceback_pairs(const gsw_scheme *s, gsw_seqset queries, gsw_seqset refs, const int32_t *h_qi,
                                   const int32_t *h_rj, int64_t n, gsw_aln *h_out, uint32_t *h_ops, int64_t ops_cap,
                                   gsw_stats *stats) {
    auto t0 = std::chrono::steady_clock::now();
    const SeqSet *qs = reinterpret_cast<const SeqSet *>(queries);
    const SeqSet *rs = reinterpret_cast<const SeqSet *>(refs);
    if (!qs || !rs) return fail(GSW_ERR_INVALID, "null sequence set");
    if (n < 0) return fail(GSW_ERR_INVALID, "negative pair count");
    if (n > 0 && (!h_qi || !h_rj || !h_out || !h_ops)) return fail(GSW_ERR_INVALID, "null argument");
    DeviceCtx *ctx = nullptr;
    GSW_TRY(ctx_get(&ctx));
    SchemeDev sd;
    GSW_TRY(scheme_prepare(s, ctx, &sd));
    if (qs->device != ctx->device || rs->device != ctx->device)
        return fail(GSW_ERR_INVALID, "sequence set belongs to another device");
    std::lock_guard<std::mutex> lk(ctx->mu);
    if (stats) *stats = gsw_stats{};
    // per-pair regions: direction rows of ceil(lq/4) words, CIGAR runs <= lq + lr
    std::vector<int64_t> dir_off(n + 1, 0), ops_off(n + 1, 0);
    int maxq = 1;
    int64_t cells = 0;
    for (int64_t k = 0; k < n; ++k) {
        const int q = h_qi[k], r = h_rj[k];
        if (q < 0 || q >= qs->n || r < 0 || r >= rs->n)
            return fail(GSW_ERR_INVALID, "pair %lld index out of range", (long long)k);
        const int lq = qs->len[q], lr = rs->len[r];
        if (!fits_int32(sd, lq, lr))
            return fail(GSW_ERR_OVERFLOW, "score range of pair %lld exceeds the int32 kernels", (long long)k);
        dir_off[k + 1] = dir_off[k] + 4 * (int64_t)((lq + 3) / 4) * lr;
        ops_off[k + 1] = ops_off[k] + lq + lr + 1;
        maxq = std::max(maxq, lq);
        cells += (int64_t)lq * lr;
    }
    if (ops_off[n] > ops_cap) return fail(GSW_ERR_LIMIT, "CIGAR buffer holds %lld runs, %lld needed",
                                          (long long)ops_cap, (long long)ops_off[n]);
    if (n == 0) return GSW_OK;
    // launches of pairs whose direction regions total <= 1 GiB (a single larger pair alone)
    const int64_t dir_cap = (int64_t)1 << 30;
    const int threads = 256;
    // queries up to 2,048 columns: the wave-parallel fill (16 lanes per pair up to 1,024
    // columns, else 32) with the smallest block of columns per lane that covers the longest
    // query, then the walk; longer queries: one thread per pair
    static constexpr int TB_R16[] = {4, 8, 12, 16, 20, 24, 32, 48, 64};
    static constexpr int TB_R32[] = {40, 48, 64};
    int tb_g = 0, tb_r = 0;
    for (int rr : TB_R16)
        if (!tb_r && 16 * rr >= maxq) { tb_g = 16; tb_r = rr; }
    for (int rr : TB_R32)
        if (!tb_r && 32 * rr >= maxq) { tb_g = 32; tb_r = rr; }
    float total_ms = 0.f;
    int launches = 0;
    for (int64_t k0 = 0; k0 < n;) {
        int64_t k1 = k0 + 1;
        while (k1 < n && dir_off[k1 + 1] - dir_off[k0] <= dir_cap) ++k1;
        const int64_t m = k1 - k0;
        const int64_t dir_bytes = dir_off[k1] - dir_off[k0];
        int64_t nthr = std::min<int64_t>(m, (int64_t)ctx->num_cus * 1024);
        nthr = (nthr + threads - 1) / threads * threads;
        const size_t ws_bytes = tb_r ? sizeof(TbEnd) * (size_t)m : 2 * sizeof(int32_t) * (size_t)(maxq + 1) * nthr;
        const size_t meta_bytes = 2 * sizeof(int32_t) * m + 2 * sizeof(int64_t) * m;
        const size_t out_bytes = sizeof(gsw_aln) * m + sizeof(uint32_t) * (ops_off[k1] - ops_off[k0]);
        GSW_TRY(ensure_buffer(&ctx->plain_ws, &ctx->plain_ws_bytes, ws_bytes + (size_t)dir_bytes + meta_bytes + out_bytes + 1024));
        char *base = static_cast<char *>(ctx->plain_ws);
        int32_t *d_ws = reinterpret_cast<int32_t *>(base);   // H and F rows, or the fill's trackers
        uint32_t *d_dir = reinterpret_cast<uint32_t *>(base + ws_bytes);
        char *mp = base + ws_bytes + ((dir_bytes + 255) / 256) * 256;
        int32_t *d_qi = reinterpret_cast<int32_t *>(mp);
        int32_t *d_rj = d_qi + m;
        // mp is 256-byte aligned and the two int32 lists take 8m bytes: the int64 lists and the
        // records (8-byte aligned) start on an 8-byte boundary
        int64_t *d_doff = reinterpret_cast<int64_t *>(d_rj + m);
        int64_t *d_ooff = d_doff + m;
        gsw_aln *d_out = reinterpret_cast<gsw_aln *>(d_ooff + m);
        uint32_t *d_ops = reinterpret_cast<uint32_t *>(d_out + m);
        std::vector<int64_t> doff(m), ooff(m);
        for (int64_t k = k0; k < k1; ++k) {
            doff[k - k0] = dir_off[k] - dir_off[k0];
            ooff[k - k0] = ops_off[k] - ops_off[k0];
        }
        GSW_HIP_TRY(hipMemcpyAsync(d_qi, h_qi + k0, sizeof(int32_t) * m, hipMemcpyHostToDevice, ctx->stream));
        GSW_HIP_TRY(hipMemcpyAsync(d_rj, h_rj + k0, sizeof(int32_t) * m, hipMemcpyHostToDevice, ctx->stream));
        GSW_HIP_TRY(hipMemcpyAsync(d_doff, doff.data(), sizeof(int64_t) * m, hipMemcpyHostToDevice, ctx->stream));
        GSW_HIP_TRY(hipMemcpyAsync(d_ooff, ooff.data(), sizeof(int64_t) * m, hipMemcpyHostToDevice, ctx->stream));
        TbArgs a{};
        a.qcodes = qs->d_codes; a.qoff = qs->d_off; a.rcodes = rs->d_codes; a.roff = rs->d_off;
        a.qi = d_qi; a.rj = d_rj; a.dir_off = d_doff; a.ops_off = d_ooff; a.n = m;
        a.sub = sd.d_sub; a.ns = sd.ns; a.mode = sd.mode; a.go = sd.go; a.ge = sd.ge;
        a.free_q = sd.free_q; a.free_r = sd.free_r;
        a.dir = d_dir; a.ws = d_ws; a.stride = nthr; a.maxq = maxq; a.out = d_out; a.ops = d_ops;
        GSW_HIP_TRY(hipEventRecord(ctx->ev0, ctx->stream));
        if (tb_r) {
            TbEnd *d_ends = reinterpret_cast<TbEnd *>(base);
            const unsigned fill_blocks = (unsigned)((m * tb_g + threads - 1) / threads);
            const bool loc = sd.mode == GSW_MODE_LOCAL;
            switch (tb_g * 1000 + tb_r) {
#define GSW_TB_FILL(GG, RR)                                                                                          \
    case GG * 1000 + RR:                                                                                             \
        if (loc)                                                                                                     \
            hipLaunchKernelGGL((traceback_fill_kernel<GG, RR, true>), dim3(fill_blocks), dim3(threads), 0,           \
                               ctx->stream, a, d_ends);                                                              \
        else                                                                                                         \
            hipLaunchKernelGGL((traceback_fill_kernel<GG, RR, false>), dim3(fill_blocks), dim3(threads), 0,          \
                               ctx->stream, a, d_ends);                                                              \
        break;
                GSW_TB_FILL(16, 4) GSW_TB_FILL(16, 8) GSW_TB_FILL(16, 12) GSW_TB_FILL(16, 16) GSW_TB_FILL(16, 20)
                GSW_TB_FILL(16, 24) GSW_TB_FILL(16, 32) GSW_TB_FILL(16, 48) GSW_TB_FILL(16, 64)
                GSW_TB_FILL(32, 40) GSW_TB_FILL(32, 48) GSW_TB_FILL(32, 64)
#undef GSW_TB_FILL
            }
            GSW_HIP_TRY(hipGetLastError());
            hipLaunchKernelGGL(traceback_walk_kernel, dim3((unsigned)((m + threads - 1) / threads)), dim3(threads), 0,
                               ctx->stream, a, d_ends);
        } else {
            hipLaunchKernelGGL(traceback_kernel, dim3((unsigned)(nthr / threads)), dim3(threads), 0, ctx->stream, a);
        }
        GSW_HIP_TRY(hipGetLastError());
        GSW_HIP_TRY(hipEventRecord(ctx->ev1, ctx->stream));
        GSW_HIP_TRY(hipMemcpyAsync(h_out + k0, d_out, sizeof(gsw_aln) * m, hipMemcpyDeviceToHost, ctx->stream));
        GSW_HIP_TRY(hipMemcpyAsync(h_ops + ops_off[k0], d_ops, sizeof(uint32_t) * (ops_off[k1] - ops_off[k0]),
                                   hipMemcpyDeviceToHost, ctx->stream));
        GSW_HIP_TRY(hipStreamSynchronize(ctx->stream));
        float ms = 0.f;
        GSW_HIP_TRY(hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
        total_ms += ms;
        ++launches;
        for (int64_t k = k0; k < k1; ++k) h_out[k].ops_offset = ops_off[k];   // offsets into h_ops
        k0 = k1;
    }
    if (stats) {
        stats->kernel_ms = total_ms;
        stats->cells = cells;
        stats->pairs = n;
        stats->launches = launches;
        stats->variant = -1;
        stats->top_variant = -1;
        stats->wall_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    }
    return GSW_OK;
}
