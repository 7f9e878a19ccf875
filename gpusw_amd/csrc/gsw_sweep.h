// Wavefront-per-pair Gotoh kernels for gfx950 (device code, included by gsw_sweep.hip).
//
// Layout.  A "group" of G lanes scores one reference against one or two queries.  Lane t
// of the group owns query columns [t*R, t*R+R) and keeps their DP state in VGPRs:
//   hup[k] = H(i-1, j)  (row above)          fn[k] = F(i, j) (vertical gap for this row)
// The reference streams down the rows: at step s lane t processes row i = s - t, so the
// group forms a diagonal pipeline along the query.  After each step the H and E values
// of every lane's last column move one lane to the right with a DPP wave_shr:1; lane 0
// of a group receives the row's left-boundary values instead.  The diagonal input of a
// lane is the H it received one step earlier.
//
// Element type V:
//   short2 (packed): two independent alignments per lane, query A in the low half and
//                    query B in the high half, both against the group's reference; the
//                    recurrence runs on v_pk_{add,sub,max}_{i16,u16} with saturation.
//   half2 (packed):  the same two alignments carried as binary16 values, which hold every
//                    integer of [-2048, 2048] exactly; the recurrence runs on v_pk_add_f16
//                    and gfx950's three-input v_pk_maximum3_f16, so H = max(diag+s, E, F)
//                    and each clamped E/F update take one instruction instead of two.
//   int:             one alignment per lane in 32-bit.
//
// Local mode uses the clamped form: E, F and H-gap are kept >= 0 (int16: unsigned
// saturating subtraction; half2: a third operand 0 of the three-input max).  Since local
// H = max(diag+s, E, F, 0), clamping E and F at 0 leaves every H unchanged and removes the
// separate max(H, 0).  In int16, a result equal to 32767 may have saturated; in half2, a
// result of 2048 or more may have been rounded.  Either is rescored by the host.
//
// Why a half2 result below 2048 is exact (local mode): every E, F and H of a real column
// is an integer in [0, best], and the only operation whose exact value can exceed best is
// diag + s.  Sums of integers are exact up to 2048 in magnitude, and rounding is monotone
// with 2048 representable, so the first inexact sum is >= 2048 after rounding, is an
// operand of that cell's H = max3(...), and makes the lane's running maximum >= 2048.
// Columns past a query's end carry the sentinel profile value (-30000): their sums round,
// but they lie right of every real column and never feed one.  Global and semiglobal
// pairs run in half2 only when the envelope bounds every value inside [-2048, 2048].
//
// Query profile.  prof(c, col) = sub[c][query[col]] (packed A|B for short2), stored in
// LDS as 16-byte chunks; lane t's four values of chunk kc of row c sit at
//     c*ROWB + kc*CH + slot*16,   CH = max(G, 16)*16,  ROWB = KC*CH,
// where slot = (g mod 16/G)*G + t for G < 16 (g = group index in the wave) and slot = t
// otherwise.  An LDS pass serves 16 lanes x 16 bytes; with this layout the 16 lanes of
// a pass always read 16 distinct 16-byte slots of one 256-byte bank row, whatever
// reference codes they are on, so the profile reads are free of bank conflicts.  The
// cross kernel therefore keeps 16/G copies of its shared profile when G < 16; the pairs
// kernel interleaves the private profiles of the 16/G groups of a pass.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <type_traits>

namespace gsw {

typedef short s2 __attribute__((ext_vector_type(2)));
typedef unsigned short u2 __attribute__((ext_vector_type(2)));
typedef _Float16 h2 __attribute__((ext_vector_type(2)));

// Profile layout constants (bytes) of a (G, R) tile.
template <int G, int R> struct ProfLayout {
    static constexpr int KC = (R + 3) / 4;              // 16-byte chunks per lane per row
    static constexpr int NSLOT = G < 16 ? 16 / G : 1;   // groups sharing an LDS pass
    static constexpr int CH = (G < 16 ? 16 : G) * 16;   // bytes between chunks kc, kc+1
    static constexpr int ROWB = KC * CH;                // bytes per profile row c
    // byte offset of lane t of group g within its profile block
    __device__ static __forceinline__ int lane_offset(int g, int t) { return ((g % NSLOT) * G + t) * 16; }
};

template <typename V> struct VOps;

template <> struct VOps<s2> {
    static constexpr int NEG = -32768;
    static constexpr int TOP = 32767;   // a result >= TOP may have saturated
    static constexpr bool PACKED = true;
    static constexpr bool MAX3 = false;
    __device__ static __forceinline__ s2 add(s2 a, s2 b) { return __builtin_elementwise_add_sat(a, b); }
    __device__ static __forceinline__ s2 subu(s2 a, s2 b) {
        return __builtin_bit_cast(s2, __builtin_elementwise_sub_sat(__builtin_bit_cast(u2, a),
                                                                    __builtin_bit_cast(u2, b)));
    }
    __device__ static __forceinline__ s2 mx(s2 a, s2 b) { return __builtin_elementwise_max(a, b); }
    __device__ static __forceinline__ s2 mx3(s2 a, s2 b, s2 c) { return mx(mx(a, b), c); }
    __device__ static __forceinline__ short clamp16(int v) { return (short)min(max(v, -32768), 32767); }
    __device__ static __forceinline__ s2 splat(int v) { short x = clamp16(v); return s2{x, x}; }
    __device__ static __forceinline__ int bits(s2 v) { return __builtin_bit_cast(int, v); }
    __device__ static __forceinline__ s2 from_bits(int b) { return __builtin_bit_cast(s2, b); }
    __device__ static __forceinline__ int lo(s2 v) { return (int)v.x; }
    __device__ static __forceinline__ int hi(s2 v) { return (int)v.y; }
    __device__ static __forceinline__ uint32_t pack(int a, int b) {
        return (uint32_t)(uint16_t)clamp16(a) | ((uint32_t)(uint16_t)clamp16(b) << 16);
    }
};

// Binary16 pairs: exact for integers in [-2048, 2048] (see the header comment).  The
// IEEE-754-2019 maximum lowers to v_pk_maximum3_f16 with no operand canonicalisation
// (maxNum would quiet each operand first: an extra instruction per input).
template <> struct VOps<h2> {
    static constexpr int NEG = -30000;   // sentinel: exact in binary16, far below any value
    static constexpr int TOP = 2048;     // a result >= TOP may have been rounded
    static constexpr bool PACKED = true;
    static constexpr bool MAX3 = true;
    __device__ static __forceinline__ h2 add(h2 a, h2 b) { return a + b; }
    __device__ static __forceinline__ h2 mx(h2 a, h2 b) { return __builtin_elementwise_maximum(a, b); }
    __device__ static __forceinline__ h2 mx3(h2 a, h2 b, h2 c) {
        return __builtin_elementwise_maximum(__builtin_elementwise_maximum(a, b), c);
    }
    __device__ static __forceinline__ _Float16 cvt(int v) { return (_Float16)min(max(v, -30000), 30000); }
    __device__ static __forceinline__ h2 splat(int v) { const _Float16 x = cvt(v); return h2{x, x}; }
    __device__ static __forceinline__ int bits(h2 v) { return __builtin_bit_cast(int, v); }
    __device__ static __forceinline__ h2 from_bits(int b) { return __builtin_bit_cast(h2, b); }
    __device__ static __forceinline__ int lo(h2 v) { return (int)v.x; }
    __device__ static __forceinline__ int hi(h2 v) { return (int)v.y; }
    __device__ static __forceinline__ uint32_t pack(int a, int b) { return (uint32_t)bits(h2{cvt(a), cvt(b)}); }
};

template <> struct VOps<int> {
    static constexpr int NEG = -(1 << 30);   // the oracle's sentinel
    static constexpr int TOP = 0x7fffffff;
    static constexpr bool PACKED = false;
    static constexpr bool MAX3 = false;
    __device__ static __forceinline__ int add(int a, int b) { return __builtin_elementwise_add_sat(a, b); }
    __device__ static __forceinline__ int subu(int a, int b) {
        return (int)__builtin_elementwise_sub_sat((unsigned)a, (unsigned)b);
    }
    __device__ static __forceinline__ int mx(int a, int b) { return max(a, b); }
    __device__ static __forceinline__ int mx3(int a, int b, int c) { return max(max(a, b), c); }
    __device__ static __forceinline__ int splat(int v) { return v; }
    __device__ static __forceinline__ int bits(int v) { return v; }
    __device__ static __forceinline__ int from_bits(int b) { return b; }
    __device__ static __forceinline__ int lo(int v) { return v; }
    __device__ static __forceinline__ int hi(int v) { return v; }
    __device__ static __forceinline__ uint32_t pack(int a, int) { return (uint32_t)a; }
};

// DPP wave_shr:1 -- lane l receives lane l-1's value, lane 0 receives 0.
__device__ __forceinline__ int wave_shr1(int v) {
    return __builtin_amdgcn_update_dpp(0, v, 0x138, 0xf, 0xf, true);
}

__device__ __forceinline__ int wave_max_i32(int v) {
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) v = max(v, __shfl_xor(v, off, 64));
    return v;
}

__device__ __forceinline__ int wave_min_i32(int v) {
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) v = min(v, __shfl_xor(v, off, 64));
    return v;
}

// Wave lanes of a lane group.  The boundary H and E of a group's lane t come from its lane
// t-1 through a DPP shift.  G = 16: a group is one 16-lane DPP row, and row_shr:1 with
// bound_ctrl zero-fills each group's lane 0.  G = 8: the two groups of a DPP row
// interleave (lane 16*(g/2) + 2t + g%2), so row_shr:2 moves each value to the next lane of
// its own group and zero-fills both groups' lane 0.  Wider groups span DPP rows
// (wave_shr:1, lane 0 of a group fixed up by a select).  The zero fill is the local-mode
// boundary (H = E = 0 at column 0), so those tiles need no select.
template <int G> struct GroupLanes {
    static constexpr bool INTERLEAVED = G == 8;
    static constexpr bool ZERO_FILL = G == 8 || G == 16;   // the shift gives lane 0 zeros
    static constexpr int STRIDE = INTERLEAVED ? 2 : 1;     // wave-lane distance of t, t+1
    __device__ static __forceinline__ int t_of(int lane) { return INTERLEAVED ? (lane & 15) >> 1 : lane % G; }
    __device__ static __forceinline__ int g_of(int lane) {
        return INTERLEAVED ? ((lane >> 4) << 1) | (lane & 1) : lane / G;
    }
    // value of the group's lane t-1 (0 for t = 0 when ZERO_FILL)
    __device__ static __forceinline__ int shift(int v) {
        if constexpr (INTERLEAVED) return __builtin_amdgcn_update_dpp(0, v, 0x112, 0xf, 0xf, true);   // row_shr:2
        else if constexpr (G == 16) return __builtin_amdgcn_update_dpp(0, v, 0x111, 0xf, 0xf, true);  // row_shr:1
        else return wave_shr1(v);
    }
    // ZERO_FILL tiles: value of lane t-1, lane 0 keeping `old` (no bound_ctrl: a lane whose
    // source is outside its DPP row is not written)
    __device__ static __forceinline__ int shift_keep(int old, int v) {
        static_assert(ZERO_FILL, "row shifts only");
        if constexpr (INTERLEAVED) return __builtin_amdgcn_update_dpp(old, v, 0x112, 0xf, 0xf, false);
        else return __builtin_amdgcn_update_dpp(old, v, 0x111, 0xf, 0xf, false);
    }
};

template <int G>
__device__ __forceinline__ int group_max_i32(int v) {
    constexpr int S = GroupLanes<G>::STRIDE;
#pragma unroll
    for (int off = G / 2; off >= 1; off >>= 1) v = max(v, __shfl_xor(v, off * S, 64));
    return v;
}

// Scalar scheme constants of one launch.
struct SweepConst {
    int mode;          // GSW_MODE_*
    int go, ge;        // gap scores (<= 0)
    int free_q, free_r;
    int ns;            // substitution table side; codes >= ns are read as ns-1
};

// Row-0 value H(0, col) and column-0 value H(row, 0), 1-based (kernel.py:85-111).
__device__ __forceinline__ int row0_value(const SweepConst &k, int col) {
    if (col <= 0) return 0;
    if (k.mode == 1 || (k.mode == 2 && !k.free_r)) return k.go + (col - 1) * k.ge;
    return 0;
}
__device__ __forceinline__ int col0_value(const SweepConst &k, int row) {
    if (row <= 0) return 0;
    if (k.mode == 1 || (k.mode == 2 && !k.free_q)) return k.go + (row - 1) * k.ge;
    return 0;
}

// Diagonal-offset form of the global/semiglobal recurrence, used by the packed types.
// Every value of cell (i, j) -- reference row i and query column j, both 1-based, j <= 0
// on front padding -- is carried as X + (i + j)*D with D = -gap_extend.  An extension
// (+gap_extend) and a step right or down (+D) then cancel, and with P = H + gap_open + D:
//   E(i,j) = max(E(i,j-1), P(i,j-1))          F(i,j) = max(F(i-1,j), P(i-1,j))
//   H(i,j) = max(H(i-1,j-1) + s + 2D, E(i,j), F(i,j))
// which is five instructions per cell instead of seven on half2 (six instead of eight on
// int16): the profile holds s + 2D, and P is shared by the E of the next column and the F
// of the next row.  Row 0 and the column-0 boundary become gap_open + D on every row or
// column (the gap ladders) or (i + j)*D (free ends, front padding); results subtract
// (i + j)*D in int32.  The host admits a pair to a packed type only when every offset value
// of its real cells is exact in it (sweep_fits_half, sweep_fits_int16).
template <typename V, bool LOCAL> struct DiagOffset {
    static constexpr bool ON = !LOCAL && VOps<V>::PACKED;
};

// Profile entry of substitution score v (front padding: v = 0), s + 2D in the offset form.
template <typename V, bool LOCAL>
__device__ __forceinline__ int prof_score(const SweepConst &k, int v) {
    return DiagOffset<V, LOCAL>::ON ? v - 2 * k.ge : v;
}

// Past-the-end sentinel of the byte-select pairs kernel: binary16 0xF700, whose low byte is
// zero like every entry that kernel selects (see pairs_tab_kernel).
constexpr int PERM_NEG = -28672;

constexpr int RES_NONE = -0x40000000 - 7;   // below every reachable result

// H(i, j-1) + gap_open for a row's first column.  The int16 local form clamps it at 0
// (unsigned saturating subtraction of the penalty); the three-input-max form clamps E
// and F instead, so it keeps the plain sum.
template <typename V, bool LOCAL>
__device__ __forceinline__ V row_hgo(V hin, V GOV, V GOP) {
    using O = VOps<V>;
    if constexpr (LOCAL && !O::MAX3) return O::subu(hin, GOP);
    else return O::add(hin, GOV);
}

// One cell with the three-input max (column kk of a lane's R; see the header comment):
//   local      E = max(E + ge, H + go, 0)   H = max(diag + s, F, E)   F = max(F + ge, H + go, 0)
//   otherwise  E = max(E + ge, H + go)      H = max(diag + s, F, E)   F = max(F + ge, H + go)
// 4 adds and 3 instructions of max per cell; local mode folds two columns' H into the
// running maximum with one more three-input max.
template <typename V, bool LOCAL, int R>
__device__ __forceinline__ void cell_max3(int kk, V xd, V &e, V &h, V &hprev, V &best, V &hup, V &hgo, V &f,
                                          V GOV, V GEV, V ZERO) {
    using O = VOps<V>;
    if (LOCAL) {
        e = O::mx3(O::add(e, GEV), hgo, ZERO);
        h = O::mx3(xd, f, e);
        if (kk & 1) best = O::mx3(best, hprev, h);
        else if (kk == R - 1) best = O::mx(best, h);
        else hprev = h;
        hup = h;
        hgo = O::add(h, GOV);
        f = O::mx3(O::add(f, GEV), hgo, ZERO);
    } else {
        e = O::mx(O::add(e, GEV), hgo);
        h = O::mx3(xd, f, e);
        hup = h;
        hgo = O::add(h, GOV);
        f = O::mx(O::add(f, GEV), hgo);
    }
}

// One cell of the offset form (DiagOffset): hgo holds P of the cell to the left on entry
// and this cell's P on return; f holds F of this cell on entry and of the cell below on
// return.  Half2: the add of xd, the three-input max and three more instructions.
template <typename V>
__device__ __forceinline__ void cell_offset(V xd, V &e, V &h, V &hup, V &hgo, V &f, V GOD) {
    using O = VOps<V>;
    e = O::mx(e, hgo);
    h = O::MAX3 ? O::mx3(xd, f, e) : O::mx(O::mx(xd, f), e);
    hup = h;
    hgo = O::add(h, GOD);
    f = O::mx(f, hgo);
}

// Per-group results after a sweep.
struct SweepOut {
    int a, b;   // scores of query A and query B (B unused for V=int)
};

// Sweep one group's reference through the group's query profile.
//   rseq/lr    : this group's reference codes and length (lr = 0: idle group)
//   nsteps     : wave-uniform step count (max lr over the wave's groups + G - 1)
//   prof_lane  : this lane's profile base (block base + ProfLayout::lane_offset)
//   lqa/lqb    : query lengths (for padding masks and result positions)
//   cbase      : wave-uniform query column (0-based) of lane 0's first register.  It is
//                negative when the query is front-padded: semiglobal with free query ends
//                tracks the last query column every row; there H(i,0) = 0, and front
//                padding columns whose profile is 0 carry exactly H = 0 and E = gap_open,
//                the column-0 boundary, so shifting the query by (R - lq mod R) mod R puts
//                its last column on the last register of a lane.
// STRIPE: the group covers query columns [cbase, cbase + G*R) of a longer query.  bh and be
// are the workgroup's (wave-uniform) boundary planes and boff the group's first row in
// them, so an access is a uniform base plus a 32-bit offset.  bh[boff+i]
// and be[boff+i] hold H and E of row i+1 at the column left of the stripe when bnd_in (else
// the column-0 boundary applies), and receive those of the stripe's last column when
// bnd_out; the
// row is read one step before lane 0 needs it and written G-1 steps after, so one buffer
// serves both.  The result then covers the stripe's columns only (the caller combines
// stripes with max).
// PF: the reference code of a row is loaded two steps ahead, from a clamped index so the
// load needs no branch and nothing waits for it until that step's table or profile
// address (a global load takes longer than one step of a narrow tile).
// TAB: no profile; prof_lane is a substitution table of trow bytes per reference code
// and toff[kk] the byte offset of column kk's (query A code, query B code) entry in a row,
// so each column costs one LDS load (see the pairs table kernel).
// PERM (with TAB): prof_lane is a table of 8 bytes per reference code, byte k the high byte
// of the binary16 entry for query code k, and toff[kk] a v_perm_b32 selector that builds
// column kk's (A, B) pair from them: one 8-byte LDS load per row, one VALU op per column.
template <typename V, int G, int R, bool LOCAL, bool STRIPE = false, bool PF_ = false, bool TAB = false,
          bool PERM = false>
__device__ __forceinline__ SweepOut sweep_group(const uint8_t *__restrict__ rseq, int lr, int nsteps,
                                                const char *prof_lane, int t, const SweepConst &k,
                                                int lqa, int lqb, int cbase, uint32_t *bh = nullptr,
                                                uint32_t *be = nullptr, bool bnd_in = false,
                                                bool bnd_out = false, const int *toff = nullptr,
                                                int trow = 0, uint32_t boff = 0) {
    constexpr bool PF = PF_ || STRIPE;   // stripes: no code-load wait right after the boundary store
    using O = VOps<V>;
    using L = ProfLayout<G, R>;
    constexpr int KC = L::KC;
    const bool semi = k.mode == 2;
    const bool track_endcol = semi && k.free_q;

    V hup[R], fn[R];
    const V GOV = O::splat(k.go), GEV = O::splat(k.ge);
    const V GOP = O::splat(-k.go), GEP = O::splat(-k.ge);
    const V ZERO = O::splat(0);
    // the offset form's D (0: plain values); H(0, col) and H(row, cbase) as carried
    constexpr bool OFS = DiagOffset<V, LOCAL>::ON;
    const int D = OFS ? -k.ge : 0;
    const V GOD = O::splat(k.go + D), NDV = O::splat(-D);
    auto row0 = [&](int col) { return row0_value(k, col) + col * D; };
    auto col0 = [&](int row) { return col0_value(k, row) + (row + cbase) * D; };
#pragma unroll
    for (int kk = 0; kk < R; ++kk) {
        if (LOCAL) {
            hup[kk] = ZERO;
            fn[kk] = ZERO;
        } else {
            const int h0 = row0(cbase + t * R + kk + 1);
            hup[kk] = O::splat(h0);
            // F(1, col)
            fn[kk] = O::splat(OFS ? max(h0 + k.go + D, O::NEG) : max(h0 + k.go, O::NEG + k.ge));
        }
    }
    V hout = LOCAL ? ZERO : O::splat(row0(cbase + (t + 1) * R));
    V hin = LOCAL ? ZERO : O::splat(t == 0 ? col0(1) : row0(cbase + t * R));
    V din = LOCAL ? ZERO : O::splat(row0(cbase));   // H(0, left column), lane 0
    V ein = LOCAL ? ZERO : O::splat(O::NEG);
    V eout = ein;
    // boundary word i of a plane: uniform base + 32-bit byte offset (SGPR-base addressing)
    auto bnd = [](uint32_t *plane, uint32_t i) -> uint32_t * {
        return reinterpret_cast<uint32_t *>(reinterpret_cast<char *>(plane) + (uint32_t)(i * 4u));
    };
    if (STRIPE && bnd_in && t == 0 && lr > 0) {
        hin = O::from_bits((int)*bnd(bh, boff));
        ein = O::from_bits((int)*bnd(be, boff));
    }
    V best = ZERO;
    V endcol = O::splat(O::NEG);
    V rowoff = O::splat((t - 1) * D);   // -i*D of the lane's row i (1-based) in this step
    int corner_a = RES_NONE, corner_b = RES_NONE, endrow_a = RES_NONE, endrow_b = RES_NONE;

    // out-of-range codes score as the sentinel
    auto code_at = [&](int r) -> int { return min((int)rseq[r], k.ns - 1); };
    auto load_row = [&](int c, V *dst) {
        if constexpr (PERM) {
            const uint2 w = *reinterpret_cast<const uint2 *>(prof_lane + c * 8);
#pragma unroll
            for (int kk = 0; kk < R; ++kk)
                dst[kk] = O::from_bits((int)__builtin_amdgcn_perm(w.y, w.x, (uint32_t)toff[kk]));
        } else if constexpr (TAB) {
            // a code is below 256 and a row fits LDS: a 24-bit multiply (full rate) is exact
            const char *pc = prof_lane + __umul24((unsigned)c, (unsigned)trow);
#pragma unroll
            for (int kk = 0; kk < R; ++kk) dst[kk] = O::from_bits(*reinterpret_cast<const int *>(pc + toff[kk]));
        } else {
            const char *pc = prof_lane + c * L::ROWB;
#pragma unroll
            for (int kc = 0; kc < KC; ++kc) {
                const int4 x = *reinterpret_cast<const int4 *>(pc + kc * L::CH);
                dst[4 * kc + 0] = O::from_bits(x.x);
                dst[4 * kc + 1] = O::from_bits(x.y);
                dst[4 * kc + 2] = O::from_bits(x.z);
                dst[4 * kc + 3] = O::from_bits(x.w);
            }
        }
    };
    // column-0 boundary H(i, 0) of lane 0's next row, kept as a running (saturating) value
    const bool ladder = !LOCAL && (k.mode == 1 || (k.mode == 2 && !k.free_q));
    // lane 0's boundary H(i, cbase) of its next row and its change per row: the plain gap
    // ladder steps by gap_extend; in the offset form the ladder is constant and free query
    // ends step by D
    V c0v = OFS ? O::splat(col0(2)) : ladder ? O::splat(col0_value(k, 2)) : ZERO;
    const bool c0_steps = OFS ? !ladder : ladder;
    // Row shifts of ZERO_FILL tiles give every lane t > 0 lane t-1's H and E of the row.
    // Lane 0 takes the column-0 boundary: local H = E = 0 (the zero fill itself); otherwise
    // E stays the sentinel it holds and H steps down the column-0 ladder by gap_extend (or
    // stays 0 with free query ends; in the offset form the ladder stays and free ends step
    // by D), so each is the kept old value of one DPP move.
    const V LADV = OFS ? (ladder ? ZERO : O::splat(D)) : ladder ? GEV : ZERO;
    auto shift_boundary = [&](V &hi, V &ei, V ho, V eo) {
        if constexpr (!GroupLanes<G>::ZERO_FILL) {
            (void)hi, (void)ei, (void)ho, (void)eo;
        } else if constexpr (LOCAL) {
            hi = O::from_bits(GroupLanes<G>::shift(O::bits(ho)));
            ei = O::from_bits(GroupLanes<G>::shift(O::bits(eo)));
        } else {
            hi = O::from_bits(GroupLanes<G>::shift_keep(O::bits(O::add(hi, LADV)), O::bits(ho)));
            ei = O::from_bits(GroupLanes<G>::shift_keep(O::bits(ei), O::bits(eo)));
        }
    };
    const int rlast = max(lr - 1, 0);   // idle groups (lr = 0) read their row 0 harmlessly
    // PF codes: stripes hold the code of the current row (c_raw) and load the next one each
    // step; other sweeps keep one register per step parity (cA even steps, cB odd ones),
    // each reloaded with the code two rows on once its step has used it, so the load stays
    // in flight through the next step with no register copy to wait for it
    int c_raw = STRIPE ? (int)rseq[min(max(-t, 0), rlast)] : 0;
    int cA = PF && !STRIPE ? (int)rseq[min(max(-t, 0), rlast)] : 0;
    int cB = PF && !STRIPE ? (int)rseq[min(max(1 - t, 0), rlast)] : 0;
    // lane 0's boundary words of its next row (stale when there is none: lane 0 is idle then)
    uint2 nb{0u, 0u};

    auto finish = [&]() -> SweepOut {
        SweepOut out;
        if (LOCAL) {
            int ba = O::lo(best), bb = O::hi(best);
            ba = group_max_i32<G>(ba);
            bb = O::PACKED ? group_max_i32<G>(bb) : ba;
            out.a = ba;
            out.b = bb;
            return out;
        }
        // non-local.  A lane stops updating after its last row, so hup[] now holds
        // H(lr, j) of the lane's columns: the corner and the last-row maximum over the
        // real columns are read once here.  (The column tests would otherwise be hoisted
        // above the step loop, where their masks spill scalar registers for the sweep.)
        asm volatile("" : "+v"(lqa), "+v"(lqb), "+v"(cbase));
        if (lr > 0) {
#pragma unroll
            for (int kk = 0; kk < R; ++kk) {
                const int col = cbase + t * R + kk;
                const int off = (lr + col + 1) * D;   // offset form: H(lr, col+1) + off
                const int va = O::lo(hup[kk]) - off;
                const int vb = O::hi(hup[kk]) - off;
                if (col == lqa - 1) corner_a = va;
                if (col == lqb - 1) corner_b = vb;
                if (col < lqa) endrow_a = max(endrow_a, va);
                if (col < lqb) endrow_b = max(endrow_b, vb);
            }
        }
        // each lane offers what it holds; the group maximum is the result
        int ca = corner_a, cb = corner_b;
        if (semi) {
            if (track_endcol) {
                // the lane whose last register is column lq-1 (none in other stripes)
                if (cbase + t * R + R - 1 == lqa - 1) ca = max(ca, O::lo(endcol) - lqa * D);
                if (cbase + t * R + R - 1 == lqb - 1) cb = max(cb, O::hi(endcol) - lqb * D);
            }
            if (k.free_r) {
                ca = max(ca, endrow_a);
                cb = max(cb, endrow_b);
            }
        }
        ca = group_max_i32<G>(ca);
        cb = O::PACKED ? group_max_i32<G>(cb) : ca;
        if (semi && k.free_q && k.free_r) {
            ca = max(ca, 0);
            cb = max(cb, 0);
        }
        out.a = ca;
        out.b = cb;
        return out;
    };
    // One row of the lane's R columns from its profile row pv.
    auto row_cells = [&](const V *pv) {
        V e = ein;
        V hgo = OFS ? O::add(hin, GOD) : row_hgo<V, LOCAL>(hin, GOV, GOP);
        V h = hin, hprev = ZERO;
        // xd = H(i-1, j-1) + s(i, j).  The sum for column kk+1 is taken before
        // hup[kk] is overwritten, so H(i, j) can replace H(i-1, j) in place.
        V xd = O::add(din, pv[0]);
#pragma unroll
        for (int kk = 0; kk < R; ++kk) {
            const V xn = (kk + 1 < R) ? O::add(hup[kk], pv[kk + 1]) : xd;
            if constexpr (OFS) {
                cell_offset<V>(xd, e, h, hup[kk], hgo, fn[kk], GOD);
            } else if constexpr (O::MAX3) {
                cell_max3<V, LOCAL, R>(kk, xd, e, h, hprev, best, hup[kk], hgo, fn[kk], GOV, GEV, ZERO);
            } else if (LOCAL) {
                e = O::mx(O::subu(e, GEP), hgo);
                h = O::mx(O::mx(xd, fn[kk]), e);   // E last: the serial chain is E -> H -> H-gap -> E
                best = O::mx(best, h);
                hup[kk] = h;
                hgo = O::subu(h, GOP);
                fn[kk] = O::mx(O::subu(fn[kk], GEP), hgo);
            } else {
                e = O::mx(O::add(e, GEV), hgo);
                h = O::mx(O::mx(xd, fn[kk]), e);
                hup[kk] = h;
                hgo = O::add(h, GOV);
                fn[kk] = O::mx(O::add(fn[kk], GEV), hgo);
            }
            xd = xn;
        }
        hout = h;
        eout = e;
        // lane's last column (offset form: H(i, j) + j*D, the row's i*D taken off)
        if (!LOCAL && track_endcol) endcol = O::mx(endcol, OFS ? O::add(h, rowoff) : h);
    };
    // One step of a tile (not a stripe): lane t on row s - t.  GUARD: rows outside [0, lr)
    // are skipped (the pipeline fill and drain); without it every lane of the wave is on a
    // real row, or in an idle group (lr = 0) whose results are never stored and whose code
    // reads stay inside the code buffer (row < the shortest active reference of the wave).
    // c: this step parity's prefetched code (PF).
    auto step = [&](int s, auto guard, int &c) {
        constexpr bool GUARD = decltype(guard)::value;
        const int row = s - t;
        if (!GUARD || (unsigned)row < (unsigned)lr) {
            V pv[KC * 4];
            if constexpr (PF) {
                load_row(min(c, k.ns - 1), pv);
                c = (int)rseq[min(max(row + 2, 0), rlast)];
            } else {
                load_row(code_at(row), pv);
            }
            row_cells(pv);
        } else if constexpr (PF) {
            c = (int)rseq[min(max(row + 2, 0), rlast)];
        }
        din = hin;
        if constexpr (OFS) rowoff = O::add(rowoff, NDV);
        if constexpr (GroupLanes<G>::ZERO_FILL) {
            shift_boundary(hin, ein, hout, eout);
        } else {
            hin = O::from_bits(GroupLanes<G>::shift(O::bits(hout)));
            ein = O::from_bits(GroupLanes<G>::shift(O::bits(eout)));
            if (t == 0) {
                hin = c0v;
                ein = LOCAL ? ZERO : O::splat(O::NEG);
            }
            if (c0_steps) c0v = O::add(c0v, LADV);
        }
    };
    // Steps from G-1 to the shortest active reference of the wave have every active lane on
    // a real row: they run unguarded.  Not in stripes, whose boundary registers leave too
    // few VGPRs for the duplicated step: the scheduler then serialises the E -> H chain
    // with a wait state per column.
    if constexpr (STRIPE) {
        // (the step written out here, not through the step lambda: with the lambda the
        // compiler schedules the global/semiglobal stripes' E -> H chain with a wait state
        // per column)
        for (int s = 0; s < nsteps; ++s) {
            const int row = s - t;
            if (bnd_in && t == 0 && row + 1 < lr)   // lane 0's next row
                nb = uint2{*bnd(bh, boff + (uint32_t)(row + 1)), *bnd(be, boff + (uint32_t)(row + 1))};
            const int c_nxt = (int)rseq[min(max(row + 1, 0), rlast)];
            if ((unsigned)row < (unsigned)lr) {
                V pv[KC * 4];
                load_row(min(c_raw, k.ns - 1), pv);
                row_cells(pv);
            }
            c_raw = c_nxt;
            // Loads and stores share one wait counter, and the compiler waits for all of
            // them once both kinds are pending: take this step's loads (empty asm reading
            // their registers; memory clobber keeps the store below it), then store lane
            // G-1's boundary row, which drains during the next step's columns.  Two planes:
            // the stores read H and E where they live (loop-carried, not rewritten until
            // late in the next step); one 64-bit store would need an adjacent copy that the
            // next step's first profile load overwrites, forcing an early wait.
            asm volatile("" ::"v"(nb.x), "v"(nb.y), "v"(c_nxt) : "memory");
            if (bnd_out && t == G - 1 && (unsigned)row < (unsigned)lr) {
                *bnd(bh, boff + (uint32_t)row) = (uint32_t)O::bits(hout);
                *bnd(be, boff + (uint32_t)row) = (uint32_t)O::bits(eout);
            }
            din = hin;
            if constexpr (OFS) rowoff = O::add(rowoff, NDV);
            hin = O::from_bits(GroupLanes<G>::shift(O::bits(hout)));
            ein = O::from_bits(GroupLanes<G>::shift(O::bits(eout)));
            if (t == 0) {
                if (bnd_in) {
                    hin = O::from_bits((int)nb.x);
                    ein = O::from_bits((int)nb.y);
                } else if (!(LOCAL && GroupLanes<G>::ZERO_FILL)) {
                    hin = c0v;
                    ein = LOCAL ? ZERO : O::splat(O::NEG);
                }
            }
            if (c0_steps) c0v = O::add(c0v, LADV);
        }
        return finish();
    }
    // Steps run in pairs from even s (cA, then cB); the unguarded pairs cover [s_lo, s_hi),
    // s_lo the first even step >= G-1.
    const int s_lo = min(G, nsteps);
    const int lrmin = __builtin_amdgcn_readfirstlane(wave_min_i32(lr > 0 ? lr : 0x7fffffff));
    const int s_hi = s_lo + ((max(s_lo, min(lrmin, nsteps)) - s_lo) & ~1);
    for (int s = 0; s < nsteps; s += 2) {
        if (s == s_lo) {
            for (; s < s_hi; s += 2) {
                step(s, std::false_type{}, cA);
                step(s + 1, std::false_type{}, cB);
            }
            if (s >= nsteps) break;
        }
        step(s, std::true_type{}, cA);
        if (s + 1 < nsteps) step(s + 1, std::true_type{}, cB);
    }

    return finish();
}

// The sweep of sweep_group for explicit pairs with private profiles, pipelined for two
// waves per SIMD: the step loop is unrolled by two with two profile buffers, so that the
// profile row of row i+1 is read from LDS during row i's columns and the reference code of
// row i+2 is read from global memory during row i; neither load is waited for until the
// step that uses it (no register copy forces an early wait).  (Unrolling by four to read
// codes three steps ahead measured 10-20 % slower.)  prow is the largest profile row (the
// last row present: codes past it read it); rows are ROWB bytes apart.
template <typename V, int G, int R, bool LOCAL>
__device__ __forceinline__ SweepOut sweep_pairs_pipe(const uint8_t *__restrict__ rseq, int lr, int nsteps,
                                                     const char *prof_lane, int t, const SweepConst &k,
                                                     int lqa, int lqb, int cbase, int prow) {
    using O = VOps<V>;
    using L = ProfLayout<G, R>;
    constexpr int KC = L::KC;
    const bool semi = k.mode == 2;
    const bool track_endcol = semi && k.free_q;

    V hup[R], fn[R];
    const V GOV = O::splat(k.go), GEV = O::splat(k.ge);
    const V GOP = O::splat(-k.go), GEP = O::splat(-k.ge);
    const V ZERO = O::splat(0);
    // the offset form's D (0: plain values); H(0, col) and H(row, cbase) as carried
    constexpr bool OFS = DiagOffset<V, LOCAL>::ON;
    const int D = OFS ? -k.ge : 0;
    const V GOD = O::splat(k.go + D), NDV = O::splat(-D);
    auto row0 = [&](int col) { return row0_value(k, col) + col * D; };
    auto col0 = [&](int row) { return col0_value(k, row) + (row + cbase) * D; };
#pragma unroll
    for (int kk = 0; kk < R; ++kk) {
        if (LOCAL) {
            hup[kk] = ZERO;
            fn[kk] = ZERO;
        } else {
            const int h0 = row0(cbase + t * R + kk + 1);
            hup[kk] = O::splat(h0);
            // F(1, col)
            fn[kk] = O::splat(OFS ? max(h0 + k.go + D, O::NEG) : max(h0 + k.go, O::NEG + k.ge));
        }
    }
    V hout = LOCAL ? ZERO : O::splat(row0(cbase + (t + 1) * R));
    V hin = LOCAL ? ZERO : O::splat(t == 0 ? col0(1) : row0(cbase + t * R));
    V din = LOCAL ? ZERO : O::splat(row0(cbase));
    V ein = LOCAL ? ZERO : O::splat(O::NEG);
    V eout = ein;
    V best = ZERO;
    V endcol = O::splat(O::NEG);
    V rowoff = O::splat((t - 1) * D);   // -i*D of the lane's row i (1-based) in this step
    const bool ladder = !LOCAL && (k.mode == 1 || (k.mode == 2 && !k.free_q));
    // lane 0's boundary H(i, cbase) of its next row and its change per row: the plain gap
    // ladder steps by gap_extend; in the offset form the ladder is constant and free query
    // ends step by D
    V c0v = OFS ? O::splat(col0(2)) : ladder ? O::splat(col0_value(k, 2)) : ZERO;
    const bool c0_steps = OFS ? !ladder : ladder;
    // Row shifts of ZERO_FILL tiles give every lane t > 0 lane t-1's H and E of the row.
    // Lane 0 takes the column-0 boundary: local H = E = 0 (the zero fill itself); otherwise
    // E stays the sentinel it holds and H steps down the column-0 ladder by gap_extend (or
    // stays 0 with free query ends; in the offset form the ladder stays and free ends step
    // by D), so each is the kept old value of one DPP move.
    const V LADV = OFS ? (ladder ? ZERO : O::splat(D)) : ladder ? GEV : ZERO;
    auto shift_boundary = [&](V &hi, V &ei, V ho, V eo) {
        if constexpr (!GroupLanes<G>::ZERO_FILL) {
            (void)hi, (void)ei, (void)ho, (void)eo;
        } else if constexpr (LOCAL) {
            hi = O::from_bits(GroupLanes<G>::shift(O::bits(ho)));
            ei = O::from_bits(GroupLanes<G>::shift(O::bits(eo)));
        } else {
            hi = O::from_bits(GroupLanes<G>::shift_keep(O::bits(O::add(hi, LADV)), O::bits(ho)));
            ei = O::from_bits(GroupLanes<G>::shift_keep(O::bits(ei), O::bits(eo)));
        }
    };
    const int rlast = max(lr - 1, 0);   // idle groups (lr = 0) read their row 0 harmlessly
    auto code = [&](int row) -> int { return min((int)rseq[min(max(row, 0), rlast)], prow); };
    auto load_row = [&](int c, V *dst) {
        const char *pc = prof_lane + c * L::ROWB;
#pragma unroll
        for (int kc = 0; kc < KC; ++kc) {
            const int4 x = *reinterpret_cast<const int4 *>(pc + kc * L::CH);
            dst[4 * kc + 0] = O::from_bits(x.x);
            dst[4 * kc + 1] = O::from_bits(x.y);
            dst[4 * kc + 2] = O::from_bits(x.z);
            dst[4 * kc + 3] = O::from_bits(x.w);
        }
    };
    V pa[KC * 4], pb[KC * 4];
    load_row(code(-t), pa);              // row 0 - t: step 0
    int cx = code(1 - t);                // row 1 - t: its profile is read during step 0
    int cy = 0;

    // One step: row = s - t with profile `use`; reads the profile of row + 1 (code cin)
    // into `nxt` and the code of row + 2 into cout.
    auto step = [&](int s, V *use, V *nxt, int cin, int &cout) {
        const int row = s - t;
        cout = code(row + 2);
        load_row(cin, nxt);
        if ((unsigned)row < (unsigned)lr) {
            V e = ein;
            V hgo = OFS ? O::add(hin, GOD) : row_hgo<V, LOCAL>(hin, GOV, GOP);
            V h = hin, hprev = ZERO;
            V xd = O::add(din, use[0]);
#pragma unroll
            for (int kk = 0; kk < R; ++kk) {
                const V xn = (kk + 1 < R) ? O::add(hup[kk], use[kk + 1]) : xd;
                if constexpr (OFS) {
                    cell_offset<V>(xd, e, h, hup[kk], hgo, fn[kk], GOD);
                } else if constexpr (O::MAX3) {
                    cell_max3<V, LOCAL, R>(kk, xd, e, h, hprev, best, hup[kk], hgo, fn[kk], GOV, GEV, ZERO);
                } else if (LOCAL) {
                    e = O::mx(O::subu(e, GEP), hgo);
                    h = O::mx(O::mx(xd, fn[kk]), e);
                    best = O::mx(best, h);
                    hup[kk] = h;
                    hgo = O::subu(h, GOP);
                    fn[kk] = O::mx(O::subu(fn[kk], GEP), hgo);
                } else {
                    e = O::mx(O::add(e, GEV), hgo);
                    h = O::mx(O::mx(xd, fn[kk]), e);
                    hup[kk] = h;
                    hgo = O::add(h, GOV);
                    fn[kk] = O::mx(O::add(fn[kk], GEV), hgo);
                }
                xd = xn;
            }
            hout = h;
            eout = e;
            // unconditional (read only when track_endcol): a branch here splits the step, and
            // the scheduler then sinks the F updates below it as dependent add/max pairs, each
            // needing a wait state between them (about 70 s_nop per step)
            if (!LOCAL) endcol = O::mx(endcol, OFS ? O::add(h, rowoff) : h);
        }
        din = hin;
        if constexpr (OFS) rowoff = O::add(rowoff, NDV);
        if constexpr (GroupLanes<G>::ZERO_FILL) {
            shift_boundary(hin, ein, hout, eout);
        } else {
            hin = O::from_bits(GroupLanes<G>::shift(O::bits(hout)));
            ein = O::from_bits(GroupLanes<G>::shift(O::bits(eout)));
            if (t == 0) {
                hin = c0v;
                ein = LOCAL ? ZERO : O::splat(O::NEG);
            }
            if (c0_steps) c0v = O::add(c0v, LADV);
        }
    };
    for (int s = 0; s < nsteps; s += 2) {   // a step past nsteps has no row in range: harmless
        step(s, pa, pb, cx, cy);
        step(s + 1, pb, pa, cy, cx);
    }

    SweepOut out;
    if (LOCAL) {
        int ba = O::lo(best), bb = O::hi(best);
        ba = group_max_i32<G>(ba);
        bb = O::PACKED ? group_max_i32<G>(bb) : ba;
        out.a = ba;
        out.b = bb;
        return out;
    }
    int corner_a = RES_NONE, corner_b = RES_NONE, endrow_a = RES_NONE, endrow_b = RES_NONE;
    // the column tests below would otherwise be hoisted above the loop, where their masks
    // (two to four per column) spill scalar registers for the whole sweep
    asm volatile("" : "+v"(lqa), "+v"(lqb), "+v"(cbase));
    if (lr > 0) {
#pragma unroll
        for (int kk = 0; kk < R; ++kk) {
            const int col = cbase + t * R + kk;
            const int off = (lr + col + 1) * D;   // offset form: H(lr, col+1) + off
            const int va = O::lo(hup[kk]) - off;
            const int vb = O::hi(hup[kk]) - off;
            if (col == lqa - 1) corner_a = va;
            if (col == lqb - 1) corner_b = vb;
            if (col < lqa) endrow_a = max(endrow_a, va);
            if (col < lqb) endrow_b = max(endrow_b, vb);
        }
    }
    int ca = corner_a, cb = corner_b;
    if (semi) {
        if (track_endcol) {
            if (cbase + t * R + R - 1 == lqa - 1) ca = max(ca, O::lo(endcol) - lqa * D);
            if (cbase + t * R + R - 1 == lqb - 1) cb = max(cb, O::hi(endcol) - lqb * D);
        }
        if (k.free_r) {
            ca = max(ca, endrow_a);
            cb = max(cb, endrow_b);
        }
    }
    ca = group_max_i32<G>(ca);
    cb = O::PACKED ? group_max_i32<G>(cb) : ca;
    if (semi && k.free_q && k.free_r) {
        ca = max(ca, 0);
        cb = max(cb, 0);
    }
    out.a = ca;
    out.b = cb;
    return out;
}

}  // namespace gsw
