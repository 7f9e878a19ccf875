// FASTA -> codes/offsets on the host: one pass to size, one pass to fill.  Records
// follow the funnel's FASTA rules (reference src/gpusw/encode.py:91-121): the text is
// split into lines at the ASCII line breaks str.splitlines() knows (\n \r \v \f \x1c
// \x1d \x1e; \r\n counts once), each line is stripped of the ASCII whitespace str.strip()
// removes, blank lines are skipped, a line starting with '>' opens a record whose id is
// the first whitespace-separated word after '>' (none: "seq<k>", made by the caller),
// text before the first header is dropped, and a record's sequence is its lines joined.
// Bytes >= 0x80 are rejected (the caller decodes such files itself).
#include <cstdint>
#include <cstring>

#include "gsw_hip.h"
#include "gsw_internal.h"

namespace {

inline bool is_break(uint8_t c) { return c == '\n' || c == '\r' || c == 0x0B || c == 0x0C || (c >= 0x1C && c <= 0x1E); }
inline bool is_space(uint8_t c) { return is_break(c) || c == ' ' || c == '\t' || c == 0x1F; }

// Calls fn(line_begin, line_end) for every stripped, non-blank line.
template <typename F> void for_lines(const uint8_t *b, int64_t n, F fn) {
    int64_t i = 0;
    while (i < n) {
        int64_t e = i;
        while (e < n && !is_break(b[e])) ++e;
        int64_t s = i, t = e;
        while (s < t && is_space(b[s])) ++s;
        while (t > s && is_space(b[t - 1])) --t;
        if (t > s) fn(s, t);
        i = e + 1;
    }
}

}  // namespace

using namespace gsw;

extern "C" int gsw_fasta_scan(const uint8_t *buf, int64_t n, int64_t *nrec, int64_t *nres) {
    if (!buf && n) return fail(GSW_ERR_INVALID, "null buffer");
    for (int64_t i = 0; i < n; ++i)
        if (buf[i] >= 0x80) return fail(GSW_ERR_INVALID, "non-ASCII byte at offset %lld", (long long)i);
    int64_t r = 0, s = 0;
    bool in = false;
    for_lines(buf, n, [&](int64_t a, int64_t b) {
        if (buf[a] == '>') { ++r; in = true; }
        else if (in) s += b - a;
    });
    *nrec = r;
    *nres = s;
    return GSW_OK;
}

extern "C" int gsw_fasta_encode(const uint8_t *buf, int64_t n, const uint8_t *lut, uint8_t *codes,
                                int64_t *offsets, int64_t *id_span) {
    if ((!buf && n) || !lut || !offsets || !id_span) return fail(GSW_ERR_INVALID, "null argument");
    int64_t r = -1, pos = 0;
    offsets[0] = 0;
    for_lines(buf, n, [&](int64_t a, int64_t b) {
        if (buf[a] == '>') {
            if (r >= 0) offsets[r + 1] = pos;
            ++r;
            int64_t s = a + 1;
            while (s < b && is_space(buf[s])) ++s;
            int64_t e = s;
            while (e < b && !is_space(buf[e])) ++e;
            id_span[2 * r] = s;          // empty span (s == e): the caller names it seq<k>
            id_span[2 * r + 1] = e - s;
        } else if (r >= 0) {
            for (int64_t i = a; i < b; ++i) codes[pos++] = lut[buf[i]];
        }
    });
    if (r >= 0) offsets[r + 1] = pos;
    return GSW_OK;
}
