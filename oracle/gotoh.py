"""Pure-Python/NumPy restatement of the reference's CPU oracle (test infrastructure).

Follows ``src/gpusw/reference.py`` of the reference:
* ``_NEG = -(1 << 30)`` sentinel (``reference.py:27``);
* ``score_codes`` restates ``_score_codes`` (``reference.py:38-125``): rows = reference,
  columns = query; F = vertical gap, E = horizontal gap; boundary ladders per mode;
  empty-sequence edges; local running max; global corner; semiglobal
  ``max(corner, endcol?, endrow?)`` then ``max(., 0)`` when both ends are free;
* ``cpu_reference_score`` / ``cpu_reference_matrix`` restate ``reference.py:128-151``.
"""
from __future__ import annotations

import numpy as np

NEG = -(1 << 30)

MODE_CODE = {"local": 0, "global": 1, "semiglobal": 2}


def score_codes(qc, rc, sub, mode: str, go: int, ge: int, free_q: bool, free_r: bool) -> int:
    """Optimal score of query codes ``qc`` vs reference codes ``rc`` (``reference.py:38-125``)."""
    n = len(qc)
    m = len(rc)
    local = mode == "local"
    semi = mode == "semiglobal"
    if n == 0 or m == 0:                                   # reference.py:49-54
        if local or (n == 0 and m == 0):
            return 0
        if n == 0:
            return 0 if (semi and free_q) else go + (m - 1) * ge
        return 0 if (semi and free_r) else go + (n - 1) * ge

    qc = [int(x) for x in qc]
    rc = [int(x) for x in rc]
    sub = [[int(v) for v in row] for row in np.asarray(sub)]

    def ladder(k):
        return go + (k - 1) * ge

    # row 0 (reference.py:56-65)
    hp = [0] * (n + 1)
    for j in range(1, n + 1):
        if local:
            hp[j] = 0
        elif mode == "global":
            hp[j] = ladder(j)
        else:
            hp[j] = 0 if free_r else ladder(j)
    f = [NEG] * (n + 1)

    best = 0
    corner = endcol = endrow = NEG
    for i in range(1, m + 1):                              # reference.py:70-112
        srow = sub[rc[i - 1]]
        if local:
            hleft = 0
        elif mode == "global":
            hleft = ladder(i)
        else:
            hleft = 0 if free_q else ladder(i)
        diag = hp[0]
        hp[0] = hleft
        eleft = NEG
        for j in range(1, n + 1):
            fj = max(hp[j] + go, f[j] + ge)
            ej = max(hleft + go, eleft + ge)
            h = max(diag + srow[qc[j - 1]], ej, fj)
            if local and h < 0:
                h = 0
            if h > best:
                best = h
            if not local:
                if i == m and j == n:
                    corner = h
                if free_q and j == n and h > endcol:
                    endcol = h
                if free_r and i == m and h > endrow:
                    endrow = h
            diag = hp[j]
            hp[j] = h
            f[j] = fj
            hleft = h
            eleft = ej

    if local:                                              # reference.py:114-125
        return best
    if mode == "global":
        return corner
    res = corner
    if free_q:
        res = max(res, endcol)
    if free_r:
        res = max(res, endrow)
    if free_q and free_r:
        res = max(res, 0)
    return res


def encode(seq, scheme, lut) -> np.ndarray:
    """Text -> codes as the reference oracle does (``reference.py:30-35``)."""
    if isinstance(seq, (bytes, bytearray)):
        seq = bytes(seq).decode("latin-1")
    s = seq.upper() if scheme.case_insensitive else seq
    return lut[np.frombuffer(s.encode("latin-1", "replace"), dtype=np.uint8)]


def scheme_args(scheme) -> dict:
    return dict(
        mode=scheme.mode, go=scheme.gap_open, ge=scheme.gap_extend,
        free_q=bool(scheme.free_end_gaps_query), free_r=bool(scheme.free_end_gaps_ref),
    )


def cpu_reference_score(query, ref, scheme) -> int:
    """``reference.py:128-134``."""
    from gpusw_amd.encode import build_lut

    lut, _ = build_lut(scheme)
    sub = scheme.substitution_table().astype(np.int64)
    return score_codes(encode(query, scheme, lut), encode(ref, scheme, lut), sub,
                       **scheme_args(scheme))


def cpu_reference_matrix(queries, refs, scheme) -> np.ndarray:
    """``reference.py:137-151``: all-pairs int64 matrix."""
    from gpusw_amd.encode import build_lut

    lut, _ = build_lut(scheme)
    sub = scheme.substitution_table().astype(np.int64)
    qcs = [encode(q, scheme, lut) for q in queries]
    rcs = [encode(r, scheme, lut) for r in refs]
    out = np.empty((len(qcs), len(rcs)), dtype=np.int64)
    kw = scheme_args(scheme)
    for a, qc in enumerate(qcs):
        for b, rc in enumerate(rcs):
            out[a, b] = score_codes(qc, rc, sub, **kw)
    return out
