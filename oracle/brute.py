"""Independent pin for the oracle: explicit enumeration of alignment paths (no DP).

An alignment is a monotone path through the (reference x query) grid made of moves
M (consume one reference and one query residue), X (reference residue against a gap in
the query) and Y (query residue against a gap in the reference).  Scoring rules, stated
from the definition of the affine-gap problem the reference solves
(``src/gpusw/reference.py:1-18`` docstring; ``docs/DESIGN.md`` gap convention):

* M scores ``sub[ref][query]``;
* an interior gap run of length L scores ``go + (L-1)*max(go, ge)``: a run may be split
  into several gaps, each opened again (the recurrence allows reopening from H), so the
  best split is taken;
* a path may begin with one boundary run along row 0 (Y moves) or column 0 (X moves);
  such a boundary run is a single gap ``go + (L-1)*ge``, or free when that end is free
  (semiglobal flags) or in local mode;
* interior X moves need a query column >= 1 and interior Y moves a reference row >= 1
  (column 0 and row 0 are the boundaries);
* local: best over every path fragment starting anywhere, floored at 0;
  global: full paths to the corner; semiglobal: the corner, plus any interior cell of
  the last column (free_q) or last row (free_r), then floored at 0 when both are free.

Exponential; meant for sequences of length <= 5.
"""
from __future__ import annotations


def _ladder(go, ge, length):
    return go + (length - 1) * ge


def brute_score(qc, rc, sub, mode, go, ge, free_q, free_r):
    n, m = len(qc), len(rc)
    local = mode == "local"
    semi = mode == "semiglobal"
    if n == 0 or m == 0:
        # closed forms for empty inputs: nothing to align, or one all-gap run
        if local or (n == 0 and m == 0):
            return 0
        if n == 0:
            return 0 if (semi and free_q) else _ladder(go, ge, m)
        return 0 if (semi and free_r) else _ladder(go, ge, n)

    gap_next = max(go, ge)
    best_at = {}

    def walk(i, j, last, score):
        if i >= 1 and j >= 1:
            key = (i, j)
            if key not in best_at or score > best_at[key]:
                best_at[key] = score
        if i < m and j < n:
            walk(i + 1, j + 1, "M", score + int(sub[rc[i]][qc[j]]))
        if i < m and j >= 1:
            walk(i + 1, j, "X", score + (gap_next if last == "X" else go))
        if j < n and i >= 1:
            walk(i, j + 1, "Y", score + (gap_next if last == "Y" else go))

    if local:
        for i in range(m + 1):
            for j in range(n + 1):
                walk(i, j, "S", 0)
        return max([0] + list(best_at.values()))

    walk(0, 0, "S", 0)
    for j in range(1, n + 1):
        walk(0, j, "B", 0 if (semi and free_r) else _ladder(go, ge, j))
    for i in range(1, m + 1):
        walk(i, 0, "B", 0 if (semi and free_q) else _ladder(go, ge, i))

    res = best_at[(m, n)]
    if semi and free_q:
        res = max([res] + [best_at[(i, n)] for i in range(1, m + 1)])
    if semi and free_r:
        res = max([res] + [best_at[(m, j)] for j in range(1, n + 1)])
    if semi and free_q and free_r:
        res = max(res, 0)
    return res
