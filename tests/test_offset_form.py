"""The diagonal-offset form of the global/semiglobal recurrence, as the packed sweeps run
it (gpusw_amd/csrc/gsw_sweep.h, DiagOffset), restated per lane group in Python and checked
against the oracle (``src/gpusw/reference.py:38-125`` restated in ``oracle/gotoh.py``).

Every value of cell (i, j) is carried as X + (i + j)*D, D = -gap_extend; the profile holds
s + 2D; row 0 and the column-0 boundary are gap_open + D (gap ladders) or (i + j)*D (free
ends and front padding); results subtract (i + j)*D.  The cases cover every mode and free-end
combination, front padding of the free-query-end layout (the last query column on a lane's
last register), padded columns past the query and gap-extend 0.
"""
import random

import pytest

from oracle.gotoh import score_codes

NEGV = -30000


def row0_value(mode, go, ge, free_r, col):
    if col <= 0:
        return 0
    if mode == "global" or (mode == "semiglobal" and not free_r):
        return go + (col - 1) * ge
    return 0


def col0_value(mode, go, ge, free_q, row):
    if row <= 0:
        return 0
    if mode == "global" or (mode == "semiglobal" and not free_q):
        return go + (row - 1) * ge
    return 0


def offs_score(qc, rc, sub, mode, go, ge, free_q, free_r, R):
    """One lane group's sweep, columns laid out with the front-padding shift, offset form."""
    lq, lr = len(qc), len(rc)
    track = mode == "semiglobal" and free_q
    shift = (R - lq % R) % R if track else 0
    cbase = -shift
    ncol = lq + shift
    ncol = ((ncol + R - 1) // R) * R + R   # a padded lane beyond the query
    D = -ge
    ladder = mode == "global" or (mode == "semiglobal" and not free_q)

    def row0(col):
        return row0_value(mode, go, ge, free_r, col) + col * D

    def col0(row):
        return col0_value(mode, go, ge, free_q, row) + (row + cbase) * D

    def prof(c, col):   # col 0-based real column (negative: front padding)
        if col >= lq:
            return NEGV
        v = 0 if col < 0 else sub[c][qc[col]]
        return v - 2 * ge

    hup = [row0(cbase + kk + 1) for kk in range(ncol)]
    fn = [max(h + go + D, NEGV) for h in hup]
    hin = col0(1)
    din = row0(cbase)
    endcol = NEGV
    for i in range(1, lr + 1):
        c = rc[i - 1]
        e = NEGV
        hgo = hin + go + D
        xd = din + prof(c, cbase + 0)
        for kk in range(ncol):
            xn = hup[kk] + prof(c, cbase + kk + 1) if kk + 1 < ncol else None
            e = max(e, hgo)
            h = max(xd, fn[kk], e)
            hup[kk] = h
            hgo = h + go + D
            fn[kk] = max(fn[kk], hgo)
            if cbase + kk == lq - 1 and track:
                endcol = max(endcol, h - i * D)
            xd = xn
        din = hin
        # next row's boundary
        if ladder:
            hin = hin  # constant go + D
        else:
            hin = hin + D
    corner = endrow = -(1 << 40)
    for kk in range(ncol):
        col = cbase + kk
        if col < 0:
            continue
        v = hup[kk] - (lr + col + 1) * D
        if col == lq - 1:
            corner = v
        if col < lq:
            endrow = max(endrow, v)
    if mode == "global":
        return corner
    res = corner
    if free_q:
        res = max(res, endcol - lq * D)
    if free_r:
        res = max(res, endrow)
    if free_q and free_r:
        res = max(res, 0)
    return res


@pytest.mark.parametrize("seed", range(6))
def test_offset_form_equals_oracle(seed):
    rng = random.Random(seed)
    for _ in range(300):
        ns = rng.choice([4, 5, 20])
        sub = [[rng.randint(-4, 5) for _ in range(ns)] for _ in range(ns)]
        go, ge = rng.randint(-12, 0), rng.randint(-4, 0)
        mode = rng.choice(["global", "semiglobal"])
        free_q, free_r = rng.random() < 0.5, rng.random() < 0.5
        qc = [rng.randrange(ns) for _ in range(rng.randint(1, 40))]
        rc = [rng.randrange(ns) for _ in range(rng.randint(1, 40))]
        R = rng.choice([3, 4, 7, 8])
        want = score_codes(qc, rc, sub, mode, go, ge, free_q, free_r)
        assert offs_score(qc, rc, sub, mode, go, ge, free_q, free_r, R) == want, (mode, free_q, free_r, go, ge, R)


def test_offset_form_catches_a_wrong_profile_shift():
    """Planted mistake: a profile of s + D instead of s + 2D must disagree with the oracle."""
    rng = random.Random(99)
    sub = [[2 if a == b else -3 for b in range(4)] for a in range(4)]
    qc = [rng.randrange(4) for _ in range(30)]
    rc = [rng.randrange(4) for _ in range(35)]
    want = score_codes(qc, rc, sub, "global", -5, -2, False, False)
    shifted = [[v + 2 for v in row] for row in sub]   # the form then adds 2D on top: s + 3D
    assert offs_score(qc, rc, shifted, "global", -5, -2, False, False, 4) != want
