"""Traceback oracle (test infrastructure): the alignment behind an oracle score.

The reference scores only (``docs/DESIGN.md:94-96``); this module defines the alignment
that the library's GPU traceback returns, on top of the same recurrence as
``oracle/gotoh.py`` (``reference.py:38-125``):

* H(i,j) = max(diag + s, E, F[, 0 local]); ties prefer zero (local start), then the
  diagonal, then E, then F.  E (query residue against a gap, op ``I``) and F (reference
  residue against a gap, op ``D``) prefer opening from H over extending on ties.
* End cell: local = first strict maximum in row-major order (none: empty alignment);
  global = (lr, lq); semiglobal = the corner, else the first best row of the last column
  (free query ends), else the first best column of the last row (free reference ends);
  with both ends free and every candidate negative the alignment is empty (score 0).
* Free end gaps are left out of the CIGAR (they show in the coordinates); other boundary
  gaps are ``I``/``D`` runs.  A gap that re-opens right after another of its kind is a
  new run, so the affine cost of the CIGAR (``cigar_score``) is always the score.

Pinned by ``tests/test_traceback_oracle.py``: score == ``gotoh.score_codes`` and
``cigar_score`` == score on random cases of every mode and free-end setting (including
gap_open > gap_extend tables), coordinates consistent with the CIGAR, and known answers.
"""
from __future__ import annotations

import numpy as np

from .gotoh import NEG

EQ, X, I, D = 0, 1, 2, 3
OPCHAR = "=XID"


def traceback_codes(qc, rc, sub, mode: str, go: int, ge: int, free_q: bool, free_r: bool) -> dict:
    """``{score, q_begin, q_end, r_begin, r_end, matches, mismatches, ins, del,
    gap_opens, ops}`` with ``ops`` a list of ``(op, length)`` runs, start to end."""
    qc = [int(x) for x in qc]
    rc = [int(x) for x in rc]
    sub = np.asarray(sub, dtype=np.int64)
    ns = sub.shape[0]
    n, m = len(qc), len(rc)
    local, semi = mode == "local", mode == "semiglobal"
    ladder_r = mode == "global" or (semi and not free_r)    # row 0 is a gap ladder
    ladder_q = mode == "global" or (semi and not free_q)    # column 0 is a gap ladder
    res = dict(score=0, q_begin=0, q_end=0, r_begin=0, r_end=0, matches=0, mismatches=0,
               ins=0, dele=0, gap_opens=0, ops=[])
    if n == 0 or m == 0:
        if local or (n == 0 and m == 0):
            return _out(res)
        if n == 0:
            res["score"] = 0 if (semi and free_q) else go + (m - 1) * ge
            if not (semi and free_q):
                res.update(ops=[(D, m)], dele=m, gap_opens=1)
        else:
            res["score"] = 0 if (semi and free_r) else go + (n - 1) * ge
            if not (semi and free_r):
                res.update(ops=[(I, n)], ins=n, gap_opens=1)
        res["q_end"] = 0 if (semi and free_r and m == 0) else n
        res["r_end"] = 0 if (semi and free_q and n == 0) else m
        return _out(res)

    H = [[0] * (n + 1) for _ in range(m + 1)]
    E = [[NEG] * (n + 1) for _ in range(m + 1)]
    F = [[NEG] * (n + 1) for _ in range(m + 1)]
    src = [[0] * (n + 1) for _ in range(m + 1)]
    ext_e = [[False] * (n + 1) for _ in range(m + 1)]
    ext_f = [[False] * (n + 1) for _ in range(m + 1)]
    for j in range(1, n + 1):
        H[0][j] = go + (j - 1) * ge if ladder_r else 0
    for i in range(1, m + 1):
        H[i][0] = go + (i - 1) * ge if ladder_q else 0
    best, bi, bj = 0, 0, 0
    ec, eci, er, erj = NEG, 0, NEG, 0
    for i in range(1, m + 1):
        for j in range(1, n + 1):
            fo, fx = H[i - 1][j] + go, F[i - 1][j] + ge
            eo, ex = H[i][j - 1] + go, E[i][j - 1] + ge
            F[i][j], E[i][j] = max(fo, fx), max(eo, ex)
            ext_f[i][j], ext_e[i][j] = fx > fo, ex > eo
            d = H[i - 1][j - 1] + int(sub[min(rc[i - 1], ns - 1)][min(qc[j - 1], ns - 1)])
            h = max(d, E[i][j], F[i][j])
            if local and h <= 0:
                h, s = 0, 0
            else:
                s = 1 if h == d else 2 if h == E[i][j] else 3
            H[i][j], src[i][j] = h, s
            if local:
                if h > best:
                    best, bi, bj = h, i, j
            else:
                if j == n and h > ec:
                    ec, eci = h, i
                if i == m and h > er:
                    er, erj = h, j
    empty = False
    if local:
        score, ei, ej = best, bi, bj
        empty = best == 0
    elif not semi:
        score, ei, ej = H[m][n], m, n
    else:
        score, ei, ej = H[m][n], m, n
        if free_q and ec > score:
            score, ei, ej = ec, eci, n
        if free_r and er > score:
            score, ei, ej = er, m, erj
        if free_q and free_r and 0 > score:
            score, empty = 0, True
    res["score"] = score
    if empty:
        return _out(res)
    runs = []                      # end to start

    def push(op, length, merge):
        if length <= 0:
            return
        if merge and runs and runs[-1][0] == op:
            runs[-1][1] += length
        else:
            runs.append([op, length])

    i, j, state, cont = ei, ej, 0, False
    while True:
        if state == 0:
            if i == 0 or j == 0:
                if i == 0 and j > 0 and ladder_r:
                    push(I, j, False)
                    res["ins"] += j
                    res["gap_opens"] += 1
                    j = 0
                elif j == 0 and i > 0 and ladder_q:
                    push(D, i, False)
                    res["dele"] += i
                    res["gap_opens"] += 1
                    i = 0
                break
            s = src[i][j]
            if s == 0:
                break
            if s == 1:
                eq = qc[j - 1] == rc[i - 1] and qc[j - 1] < ns - 1
                push(EQ if eq else X, 1, True)
                res["matches" if eq else "mismatches"] += 1
                i, j = i - 1, j - 1
            else:
                state, cont = (1 if s == 2 else 2), False
                res["gap_opens"] += 1
        elif state == 1:
            push(I, 1, cont)
            res["ins"] += 1
            cont = ext_e[i][j]
            if not cont:
                state = 0
            j -= 1
        else:
            push(D, 1, cont)
            res["dele"] += 1
            cont = ext_f[i][j]
            if not cont:
                state = 0
            i -= 1
    res.update(q_begin=j, r_begin=i, q_end=ej, r_end=ei, ops=[(op, ln) for op, ln in reversed(runs)])
    return _out(res)


def _out(res: dict) -> dict:
    res["del"] = res.pop("dele")
    return res


def cigar_string(ops) -> str:
    return "".join(f"{ln}{OPCHAR[op]}" for op, ln in ops)


def cigar_score(qc, rc, sub, go: int, ge: int, aln: dict) -> int:
    """Affine cost of an alignment's CIGAR over its coordinates: substitution scores of
    the =/X columns plus go + (len-1)*ge per I or D run."""
    sub = np.asarray(sub, dtype=np.int64)
    ns = sub.shape[0]
    i, j, total = aln["r_begin"], aln["q_begin"], 0
    for op, ln in aln["ops"]:
        if op in (EQ, X):
            for _ in range(ln):
                total += int(sub[min(int(rc[i]), ns - 1)][min(int(qc[j]), ns - 1)])
                i, j = i + 1, j + 1
        elif op == I:
            total += go + (ln - 1) * ge
            j += ln
        else:
            total += go + (ln - 1) * ge
            i += ln
    assert (i, j) == (aln["r_end"], aln["q_end"]), "CIGAR does not span the coordinates"
    return total
