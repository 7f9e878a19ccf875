"""Pins the oracle before it is trusted as the parity target.

* known answers from the reference's own tests and docs (tests/golden/known_answers.json);
* exhaustive path enumeration (oracle/brute.py) over every mode and free-end flag pair,
  including gap_open > gap_extend and asymmetric matrices;
* closed forms: empty sequences, identical sequences in local mode, global all-gap ladders;
* invariants: local >= 0, swap symmetry (query<->reference with swapped flags and a
  transposed table), reversal symmetry;
* the C restatement equals the Python restatement.
"""
import json
from pathlib import Path

import numpy as np
import pytest

from gpusw_amd import Matrix, Scheme
from oracle import native
from oracle.brute import brute_score
from oracle.gotoh import cpu_reference_score, score_codes

GOLDEN = json.loads((Path(__file__).parent / "golden" / "known_answers.json").read_text())
MODES = [("local", False, False), ("global", False, False)] + [
    ("semiglobal", fq, fr) for fq in (False, True) for fr in (False, True)]


@pytest.mark.parametrize("case", GOLDEN["scores"], ids=lambda c: c["src"])
def test_known_answers(case):
    assert cpu_reference_score(case["query"], case["ref"], Scheme(**case["scheme"])) == case["score"]


def test_known_pairs():
    dna = Scheme()
    for case in GOLDEN["pairs"]:
        pairs = case["pairs"] or list(zip(range(len(case["queries"])), range(len(case["refs"]))))
        got = [cpu_reference_score(case["queries"][i], case["refs"][j], dna) for i, j in pairs]
        assert got == case["scores"], case["src"]


def test_global_below_semiglobal():
    c = GOLDEN["global_lt_semiglobal"]
    g = cpu_reference_score(c["query"], c["ref"], Scheme(mode="global"))
    s = cpu_reference_score(c["query"], c["ref"], Scheme(mode="semiglobal"))
    assert g < s == 4


@pytest.mark.parametrize("mode,fq,fr", MODES)
def test_brute_force(mode, fq, fr):
    rng = np.random.default_rng(17)
    for trial in range(120):
        n, m = int(rng.integers(0, 5)), int(rng.integers(0, 5))
        q = rng.integers(0, 5, n).tolist()
        r = rng.integers(0, 5, m).tolist()
        go = int(rng.choice([0, -1, -3, -5, -11]))
        ge = int(rng.choice([0, -1, -2, -4, -7]))
        if trial % 2:
            sub = rng.integers(-6, 7, (5, 5))
        else:
            sub = np.where(np.eye(5, dtype=bool), 2, -3)
        assert score_codes(q, r, sub, mode, go, ge, fq, fr) == brute_score(q, r, sub, mode, go, ge, fq, fr)


def test_brute_force_blosum():
    sub = Scheme.preset("blosum62").substitution_table()
    rng = np.random.default_rng(23)
    for _ in range(60):
        q = rng.integers(0, 25, int(rng.integers(1, 5))).tolist()
        r = rng.integers(0, 25, int(rng.integers(1, 5))).tolist()
        for mode, fq, fr in MODES:
            assert score_codes(q, r, sub, mode, -11, -1, fq, fr) == brute_score(q, r, sub, mode, -11, -1, fq, fr)


@pytest.mark.parametrize("mode,fq,fr", MODES)
def test_empty_edges_closed_form(mode, fq, fr):
    sub = np.where(np.eye(5, dtype=bool), 2, -3)
    go, ge = -5, -2
    semi = mode == "semiglobal"
    for n in range(0, 6):
        for m in range(0, 6):
            if n and m:
                continue
            got = score_codes([0] * n, [1] * m, sub, mode, go, ge, fq, fr)
            if mode == "local" or n == m == 0:
                want = 0
            elif n == 0:
                want = 0 if (semi and fq) else go + (m - 1) * ge
            else:
                want = 0 if (semi and fr) else go + (n - 1) * ge
            assert got == want


def test_identical_local_is_match_times_length():
    rng = np.random.default_rng(2)
    sub = np.where(np.eye(5, dtype=bool), 2, -3)
    for L in (1, 7, 31):
        s = rng.integers(0, 4, L).tolist()
        assert score_codes(s, s, sub, "local", -5, -2, True, True) == 2 * L


def test_global_ladder_against_gaps_only():
    # query of one residue that mismatches everything: best global = one substitution
    # plus one gap run, or two gap runs; checked against the brute force and a formula
    sub = np.full((5, 5), -10)
    for m in range(1, 5):
        got = score_codes([4], [0] * m, sub, "global", -5, -2, False, False)
        gap_only = (-5) + (-5 + (m - 1) * -2)        # query residue as gap + ref run
        with_sub = -10 + ((-5 + (m - 2) * -2) if m > 1 else 0)
        assert got == max(gap_only, with_sub)


@pytest.mark.parametrize("mode,fq,fr", MODES)
def test_invariants(mode, fq, fr):
    rng = np.random.default_rng(31)
    sub = rng.integers(-5, 6, (5, 5))
    for _ in range(40):
        q = rng.integers(0, 5, int(rng.integers(0, 12))).tolist()
        r = rng.integers(0, 5, int(rng.integers(0, 12))).tolist()
        s = score_codes(q, r, sub, mode, -4, -1, fq, fr)
        if mode == "local":
            assert s >= 0
        # swapping query and reference swaps the flags and transposes the table
        assert score_codes(r, q, sub.T, mode, -4, -1, fr, fq) == s
        # reversing both sequences keeps the optimum
        assert score_codes(q[::-1], r[::-1], sub, mode, -4, -1, fq, fr) == s


@pytest.mark.parametrize("mode,fq,fr", MODES)
def test_c_oracle_equals_python(mode, fq, fr):
    rng = np.random.default_rng(77)
    sub = rng.integers(-5, 6, (6, 6))
    n = 60
    lq, lr = rng.integers(0, 30, n), rng.integers(0, 30, n)
    qoff = np.concatenate([[0], np.cumsum(lq)]).astype(np.int32)
    roff = np.concatenate([[0], np.cumsum(lr)]).astype(np.int32)
    qc = rng.integers(0, 6, qoff[-1]).astype(np.uint8)
    rc = rng.integers(0, 6, roff[-1]).astype(np.uint8)
    qi = np.arange(n, dtype=np.int32)
    rj = rng.permutation(n).astype(np.int32)
    c = native.score_pairs(qc, qoff, rc, roff, qi, rj, sub, mode, -5, -2, fq, fr, nthreads=2)
    p = [score_codes(qc[qoff[a]:qoff[a + 1]], rc[roff[b]:roff[b + 1]], sub, mode, -5, -2, fq, fr)
         for a, b in zip(qi, rj)]
    assert c.tolist() == p
    cross = native.score_cross(qc, qoff, rc, roff, qi[:7], sub, mode, -5, -2, fq, fr, nthreads=2)
    for a in range(7):
        for b in range(n):
            assert cross[a, b] == score_codes(qc[qoff[a]:qoff[a + 1]], rc[roff[b]:roff[b + 1]], sub,
                                              mode, -5, -2, fq, fr)


def test_unknown_sentinel_blosum62():
    s = Scheme.preset("blosum62")
    # 'U' is outside the BLOSUM62 alphabet: it scores the sentinel (-4, the table minimum
    # off-diagonal) against everything, itself included.
    assert cpu_reference_score("U", "U", s) == 0
    t = s.substitution_table()
    assert t[-1].tolist() == [-4] * t.shape[0]
    assert cpu_reference_score("WU", "WU", s.replace(mode="global")) == 11 - 4
