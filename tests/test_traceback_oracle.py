"""Pins the traceback oracle (oracle/traceback.py) without a GPU: its score is the Gotoh
oracle's score, the affine cost of its CIGAR over its coordinates is that score, its
counts agree with its CIGAR, and hand-checked alignments come out as expected."""
import itertools

import numpy as np
import pytest

from gpusw_amd import Scheme
from oracle.gotoh import score_codes
from oracle.traceback import EQ, I, D, X, cigar_score, cigar_string, traceback_codes
from tests.helpers import stable_seed

SETTINGS = [("local", False, False), ("global", False, False)] + [
    ("semiglobal", fq, fr) for fq, fr in itertools.product([False, True], repeat=2)]


def _check(qc, rc, sub, mode, go, ge, fq, fr):
    aln = traceback_codes(qc, rc, sub, mode, go, ge, fq, fr)
    assert aln["score"] == score_codes(qc, rc, sub, mode, go, ge, fq, fr)
    # the CIGAR spans the coordinates (asserted inside cigar_score) and costs the score
    assert cigar_score(qc, rc, sub, go, ge, aln) == aln["score"]
    cols = {EQ: 0, X: 0, I: 0, D: 0}
    for op, ln in aln["ops"]:
        assert ln > 0
        cols[op] += ln
    assert (cols[EQ], cols[X], cols[I], cols[D]) == (aln["matches"], aln["mismatches"], aln["ins"], aln["del"])
    assert aln["q_end"] - aln["q_begin"] == cols[EQ] + cols[X] + cols[I]
    assert aln["r_end"] - aln["r_begin"] == cols[EQ] + cols[X] + cols[D]
    if mode == "global":
        assert (aln["q_begin"], aln["q_end"], aln["r_begin"], aln["r_end"]) == (0, len(qc), 0, len(rc))
    return aln


@pytest.mark.parametrize("mode,fq,fr", SETTINGS)
def test_random_small_cases(mode, fq, fr):
    rng = np.random.default_rng(stable_seed(f"{mode}/{fq}/{fr}"))
    for trial in range(150):
        k = int(rng.integers(2, 6))
        sub = rng.integers(-4, 5, (k + 1, k + 1))
        sub[k, :] = sub[:, k] = int(sub[:k, :k].min())
        go, ge = sorted(rng.integers(-6, 0, 2).tolist())
        if trial % 3 == 0:
            go, ge = ge, go                      # gap_open > gap_extend tables too
        qc = rng.integers(0, k + 1, int(rng.integers(0, 10)))
        rc = rng.integers(0, k + 1, int(rng.integers(0, 10)))
        _check(qc, rc, sub, mode, go, ge, fq, fr)


def test_presets_medium():
    rng = np.random.default_rng(7)
    for name in ("dna", "blosum62"):
        s = Scheme.preset(name)
        sub = s.substitution_table()
        k = len(s.eff_alphabet)
        for mode, fq, fr in SETTINGS:
            for _ in range(4):
                qc = rng.integers(0, k, int(rng.integers(20, 60)))
                rc = rng.integers(0, k, int(rng.integers(20, 80)))
                _check(qc, rc, sub, mode, s.gap_open, s.gap_extend, fq, fr)


def _enc(text):
    return [("ACGT".index(c)) for c in text]


def test_hand_checked():
    sub = Scheme.preset("dna").substitution_table()
    # identical sequences: one run of matches
    a = traceback_codes(_enc("ACGTACGTAC"), _enc("ACGTACGTAC"), sub, "local", -5, -2, True, True)
    assert (a["score"], cigar_string(a["ops"]), a["q_begin"], a["q_end"]) == (20, "10=", 0, 10)
    # semiglobal, free query ends: the query sits inside the reference (reference tests:93-94)
    a = traceback_codes(_enc("ACGT"), _enc("TTACGTTT"), sub, "semiglobal", -5, -2, True, False)
    assert (a["score"], cigar_string(a["ops"]), a["r_begin"], a["r_end"]) == (8, "4=", 2, 6)
    # global with a 2-residue gap in the query
    a = traceback_codes(_enc("ACAC"), _enc("ACGTAC"), sub, "global", -5, -2, False, False)
    assert a["score"] == 1 and cigar_string(a["ops"]) == "2=2D2="
    # local with a mismatch kept inside
    a = traceback_codes(_enc("ACGTTCGT"), _enc("ACGTACGT"), sub, "local", -5, -2, True, True)
    assert a["score"] == 2 * 7 - 3 and cigar_string(a["ops"]) == "4=1X3="
