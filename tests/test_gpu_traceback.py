"""GPU traceback (gsw_traceback_pairs through Aligner.align) == the traceback oracle,
field by field and CIGAR by CIGAR, in every mode and free-end setting, on the reference's
scheme set, empty and long sequences, and for the hits of a top-k."""
import itertools

import numpy as np
import pytest

from gpusw_amd import Aligner, Scheme
from gpusw_amd.cli import main
from gpusw_amd.encode import funnel
from oracle.traceback import cigar_string, traceback_codes
from tests.helpers import PROT, rand_seqs, schemes, stable_seed

pytestmark = pytest.mark.gpu

SCHEMES = schemes()


def _oracle(al, qi, rj):
    s = al.scheme
    sub = s.substitution_table()
    q, r = al._queries, al._refs
    out = []
    for a, b in zip(qi, rj):
        qc = q.codes[q.offsets[a]:q.offsets[a + 1]]
        rc = r.codes[r.offsets[b]:r.offsets[b + 1]]
        out.append(traceback_codes(qc, rc, sub, s.mode, s.gap_open, s.gap_extend,
                                   s.free_end_gaps_query, s.free_end_gaps_ref))
    return out


def _compare(got, want):
    for g, w in zip(got, want):
        assert (g.score, g.q_begin, g.q_end, g.r_begin, g.r_end) == \
               (w["score"], w["q_begin"], w["q_end"], w["r_begin"], w["r_end"])
        assert g.cigar == cigar_string(w["ops"])
        assert (g.matches, g.mismatches, g.insertions, g.deletions, g.gap_opens) == \
               (w["matches"], w["mismatches"], w["ins"], w["del"], w["gap_opens"])


@pytest.mark.parametrize("name", list(SCHEMES))
def test_random_pairs_equal_oracle(name):
    scheme, alpha = SCHEMES[name]
    rng = np.random.default_rng(stable_seed("tb" + name))
    queries = rand_seqs(rng, 25, 0, 70, alpha) + [""]
    refs = rand_seqs(rng, 25, 0, 90, alpha) + [""]
    # related pairs (a mutated copy) as well as random ones
    refs += [q[3:] + "".join(rng.choice(list(alpha), 5)) for q in queries[:10]]
    al = Aligner(scheme, backend="hip").index(refs).set_queries(queries)
    qi = rng.integers(0, len(queries), 300).astype(np.int32)
    rj = rng.integers(0, len(refs), 300).astype(np.int32)
    qi[-10:] = np.arange(10)
    rj[-10:] = np.arange(len(refs) - 10, len(refs))
    got = al.align(qi, rj)
    _compare(got, _oracle(al, qi, rj))
    assert [a.score for a in got] == al.score_pairs(qi, rj).tolist()


@pytest.mark.parametrize("fq,fr", list(itertools.product([False, True], repeat=2)))
def test_semiglobal_free_ends_and_open_below_extend(fq, fr):
    rng = np.random.default_rng(31 + 2 * fq + fr)
    scheme = Scheme(mode="semiglobal", match=3, mismatch=-2, gap_open=-1, gap_extend=-3,
                    free_end_gaps_query=fq, free_end_gaps_ref=fr)
    queries = rand_seqs(rng, 20, 0, 40, "ACGT")
    refs = rand_seqs(rng, 20, 0, 60, "ACGT")
    al = Aligner(scheme, backend="hip").index(refs).set_queries(queries)
    qi, rj = np.repeat(np.arange(20), 20).astype(np.int32), np.tile(np.arange(20), 20).astype(np.int32)
    _compare(al.align(qi, rj), _oracle(al, qi, rj))


def test_long_protein_pairs():
    rng = np.random.default_rng(32)
    scheme = Scheme.preset("blosum62")
    base = rand_seqs(rng, 4, 900, 1200, PROT)
    queries = base + [b[100:700] for b in base]
    refs = [b[::-1] for b in base] + [b[50:] for b in base]
    al = Aligner(scheme, backend="hip").index(refs).set_queries(queries)
    qi = np.arange(8, dtype=np.int32)
    rj = np.array([4, 5, 6, 7, 4, 5, 6, 7], dtype=np.int32)
    got = al.align(qi, rj)
    _compare(got, _oracle(al, qi, rj))
    assert got[4].identity == 1.0 and got[4].length == 600   # base[k][100:700] inside base[k][50:]


@pytest.mark.parametrize("qlen", [1, 63, 64, 65, 257, 1024, 1025, 2048, 2049])
@pytest.mark.parametrize("mode", ["local", "global", "semiglobal"])
def test_fill_paths_at_their_limits(qlen, mode):
    """Queries up to 2,048 columns take the wave-parallel fill (16 lanes per pair up to
    1,024 columns, else 32, with the smallest block of columns per lane that covers the
    longest query); longer ones the one-thread fill.  All equal the oracle at the lengths
    where the lane count or block size changes."""
    rng = np.random.default_rng(qlen * 7 + len(mode))
    scheme = Scheme.preset("dna").replace(mode=mode)
    base = ["".join(rng.choice(list("ACGT"), qlen)) for _ in range(3)]
    queries = base + [b[: max(1, qlen // 3)] for b in base]
    refs = [b[max(0, qlen // 2 - 40): qlen // 2 + 40] + "".join(rng.choice(list("ACGT"), 30)) for b in base] \
        + rand_seqs(rng, 3, 20, 120, "ACGT")
    al = Aligner(scheme, backend="hip").index(refs).set_queries(queries)
    qi = np.repeat(np.arange(len(queries)), 2).astype(np.int32)
    rj = np.tile([0, 4], len(queries)).astype(np.int32)
    rj[::3] = np.arange(len(rj[::3])) % len(refs)
    got = al.align(qi, rj)
    _compare(got, _oracle(al, qi, rj))


def test_top_k_alignments_and_cli(tmp_path):
    rng = np.random.default_rng(33)
    refs = [(f"r{i}", s) for i, s in enumerate(rand_seqs(rng, 60, 80, 200, "ACGT"))]
    queries = [(f"q{i}", refs[i * 7][1][10:70]) for i in range(5)]
    al = Aligner("dna", backend="hip").index(refs).set_queries(queries)
    res = al.top_k_alignments(k=3)
    tk = al.top_k(k=3)
    for qn, (qid, alns) in enumerate(res):
        assert qid == f"q{qn}" and [a.score for a in alns] == tk.topk_scores[qn].tolist()
        best = alns[0]
        assert best.reference_id == f"r{qn * 7}" and best.cigar == "60=" and best.r_begin == 10
    qf, rf = tmp_path / "q.fa", tmp_path / "r.fa"
    qf.write_text("".join(f">{i}\n{s}\n" for i, s in queries))
    rf.write_text("".join(f">{i}\n{s}\n" for i, s in refs))
    out = tmp_path / "hits.csv"
    assert main(["topk", "--queries", str(qf), "--refs", str(rf), "-k", "2", "--traceback", "-o", str(out)]) == 0
    rows = [l.split(",") for l in out.read_text().splitlines()]
    assert rows[0][-1] == "cigar" and rows[1][2] == "r0" and rows[1][9] == "60=" and float(rows[1][4]) == 1.0


@pytest.mark.parametrize("npairs", [1, 7, 301])
def test_odd_pair_counts(npairs):
    """Odd launch sizes: the per-pair offset lists and records after the two int32 index
    lists keep their 8-byte alignment."""
    scheme, alpha = SCHEMES["dna_local"]
    rng = np.random.default_rng(stable_seed(f"odd{npairs}"))
    queries = rand_seqs(rng, 13, 1, 120, alpha)
    refs = rand_seqs(rng, 11, 1, 150, alpha)
    al = Aligner(scheme, backend="hip").index(refs).set_queries(queries)
    qi = rng.integers(0, len(queries), npairs).astype(np.int32)
    rj = rng.integers(0, len(refs), npairs).astype(np.int32)
    _compare(al.align(qi, rj), _oracle(al, qi, rj))
