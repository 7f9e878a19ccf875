"""GPU scores == oracle, bit-exact, through the C ABI (ctypes backends).

Mirrors the reference's tests/test_gpu_bitexact.py (schemes :20-30, random and edge
cross products :33-52, 1:1 and explicit pairs :55-76, int32 promotion :79-102, known
answers :114-118) for both the batched wavefront backend ("hip") and the
one-thread-per-pair comparator ("hip-plain"), and adds cases the batched path needs:
every compiled tile width, long queries, ragged lengths, and int16 saturation.
"""
import json
from pathlib import Path

import numpy as np
import pytest

from gpusw_amd import Aligner, Scheme, align_pairs, align_score
from gpusw_amd import synth
from gpusw_amd.encode import funnel
from oracle import native
from oracle.gotoh import cpu_reference_matrix, cpu_reference_score
from tests.helpers import rand_seqs, schemes, stable_seed

pytestmark = pytest.mark.gpu

BACKENDS = ["hip", "hip-plain"]
SCHEMES = schemes()
GOLDEN = json.loads((Path(__file__).parent / "golden" / "known_answers.json").read_text())


@pytest.mark.parametrize("backend", BACKENDS)
@pytest.mark.parametrize("name", list(SCHEMES))
def test_cross_random(name, backend):
    scheme, alpha = SCHEMES[name]
    rng = np.random.default_rng(stable_seed(name))
    queries = rand_seqs(rng, 6, 5, 28, alpha)
    refs = rand_seqs(rng, 9, 5, 32, alpha)
    got = align_score(queries, refs, scheme=scheme, backend=backend, return_ids=False)
    assert got.dtype == np.int32
    assert np.array_equal(got.astype(np.int64), cpu_reference_matrix(queries, refs, scheme))


@pytest.mark.parametrize("backend", BACKENDS)
@pytest.mark.parametrize("name", list(SCHEMES))
def test_cross_edges(name, backend):
    scheme, alpha = SCHEMES[name]
    c0 = alpha[0]
    queries = ["", c0, c0 * 30, "".join(alpha[:4]), alpha[:1]]
    refs = ["", c0, "".join(reversed(alpha)) * 2, c0 * 25]
    got = align_score(queries, refs, scheme=scheme, backend=backend, return_ids=False)
    assert np.array_equal(got.astype(np.int64), cpu_reference_matrix(queries, refs, scheme))


@pytest.mark.parametrize("backend", BACKENDS)
@pytest.mark.parametrize("name", list(SCHEMES))
def test_pairs_random(name, backend):
    scheme, alpha = SCHEMES[name]
    rng = np.random.default_rng(stable_seed(name) ^ 0x5151)
    queries = rand_seqs(rng, 12, 5, 28, alpha)
    refs = rand_seqs(rng, 12, 5, 32, alpha)
    got = align_pairs(queries, refs, scheme=scheme, backend=backend, return_ids=False)
    want = [cpu_reference_score(q, r, scheme) for q, r in zip(queries, refs)]
    assert got.astype(np.int64).tolist() == want


@pytest.mark.parametrize("backend", BACKENDS)
@pytest.mark.parametrize("name", list(SCHEMES))
def test_pairs_edges(name, backend):
    scheme, alpha = SCHEMES[name]
    c0 = alpha[0]
    qs = ["", c0, c0 * 30, "".join(alpha[:4])]
    rs = ["", c0, "".join(reversed(alpha)) * 2, c0 * 25]
    pairs = [(i, j) for i in range(len(qs)) for j in range(len(rs))]
    got = align_pairs(qs, rs, pairs=pairs, scheme=scheme, backend=backend, return_ids=False)
    assert got.astype(np.int64).tolist() == [cpu_reference_score(qs[i], rs[j], scheme) for i, j in pairs]


@pytest.mark.parametrize("backend", BACKENDS)
def test_known_answers(backend):
    for case in GOLDEN["scores"]:
        s = Scheme(**case["scheme"])
        got = align_score([case["query"]], [case["ref"]], scheme=s, backend=backend, return_ids=False)
        assert int(got[0, 0]) == case["score"], case["src"]
    for case in GOLDEN["pairs"]:
        res = align_pairs(case["queries"], case["refs"], pairs=case["pairs"], scheme="dna",
                          backend=backend, return_ids=False)
        assert res.tolist() == case["scores"], case["src"]
    for case in GOLDEN["topk"]:
        res = align_score(case["queries"], case["refs"], scheme="dna", backend=backend)
        assert [[q, [tuple(h) for h in hits]] for q, hits in res.top_k(case["k"])] == \
            [[q, [tuple(h) for h in hits]] for q, hits in case["expect"]]


@pytest.mark.parametrize("backend", BACKENDS)
def test_int32_promotion(backend):
    scheme = Scheme(mode="local", match=1000, mismatch=-3, gap_open=-5, gap_extend=-2)
    qs = ["ACGT" * 15, "ACGTAC" * 5, "AAAA"]
    rs = ["ACGT" * 15, "ACGTAC" * 5, "TTTT", "ACGT" * 15]
    pairs = [(0, 0), (1, 1), (2, 2), (0, 3)]
    res = align_pairs(qs, rs, pairs=pairs, scheme=scheme, backend=backend, return_ids=False)
    cpu = cpu_reference_matrix(qs, rs, scheme)
    assert res.dtype == np.int32
    assert res.tolist() == [int(cpu[i, j]) for i, j in pairs]
    full = align_score(qs, rs, scheme=scheme, backend=backend, return_ids=False)
    assert np.array_equal(full.astype(np.int64), cpu)



@pytest.mark.parametrize("mode", ["local", "global", "semi_qr", "semi_q", "semi_r"])
def test_cross_all_tile_widths(mode):
    """Query lengths 1..700 across the compiled tile widths (a call this small runs its
    single-stripe queries on the widest chosen tile; every tile on its own is covered by
    the forced-tile tests), stripes beyond; references are ragged."""
    kw = {"local": dict(mode="local"), "global": dict(mode="global"),
          "semi_qr": dict(mode="semiglobal"),
          "semi_q": dict(mode="semiglobal", free_end_gaps_query=True, free_end_gaps_ref=False),
          "semi_r": dict(mode="semiglobal", free_end_gaps_query=False, free_end_gaps_ref=True)}[mode]
    scheme = Scheme(**kw)
    rng = np.random.default_rng(stable_seed("widths" + mode))
    qlens = list(range(1, 80)) + list(rng.integers(80, 701, 40))
    queries = ["".join(rng.choice(list("ACGTN"), int(L))) for L in qlens]
    refs = rand_seqs(rng, 37, 1, 400, "ACGT")
    got = align_score(queries, refs, scheme=scheme, backend="hip", return_ids=False)
    qe, re_ = funnel(queries, scheme), funnel(refs, scheme)
    want = native.score_encoded_cross(qe, re_, scheme)
    bad = np.argwhere(got.astype(np.int64) != want)
    assert bad.size == 0, f"{len(bad)} mismatches, first {bad[:5].tolist()}"


@pytest.mark.parametrize("mode", ["local", "global", "semiglobal"])
def test_pairs_ragged(mode):
    scheme = Scheme(mode=mode)
    rng = np.random.default_rng(stable_seed("ragged" + mode))
    queries = rand_seqs(rng, 300, 1, 500, "ACGT")
    refs = rand_seqs(rng, 50, 1, 600, "ACGT")
    qi = rng.integers(0, 300, 3000).astype(np.int32)
    rj = rng.integers(0, 50, 3000).astype(np.int32)
    al = Aligner(scheme, backend="hip").index(refs).set_queries(queries)
    got = al.score_pairs(qi, rj)
    want = native.score_encoded_pairs(al._queries, al._refs, qi, rj, scheme)
    bad = np.flatnonzero(got.astype(np.int64) != want)
    assert bad.size == 0, f"{bad.size} mismatches, first {bad[:5].tolist()}"


def test_protein_long_queries():
    """Queries longer than the widest tile run as stripes and stay exact."""
    scheme = Scheme.preset("blosum62")
    rng = np.random.default_rng(11)
    queries = rand_seqs(rng, 3, 2000, 2600, "ARNDCQEGHILKMFPSTWYV")
    refs = rand_seqs(rng, 5, 100, 900, "ARNDCQEGHILKMFPSTWYV")
    got = align_score(queries, refs, scheme=scheme, backend="hip", return_ids=False)
    want = native.score_encoded_cross(funnel(queries, scheme), funnel(refs, scheme), scheme)
    assert np.array_equal(got.astype(np.int64), want)


def test_int16_saturation_rescored():
    """Local int16 results that reach 32767 are rescored in int32."""
    scheme = Scheme(mode="local", match=300, mismatch=-3, gap_open=-5, gap_extend=-2)
    rng = np.random.default_rng(5)
    base = rng.choice(list("ACGT"), 250)
    q = "".join(base)
    refs = [q, q[:100], "".join(rng.choice(list("ACGT"), 300)), q + q]
    got = align_score([q, q[:50]], refs, scheme=scheme, backend="hip", return_ids=False)
    want = cpu_reference_matrix([q, q[:50]], refs, scheme)
    assert int(want.max()) > 32767
    assert np.array_equal(got.astype(np.int64), want)


def test_plain_dna_config_bitexact():
    """64 x 256 DNA local int32 (seed 0): comparator and batched == oracle."""
    (qc, qo), (rc, ro) = synth.plain_dna(seed=0)
    scheme = Scheme.preset("dna").replace(dtype="int32")
    want = native.score_cross(qc, qo, rc, ro, np.arange(64, dtype=np.int32),
                              scheme.substitution_table(), "local", -5, -2, True, True)
    for backend in BACKENDS:
        al = Aligner(scheme, backend=backend).index((rc, ro)).set_queries((qc, qo))
        got = al.score_cross()
        assert np.array_equal(got.astype(np.int64), want), backend
        assert al.gcups > 0


def test_topk_ties_to_lower_index():
    rng = np.random.default_rng(3)
    refs = ["ACGT" * 5] * 7 + rand_seqs(rng, 30, 5, 40, "ACGT") + ["ACGT" * 5] * 5
    queries = ["ACGT" * 5, "TTTT", "ACGTTGCA"]
    al = Aligner("dna", backend="hip").index(refs).set_queries(queries)
    full = al.score_cross()
    tk = al.top_k(k=9)
    order = np.argsort(-full, axis=1, kind="stable")[:, :9]
    assert np.array_equal(tk.topk_idx, order)
    assert np.array_equal(tk.topk_scores, np.take_along_axis(full, order, axis=1))
    assert tk.top_k(9) == al.score_cross(return_ids=True).top_k(9)
