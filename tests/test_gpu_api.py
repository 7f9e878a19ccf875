"""The engine and API anchors of the reference's own tests, on the HIP backend.

* launch-cap chunking of cross products, pairs and top-k (``tests/test_metal_bitexact.py:
  132-152``): a tiny ``max_threads_per_launch`` forces many launches, results equal the
  oracle;
* ``score_cross(query_batch=...)`` (``src/gpusw/aligner.py:113-141``);
* dict inputs give ids and shapes, ``return_ids=False`` gives an array, the memory-bounded
  ``top_k`` keeps no score matrix, a 1:1 count mismatch raises ``ValueError``, the one-call
  modes and the GCUPS/source introspection (``tests/test_api.py:13-79``).
"""
import numpy as np
import pytest

from gpusw_amd import Aligner, Scheme, align_pairs, align_score
from gpusw_amd.backends.hip import HipBackend
from oracle.gotoh import cpu_reference_matrix, cpu_reference_score
from tests.helpers import rand_seqs

pytestmark = pytest.mark.gpu

QUERIES = {"q_a": "ACGTACGTACGT", "q_b": "TTTTGGGG", "q_c": "ACGGACGT"}
REFS = {"r1": "ACGTACGTACGT", "r2": "ACGTAAAA", "r3": "GGGGTTTT", "r4": "TACGGACGTA"}


def test_launch_chunking_under_thread_cap(monkeypatch):
    monkeypatch.setattr(HipBackend, "max_threads_per_launch", 64)
    scheme = Scheme()
    rng = np.random.default_rng(7)
    queries = rand_seqs(rng, 9, 6, 20, "ACGT")
    refs = rand_seqs(rng, 11, 6, 24, "ACGT")          # 9 x 11 = 99 > 64: several launches
    al = Aligner(scheme, backend="hip").index(refs).set_queries(queries)
    cpu = cpu_reference_matrix(queries, refs, scheme)
    launches = []
    al._b.history = launches
    assert np.array_equal(al.score_cross(return_ids=False).astype(np.int64), cpu)
    assert len(launches) == 2                          # 5 rows x 11 refs <= 64, then 4 rows
    pairs = [(i, j) for i in range(len(queries)) for j in range(len(refs))]
    gp = align_pairs(queries, refs, pairs=pairs, scheme=scheme, backend="hip", return_ids=False)
    assert gp.astype(np.int64).tolist() == [cpu_reference_score(queries[i], refs[j], scheme)
                                            for i, j in pairs]
    tk = al.top_k(k=3)
    full = al.score_cross(return_ids=True)
    assert tk.top_k(3) == full.top_k(3)                # same ties: lower reference index first
    al._b.history = None


def test_pairs_chunked_calls():
    """max_threads_per_launch chunks one score_pairs call into library calls of that size."""
    rng = np.random.default_rng(8)
    queries = rand_seqs(rng, 30, 5, 60, "ACGT")
    refs = rand_seqs(rng, 20, 5, 80, "ACGT")
    qi = rng.integers(0, 30, 1000).astype(np.int32)
    rj = rng.integers(0, 20, 1000).astype(np.int32)
    al = Aligner("dna", backend="hip").index(refs).set_queries(queries)
    whole = al.score_pairs(qi, rj)
    al._b.max_threads_per_launch = 97
    calls = []
    al._b.history = calls
    chunked = al.score_pairs(qi, rj)
    al._b.history = None
    assert len(calls) == 11 and np.array_equal(whole, chunked)
    cpu = cpu_reference_matrix(queries, refs, Scheme())
    assert np.array_equal(whole.astype(np.int64), cpu[qi, rj])


def test_score_cross_query_batch():
    rng = np.random.default_rng(9)
    queries = rand_seqs(rng, 13, 5, 70, "ACGT")
    refs = rand_seqs(rng, 17, 5, 90, "ACGT")
    al = Aligner("dna", backend="hip").index(refs).set_queries(queries)
    whole = al.score_cross()
    for qb in (1, 4, 13, 50):
        assert np.array_equal(al.score_cross(query_batch=qb), whole)
    assert np.array_equal(whole.astype(np.int64), cpu_reference_matrix(queries, refs, Scheme()))


def test_align_score_shapes_and_ids():
    res = align_score(QUERIES, REFS, scheme="dna", backend="hip")
    assert res.scores.shape == (3, 4)
    assert res.scores.dtype == np.int32
    assert res.query_ids == ["q_a", "q_b", "q_c"]
    assert res.reference_ids == ["r1", "r2", "r3", "r4"]
    assert np.array_equal(res.scores.astype(np.int64),
                          cpu_reference_matrix(list(QUERIES.values()), list(REFS.values()), Scheme()))


def test_return_ids_false_is_ndarray():
    arr = align_score(QUERIES, REFS, scheme="dna", backend="hip", return_ids=False)
    assert isinstance(arr, np.ndarray) and arr.shape == (3, 4)


def test_top_k_matches_full_sort():
    al = Aligner("dna", backend="hip").index(REFS).set_queries(QUERIES)
    assert al.top_k(k=2).top_k(2) == al.score_cross(return_ids=True).top_k(2)


def test_topk_memory_bounded_result():
    al = Aligner("dna", backend="hip").index(REFS)
    tk = al.top_k(queries=QUERIES, k=2)
    assert tk.scores is None
    assert tk.topk_idx.shape == (3, 2)


def test_align_pairs_length_mismatch_raises():
    with pytest.raises(ValueError):
        align_pairs(["ACGT", "TT"], ["ACGT"], scheme="dna", backend="hip")


def test_one_call_modes():
    g = align_score(["AC"], ["TTACTT"], scheme="dna", mode="global", backend="hip", return_ids=False)
    s = align_score(["AC"], ["TTACTT"], scheme="dna", mode="semiglobal", backend="hip", return_ids=False)
    assert s[0, 0] == 4 and g[0, 0] < s[0, 0]


def test_gcups_and_source():
    al = Aligner("dna", backend="hip").index(REFS).set_queries(QUERIES)
    al.score_cross()
    assert al.gcups > 0
    src = al.kernel_source()
    assert "MODE=local" in src and "libgsw_hip.so" in src
