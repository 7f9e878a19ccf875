"""Explicit pairs through the graph-replayed batches (131,072 pairs per replay).

Each call replays the two graph slots several times; the results must equal the eager
one-thread-per-pair kernel on every pair and the oracle on a random sample, and a second
call after unrelated eager work must reproduce them exactly."""
import numpy as np
import pytest

from gpusw_amd import Aligner, Scheme, align_score, synth
from oracle import native

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("mode", ["semiglobal", "global", "local"])
def test_pair_batches_match_eager_and_oracle(mode):
    (rc, ro), (wc, wo), qi, rj = synth.read_pairs(seed=3, n_windows=400, n_reads=300000, read_len=60)
    scheme = Scheme.preset("dna").replace(mode=mode)
    al = Aligner(scheme, backend="hip").index((wc, wo)).set_queries((rc, ro))
    first = al.score_pairs(qi, rj)
    assert al._b.last_stats.launches >= 3          # three graph replays (and no comparator)
    assert al._b.last_stats.plain_pairs == 0
    eager = Aligner(scheme, backend="hip-plain").index((wc, wo)).set_queries((rc, ro)).score_pairs(qi, rj)
    assert np.array_equal(first, eager)
    rng = np.random.default_rng(0)
    pick = rng.choice(qi.size, 3000, replace=False)
    want = native.score_encoded_pairs(al._queries, al._refs, qi[pick], rj[pick], scheme)
    assert np.array_equal(first[pick].astype(np.int64), want)
    # unrelated eager work, then the same call again
    align_score(["ACGTTGCA"] * 3, ["ACGT", "TTTT"], scheme=scheme, backend="hip")
    again = al.score_pairs(qi, rj)
    assert np.array_equal(first, again)
    # a shuffled order of the same pairs gives the same scores per pair
    perm = rng.permutation(qi.size)
    shuffled = al.score_pairs(qi[perm], rj[perm])
    assert np.array_equal(shuffled, first[perm])
