"""Parity at the benchmark workloads' full sizes: 16S-shape DNA cross (2,000 x 20,000),
protein top-k (1,000 x 100,000) and 1,000,000 read/window pairs.

The oracle cannot score whole matrices of this size in a test, so each workload is
checked through properties that hold at any size: the oracle on a random sample of
cells, the one-thread-per-pair comparator on whole rows or on every pair, the device
top-k against a stable sort of full rows, and the transpose law score(q, r) = score(r, q)
of a symmetric scheme in local mode, which runs references as queries (1,300-3,000
columns: the stripe kernel) against queries as references.
"""
import numpy as np
import pytest

from gpusw_amd import Aligner, Scheme, synth
from oracle import native

pytestmark = pytest.mark.gpu


def _subset(codes_off, idx):
    seqs = synth.unpack(*codes_off)
    return synth.pack([seqs[int(i)] for i in idx])


def _topk_rows(rows, k):
    order = np.argsort(-rows, axis=1, kind="stable")[:, :k]
    return order, np.take_along_axis(rows, order, axis=1)


def test_dna_16s_cross_full_size():
    (qc, qo), (rc, ro) = synth.dna_16s_cross(seed=0)
    scheme = Scheme.preset("dna")
    al = Aligner(scheme, backend="hip").index((rc, ro)).set_queries((qc, qo))
    full = al.score_cross()
    assert full.shape == (2000, 20000)
    rng = np.random.default_rng(101)
    qi = rng.integers(0, 2000, 3000).astype(np.int32)
    rj = rng.integers(0, 20000, 3000).astype(np.int32)
    want = native.score_encoded_pairs(al._queries, al._refs, qi, rj, scheme)
    assert np.array_equal(full[qi, rj].astype(np.int64), want)
    # whole rows from the comparator
    rows = np.sort(rng.choice(2000, 8, replace=False))
    plain = Aligner(scheme, backend="hip-plain").index((rc, ro)).set_queries(_subset((qc, qo), rows))
    assert np.array_equal(plain.score_cross(), full[rows])
    # references as queries (stripes) against queries as references
    refs = np.sort(rng.choice(20000, 48, replace=False))
    tr = Aligner(scheme, backend="hip").index((qc, qo)).set_queries(_subset((rc, ro), refs))
    assert np.array_equal(tr.score_cross(), full[:, refs].T)
    # device top-k of every row against a stable sort of sampled full rows
    tk = al.top_k(k=5)
    sample = rng.choice(2000, 64, replace=False)
    idx, sc = _topk_rows(full[sample], 5)
    assert np.array_equal(tk.topk_idx[sample], idx)
    assert np.array_equal(tk.topk_scores[sample], sc)


def test_protein_topk_full_size():
    (qc, qo), (rc, ro) = synth.protein_topk(seed=0)
    scheme = Scheme.preset("blosum62")
    al = Aligner(scheme, backend="hip").index((rc, ro)).set_queries((qc, qo))
    tk = al.top_k(k=5)
    assert tk.topk_idx.shape == (1000, 5)
    rng = np.random.default_rng(202)
    rows = np.sort(rng.choice(1000, 24, replace=False))
    sub = Aligner(scheme, backend="hip").index((rc, ro)).set_queries(_subset((qc, qo), rows))
    block = sub.score_cross()
    idx, sc = _topk_rows(block, 5)
    assert np.array_equal(tk.topk_idx[rows], idx)
    assert np.array_equal(tk.topk_scores[rows], sc)
    # the block against the oracle on sampled cells
    qi = rng.integers(0, len(rows), 1500).astype(np.int32)
    rj = rng.integers(0, 100000, 1500).astype(np.int32)
    want = native.score_encoded_pairs(sub._queries, sub._refs, qi, rj, scheme)
    assert np.array_equal(block[qi, rj].astype(np.int64), want)
    # transpose law on 16 references (up to 3,000 aa: stripes) x all queries
    refs = np.sort(rng.choice(100000, 16, replace=False))
    fwd = Aligner(scheme, backend="hip").index(_subset((rc, ro), refs)).set_queries((qc, qo)).score_cross()
    rev = Aligner(scheme, backend="hip").index((qc, qo)).set_queries(_subset((rc, ro), refs)).score_cross()
    assert np.array_equal(rev, fwd.T)


@pytest.mark.parametrize("mode", ["semiglobal", "global"])
def test_read_pairs_full_size(mode):
    (rc, ro), (wc, wo), qi, rj = synth.read_pairs(seed=0)
    scheme = Scheme.preset("dna").replace(mode=mode, dtype="int16")
    al = Aligner(scheme, backend="hip").index((wc, wo)).set_queries((rc, ro))
    got = al.score_pairs(qi, rj)
    assert got.shape == (1_000_000,) and al._b.last_stats.plain_pairs == 0
    plain = Aligner(scheme, backend="hip-plain").index((wc, wo)).set_queries((rc, ro)).score_pairs(qi, rj)
    assert np.array_equal(got, plain)
    rng = np.random.default_rng(303)
    pick = rng.choice(qi.size, 5000, replace=False)
    want = native.score_encoded_pairs(al._queries, al._refs, qi[pick], rj[pick], scheme)
    assert np.array_equal(got[pick].astype(np.int64), want)
    perm = rng.permutation(qi.size)
    assert np.array_equal(al.score_pairs(qi[perm], rj[perm]), got[perm])
