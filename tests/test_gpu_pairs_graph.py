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


def test_many_batches_two_tiles_slot_reuse():
    """1.3M pairs of two read lengths: one tile gets 9 batches (more than its 8 slots), so planning
    threads reuse pinned slots after their earlier replays finish."""
    rng = np.random.default_rng(17)
    nw = 300
    wlen = rng.integers(200, 401, nw)
    windows = [rng.integers(0, 4, int(L)).astype(np.uint8) for L in wlen]
    n = 1_300_000
    rj = rng.integers(0, nw, n).astype(np.int32)
    rlen = np.where(rng.random(n) < 0.15, 40, 150)   # ~1.1M of 150 nt: 9 batches of one tile
    reads = []
    for j, L in zip(rj.tolist(), rlen.tolist()):
        w = windows[j]
        s0 = int(rng.integers(0, w.size - L + 1))
        reads.append(w[s0:s0 + L])
    qi = np.arange(n, dtype=np.int32)
    scheme = Scheme.preset("dna").replace(mode="semiglobal")
    al = Aligner(scheme, backend="hip").index(synth.pack(windows)).set_queries(synth.pack(reads))
    got = al.score_pairs(qi, rj)
    assert al._b.last_stats.launches >= 10 and al._b.last_stats.plain_pairs == 0
    # cells counted by the planning threads == sum of lq * lr
    assert al._b.last_stats.cells == int((rlen.astype(np.int64) * wlen[rj].astype(np.int64)).sum())
    eager = Aligner(scheme, backend="hip-plain").index(synth.pack(windows)).set_queries(synth.pack(reads)).score_pairs(qi, rj)
    assert np.array_equal(got, eager)
    pick = rng.choice(n, 2000, replace=False)
    want = native.score_encoded_pairs(al._queries, al._refs, qi[pick], rj[pick], scheme)
    assert np.array_equal(got[pick].astype(np.int64), want)


@pytest.mark.parametrize("mode", ["local", "global", "semiglobal"])
def test_head_and_rest_segments_mixed(mode):
    """600,000 pairs (a head segment of 344,064 pairs planned before the rest is classified)
    mixing empty sequences, int16 and int32 pairs of several tiles, queries longer than any
    pairs tile against shorter references (run transposed) and pairs longer than every tile
    on both sides (int32 stripe kernels): every pair equals the comparator, a sample equals the oracle,
    and a shuffled order (other pairs in the head) gives the same score per pair."""
    rng = np.random.default_rng(23)
    qlen = rng.integers(1, 600, 2000)
    qlen[:20] = 0
    qlen[20:30] = 2100
    rlen = rng.integers(1, 500, 500)
    rlen[:5] = 0
    rlen[5:8] = 2200
    queries = [rng.integers(0, 4, int(L)).astype(np.uint8) for L in qlen]
    refs = [rng.integers(0, 4, int(L)).astype(np.uint8) for L in rlen]
    n = 600_000
    qi = rng.integers(0, 2000, n).astype(np.int32)
    rj = rng.integers(0, 500, n).astype(np.int32)
    # match 100: pairs whose shorter side passes 327 need int32
    scheme = Scheme(mode=mode, match=100, mismatch=-3, gap_open=-5, gap_extend=-2)
    al = Aligner(scheme, backend="hip").index(synth.pack(refs)).set_queries(synth.pack(queries))
    got = al.score_pairs(qi, rj)
    st = al._b.last_stats
    both_long = int(((qlen[qi] > 2048) & (rlen[rj] > 2048)).sum())
    assert both_long > 0 and st.plain_pairs == 0 and st.launches >= 8
    assert st.cells == int((qlen[qi].astype(np.int64) * rlen[rj].astype(np.int64)).sum())
    eager = Aligner(scheme, backend="hip-plain").index(synth.pack(refs)).set_queries(
        synth.pack(queries)).score_pairs(qi, rj)
    assert np.array_equal(got, eager)
    pick = rng.choice(n, 2000, replace=False)
    want = native.score_encoded_pairs(al._queries, al._refs, qi[pick], rj[pick], scheme)
    assert np.array_equal(got[pick].astype(np.int64), want)
    perm = rng.permutation(n)
    assert np.array_equal(al.score_pairs(qi[perm], rj[perm]), got[perm])


@pytest.mark.parametrize("mode", ["local", "global", "semiglobal"])
def test_protein_pairs_table_kernel(mode):
    """Protein queries past 160 columns have no profile tile that fits LDS: the table
    kernel scores them (no comparator), bit-exact."""
    from tests.helpers import PROT, rand_seqs

    rng = np.random.default_rng(41)
    scheme = Scheme.preset("blosum62").replace(mode=mode)
    queries = rand_seqs(rng, 60, 150, 900, PROT)
    refs = rand_seqs(rng, 40, 50, 700, PROT) + [queries[0][20:500]]
    qi = rng.integers(0, len(queries), 2000).astype(np.int32)
    rj = rng.integers(0, len(refs), 2000).astype(np.int32)
    qi[0], rj[0] = 0, len(refs) - 1
    al = Aligner(scheme, backend="hip").index(refs).set_queries(queries)
    got = al.score_pairs(qi, rj)
    assert al._b.last_stats.plain_pairs == 0
    want = native.score_encoded_pairs(al._queries, al._refs, qi, rj, scheme)
    assert np.array_equal(got.astype(np.int64), want)


@pytest.mark.parametrize("mode", ["local", "global", "semiglobal"])
@pytest.mark.parametrize("tab", ["0", "1", None])
@pytest.mark.parametrize("ref_codes", [4, 5])
def test_dna_pairs_both_kernels(mode, tab, ref_codes, monkeypatch):
    """DNA pairs through the table kernel (GSW_PAIRS_TAB=1), through per-lane profiles
    (GSW_PAIRS_TAB=0) and the default choice, with references of ACGT only (profiles of four
    rows) and with N (all five rows); the queries hold N either way."""
    if tab is None:
        monkeypatch.delenv("GSW_PAIRS_TAB", raising=False)
    else:
        monkeypatch.setenv("GSW_PAIRS_TAB", tab)
    rng = np.random.default_rng(42)
    scheme = Scheme.preset("dna").replace(mode=mode)
    queries = [rng.integers(0, 5, int(L)).astype(np.uint8) for L in rng.integers(1, 400, 300)]
    refs = [rng.integers(0, ref_codes, int(L)).astype(np.uint8) for L in rng.integers(1, 500, 80)]
    qi = rng.integers(0, 300, 5000).astype(np.int32)
    rj = rng.integers(0, 80, 5000).astype(np.int32)
    al = Aligner(scheme, backend="hip").index(synth.pack(refs)).set_queries(synth.pack(queries))
    got = al.score_pairs(qi, rj)
    # default: the table kernel (4: its byte-select form, on half2 tiles up to 192 columns)
    assert al._b.last_stats.top_striped in ((2,) if tab == "0" else (3, 4))
    want = native.score_encoded_pairs(al._queries, al._refs, qi, rj, scheme)
    assert np.array_equal(got.astype(np.int64), want)


@pytest.mark.parametrize("n", [1000, 200_000, 600_000])
def test_pair_index_out_of_range_raises_index_error(n):
    """The library checks pair indices while it plans (single- and multi-threaded, and in
    the rest of a call whose head already runs) and the host maps its error to IndexError,
    as indexing the sequence sets would."""
    rng = np.random.default_rng(5)
    refs = [rng.integers(0, 4, 50).astype(np.uint8) for _ in range(20)]
    qs = [rng.integers(0, 4, 30).astype(np.uint8) for _ in range(10)]
    al = Aligner(Scheme.preset("dna"), backend="hip").index(synth.pack(refs)).set_queries(synth.pack(qs))
    qi = rng.integers(0, 10, n).astype(np.int32)
    rj = rng.integers(0, 20, n).astype(np.int32)
    assert al.score_pairs(qi, rj).shape == (n,)
    for at in sorted({n // 8, n - 5}):
        for bad_q, bad_r in ((10, 0), (-1, 0), (0, 20), (0, -3)):
            q2, r2 = qi.copy(), rj.copy()
            q2[at], r2[at] = bad_q, bad_r
            with pytest.raises(IndexError):
                al.score_pairs(q2, r2)
    # the device state is usable after the errors
    assert np.array_equal(al.score_pairs(qi, rj), al.score_pairs(qi, rj))
