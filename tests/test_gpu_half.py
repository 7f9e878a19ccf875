"""half2 tiles: two alignments per lane carried as binary16 values (exact for integers in
[-2048, 2048]) on gfx950's three-input packed max.

Local pairs run in half2 wherever int16 would; a result of 2048 or more may have been
rounded and is listed and rescored in int32 (the listed ones), or, when more are listed
than the list holds (2^20), every pair that ran in half2 runs again on the int16 tiles.
Global and semiglobal pairs run in half2 only where the envelope (``scheme.py:116-141``
with the bound 2048) keeps every DP value exact.  Every result is compared with the oracle;
the cases sit on both sides of 2048, including odd scores past it that binary16 cannot hold.
"""
import numpy as np
import pytest

from gpusw_amd import Aligner, Matrix, Scheme, _hip
from oracle import native
from tests.helpers import call_cross, call_pairs, rand_seqs, stable_seed

pytestmark = pytest.mark.gpu

_EYE = np.eye(4, dtype=int)
# match 3: self scores 3 * length straddle 2048 at 682 / 683 nt (2046 / 2049)
M3 = Scheme(mode="local", matrix=Matrix.from_array("ACGT", 3 * _EYE - 2 * (1 - _EYE)), gap_open=-5,
            gap_extend=-2)
LENS = (681, 682, 683, 684, 700)


def _kind(stats) -> str:
    return _hip.variant_info(stats.variant)[2]


def _around_2048(seed):
    rng = np.random.default_rng(seed)
    base = "".join(rng.choice(list("ACGT"), 1700))
    queries = [base[:n] for n in LENS] + rand_seqs(rng, 6, 100, 750, "ACGT") + [base[:1500]]
    refs = [base[:n] for n in LENS] + [base[100:1700]] + rand_seqs(rng, 40, 50, 900, "ACGT")
    return queries, refs


def test_local_cross_around_2048_exact():
    queries, refs = _around_2048(stable_seed("half/cross"))
    al = Aligner(M3, backend="hip").index(refs).set_queries(queries)
    got = al.score_cross()
    st = al._b.last_stats
    want = native.score_encoded_cross(al._queries, al._refs, M3)
    assert np.array_equal(got.astype(np.int64), want)
    assert _kind(st) == "half2" and st.plain_pairs == 0
    assert ((want >= 2040) & (want < 2048)).any() and (want % 2 == 1)[want > 2048].any()
    assert st.rescored == int((want >= 2048).sum()) > 0


def test_local_pairs_around_2048_exact():
    queries, refs = _around_2048(stable_seed("half/pairs"))
    al = Aligner(M3, backend="hip").index(refs).set_queries(queries)
    rng = np.random.default_rng(3)
    qi = rng.integers(0, len(queries), 3000).astype(np.int32)
    rj = rng.integers(0, len(refs), 3000).astype(np.int32)
    qi[:5] = rj[:5] = np.arange(5)                       # self pairs 2043 ... 2100
    got = al.score_pairs(qi, rj)
    st = al._b.last_stats
    want = native.score_encoded_pairs(al._queries, al._refs, qi, rj, M3)
    assert list(want[:5]) == [3 * n for n in LENS]
    assert np.array_equal(got.astype(np.int64), want)
    assert _kind(st) == "half2"
    assert st.rescored == int((want >= 2048).sum()) > 0


# match 20 on 150-nt reads: an identical pair scores 3,000, past half2, within int16
MID = Scheme(mode="local", match=20, mismatch=-3, gap_open=-5, gap_extend=-2)


def test_half_overflow_reruns_on_int16():
    """More than 2^20 half2 results at the top: every pair that ran in half2 runs again on
    the int16 tiles, where none saturates."""
    rng = np.random.default_rng(stable_seed("half/overflow"))
    seqs = rand_seqs(rng, 64, 150, 151, "ACGT")
    others = rand_seqs(rng, 32, 120, 181, "ACGT")
    al = Aligner(MID, backend="hip").index(seqs + others).set_queries(seqs)
    n = 1_150_000
    qi = (np.arange(n) % 64).astype(np.int32)
    rj = qi.copy()
    mixed = rng.random(n) < 0.05
    rj[mixed] = rng.integers(0, 96, int(mixed.sum())).astype(np.int32)
    got = al.score_pairs(qi, rj)
    st = al._b.last_stats
    assert np.all(got[~mixed] == 20 * 150)
    pick = np.concatenate([rng.choice(np.flatnonzero(mixed), 500, replace=False), rng.choice(n, 500, replace=False)])
    want = native.score_encoded_pairs(al._queries, al._refs, qi[pick], rj[pick], MID)
    assert np.array_equal(got[pick].astype(np.int64), want)
    assert st.plain_pairs == 0 and st.rescored == n


def test_half_cross_overflow_reruns_on_int16():
    rng = np.random.default_rng(stable_seed("half/cross-overflow"))
    s0 = rand_seqs(rng, 1, 150, 151, "ACGT")[0]
    al = Aligner(MID, backend="hip").index([s0] * 1000).set_queries([s0] * 1100)
    got = al.score_cross()
    st = al._b.last_stats
    assert np.all(got == 20 * 150)
    assert st.plain_pairs == 0 and st.rescored == 1100 * 1000


@pytest.mark.parametrize("mode,fq,fr", [("global", True, True), ("semiglobal", True, True),
                                        ("semiglobal", True, False), ("semiglobal", False, True)])
def test_nonlocal_half_by_envelope(mode, fq, fr):
    """Short pairs (every value within +/-2048) run in half2, long ones on int16."""
    scheme = Scheme(mode=mode, free_end_gaps_query=fq, free_end_gaps_ref=fr)
    rng = np.random.default_rng(stable_seed(f"half/{mode}/{fq}/{fr}"))
    for (qlo, qhi, rlo, rhi), kind in (((100, 151, 200, 351), "half2"), ((500, 701, 600, 901), "int16x2")):
        queries = rand_seqs(rng, 24, qlo, qhi, "ACGT")
        refs = rand_seqs(rng, 24, rlo, rhi, "ACGT")
        refs[0] = queries[0][10:] + refs[0][:50]
        al = Aligner(scheme, backend="hip").index(refs).set_queries(queries)
        qi = rng.integers(0, 24, 600).astype(np.int32)
        rj = rng.integers(0, 24, 600).astype(np.int32)
        qi[0], rj[0] = 0, 0
        got = al.score_pairs(qi, rj)
        st = al._b.last_stats
        want = native.score_encoded_pairs(al._queries, al._refs, qi, rj, scheme)
        assert np.array_equal(got.astype(np.int64), want)
        assert _kind(st) == kind and st.rescored == 0
        gotx = al.score_cross()
        assert _kind(al._b.last_stats) == kind
        assert np.array_equal(gotx.astype(np.int64), native.score_encoded_cross(al._queries, al._refs, scheme))


def test_half_off_by_environment(monkeypatch):
    monkeypatch.setenv("GSW_HALF", "0")
    rng = np.random.default_rng(11)
    queries = rand_seqs(rng, 10, 200, 300, "ACGT")
    refs = rand_seqs(rng, 50, 300, 900, "ACGT")
    al = Aligner(Scheme.preset("dna"), backend="hip").index(refs).set_queries(queries)
    got = al.score_cross()
    assert _kind(al._b.last_stats) == "int16x2"
    monkeypatch.delenv("GSW_HALF")
    assert np.array_equal(al.score_cross(), got)
    assert _kind(al._b.last_stats) == "half2"
    assert np.array_equal(got.astype(np.int64), native.score_encoded_cross(al._queries, al._refs, al.scheme))


def test_half_forced_int16_policy_lists_from_2048():
    """A C caller vouching for int16 (GSW_DTYPE_INT16) gets half2 too: every local result
    from 2048 up is listed and rescored, exact either way."""
    queries, refs = _around_2048(stable_seed("half/int16-policy"))
    al = Aligner(M3, backend="hip").index(refs).set_queries(queries)
    got, st = call_cross(al, _hip.DTYPE_INT16)
    want = native.score_encoded_cross(al._queries, al._refs, M3)
    assert np.array_equal(got.astype(np.int64), want)
    assert st.rescored == int((want >= 2048).sum()) > 0
    qi = np.arange(len(queries), dtype=np.int32) % len(refs)
    gp, sp = call_pairs(al, qi, qi.copy(), _hip.DTYPE_INT16)
    assert np.array_equal(gp.astype(np.int64), native.score_encoded_pairs(al._queries, al._refs, qi, qi, M3))
