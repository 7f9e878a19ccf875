"""int16 / int32 routing by the score envelope, and saturated int16 results rescored on
the batched int32 tiles (never on the one-thread-per-pair comparator, never an error).

The reference resolves one element type per call from the longest query and reference
(``src/gpusw/aligner.py:90-93`` -> ``scheme.py:116-141``) and launches int32 when a score
could leave +/-30,000 (``backends/cuda.py:68-85``).  ``dtype="auto"`` here applies that
envelope to every pair; ``GSW_DTYPE_INT16`` (a C caller vouching for int16) runs every
local pair packed and rescores the results that reach 32767 -- the listed ones, or, when
more than the list holds (2^20) saturate, every pair that ran packed.  The int16 tests here
turn the half2 tiles off (``GSW_HALF=0``), which would otherwise take these pairs first;
``test_gpu_half.py`` covers that route.
"""
import ctypes

import numpy as np
import pytest

from gpusw_amd import Aligner, Matrix, Scheme, _hip
from gpusw_amd.backends.hip import scheme_struct_cached
from oracle import native
from tests.helpers import rand_seqs, stable_seed

pytestmark = pytest.mark.gpu

# match 300 on 150-nt reads: an identical pair scores 45,000, past int16 (and past the
# reference's 30,000 envelope, which sends such a call to int32)
HEAVY = Scheme(mode="local", match=300, mismatch=-3, gap_open=-5, gap_extend=-2)
N_OVER = 1_150_000   # > 2^20 = 1,048,576 saturating pairs


def _call_pairs(al, qi, rj, policy):
    """gsw_score_pairs through the C ABI with an explicit dtype policy."""
    qset = _hip.seqset_for(al._d_qbuf, al._d_qoff)
    rset = _hip.seqset_for(al._d_rbuf, al._d_roff)
    s, _table = scheme_struct_cached(al.scheme)
    n = int(qi.size)
    out = _hip.DeviceBuffer(4 * n, np.int32, (n,))
    stats = _hip.GswStats()
    _hip.call("gsw_score_pairs", ctypes.byref(s), policy, qset.handle, rset.handle, _hip.ptr(qi),
              _hip.ptr(rj), n, out.ptr, ctypes.byref(stats))
    return out.to_host(), stats


def _call_cross(al, policy):
    qset = _hip.seqset_for(al._d_qbuf, al._d_qoff)
    rset = _hip.seqset_for(al._d_rbuf, al._d_roff)
    s, _table = scheme_struct_cached(al.scheme)
    nq, nr = al._queries.n, al._refs.n
    qlist = np.arange(nq, dtype=np.int32)
    out = _hip.DeviceBuffer(4 * nq * nr, np.int32, (nq, nr))
    stats = _hip.GswStats()
    _hip.call("gsw_score_cross", ctypes.byref(s), policy, qset.handle, _hip.ptr(qlist), nq,
              rset.handle, out.ptr, ctypes.byref(stats))
    return out.to_host(), stats


@pytest.fixture(scope="module")
def heavy_pairs():
    rng = np.random.default_rng(4242)
    seqs = rand_seqs(rng, 64, 150, 151, "ACGT")
    others = rand_seqs(rng, 32, 120, 181, "ACGT")
    al = Aligner(HEAVY, backend="hip").index(seqs + others).set_queries(seqs)
    k = np.arange(N_OVER, dtype=np.int64)
    qi = (k % 64).astype(np.int32)
    rj = qi.copy()
    mixed = rng.random(N_OVER) < 0.05          # 5 % against any reference
    rj[mixed] = rng.integers(0, 96, int(mixed.sum())).astype(np.int32)
    return al, qi, rj, ~mixed


def _check_heavy(al, qi, rj, same, got):
    assert got.shape == qi.shape
    assert np.all(got[same] == 300 * 150)                 # closed form: match x length
    rng = np.random.default_rng(7)
    pick = np.concatenate([rng.choice(np.flatnonzero(~same), 400, replace=False),
                           rng.choice(qi.size, 400, replace=False)])
    want = native.score_encoded_pairs(al._queries, al._refs, qi[pick], rj[pick], al.scheme)
    assert np.array_equal(got[pick].astype(np.int64), want)


def test_heavy_pairs_auto_route_to_int32(heavy_pairs):
    """>2^20 local pairs whose scores exceed int16: the envelope sends them to the int32
    tiles; nothing saturates, nothing is rescored, nothing runs on the comparator."""
    al, qi, rj, same = heavy_pairs
    assert HEAVY.resolve_dtype(150, 180) == "int32"
    got = al.score_pairs(qi, rj)
    st = al._b.last_stats
    assert st.plain_pairs == 0 and st.rescored == 0
    _check_heavy(al, qi, rj, same, got)


def test_heavy_pairs_int16_overflow_rescored(heavy_pairs, monkeypatch):
    """The same pairs forced to int16 through the C ABI (half2 tiles off): every identical
    pair saturates, more than the saturation list holds, so every packed pair is rescored
    in int32."""
    monkeypatch.setenv("GSW_HALF", "0")
    al, qi, rj, same = heavy_pairs
    got, st = _call_pairs(al, qi, rj, _hip.DTYPE_INT16)
    assert st.plain_pairs == 0
    assert st.rescored == qi.size                          # every pair ran packed
    _check_heavy(al, qi, rj, same, got)


def test_pairs_int16_listed_saturation_rescored(monkeypatch):
    """A few saturating pairs among many (int16 tiles, half2 off): only the listed ones are
    rescored."""
    monkeypatch.setenv("GSW_HALF", "0")
    rng = np.random.default_rng(99)
    seqs = rand_seqs(rng, 40, 100, 160, "ACGT")
    al = Aligner(HEAVY, backend="hip").index(seqs).set_queries(seqs)
    qi = rng.integers(0, 40, 6000).astype(np.int32)
    rj = rng.integers(0, 40, 6000).astype(np.int32)
    rj[:300] = qi[:300]                                   # identical: 300 x len > 32767
    got, st = _call_pairs(al, qi, rj, _hip.DTYPE_INT16)
    want = native.score_encoded_pairs(al._queries, al._refs, qi, rj, HEAVY)
    assert np.array_equal(got.astype(np.int64), want)
    assert st.plain_pairs == 0
    assert st.rescored == int((want >= 32767).sum()) > 0


def test_cross_envelope_int32_no_comparator():
    """A cross product whose envelope says int32 runs on the int32 tiles."""
    rng = np.random.default_rng(5)
    queries = rand_seqs(rng, 40, 140, 160, "ACGT")
    refs = queries[:10] + rand_seqs(rng, 290, 90, 400, "ACGT")
    assert HEAVY.resolve_dtype(160, 400) == "int32"
    al = Aligner(HEAVY, backend="hip").index(refs).set_queries(queries)
    got = al.score_cross()
    st = al._b.last_stats
    assert st.plain_pairs == 0 and st.rescored == 0
    want = native.score_encoded_cross(al._queries, al._refs, HEAVY)
    assert int(want.max()) > 32767
    assert np.array_equal(got.astype(np.int64), want)


def test_cross_int16_listed_and_overflow(monkeypatch):
    monkeypatch.setenv("GSW_HALF", "0")
    rng = np.random.default_rng(6)
    # listed: a few saturating cells
    queries = rand_seqs(rng, 30, 120, 160, "ACGT")
    refs = queries[:4] + rand_seqs(rng, 200, 60, 300, "ACGT")
    al = Aligner(HEAVY, backend="hip").index(refs).set_queries(queries)
    got, st = _call_cross(al, _hip.DTYPE_INT16)
    want = native.score_encoded_cross(al._queries, al._refs, HEAVY)
    assert np.array_equal(got.astype(np.int64), want)
    assert st.plain_pairs == 0 and st.rescored == int((want >= 32767).sum()) > 0
    # overflow: 1,100 x 1,000 identical 150-nt pairs all saturate (> 2^20)
    s0 = rand_seqs(rng, 1, 150, 151, "ACGT")[0]
    al = Aligner(HEAVY, backend="hip").index([s0] * 1000).set_queries([s0] * 1100)
    got, st = _call_cross(al, _hip.DTYPE_INT16)
    assert np.all(got == 300 * 150)
    assert st.plain_pairs == 0 and st.rescored == 1100 * 1000


@pytest.mark.parametrize("fq,fr", [(True, True), (True, False), (False, True)])
@pytest.mark.parametrize("mode", ["local", "global", "semiglobal"])
def test_long_query_pairs_run_transposed(mode, fq, fr):
    """Explicit pairs whose query is wider than every pairs tile (2,048 columns) but whose
    reference fits one run with the roles exchanged under the transposed scheme (free-end
    flags swapped): batched, no comparator, equal to the oracle."""
    if mode != "semiglobal" and (fq, fr) != (True, True):
        pytest.skip("free-end flags only matter in semiglobal mode")
    m = np.array([[2, -3, -1, -4], [-2, 3, -3, 0], [-1, -2, 4, -3], [-4, 1, -2, 2]])   # asymmetric
    scheme = Scheme(mode=mode, matrix=Matrix.from_array("ACGT", m), gap_open=-4, gap_extend=-1,
                    free_end_gaps_query=fq, free_end_gaps_ref=fr)
    rng = np.random.default_rng(stable_seed(f"transposed/{mode}/{fq}/{fr}"))
    queries = rand_seqs(rng, 6, 2049, 3200, "ACGT") + rand_seqs(rng, 4, 10, 300, "ACGT")
    refs = rand_seqs(rng, 12, 1, 600, "ACGT") + [queries[0][100:700]] + rand_seqs(rng, 2, 2100, 2300, "ACGT")
    al = Aligner(scheme, backend="hip").index(refs).set_queries(queries)
    qi = rng.integers(0, len(queries), 400).astype(np.int32)
    rj = rng.integers(0, len(refs) - 2, 400).astype(np.int32)
    qi[:3] = 0
    rj[:3] = [12, len(refs) - 2, len(refs) - 1]            # two pairs long on both sides
    got = al.score_pairs(qi, rj)
    st = al._b.last_stats
    want = native.score_encoded_pairs(al._queries, al._refs, qi, rj, scheme)
    assert np.array_equal(got.astype(np.int64), want)
    assert st.plain_pairs == 0   # the two long pairs run on the int32 stripe kernels


def test_long_query_saturation_rescored_transposed(monkeypatch):
    """A saturated long-query pair (forced int16, half2 off) is rescored on the int32 tiles
    through the transposed route."""
    monkeypatch.setenv("GSW_HALF", "0")
    rng = np.random.default_rng(77)
    ref = rand_seqs(rng, 1, 150, 151, "ACGT")[0]
    queries = ["".join(rng.choice(list("ACGT"), 1500)) + ref + "".join(rng.choice(list("ACGT"), 1000))]
    refs = [ref] + rand_seqs(rng, 5, 100, 400, "ACGT")
    al = Aligner(HEAVY, backend="hip").index(refs).set_queries(queries)
    qi = np.zeros(6, np.int32)
    rj = np.arange(6, dtype=np.int32)
    got, st = _call_pairs(al, qi, rj, _hip.DTYPE_INT16)
    want = native.score_encoded_pairs(al._queries, al._refs, qi, rj, HEAVY)
    assert want[0] == 300 * 150
    assert np.array_equal(got.astype(np.int64), want)
    assert st.plain_pairs == 0 and st.rescored >= 1
