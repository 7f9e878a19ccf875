"""gsw_score_pairs_noalloc (explicit pairs with every buffer in the library's file-scope
device array) against gsw_score_pairs on the same inputs, through the C ABI: two calls in
a row, the second smaller than the first (the array keeps the first call's contents), then
a forced-int16 call whose saturated results are rescored through the same array."""
import ctypes

import numpy as np
import pytest

from gpusw_amd import Aligner, Scheme, _hip
from gpusw_amd.backends.hip import scheme_struct_cached
from oracle import native
from tests.helpers import rand_seqs

MAX_PAIRS = 1 << 18


def _call(name, al, qi, rj, policy=_hip.DTYPE_AUTO):
    qset = _hip.seqset_for(al._d_qbuf, al._d_qoff)
    rset = _hip.seqset_for(al._d_rbuf, al._d_roff)
    s, _table = scheme_struct_cached(al.scheme)
    n = int(qi.size)
    out = _hip.DeviceBuffer(4 * max(n, 1), np.int32, (max(n, 1),))
    _hip.call("gsw_memset", out.ptr, 0x7F, 4 * max(n, 1))     # stale values must be overwritten
    stats = _hip.GswStats()
    _hip.call(name, ctypes.byref(s), policy, qset.handle, rset.handle, _hip.ptr(qi), _hip.ptr(rj), n,
              out.ptr, ctypes.byref(stats))
    return out.to_host()[:n], stats


def test_noalloc_limit_checked_before_device_work():
    """Beyond the pair limit the call fails with GSW_ERR_LIMIT before touching a device."""
    lib = _hip.load()
    qi = np.zeros(1, np.int32)
    rc = lib.gsw_score_pairs_noalloc(None, 0, None, None, _hip.ptr(qi), _hip.ptr(qi), MAX_PAIRS + 1, None, None)
    assert rc == _hip.GSW_ERR_LIMIT
    assert "at most 262144 pairs" in _hip.last_error()


@pytest.mark.gpu
@pytest.mark.parametrize("mode", ["semiglobal", "local"])
def test_noalloc_equals_batched_two_calls(mode):
    rng = np.random.default_rng(31 if mode == "local" else 32)
    scheme = Scheme.preset("dna").replace(mode=mode)
    reads = rand_seqs(rng, 3000, 0, 400, "ACGTN") + [""]
    windows = rand_seqs(rng, 900, 0, 700, "ACGT") + [""]
    al = Aligner(scheme, backend="hip").index(windows).set_queries(reads)
    for n in (200_000, 3_001):                      # second call smaller than the first
        qi = rng.integers(0, len(reads), n).astype(np.int32)
        rj = rng.integers(0, len(windows), n).astype(np.int32)
        got, st = _call("gsw_score_pairs_noalloc", al, qi, rj)
        want, _ = _call("gsw_score_pairs", al, qi, rj)
        assert np.array_equal(got, want), f"n={n}: {int((got != want).sum())} differ"
        assert st.pairs == n and st.plain_pairs == 0
        pick = rng.choice(n, 1500, replace=False)
        ora = native.score_encoded_pairs(al._queries, al._refs, qi[pick], rj[pick], scheme)
        assert np.array_equal(got[pick].astype(np.int64), ora)


@pytest.mark.gpu
def test_noalloc_int16_saturation_rescored(monkeypatch):
    monkeypatch.setenv("GSW_HALF", "0")   # both entry points on the int16 tiles
    heavy = Scheme(mode="local", match=300, mismatch=-3, gap_open=-5, gap_extend=-2)
    rng = np.random.default_rng(33)
    seqs = rand_seqs(rng, 50, 100, 170, "ACGT")
    al = Aligner(heavy, backend="hip").index(seqs).set_queries(seqs)
    qi = rng.integers(0, 50, 5000).astype(np.int32)
    rj = rng.integers(0, 50, 5000).astype(np.int32)
    rj[:400] = qi[:400]
    got, st = _call("gsw_score_pairs_noalloc", al, qi, rj, _hip.DTYPE_INT16)
    want, _ = _call("gsw_score_pairs", al, qi, rj, _hip.DTYPE_INT16)
    ora = native.score_encoded_pairs(al._queries, al._refs, qi, rj, heavy)
    assert np.array_equal(got, want) and np.array_equal(got.astype(np.int64), ora)
    assert st.rescored == int((ora >= 32767).sum()) > 0
