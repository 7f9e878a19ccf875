"""The measured autotuner installs tile rules; scores stay bit-exact under them and
under every forced tile."""
import numpy as np
import pytest

from gpusw_amd import Aligner, Scheme, _hip, autotune
from gpusw_amd.encode import funnel
from oracle import native
from tests.helpers import rand_seqs

pytestmark = pytest.mark.gpu


def test_tune_installs_rules_and_stays_exact(tmp_path):
    rng = np.random.default_rng(12)
    refs = rand_seqs(rng, 300, 100, 700, "ACGT")
    scheme = Scheme.preset("dna")
    try:
        rules = autotune.tune(refs, scheme, bins=[(20, 60), (240, 260)], n_queries=16, n_refs=200,
                              reps=1, cache=tmp_path / "tiles.json")
        stripes = set(_hip.list_stripe_variants())
        assert rules and all(r["G"] * r["R"] >= r["hi"] or (r["G"], r["R"]) in stripes for r in rules)
        assert (tmp_path / "tiles.json").exists()
        queries = rand_seqs(rng, 40, 20, 61, "ACGT") + rand_seqs(rng, 20, 240, 261, "ACGT")
        got = Aligner(scheme, backend="hip").index(refs).score_cross(queries)
        want = native.score_encoded_cross(funnel(queries, scheme), funnel(refs, scheme), scheme)
        assert np.array_equal(got.astype(np.int64), want)
    finally:
        autotune.clear()


@pytest.mark.parametrize("mode", ["local", "global", "semiglobal"])
def test_every_forced_tile_exact(mode):
    rng = np.random.default_rng(13)
    scheme = Scheme(mode=mode)
    refs = rand_seqs(rng, 40, 1, 300, "ACGT")
    queries = rand_seqs(rng, 12, 1, 64, "ACGT")
    want = native.score_encoded_cross(funnel(queries, scheme), funnel(refs, scheme), scheme)
    for G, R in _hip.list_variants():
        _hip.force_variant(G, R)
        try:
            got = Aligner(scheme, backend="hip").index(refs).score_cross(queries)
        finally:
            _hip.force_variant(0, 0)
        assert np.array_equal(got.astype(np.int64), want), (G, R)


@pytest.mark.parametrize("mode", ["local", "global", "semiglobal"])
@pytest.mark.parametrize("tab", ["1", "0"])
def test_every_forced_tile_exact_pairs(mode, tab, monkeypatch):
    """Explicit pairs on every compiled tile, with the table kernel (GSW_PAIRS_TAB=1) and
    with per-lane profiles (GSW_PAIRS_TAB=0, where the profiles fit LDS)."""
    monkeypatch.setenv("GSW_PAIRS_TAB", tab)
    rng = np.random.default_rng(29)
    scheme = Scheme(mode=mode)
    refs = rand_seqs(rng, 60, 1, 300, "ACGT")
    queries = rand_seqs(rng, 50, 1, 64, "ACGT")
    qi = rng.integers(0, len(queries), 3000).astype(np.int32)
    rj = rng.integers(0, len(refs), 3000).astype(np.int32)
    want = native.score_encoded_pairs(funnel(queries, scheme), funnel(refs, scheme), qi, rj, scheme)
    for G, R in _hip.list_variants():
        _hip.force_variant(G, R)
        try:
            al = Aligner(scheme, backend="hip").index(refs).set_queries(queries)
            got = al.score_pairs(qi, rj)
        finally:
            _hip.force_variant(0, 0)
        assert np.array_equal(got.astype(np.int64), want), (G, R)
