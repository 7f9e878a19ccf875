"""Queries wider than a tile run as stripes of G*R columns (stripe kernel): scores stay
bit-exact with the oracle in every mode, in int16 and int32, when striped and unstriped
queries are mixed and paired, when int16 saturates in a late stripe, and when served."""
import ctypes

import numpy as np
import pytest

from gpusw_amd import Aligner, Matrix, Scheme, _hip
from gpusw_amd.backends.hip import scheme_struct_cached
from gpusw_amd.encode import funnel
from gpusw_amd.serve import Server
from oracle import native
from tests.helpers import PROT, rand_seqs, schemes, stable_seed

pytestmark = pytest.mark.gpu

SCHEMES = schemes()
STRIPE_TILE = (8, 32)   # W = 256 columns per stripe


def _oracle(queries, refs, scheme):
    return native.score_encoded_cross(funnel(queries, scheme), funnel(refs, scheme), scheme)


def _score(queries, refs, scheme, tile=None):
    if tile:
        _hip.force_variant(*tile)
    try:
        al = Aligner(scheme, backend="hip").index(refs)
        got = al.score_cross(queries)
        return got, al._b.last_stats
    finally:
        if tile:
            _hip.force_variant(0, 0)


@pytest.mark.parametrize("dtype", ["auto", "int32"])
@pytest.mark.parametrize("name", list(SCHEMES))
def test_forced_stripes_every_scheme(name, dtype):
    scheme, alpha = SCHEMES[name]
    scheme = scheme.replace(dtype=dtype)
    rng = np.random.default_rng(stable_seed("stripes" + name + dtype))
    # stripe edges (256, 512), lengths 1 past them, a ragged mix, and short queries that
    # take the unstriped kernel in the same call; 11 queries leave one unpaired
    qlens = [255, 256, 257, 300, 511, 512, 513, 777, 1030, 40, 129]
    queries = ["".join(rng.choice(list(alpha), L)) for L in qlens]
    refs = rand_seqs(rng, 45, 1, 420, alpha) + ["", queries[3][:150]]
    got, st = _score(queries, refs, scheme, STRIPE_TILE)
    want = _oracle(queries, refs, scheme)
    bad = np.argwhere(got.astype(np.int64) != want)
    assert bad.size == 0, f"{len(bad)} mismatches, first {bad[:5].tolist()}"
    assert st.plain_pairs == 0


@pytest.mark.parametrize("tile", [(8, 26), (8, 28), (8, 30), (8, 34), (8, 36), (8, 40), (8, 44)])
@pytest.mark.parametrize("preset", ["dna", "blosum62"])
def test_forced_half2_local_stripe_widths(tile, preset):
    """The local stripe widths compiled for half2 only (208 to 352 columns): queries at, one
    past and well beyond the stripe edges, paired with short unstriped queries."""
    scheme = Scheme.preset(preset).replace(mode="local")
    alpha = PROT if preset == "blosum62" else "ACGT"
    W = tile[0] * tile[1]
    rng = np.random.default_rng(stable_seed("h2stripes" + preset + str(W)))
    qlens = [W - 1, W, W + 1, 2 * W + 3, 3 * W - 7, 700, 1301, 33, 150]
    queries = ["".join(rng.choice(list(alpha), L)) for L in qlens]
    refs = rand_seqs(rng, 40, 1, 500, alpha) + ["", queries[4][100:600]]
    got, st = _score(queries, refs, scheme, tile)
    want = _oracle(queries, refs, scheme)
    bad = np.argwhere(got.astype(np.int64) != want)
    assert bad.size == 0, f"{len(bad)} mismatches, first {bad[:5].tolist()}"
    # the protein query against its own 500-residue stretch scores above 2048: that one
    # listed result is rescored in int32 as an explicit pair of 713 x 500 residues, wider
    # than the forced tile, whose int32 rows have no stripe kernel: the comparator takes it
    listed = 1 if preset == "blosum62" else 0
    assert st.plain_pairs == listed


@pytest.mark.parametrize("tile", [(16, 32), (32, 32)])
@pytest.mark.parametrize("mode,fq,fr", [("local", False, False), ("global", False, False),
                                        ("semiglobal", True, False), ("semiglobal", True, True)])
def test_forced_int32_wide_stripes(tile, mode, fq, fr):
    """The int32 stripe kernels of the 16- and 32-lane tiles (512 and 1,024 columns):
    queries at, one past and beyond the stripe edges beside unstriped ones."""
    scheme = Scheme(mode=mode, dtype="int32", free_end_gaps_query=fq, free_end_gaps_ref=fr)
    W = tile[0] * tile[1]
    rng = np.random.default_rng(stable_seed(f"wide/{W}/{mode}/{fq}/{fr}"))
    qlens = [W - 1, W, W + 1, 2 * W + 5, 300, 40]
    queries = ["".join(rng.choice(list("ACGT"), L)) for L in qlens]
    refs = rand_seqs(rng, 30, 1, 400, "ACGT") + ["", queries[3][:350]]
    got, st = _score(queries, refs, scheme, tile)
    want = _oracle(queries, refs, scheme)
    bad = np.argwhere(got.astype(np.int64) != want)
    assert bad.size == 0, f"{len(bad)} mismatches, first {bad[:5].tolist()}"
    assert st.plain_pairs == 0


@pytest.mark.parametrize("mode", ["local", "global", "semiglobal"])
def test_protein_beyond_widest_tile(mode):
    """No comparator fallback: 1500-3300 aa protein queries are striped by the cost model."""
    scheme = Scheme.preset("blosum62").replace(mode=mode)
    rng = np.random.default_rng(stable_seed("longprot" + mode))
    queries = rand_seqs(rng, 5, 1500, 3300, PROT) + rand_seqs(rng, 4, 100, 400, PROT)
    refs = rand_seqs(rng, 70, 20, 900, PROT)
    # a reference that shares a long stretch with query 0: a high score across stripes
    refs.append(queries[0][200:1400])
    got, st = _score(queries, refs, scheme)
    want = _oracle(queries, refs, scheme)
    assert np.array_equal(got.astype(np.int64), want)
    assert st.plain_pairs == 0


def test_saturation_in_late_stripe_rescored(monkeypatch):
    """A local score past int16 that builds up over several stripes.  With dtype "auto" the
    envelope sends the pair to the int32 stripes; forced to int16 through the C ABI (half2
    tiles off), the int16 result reaches 32767 only in a late stripe and is rescored on the
    int32 tiles."""
    monkeypatch.setenv("GSW_HALF", "0")
    scheme = Scheme.preset("blosum62")
    rng = np.random.default_rng(21)
    w = "W" * 3100                      # self score 11 per residue: 34100 > 32767
    queries = [w, "".join(rng.choice(list(PROT), 2000)), w[:600]]
    refs = [w, "W" * 200, "".join(rng.choice(list(PROT), 700)), w[:2990] + "A" * 20]
    want = _oracle(queries, refs, scheme)
    assert int(want.max()) > 32767
    got, st = _score(queries, refs, scheme)
    assert np.array_equal(got.astype(np.int64), want)
    assert st.rescored == 0 and st.plain_pairs == 0
    al = Aligner(scheme, backend="hip").index(refs).set_queries(queries)
    qset = _hip.seqset_for(al._d_qbuf, al._d_qoff)
    rset = _hip.seqset_for(al._d_rbuf, al._d_roff)
    s, _table = scheme_struct_cached(scheme)
    out = _hip.DeviceBuffer(4 * want.size, np.int32, want.shape)
    st = _hip.GswStats()
    qlist = np.arange(len(queries), dtype=np.int32)
    _hip.call("gsw_score_cross", ctypes.byref(s), _hip.DTYPE_INT16, qset.handle, _hip.ptr(qlist),
              len(queries), rset.handle, out.ptr, ctypes.byref(st))
    assert np.array_equal(out.to_host().astype(np.int64), want)
    # both saturated pairs are longer than every explicit-pairs tile (2,048 columns) on both
    # sides: they are rescored as explicit pairs on the int32 stripe kernels
    assert st.rescored == int((want >= 32767).sum()) == 2 and st.plain_pairs == 0


def test_served_long_queries_equal_aligner():
    scheme = Scheme.preset("blosum62")
    rng = np.random.default_rng(22)
    refs = rand_seqs(rng, 300, 30, 700, PROT)
    queries = rand_seqs(rng, 6, 1600, 2600, PROT) + rand_seqs(rng, 5, 50, 300, PROT)
    want = Aligner(scheme, backend="hip").index(refs).top_k(queries, k=7)
    with Server(memory_budget=1 << 30, launch_cells=2e8) as srv:
        db = srv.add_database(refs, scheme)
        ids = [srv.submit(db, queries, k=7), srv.submit(db, queries[:3], k=7)]
        res = [srv.result(i) for i in ids]
    assert np.array_equal(res[0].result.topk_idx, want.topk_idx)
    assert np.array_equal(res[0].result.topk_scores, want.topk_scores)
    assert np.array_equal(res[1].result.topk_scores, want.topk_scores[:3])


@pytest.mark.parametrize("mode,fq,fr", [("local", False, False), ("global", False, False),
                                        ("semiglobal", False, False), ("semiglobal", True, False),
                                        ("semiglobal", False, True), ("semiglobal", True, True)])
def test_long_explicit_pairs_on_int32_stripes(mode, fq, fr):
    """Explicit pairs longer than every pairs tile (2,048 columns) on both sides run on the
    int32 stripe kernels as per-query lists of references: one query with many references
    (several items), queries of several stripe counts, a repeated pair, shuffled order and
    shorter pairs in the same call; no comparator, every pair equal to the oracle."""
    m = np.array([[2, -3, -1, -4], [-2, 3, -3, 0], [-1, -2, 4, -3], [-4, 1, -2, 2]])   # asymmetric
    scheme = Scheme(mode=mode, matrix=Matrix.from_array("ACGT", m), gap_open=-4, gap_extend=-1,
                    free_end_gaps_query=fq, free_end_gaps_ref=fr)
    rng = np.random.default_rng(stable_seed(f"longpairs/{mode}/{fq}/{fr}"))
    queries = rand_seqs(rng, 5, 2049, 3600, "ACGT") + rand_seqs(rng, 3, 20, 300, "ACGT")
    refs = rand_seqs(rng, 30, 2049, 2500, "ACGT") + rand_seqs(rng, 10, 1, 400, "ACGT")
    refs[3] = queries[1][300:2700]                           # a high-scoring long pair
    qi = np.concatenate([np.zeros(30, np.int32), rng.integers(1, 5, 40).astype(np.int32),
                         np.array([1, 1], np.int32), rng.integers(0, 8, 60).astype(np.int32)])
    rj = np.concatenate([np.arange(30, dtype=np.int32), rng.integers(0, 30, 40).astype(np.int32),
                         np.array([3, 3], np.int32), rng.integers(0, 40, 60).astype(np.int32)])
    perm = rng.permutation(len(qi))
    qi, rj = qi[perm], rj[perm]
    al = Aligner(scheme, backend="hip").index(refs).set_queries(queries)
    got = al.score_pairs(qi, rj)
    st = al._b.last_stats
    want = native.score_encoded_pairs(al._queries, al._refs, qi, rj, scheme)
    bad = np.flatnonzero(got.astype(np.int64) != want)
    assert bad.size == 0, f"{bad.size} mismatches, first {[(int(qi[i]), int(rj[i])) for i in bad[:5]]}"
    assert st.plain_pairs == 0
