"""Global/semiglobal pairs on the packed tiles, which carry values in the diagonal-offset
form X + (i + j)*D (D = -gap_extend): results exact on both sides of the half2 envelope.

For the ``dna`` global scheme (match 2, gap -5/-2) the envelope admits a 150-nt query to
half2 against references of up to 648 nt (every offset value of its real cells at most
2*150 + (150 + 648 + 64)*2 + 6 + 16 = 2046) and sends 649 nt and longer to int16.  Identical
prefixes make the largest offset values of a pair reach that bound.  Every score is compared
with the oracle (``src/gpusw/reference.py:38-125``).
"""
import numpy as np
import pytest

from gpusw_amd import Aligner, Scheme, _hip
from oracle import native
from tests.helpers import rand_seqs, stable_seed

pytestmark = pytest.mark.gpu


def _kind(stats) -> str:
    return _hip.variant_info(stats.variant)[2]


@pytest.mark.parametrize("mode,fq,fr", [("global", True, True), ("semiglobal", True, True),
                                        ("semiglobal", True, False), ("semiglobal", False, True),
                                        ("semiglobal", False, False)])
@pytest.mark.parametrize("lr_lo,lr_hi,kind", [(630, 649, "half2"), (649, 670, "int16x2")])
def test_offset_envelope_boundary(mode, fq, fr, lr_lo, lr_hi, kind):
    scheme = Scheme(mode=mode, free_end_gaps_query=fq, free_end_gaps_ref=fr)
    rng = np.random.default_rng(stable_seed(f"offset/{mode}/{fq}/{fr}/{lr_lo}"))
    base = "".join(rng.choice(list("ACGT"), 800))
    queries = [base[:150], base[5:155]] + rand_seqs(rng, 14, 150, 151, "ACGT")
    refs = [base[:n] for n in range(lr_lo, lr_hi)] + rand_seqs(rng, 12, lr_lo, lr_hi, "ACGT")
    al = Aligner(scheme, backend="hip").index(refs).set_queries(queries)
    nq, nr = len(queries), len(refs)
    qi = np.repeat(np.arange(nq, dtype=np.int32), nr)
    rj = np.tile(np.arange(nr, dtype=np.int32), nq)
    got = al.score_pairs(qi, rj)
    st = al._b.last_stats
    want = native.score_encoded_pairs(al._queries, al._refs, qi, rj, scheme)
    assert np.array_equal(got.astype(np.int64), want)
    assert _kind(st) == kind and st.plain_pairs == 0
    gotx = al.score_cross()
    assert _kind(al._b.last_stats) == kind
    assert np.array_equal(gotx.astype(np.int64), native.score_encoded_cross(al._queries, al._refs, scheme))


@pytest.mark.parametrize("ge", [0, -1, -7, -20])
def test_offset_form_gap_extend_range(ge):
    """D from 0 to 20: short pairs stay in half2 or int16, every score exact."""
    scheme = Scheme(mode="global", gap_open=-9, gap_extend=ge)
    rng = np.random.default_rng(stable_seed(f"offset/ge/{ge}"))
    queries = rand_seqs(rng, 20, 1, 90, "ACGT")
    refs = rand_seqs(rng, 30, 1, 120, "ACGT")
    refs[0] = queries[0] + refs[0]
    al = Aligner(scheme, backend="hip").index(refs).set_queries(queries)
    got = al.score_cross()
    assert al._b.last_stats.plain_pairs == 0
    assert np.array_equal(got.astype(np.int64), native.score_encoded_cross(al._queries, al._refs, scheme))


@pytest.mark.parametrize("mode,fq,fr", [("local", True, True), ("global", True, True),
                                        ("semiglobal", True, False), ("semiglobal", False, True)])
@pytest.mark.parametrize("match", [2, 9])
def test_pairs_byte_select_and_lds_tables_agree(monkeypatch, mode, fq, fr, match):
    """half2 read pairs on byte-select tables (every binary16 entry with a zero low byte:
    match 2) and on LDS tables (GSW_PAIRS_PERM=0, or match 9, whose entry 9 + 4 = 13 needs
    more than three significant bits): the same scores, equal to the oracle."""
    scheme = Scheme(mode=mode, match=match, free_end_gaps_query=fq, free_end_gaps_ref=fr)
    rng = np.random.default_rng(stable_seed(f"perm/{mode}/{fq}/{fr}/{match}"))
    reads = rand_seqs(rng, 300, 140, 161, "ACGTN")
    windows = rand_seqs(rng, 60, 180, 330, "ACGTN")
    al = Aligner(scheme, backend="hip").index(windows).set_queries(reads)
    qi = rng.integers(0, 300, 9000).astype(np.int32)
    rj = rng.integers(0, 60, 9000).astype(np.int32)
    got = al.score_pairs(qi, rj)
    if match == 2 or mode == "local":   # global match 9 leaves the half2 envelope: int16
        assert _kind(al._b.last_stats) == "half2"
    monkeypatch.setenv("GSW_PAIRS_PERM", "0")
    lds = al.score_pairs(qi, rj)
    assert np.array_equal(got, lds)
    want = native.score_encoded_pairs(al._queries, al._refs, qi, rj, scheme)
    assert np.array_equal(got.astype(np.int64), want)
