"""Scheme validation and dtype envelope, matrices, and the input funnel (CPU).

Behaviour pinned by the reference's tests/test_scheme_dtype.py, tests/test_matrix.py
and tests/test_encode.py; the checks below are this package's own restatement.
"""
import gzip

import numpy as np
import pytest

from gpusw_amd import Matrix, Scheme, available_matrices, schemes
from gpusw_amd.encode import build_lut, funnel, read_fasta
from gpusw_amd.errors import EncodeError, OverflowRiskError, SchemeError

DNA = Scheme(mode="local")


# ---------------------------------------------------------------- scheme
def test_scheme_defaults():
    s = Scheme()
    assert (s.mode, s.match, s.mismatch, s.gap_open, s.gap_extend) == ("local", 2, -3, -5, -2)
    assert s.eff_alphabet == "ACGT" and s.ns == 5


@pytest.mark.parametrize("kw", [dict(mode="banana"), dict(gap_open=5), dict(gap_extend=1),
                                dict(unknown="sometimes"), dict(dtype="int8"),
                                dict(alphabet="AACG"), dict(match=1.5)])
def test_scheme_rejects(kw):
    with pytest.raises(SchemeError):
        Scheme(**kw)


def test_substitution_table():
    t = Scheme(match=2, mismatch=-3).substitution_table()
    assert t.shape == (5, 5) and t.dtype == np.int32
    assert np.diag(t)[:4].tolist() == [2, 2, 2, 2]
    assert (t[4] == -3).all() and (t[:, 4] == -3).all()
    b = Scheme(mode="local", matrix=Matrix.named("BLOSUM62"), gap_open=-11, gap_extend=-1)
    assert b.substitution_table().shape == (25, 25)


@pytest.mark.parametrize("args,want", [((500, 1500), "int16"), ((20000, 20000), "int32"),
                                       ((15000, 15000), "int16"), ((15001, 15001), "int32")])
def test_dtype_envelope_local(args, want):
    assert Scheme(match=2, mismatch=-3).resolve_dtype(*args) == want


def test_dtype_forced():
    with pytest.raises(OverflowRiskError):
        Scheme(dtype="int16").resolve_dtype(20000, 20000)
    assert Scheme(dtype="int32").resolve_dtype(10, 10) == "int32"
    assert Scheme(mode="global").resolve_dtype(20000, 20000) == "int32"
    assert Scheme(mode="global").resolve_dtype(5000, 6000) == "int16"


def test_presets_and_hash():
    assert Scheme.preset("dna") is schemes.DNA
    assert Scheme.preset("BLOSUM62").gap_open == -11
    with pytest.raises(SchemeError):
        Scheme.preset("nope")
    a = Scheme(unknown="mismatch").table_hash()
    assert a != Scheme(unknown="zero").table_hash()
    assert schemes.DNA_SEMIGLOBAL.module_fields() == {
        "mode": "semiglobal", "gap_open": -5, "gap_extend": -2, "free_q": True, "free_r": True, "ns": 5}


# ---------------------------------------------------------------- matrix
def test_matrix_builders():
    m = Matrix.from_match_mismatch(2, -3, "ACGT")
    assert np.diag(m.array).tolist() == [2] * 4
    assert (m.array[~np.eye(4, dtype=bool)] == -3).all()
    with pytest.raises(SchemeError):
        Matrix("ACGT", np.zeros((3, 3), dtype=int))


def test_named_tables():
    m = Matrix.named("BLOSUM62")
    assert np.array_equal(m.array, m.array.T)
    assert m.array[m.alphabet.index("W"), m.alphabet.index("W")] == 11
    assert m.alphabet == "ARNDCQEGHILKMFPSTWYVBZX*"
    assert Matrix.named("blosum-62").alphabet == m.alphabet
    e = Matrix.named("ednafull")
    assert e.alphabet == "ATGCSWRYKMBVHDN" and e.array[0, 0] == 5 and e.array[0, 1] == -4
    assert {"BLOSUM62", "BLOSUM80", "PAM250", "EDNAFULL"} <= set(available_matrices())
    with pytest.raises(SchemeError):
        Matrix.named("NOPE")


def test_ncbi_roundtrip_and_packing():
    text = "# tiny\n  A  C  G  T\nA  1 -1 -1 -1\nC -1  1 -1 -1\nG -1 -1  1 -1\nT -1 -1 -1  1\n"
    m = Matrix.from_ncbi(text)
    assert m.alphabet == "ACGT" and np.diag(m.array).tolist() == [1, 1, 1, 1]
    _, pk = Matrix.from_match_mismatch(2, -3).packed("mismatch")
    assert (pk[4] == -3).all() and (pk[:, 4] == -3).all()
    _, pz = Matrix.from_match_mismatch(2, -3).packed("zero")
    assert (pz[4] == 0).all() and (pz[:, 4] == 0).all()
    assert Matrix.named("BLOSUM62").packed("mismatch")[1][-1, 0] == -4
    assert Matrix.from_match_mismatch(2, -3).hash() != Matrix.from_match_mismatch(2, -4).hash()


# ---------------------------------------------------------------- encode
def _codes(enc, i):
    return enc.codes[enc.offsets[i]:enc.offsets[i + 1]].tolist()


def test_funnel_forms_agree():
    a = funnel(["ACGT", "TTGGCC", "A"], DNA)
    b = funnel({"q0": "ACGT", "q1": "TTGGCC", "q2": "A"}, DNA)
    c = funnel([("q0", "ACGT"), ("q1", "TTGGCC"), ("q2", "A")], DNA)
    assert a.ids == b.ids == c.ids == ["q0", "q1", "q2"]
    for i in range(3):
        assert _codes(a, i) == _codes(b, i) == _codes(c, i)
    assert a.offsets.tolist() == [0, 4, 10, 11] and a.lengths.tolist() == [4, 6, 1]


def test_lut_and_case():
    lut, sent = build_lut(DNA)
    assert sent == 4 and [lut[ord(ch)] for ch in "ACGT"] == [0, 1, 2, 3]
    assert lut[ord("N")] == 4 and lut[ord("a")] == 0
    e = funnel(["acgt", "AcGt"], DNA)
    assert _codes(e, 0) == _codes(e, 1) == [0, 1, 2, 3]
    e2 = funnel(["acgt"], Scheme(case_insensitive=False))
    assert _codes(e2, 0) == [4, 4, 4, 4]


def test_unknown_error_policy():
    with pytest.raises(EncodeError):
        funnel(["ACGTN"], Scheme(unknown="error"))
    with pytest.raises(EncodeError):
        funnel(b"ACG\xffT", Scheme(unknown="error"))


def test_fasta(tmp_path):
    p = tmp_path / "q.fasta"
    p.write_text(">seqA desc\nACGT\nACGT\n>seqB\nTTTT\n")
    e = funnel(str(p), DNA)
    assert e.ids == ["seqA", "seqB"] and e.lengths.tolist() == [8, 4]
    g = tmp_path / "q.fasta.gz"
    with gzip.open(g, "wt") as fh:
        fh.write(">s1\nACGT\n")
    assert read_fasta(str(g)) == [("s1", "ACGT")]
    t = funnel(">x\nACGT", DNA)
    assert t.ids == ["x"] and _codes(t, 0) == [0, 1, 2, 3]
    r = funnel("ACGT", DNA)
    assert r.ids == ["q0"]
    with pytest.raises(FileNotFoundError):
        read_fasta("/no/such/file.fasta")
    with pytest.raises(EncodeError):
        funnel("nonexistent_refs.fasta", DNA)


def test_preencoded_and_records():
    e = funnel((np.array([0, 1, 2, 3, 3, 2], np.uint8), np.array([0, 4, 6], np.int32)), DNA)
    assert e.n == 2 and _codes(e, 1) == [3, 2]
    codes = np.array([0, 1, 2, 3], np.uint8)
    for off in ([0, 4, 1000], [0, -1, 4], [0, 4, 2]):
        with pytest.raises(EncodeError):
            funnel((codes, np.array(off, np.int32)), DNA)

    class Rec:
        def __init__(self, i, s):
            self.id, self.seq = i, s

    r = funnel([Rec("a", "ACGT"), Rec("b", "TT")], DNA)
    assert r.ids == ["a", "b"] and r.lengths.tolist() == [4, 2]


def test_empty_and_non_ascii():
    e = funnel(["", "ACGT", ""], DNA)
    assert e.lengths.tolist() == [0, 4, 0] and e.offsets.tolist() == [0, 0, 4, 4]
    b = funnel(b"ACG\xffT", DNA)
    assert b.lengths.tolist() == [5] and _codes(b, 0) == [0, 1, 2, 4, 3]
    u = funnel(["ACGµT"], DNA)
    assert u.lengths.tolist() == [5] and _codes(u, 0)[3] == 4


def test_vectorised_encode_matches_per_sequence():
    rng = np.random.default_rng(4)
    seqs = ["".join(rng.choice(list("ACGTNacgtn-*"), int(rng.integers(0, 50)))) for _ in range(200)]
    e = funnel(seqs, DNA)
    lut, _ = build_lut(DNA)
    for i, s in enumerate(seqs):
        assert _codes(e, i) == lut[np.frombuffer(s.upper().encode("latin-1"), np.uint8)].tolist()
