"""The native FASTA funnel (gsw_fasta_scan/encode) gives exactly the records of the
record-by-record reader (reference src/gpusw/encode.py:91-121) on awkward inputs, and
the on-disk sequence database round-trips.  Host code only: no GPU is touched."""
import gzip

import numpy as np
import pytest

from gpusw_amd import Scheme, _hip
from gpusw_amd.encode import encode_fasta, funnel, load_db, read_fasta, save_db
from gpusw_amd.errors import EncodeError

needs_lib = pytest.mark.skipif(not _hip.LIB_PATH.exists(), reason="libgsw_hip.so not built")

CASES = {
    "plain": b">a first\nACGT\nAC\n>b\nGG\n",
    "crlf": b">a\r\nACGT\r\nAC\r\n>b x\r\nGG\r\n",
    "lone_cr": b">a\rAC\rGT\r>b\rT",
    "blank_lines": b"\n\n>a\n\nAC\n\n\nGT\n>b\n\n",
    "before_first_header": b"JUNK\nACGT\n>a\nAC\n",
    "header_leading_space": b"  >a desc\nAC\n   >b\nGT\n",
    "strip_ends": b">a\n  AC  \n\tGT\t\n",
    "inner_space_kept": b">a\nA C\nG\tT\n",
    "empty_header": b">\nAC\n>   \nGT\n>c\n",
    "vt_ff_breaks": b">a\x0bAC\x0cGT\x1cTT\x1dA\x1eC\n",
    "unit_sep_stripped": b">a\n\x1fAC\x1f\n",
    "no_final_newline": b">a\nACGT",
    "lowercase_unknown": b">a\nacgtnXZ*\n>b\n-\n",
    "id_tabs": b">\tid1\tdesc\nAC\n",
    "only_headers": b">a\n>b\n>c\n",
    "empty": b"",
}


@needs_lib
@pytest.mark.parametrize("name", list(CASES))
@pytest.mark.parametrize("scheme_name", ["dna", "blosum62"])
def test_native_equals_record_reader(tmp_path, name, scheme_name):
    scheme = Scheme.preset(scheme_name)
    p = tmp_path / f"{name}.fa"
    p.write_bytes(CASES[name])
    got = encode_fasta(p, scheme)
    assert got is not None
    recs = read_fasta(str(p))
    want = funnel(recs, scheme) if recs else None
    assert got.ids == [r[0] for r in recs]
    if want is not None:
        assert np.array_equal(got.offsets, want.offsets)
        assert np.array_equal(got.codes, want.codes)
    else:
        assert got.n == 0 and got.codes.size == 0


@needs_lib
def test_gzip_and_funnel_path(tmp_path):
    scheme = Scheme.preset("dna")
    rng = np.random.default_rng(0)
    recs = [(f"r{i}", "".join(rng.choice(list("ACGTN"), int(rng.integers(0, 200))))) for i in range(300)]
    text = "".join(f">{i} d\n" + "\n".join(s[j:j + 60] for j in range(0, len(s), 60)) + "\n" for i, s in recs)
    p = tmp_path / "x.fa.gz"
    with gzip.open(p, "wt") as fh:
        fh.write(text)
    enc = funnel(str(p), scheme)               # the funnel takes the native path for files
    want = funnel(recs, scheme)
    assert enc.ids == want.ids and np.array_equal(enc.codes, want.codes) and np.array_equal(enc.offsets, want.offsets)


@needs_lib
def test_non_ascii_takes_record_path(tmp_path):
    scheme = Scheme.preset("dna")
    p = tmp_path / "u.fa"
    p.write_bytes(">aé\nACGT\n".encode("utf-8"))
    assert encode_fasta(p, scheme) is None
    enc = funnel(str(p), scheme)               # still encoded, by the record reader
    assert enc.n == 1 and enc.lengths.tolist() == [4]


@needs_lib
def test_unknown_error_policy(tmp_path):
    scheme = Scheme.preset("dna").replace(unknown="error")
    p = tmp_path / "e.fa"
    p.write_bytes(b">ok\nACGT\n>bad\nACXT\n")
    with pytest.raises(EncodeError, match="bad"):
        encode_fasta(p, scheme)


def test_db_round_trip(tmp_path):
    scheme = Scheme.preset("blosum62")
    recs = [("p1", "MKV"), ("p2", ""), ("p3", "WWXY")]
    enc = funnel(recs, scheme)
    save_db(tmp_path / "db", enc)
    back = load_db(tmp_path / "db", scheme)
    assert back.ids == enc.ids and np.array_equal(back.codes, enc.codes) and np.array_equal(back.offsets, enc.offsets)
    again = funnel(str(tmp_path / "db"), scheme)     # a database directory is a funnel input
    assert again.ids == enc.ids and np.array_equal(again.codes, enc.codes)
    with pytest.raises(EncodeError, match="alphabet"):
        load_db(tmp_path / "db", Scheme.preset("dna"))
