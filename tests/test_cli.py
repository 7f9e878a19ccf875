"""Command line: argument handling, info and index without a GPU (reference
tests/test_cli.py:20-36,77-80); the scoring subcommands and the HTTP server on the GPU
(reference tests/test_cli.py:39-75 known answers)."""
import csv
import json
import threading
import urllib.error
import urllib.request

import numpy as np
import pytest

from gpusw_amd import Aligner, Scheme, __version__
from gpusw_amd.cli import _parse_db, main, make_http_server
from gpusw_amd.encode import funnel, load_db


def _fasta(tmp_path, name, recs):
    p = tmp_path / name
    p.write_text("".join(f">{i}\n{s}\n" for i, s in recs))
    return str(p)


def _read_csv(path):
    with open(path) as fh:
        return list(csv.reader(fh))


# ---- no GPU ---------------------------------------------------------------------
def test_info_runs(capsys):
    assert main(["info"]) == 0
    out = capsys.readouterr().out
    assert __version__ in out
    assert "bundled matrices:" in out and "BLOSUM62" in out


def test_missing_required_arg_exits_2():
    with pytest.raises(SystemExit) as e:
        main(["cross", "--refs", "r.fasta"])
    assert e.value.code == 2


def test_no_subcommand_exits():
    with pytest.raises(SystemExit):
        main([])


def test_missing_file_errors(tmp_path):
    r = _fasta(tmp_path, "r.fasta", [("ra", "ACGT")])
    with pytest.raises(FileNotFoundError):
        main(["cross", "--queries", str(tmp_path / "nope.fasta"), "--refs", r, "--scheme", "dna"])


def test_backend_choices():
    with pytest.raises(SystemExit) as e:
        main(["cross", "--queries", "q", "--refs", "r", "--backend", "cuda"])
    assert e.value.code == 2


def test_index_writes_database(tmp_path):
    r = _fasta(tmp_path, "r.fasta", [("p1", "MKVL"), ("p2", "WW"), ("p3", "")])
    out = tmp_path / "r.gswdb"
    assert main(["index", "--refs", r, "--scheme", "blosum62", "--out", str(out)]) == 0
    db = load_db(out, Scheme.preset("blosum62"))
    want = funnel(r, Scheme.preset("blosum62"))
    assert db.ids == ["p1", "p2", "p3"] and np.array_equal(db.codes, want.codes)


def test_parse_db_spec():
    assert _parse_db("prot=/x/y.gswdb:blosum62", "dna") == ("prot", "/x/y.gswdb", "blosum62")
    assert _parse_db("d=r.fa", "dna") == ("d", "r.fa", "dna")


# ---- GPU ------------------------------------------------------------------------
@pytest.mark.gpu
@pytest.mark.parametrize("backend", ["hip", "hip-plain"])
def test_cli_cross(tmp_path, backend):
    q = _fasta(tmp_path, "q.fasta", [("qa", "ACGTACGT"), ("qb", "TTTT")])
    r = _fasta(tmp_path, "r.fasta", [("ra", "ACGTACGT"), ("rb", "ACGT")])
    out = str(tmp_path / "cross.csv")
    assert main(["cross", "--queries", q, "--refs", r, "--scheme", "dna", "--backend", backend, "-o", out]) == 0
    rows = _read_csv(out)
    assert rows[0] == ["query", "reference", "score"] and len(rows) == 1 + 2 * 2
    scores = {(a, b): int(s) for a, b, s in rows[1:]}
    assert scores[("qa", "ra")] == 16


@pytest.mark.gpu
def test_cli_topk_from_index(tmp_path):
    q = _fasta(tmp_path, "q.fasta", [("qa", "ACGTACGT")])
    r = _fasta(tmp_path, "r.fasta", [("ra", "ACGTACGT"), ("rb", "TTTTTTTT"), ("rc", "ACGT")])
    db = str(tmp_path / "r.gswdb")
    assert main(["index", "--refs", r, "--scheme", "dna", "--out", db]) == 0
    out = str(tmp_path / "topk.csv")
    assert main(["topk", "--queries", q, "--refs", db, "--scheme", "dna", "-k", "2", "-o", out]) == 0
    rows = _read_csv(out)
    assert rows[0] == ["query", "rank", "reference", "score"] and len(rows) == 1 + 2
    assert rows[1][2] == "ra"


@pytest.mark.gpu
def test_cli_pairs(tmp_path):
    q = _fasta(tmp_path, "q.fasta", [("qa", "ACGT"), ("qb", "TTTT")])
    r = _fasta(tmp_path, "r.fasta", [("ra", "ACGT"), ("rb", "TTTA")])
    out = str(tmp_path / "pairs.csv")
    assert main(["pairs", "--queries", q, "--refs", r, "--scheme", "dna", "-o", out]) == 0
    rows = _read_csv(out)
    assert rows[0] == ["query", "reference", "score"]
    assert [int(x[2]) for x in rows[1:]] == [8, 6]


@pytest.mark.gpu
def test_http_serve_matches_aligner():
    rng = np.random.default_rng(4)
    alpha = np.array(list("ARNDCQEGHILKMFPSTWYV"))
    refs = [(f"r{i}", "".join(rng.choice(alpha, int(rng.integers(30, 400))))) for i in range(200)]
    queries = [(f"q{i}", "".join(rng.choice(alpha, int(rng.integers(20, 300))))) for i in range(12)]
    scheme = Scheme.preset("blosum62")
    want = Aligner(scheme, backend="hip").index(refs).top_k(queries, k=3)
    httpd, srv = make_http_server({"prot": (refs, scheme)}, port=0, memory_budget=1 << 30)
    th = threading.Thread(target=httpd.serve_forever, daemon=True)
    th.start()
    try:
        base = f"http://127.0.0.1:{httpd.server_address[1]}"

        def post(obj):
            req = urllib.request.Request(base + "/topk", data=json.dumps(obj).encode(),
                                         headers={"Content-Type": "application/json"})
            with urllib.request.urlopen(req, timeout=60) as resp:
                return resp.status, json.loads(resp.read())

        results = [None] * 3
        def worker(i):   # concurrent requests share launches
            results[i] = post({"db": "prot", "k": 3, "queries": queries[i * 4:(i + 1) * 4]})
        ts = [threading.Thread(target=worker, args=(i,)) for i in range(3)]
        for t in ts:
            t.start()
        for t in ts:
            t.join()
        got = [h for _, body in results for h in body["hits"]]
        assert [h["query"] for h in got] == [q[0] for q in queries]
        for qi, h in enumerate(got):
            assert [tuple(x) for x in h["hits"]] == [(refs[j][0], int(s)) for j, s in
                                                      zip(want.topk_idx[qi], want.topk_scores[qi])]
        # FASTA body, error paths, stats
        status, body = post({"db": "prot", "k": 1, "fasta": f">x\n{queries[0][1]}\n"})
        assert status == 200 and body["hits"][0]["hits"][0][1] == int(want.topk_scores[0][0])
        with pytest.raises(urllib.error.HTTPError) as e:
            post({"db": "nope", "queries": queries[:1]})
        assert e.value.code == 400
        with urllib.request.urlopen(base + "/stats", timeout=30) as resp:
            st = json.loads(resp.read())
        assert st["completed"] >= 4 and st["databases"] == ["prot"]
    finally:
        httpd.shutdown()
        httpd.server_close()
        srv.close()
