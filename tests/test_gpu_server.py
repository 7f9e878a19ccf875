"""Serving loop on the GPU: every served top-k equals Aligner.top_k of the same queries
alone, whether the request runs alone or mixed with concurrent requests of other sizes
and another database; memory-budget admission queues and rejects as specified."""
import threading

import numpy as np
import pytest

from gpusw_amd import Aligner, GpuSWError, Scheme
from gpusw_amd.serve import Server
from tests.helpers import PROT, rand_seqs

pytestmark = pytest.mark.gpu


def _expected(scheme, refs, queries, k):
    al = Aligner(scheme, backend="hip").index(refs)
    r = al.top_k(queries, k=k)
    return r.topk_idx, r.topk_scores


@pytest.fixture(scope="module")
def dbs():
    rng = np.random.default_rng(8)
    dna_refs = rand_seqs(rng, 700, 80, 400, "ACGT") + ["", "ACGT"]
    prot_refs = rand_seqs(rng, 400, 30, 500, PROT)
    return rng, dna_refs, prot_refs


def test_alone_and_in_mix_identical(dbs):
    rng, dna_refs, prot_refs = dbs
    dna, prot = Scheme.preset("dna"), Scheme.preset("blosum62")
    probe = rand_seqs(rng, 37, 20, 260, "ACGT") + [""]
    others = [(0, rand_seqs(rng, int(n), 10, 300, "ACGT")) for n in rng.integers(1, 80, 10)]
    others += [(1, rand_seqs(rng, int(n), 10, 600, PROT)) for n in rng.integers(1, 40, 6)]
    want_idx, want_sc = _expected(dna, dna_refs, probe, 5)
    with Server(memory_budget=1 << 30, launch_cells=5e8) as srv:
        d0 = srv.add_database(dna_refs, dna)
        d1 = srv.add_database(prot_refs, prot)
        alone = srv.result(srv.submit(d0, probe, k=5))
        ids = []
        for i, (db, qs) in enumerate(others):
            ids.append((db, qs, srv.submit(d0 if db == 0 else d1, qs, k=5)))
            if i == 5:
                mixed_id = srv.submit(d0, probe, k=5)
        mixed = srv.result(mixed_id)
        for db, qs, rid in ids:
            res = srv.result(rid)
            w_idx, w_sc = _expected(dna if db == 0 else prot, dna_refs if db == 0 else prot_refs, qs, 5)
            assert np.array_equal(res.result.topk_scores, w_sc)
            assert np.array_equal(res.result.topk_idx, w_idx)
        st = srv.stats()
    for r in (alone, mixed):
        assert np.array_equal(r.result.topk_idx, want_idx)
        assert np.array_equal(r.result.topk_scores, want_sc)
        assert r.latency_ms > 0 and r.done_ms >= r.admit_ms >= r.submit_ms
    assert st["completed"] == 2 + len(others) and st["rejected"] == 0


def test_concurrent_submitters(dbs):
    rng, dna_refs, _ = dbs
    dna = Scheme.preset("dna")
    batches = [rand_seqs(rng, int(n), 30, 250, "ACGT") for n in rng.integers(1, 30, 12)]
    out = {}
    with Server(memory_budget=1 << 30, launch_cells=2e8) as srv:
        d0 = srv.add_database(dna_refs, dna)

        def client(i):
            out[i] = srv.result(srv.submit(d0, batches[i], k=3))

        threads = [threading.Thread(target=client, args=(i,)) for i in range(len(batches))]
        for t in threads:
            t.start()
        for t in threads:
            t.join()
    for i, qs in enumerate(batches):
        w_idx, w_sc = _expected(dna, dna_refs, qs, 3)
        assert np.array_equal(out[i].result.topk_idx, w_idx)
        assert np.array_equal(out[i].result.topk_scores, w_sc)


def test_budget_admission_and_rejection(dbs):
    """Requests whose score rows exceed what is free wait for memory (FIFO) and still
    finish exact; a request larger than the whole budget is rejected."""
    rng, dna_refs, _ = dbs
    dna = Scheme.preset("dna")
    nr = len(dna_refs)
    budget = 64 << 20                      # 4 MiB of query codes + 60 MiB of score rows
    per_req = 9000                         # 9000 x nr x 4 B ~ 25 MB: two fit at a time
    with Server(memory_budget=budget, launch_cells=2e8) as srv:
        d0 = srv.add_database(dna_refs, dna)
        reqs = [rand_seqs(rng, per_req, 20, 60, "ACGT") for _ in range(4)]
        ids = [srv.submit(d0, q, k=4) for q in reqs]
        too_big = srv.submit(d0, rand_seqs(rng, (60 << 20) // (4 * nr) + 50, 20, 30, "ACGT"), k=4)
        results = [srv.result(rid) for rid in ids]
        with pytest.raises(GpuSWError):
            srv.result(too_big)
        st = srv.stats()
    assert st["rejected"] == 1 and st["completed"] == 4
    # later requests were admitted only after earlier ones released memory
    assert results[3].admit_ms >= min(r.done_ms for r in results[:2])
    for q, res in zip(reqs, results):
        w_idx, w_sc = _expected(dna, dna_refs, q, 4)
        assert np.array_equal(res.result.topk_idx, w_idx)
        assert np.array_equal(res.result.topk_scores, w_sc)


def test_serving_mix_at_full_database_sizes():
    """The serving mix at its workload: the 16S-shape DNA database (20,000 references) and
    the protein database (100,000 references) resident in one server, 24 Poisson-timed
    requests of 1-64 queries (70 % DNA, 30 % protein).  Every served top-k equals
    Aligner.top_k of the same queries alone (src/gpusw/aligner.py:169-195), and the oracle
    agrees with a sample of the served hits."""
    import time

    from gpusw_amd import synth
    from oracle import native

    (dqc, dqo), (drc, dro) = synth.dna_16s_cross(seed=0)
    (pqc, pqo), (prc, pro) = synth.protein_topk(seed=0)
    pools = [(synth.unpack(dqc, dqo), Scheme.preset("dna"), (drc, dro)),
             (synth.unpack(pqc, pqo), Scheme.preset("blosum62"), (prc, pro))]
    rng = np.random.default_rng(404)
    n = 24
    kinds = (rng.random(n) >= 0.7).astype(int)
    sizes = np.floor(np.exp(rng.uniform(0, np.log(65), n))).astype(int)
    gaps = rng.exponential(1 / 20.0, n)
    reqs = [(int(k), rng.integers(0, len(pools[k][0]), int(m))) for k, m in zip(kinds, sizes)]
    with Server(memory_budget=16 << 30) as srv:
        dbs = [srv.add_database(p[2], p[1]) for p in pools]
        ids = []
        for (k, pick), g in zip(reqs, gaps):
            time.sleep(g)
            ids.append(srv.submit(dbs[k], synth.pack([pools[k][0][j] for j in pick]), k=5))
        served = [srv.result(i) for i in ids]
        st = srv.stats()
    assert st["completed"] == n and st["rejected"] == 0
    alone = [Aligner(p[1], backend="hip").index(p[2]) for p in pools]
    hits = [[], []]
    for (k, pick), r in zip(reqs, served):
        want = alone[k].top_k(synth.pack([pools[k][0][j] for j in pick]), k=5)
        assert np.array_equal(r.result.topk_idx, want.topk_idx)
        assert np.array_equal(r.result.topk_scores, want.topk_scores)
        hits[k] += [(int(q), int(i), int(s)) for q, i, s in
                    zip(pick, r.result.topk_idx[:, 0], r.result.topk_scores[:, 0])]
    for k, (qc, qo) in enumerate([(dqc, dqo), (pqc, pqo)]):
        if not hits[k]:
            continue
        h = np.array(hits[k][:60], dtype=np.int64)
        rc, ro = pools[k][2]
        sch = pools[k][1]
        want = native.score_pairs(qc, qo, rc, ro, h[:, 0].astype(np.int32), h[:, 1].astype(np.int32),
                                  sch.substitution_table(), sch.mode, sch.gap_open, sch.gap_extend,
                                  sch.free_end_gaps_query, sch.free_end_gaps_ref)
        assert np.array_equal(want, h[:, 2])


def test_listed_half2_results_rerun_exact(dbs):
    """Protein queries that are copies of long references score 2,048 or more: their half2
    results are listed, the rows run again on the int16 tiles before the top-k, and every
    served row equals the batched top-k (and the oracle on its best hit), alone and beside
    other requests."""
    from oracle import native
    from gpusw_amd.encode import funnel

    rng, _, prot_refs = dbs
    prot = Scheme.preset("blosum62")
    heavy = rand_seqs(rng, 6, 700, 1100, PROT)   # self scores ~5 x length
    refs = prot_refs + heavy
    queries = heavy[:3] + [h[50:650] for h in heavy[3:]] + rand_seqs(rng, 10, 40, 500, PROT)
    want_idx, want_sc = _expected(prot, refs, queries, 4)
    assert int(want_sc[:, 0].max()) >= 2048 and int(want_sc[:, 0].min()) < 2048
    with Server(memory_budget=1 << 30, launch_cells=3e8) as srv:
        d1 = srv.add_database(refs, prot)
        side = srv.submit(d1, rand_seqs(rng, 30, 30, 600, PROT), k=4)
        res = srv.result(srv.submit(d1, queries, k=4))
        srv.result(side)
        st = srv.stats()
    assert np.array_equal(res.result.topk_idx, want_idx)
    assert np.array_equal(res.result.topk_scores, want_sc)
    assert st["rescored"] > 0
    qi = np.arange(len(queries), dtype=np.int32)
    best = native.score_encoded_pairs(funnel(queries, prot), funnel(refs, prot), qi,
                                      want_idx[:, 0].astype(np.int32), prot)
    assert np.array_equal(best, want_sc[:, 0].astype(np.int64))
