#!/usr/bin/env python
"""Benchmark: bit-exact GCUPS of the batched HIP scoring path on one MI355X.

Default workload (one JSON line on stdout): the 16S-shape DNA cross -- 2,000 queries of
240-260 nt against 20,000 references of ~1,450 nt (synthetic, seed 0), local ``dna``
scheme, int16 with int32 fallback.  One step = the full (2000, 20000) score matrix
computed into HBM plus a device top_k(5) of every row.  Inputs are resident in HBM
before the timed region.  Also reported:

* ``roofline``: packed-int16 VALU roofline of the cross kernel -- algorithmic int16 ops
  (cells x 10 for local mode) over the kernel's device time measured with HIP events on
  the library stream, against the gfx950 peak of 78.6 T int16-ops/s;
* ``plain``: the one-thread-per-pair comparator on the same GPU (a query subsample);
* ``cpu_baseline``: the oracle's C restatement on the host cores over a bounded sample
  of the same pairs (those pairs are also checked bit-exact against the GPU scores).

``--workload`` selects another configuration for extra lines: ``plain_dna`` (the 64 x 256
int32 case, oracle-checked in full), ``protein_topk`` (BLOSUM62
top-5 of 1,000 queries against 100,000 references), ``read_pairs`` (1,000,000 explicit
read/window pairs, semiglobal then global, graph-replayed batches), ``serving`` (Poisson
request mix through the continuous-batching server, p50/p99 latency).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

METRIC = "bit-exact GCUPS on 1 MI355X (batched vs plain); p50/p99 latency when served"
PEAK_INT16_TOPS = 78.6          # 256 CU x 64 lanes x 2 (packed) x 2.4 GHz, int16 ops/s / 1e12
OPS_PER_CELL = {"local": 10, "global": 8, "semiglobal": 8}
# int16 ops the kernels issue per cell (a packed instruction is 2 int16 ops for 2 cells,
# one per query of the pair).  Local folds max(H, 0) into saturating E/F updates and
# issues 9, so its algorithmic fraction can reach 10/9 at full VALU issue.
ISSUED_OPS_PER_CELL = {10: 9, 8: 8}


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def cpu_baseline(qenc, renc, scheme, gpu_scores, rng, target_s: float = 12.0):
    """Time the oracle's C restatement on a bounded random sample of (query, ref) pairs
    and check the GPU scores of those pairs bit-exactly."""
    from oracle import native

    threads = len(os.sched_getaffinity(0))
    qlen = qenc.lengths.astype(np.int64)
    rlen = renc.lengths.astype(np.int64)
    # calibration sample (large enough that thread start-up does not dominate its time)
    n = 2048
    qi = rng.integers(0, qenc.n, n).astype(np.int32)
    rj = rng.integers(0, renc.n, n).astype(np.int32)
    t0 = time.perf_counter()
    native.score_encoded_pairs(qenc, renc, qi, rj, scheme, nthreads=threads)
    dt = time.perf_counter() - t0
    rate = float((qlen[qi] * rlen[rj]).sum()) / max(dt, 1e-6)
    n = int(max(256, min(2_000_000, target_s * rate / float(qlen.mean() * rlen.mean()))))
    qi = rng.integers(0, qenc.n, n).astype(np.int32)
    rj = rng.integers(0, renc.n, n).astype(np.int32)
    t0 = time.perf_counter()
    want = native.score_encoded_pairs(qenc, renc, qi, rj, scheme, nthreads=threads)
    dt = time.perf_counter() - t0
    cells = float((qlen[qi] * rlen[rj]).sum())
    got = gpu_scores[qi, rj].astype(np.int64)
    mism = int((got != want).sum())
    return {
        "value": round(cells / dt / 1e9, 4), "unit": "GCUPS", "cores": threads, "kind": "port",
        "sample": f"{n} random (query, ref) pairs of the same workload, {cells:.3e} cells, {dt:.1f} s",
    }, mism, n


def _tile_name(variant: int) -> str | None:
    from gpusw_amd import _hip

    if variant < 0:
        return None
    g, r = _hip.list_variants()[variant // 4]
    kind = ("int16x2" if variant % 4 < 2 else "int32") + ("/local" if variant % 2 == 0 else "/nonlocal")
    return f"G={g} lanes x R={r} columns ({kind})"


def _kernel_name(variant: int, striped: int) -> str | None:
    """The kernel's name as rocprofv3 prints it, for a variant id of the batched path."""
    from gpusw_amd import _hip

    if variant < 0:
        return None
    g, r = _hip.list_variants()[variant // 4]
    elem = "short __vector(2)" if variant % 4 < 2 else "int"
    local = "true" if variant % 2 == 0 else "false"
    kind, waves = {0: ("cross", 8), 1: ("stripe", 8), 2: ("pairs", 8), 3: ("pairs_tab", 4)}[int(striped)]
    return f"void gsw::{kind}_kernel<{elem}, {g}, {r}, {local}, {waves}>(gsw::KArgs)"


def _measured_traffic(workload: str, kernel: str | None):
    """HBM bytes per launch of `kernel` on `workload` from the committed counter summaries
    (rows ``| traffic | <workload> | `<kernel>` | <bytes> |`` under profiles/), or None."""
    if kernel is None:
        return None, None
    for f in sorted((ROOT / "profiles").glob("*.md")):
        for line in f.read_text().splitlines():
            cells = [c.strip() for c in line.strip().strip("|").split("|")]
            if len(cells) == 4 and cells[0] == "traffic" and cells[1] == workload and cells[2].strip("`") == kernel:
                return float(cells[3]), f"profiles/{f.name}"
    return None, None


def _roofline(history, ops_per_cell: int, workload: str) -> dict:
    """Roofline of the dominant kernel: the batched launch with the most cells in each timed
    call (its cells x ops over its device time from HIP events on the stream it ran on)."""
    top = [h for h in history if h.top_variant >= 0 and h.top_ms > 0]
    if not top:
        return {"bound": "valu", "achieved": None, "peak": PEAK_INT16_TOPS, "unit": "TOP/s (int16)",
                "frac": None, "traffic": None}
    # the most frequent (variant, striped) among the calls' dominant launches (most cells)
    keys = [(h.top_variant, h.top_striped) for h in top]
    var, striped = max(set(keys), key=keys.count)
    sel = [h for h in top if (h.top_variant, h.top_striped) == (var, striped)]
    ms = float(np.mean([h.top_ms for h in sel]))
    cells = float(np.mean([h.top_cells for h in sel]))
    achieved = cells * ops_per_cell / (ms * 1e-3) / 1e12
    kernel = _kernel_name(var, striped)
    traffic, source = _measured_traffic(workload, kernel)
    out = {"bound": "valu", "achieved": round(achieved, 3), "peak": PEAK_INT16_TOPS, "unit": "TOP/s (int16)",
           "frac": round(achieved / PEAK_INT16_TOPS, 4), "traffic": traffic,
           "kernel": kernel, "tile": _tile_name(var), "launch_ms": round(ms, 3), "launch_cells": int(cells),
           "ops_per_cell": ops_per_cell, "issued_ops_per_cell": ISSUED_OPS_PER_CELL[ops_per_cell],
           "issued_frac": round(achieved * ISSUED_OPS_PER_CELL[ops_per_cell] / ops_per_cell / PEAK_INT16_TOPS, 4),
           "launches_timed": len(sel)}
    if traffic is not None:
        out["traffic_unit"] = "bytes/launch (rocprofv3 FETCH_SIZE, WRITE_SIZE, copy-calibrated)"
        out["traffic_source"] = source
        out["hbm_GBps"] = round(traffic / (ms * 1e-3) / 1e9, 2)   # secondary: far below 8 TB/s
    return out


def _call_traffic(workload: str):
    """HBM bytes of all batched launches of one call: the sum of the workload's traffic rows
    (one dispatch of each kernel per call) under profiles/, or None."""
    total, src = 0.0, None
    for f in sorted((ROOT / "profiles").glob("*.md")):
        for line in f.read_text().splitlines():
            cells = [c.strip() for c in line.strip().strip("|").split("|")]
            if len(cells) == 4 and cells[0] == "traffic" and cells[1] == workload:
                total += float(cells[3])
                src = f"profiles/{f.name}"
    return (total, src) if src else (None, None)


def _call_roofline(history, cells: int, kernel_s: float, steps: int) -> dict:
    """Roofline of a protein top-k call as a whole.  Its tile launches (19) run on two
    streams and overlap (each launch's workgroups take CUs as the other stream's finish), so
    no single launch owns the GPU for its interval: the dominant unit is the call's set of
    batched launches, timed from the first launch's start to the last one's end (HIP events
    on the library stream, joined with the second stream)."""
    achieved = cells * OPS_PER_CELL["local"] / kernel_s / 1e12
    calls = max(1, len(history) // max(steps, 1))
    traffic, src = _call_traffic("protein")
    out = {"bound": "valu", "achieved": round(achieved, 3), "peak": PEAK_INT16_TOPS, "unit": "TOP/s (int16)",
           "frac": round(achieved / PEAK_INT16_TOPS, 4),
           "traffic": None if traffic is None else round(traffic * calls, 1),
           "kernel": "all batched launches of a top_k call (cross_kernel + stripe_kernel tiles, two streams)",
           "ops_per_cell": OPS_PER_CELL["local"], "issued_ops_per_cell": ISSUED_OPS_PER_CELL[OPS_PER_CELL["local"]],
           "issued_frac": round(achieved * ISSUED_OPS_PER_CELL[OPS_PER_CELL["local"]] / OPS_PER_CELL["local"]
                                / PEAK_INT16_TOPS, 4),
           "kernels_ms_per_step": round(kernel_s * 1e3, 2), "calls_per_step": calls,
           "largest_launch": _roofline(history, OPS_PER_CELL["local"], "protein")}
    if traffic is not None:
        out["traffic_unit"] = "bytes per step (sum of the per-dispatch rocprofv3 rows of every batched kernel)"
        out["traffic_source"] = src
        out["hbm_GBps"] = round(traffic * calls / kernel_s / 1e9, 2)
    return out


def run_dna_cross(args):
    import gpusw_amd
    from gpusw_amd import Aligner, Scheme, synth

    t0 = time.perf_counter()
    (qc, qo), (rc, ro) = synth.dna_16s_cross(seed=0)
    log(f"workload generated in {time.perf_counter() - t0:.1f}s")
    scheme = Scheme.preset("dna")
    al = Aligner(scheme, backend="hip").index((rc, ro)).set_queries((qc, qo))
    b = al._b
    nq, nr = al._queries.n, al._refs.n
    qlist = np.arange(nq, dtype=np.int32)
    cells = int(al._queries.lengths.astype(np.int64).sum()) * int(al._refs.lengths.astype(np.int64).sum())
    log(f"device: {gpusw_amd._hip.device_name(0)}; {nq} x {nr} pairs, {cells:.3e} cells")

    def step():
        scores, dt, _src = b.run_cross(scheme, al._queries.max_len, "int16", 128, al._d_qbuf,
                                       al._d_qoff, al._d_rbuf, al._d_roff, qlist, nq, nr)
        idx, sc = b.topk(scores, 5)
        return scores, dt, idx, sc
    history = []

    for w in range(args.warmup):
        s = time.perf_counter()
        step()
        log(f"warmup {w + 1}/{args.warmup}: {time.perf_counter() - s:.2f}s")
    gpusw_amd._hip.call("gsw_synchronize")
    kernel_s = []
    t0 = time.perf_counter()
    for k in range(args.steps):
        scores, dt, idx, sc = step()
        kernel_s.append(dt)
        history.append(b.last_stats)
        log(f"step {k + 1}/{args.steps}: kernel {dt * 1e3:.1f} ms")
    gpusw_amd._hip.call("gsw_synchronize")
    wall = time.perf_counter() - t0
    ms_per_step = wall / args.steps * 1e3
    value = cells / (wall / args.steps) / 1e9
    stats = b.last_stats
    kavg = float(np.mean(kernel_s))
    roof = _roofline(history, OPS_PER_CELL["local"], "dna16s")
    roof.update({"kernels_ms_per_step": round(kavg * 1e3, 3), "launches_per_step": int(stats.launches)})

    # ---- correctness of the timed result (outside the timed region)
    host = scores.to_host()
    order = np.argsort(-host, axis=1, kind="stable")[:, :5]
    topk_ok = bool(np.array_equal(order, idx) and np.array_equal(np.take_along_axis(host, order, 1), sc))

    # ---- plain comparator on a query subsample
    rng = np.random.default_rng(1)
    nplain = args.plain_queries
    psel = np.sort(rng.choice(nq, nplain, replace=False)).astype(np.int32)
    pb = gpusw_amd.backends.resolve_backend("hip-plain")
    pscores, pdt, _ = pb.run_cross(scheme, al._queries.max_len, "int16", 128, al._d_qbuf, al._d_qoff,
                                   al._d_rbuf, al._d_roff, psel, len(psel), nr)
    phost = pscores.to_host()
    pcells = int(al._queries.lengths[psel].astype(np.int64).sum()) * int(al._refs.lengths.astype(np.int64).sum())
    plain_gcups = pcells / pdt / 1e9
    plain_match = bool(np.array_equal(phost, host[psel]))
    log(f"plain: {plain_gcups:.1f} GCUPS over {len(psel)} queries, equal to batched: {plain_match}")

    # ---- traceback of the top-5 hits of 512 queries (outside the timed region)
    tq = np.repeat(np.arange(512, dtype=np.int32), 5)
    tr = np.ascontiguousarray(idx[:512].reshape(-1), dtype=np.int32)
    t0 = time.perf_counter()
    alns = al.align(tq, tr)
    tb_wall = time.perf_counter() - t0
    tb_ok = [a.score for a in alns] == np.take_along_axis(host[:512], idx[:512], 1).reshape(-1).tolist()
    traceback = {"pairs": int(tq.size), "wall_ms": round(tb_wall * 1e3, 2),
                 "kernel_ms": round(al._last_seconds * 1e3, 2), "scores_equal_matrix": bool(tb_ok),
                 "mean_identity": round(float(np.mean([a.identity for a in alns])), 4)}
    log(f"traceback: {tq.size} alignments in {tb_wall * 1e3:.1f} ms, scores equal: {tb_ok}")

    cpu = None
    mism = 0
    if not args.no_cpu:
        cpu, mism, nsample = cpu_baseline(al._queries, al._refs, scheme, host, rng)
        log(f"cpu baseline: {cpu['value']} GCUPS on {cpu['cores']} threads; {mism} mismatches in {nsample}")

    line = {
        "metric": METRIC,
        "value": round(value, 2),
        "unit": "GCUPS",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 2),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int16",
        "data": "synthetic",
        "config": {
            "workload": "16S-shape DNA cross: 2000 queries (240-260 nt) x 20000 references "
                        "(~1450 nt), dna local scheme, int16 with int32 fallback, full score "
                        "matrix in HBM + top_k(5)",
            "queries": nq, "references": nr, "cells": cells, "scheme": "dna", "mode": "local",
            "parallelism": "single GPU",
        },
        "roofline": roof,
        "plain": {"gcups": round(plain_gcups, 2), "queries": int(len(psel)),
                  "speedup": round(value / plain_gcups, 2), "equal_to_batched": plain_match},
        "checks": {"topk_equals_stable_sort": topk_ok, "plain_equals_batched": plain_match,
                   "oracle_mismatches": mism, "rescored": int(stats.rescored)},
        "traceback": traceback,
        "cpu_baseline": cpu,
    }
    return line


def run_plain_dna(args):
    """64 x 256 DNA local int32 (lengths 100-300, seed 0): the reference's CPU-runnable case.
    Batched and plain kernels are both timed; the oracle scores all 16,384 pairs (the CPU
    baseline here is the whole workload) and both GPU matrices must equal it."""
    import gpusw_amd
    from gpusw_amd import Aligner, Scheme, synth
    from oracle import native

    (qc, qo), (rc, ro) = synth.plain_dna(seed=0)
    scheme = Scheme.preset("dna").replace(dtype="int32")
    res = {}
    # a call is under a millisecond: at least three untimed calls (the first calls of a new
    # aligner still settle kernel attributes, events and buffers)
    warmup = max(args.warmup, 3)
    for backend in ("hip", "hip-plain"):
        al = Aligner(scheme, backend=backend).index((rc, ro)).set_queries((qc, qo))
        for _ in range(warmup):
            al.score_cross()
        gpusw_amd._hip.call("gsw_synchronize")
        t0 = time.perf_counter()
        for _ in range(args.steps):
            got = al.score_cross()
        wall = (time.perf_counter() - t0) / args.steps
        res[backend] = (got, wall, al)
    al = res["hip"][2]
    cells = int(al._queries.lengths.astype(np.int64).sum()) * int(al._refs.lengths.astype(np.int64).sum())
    threads = len(os.sched_getaffinity(0))
    t0 = time.perf_counter()
    want = native.score_cross(qc, qo, rc, ro, np.arange(64, dtype=np.int32), scheme.substitution_table(),
                              "local", scheme.gap_open, scheme.gap_extend, True, True, nthreads=threads)
    cpu_s = time.perf_counter() - t0
    hip_got, hip_wall, _ = res["hip"]
    plain_got, plain_wall, _ = res["hip-plain"]
    value = cells / hip_wall / 1e9
    return {
        "metric": METRIC, "value": round(value, 2), "unit": "GCUPS", "n_gpus": 1, "steps": args.steps,
        "warmup": warmup, "ms_per_step": round(hip_wall * 1e3, 3), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "int32", "data": "synthetic",
        "config": {"workload": "plain DNA: 64 x 256 (lengths 100-300, seed 0), dna local, int32; host-side "
                               "call including the 64 KB result copy", "cells": cells},
        "plain": {"gcups": round(cells / plain_wall / 1e9, 2), "ms_per_step": round(plain_wall * 1e3, 3)},
        "checks": {"batched_equals_oracle": bool(np.array_equal(hip_got.astype(np.int64), want)),
                   "plain_equals_oracle": bool(np.array_equal(plain_got.astype(np.int64), want))},
        "cpu_baseline": {"value": round(cells / cpu_s / 1e9, 4), "unit": "GCUPS", "cores": threads, "kind": "port",
                         "sample": f"all 16,384 pairs ({cells:.3e} cells), {cpu_s:.2f} s"},
    }


def _plain_subsample_cross(al, scheme, qsel):
    """GCUPS of the one-thread-per-pair kernel on a query subsample (same GPU)."""
    import gpusw_amd

    pb = gpusw_amd.backends.resolve_backend("hip-plain")
    nr = al._refs.n
    sc, dt, _ = pb.run_cross(scheme, al._queries.max_len, "int32", 128, al._d_qbuf, al._d_qoff,
                             al._d_rbuf, al._d_roff, qsel, len(qsel), nr)
    cells = int(al._queries.lengths[qsel].astype(np.int64).sum()) * int(al._refs.lengths.astype(np.int64).sum())
    return sc.to_host(), cells / dt / 1e9


def run_protein_topk(args):
    """Protein top-k: 1,000 lognormal queries vs 100,000 lognormal references, BLOSUM62
    local -11/-1, device top_k(5) per query batch of 512 (the Aligner.top_k loop)."""
    import gpusw_amd
    from gpusw_amd import Aligner, Scheme, synth

    (qc, qo), (rc, ro) = synth.protein_topk(seed=0)
    scheme = Scheme.preset("blosum62")
    al = Aligner(scheme, backend="hip").index((rc, ro)).set_queries((qc, qo))
    nq, nr = al._queries.n, al._refs.n
    cells = int(al._queries.lengths.astype(np.int64).sum()) * int(al._refs.lengths.astype(np.int64).sum())
    log(f"protein top-k: {nq} x {nr}, {cells:.3e} cells")
    for _ in range(args.warmup):
        al.top_k(k=5)
    gpusw_amd._hip.call("gsw_synchronize")
    al._b.history = []          # the backend appends each call's stats
    t0 = time.perf_counter()
    for _ in range(args.steps):
        res = al.top_k(k=5)
    wall = (time.perf_counter() - t0) / args.steps
    history, al._b.history = al._b.history, None
    kernel_s = sum(h.kernel_ms for h in history) / 1e3 / args.steps
    # checks: top-k of a query subsample against the comparator's full rows
    rng = np.random.default_rng(2)
    qsel = np.sort(rng.choice(nq, 8, replace=False)).astype(np.int32)
    prow, plain_gcups = _plain_subsample_cross(al, scheme, qsel)
    order = np.argsort(-prow, axis=1, kind="stable")[:, :5]
    ok = bool(np.array_equal(order, res.topk_idx[qsel]) and
              np.array_equal(np.take_along_axis(prow, order, 1), res.topk_scores[qsel]))
    value = cells / wall / 1e9
    return {
        "metric": METRIC, "value": round(value, 2), "unit": "GCUPS", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(wall * 1e3, 2), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "int16", "data": "synthetic",
        "config": {"workload": "protein top-k: 1000 queries (lognormal median 350, 30-2000) x 100000 "
                               "references (lognormal median 300, 30-3000), BLOSUM62 local -11/-1, top_k(5)",
                   "cells": cells},
        "roofline": _call_roofline(history, cells, kernel_s, args.steps),
        "plain": {"gcups": round(plain_gcups, 2), "queries": int(len(qsel)), "speedup": round(value / plain_gcups, 2)},
        "checks": {"topk_equals_comparator_rows": ok},
    }


def run_read_pairs(args):
    """1,000,000 explicit (read, window) pairs, dna scheme, semiglobal then global, int16."""
    import gpusw_amd
    from gpusw_amd import Aligner, Scheme, synth
    from oracle import native

    (rc, ro), (wc, wo), qi, rj = synth.read_pairs(seed=0)
    # both modes are timed before any check runs (the CPU oracle's threads and the comparator
    # stay out of the timed calls)
    timed = []
    # a call is ~10 ms: at least three untimed calls per mode (the first calls of a new
    # aligner still settle host buffers and planning threads)
    warmup = max(args.warmup, 3)
    for mode in ("semiglobal", "global"):
        scheme = Scheme.preset("dna").replace(mode=mode, dtype="int16")
        al = Aligner(scheme, backend="hip").index((wc, wo)).set_queries((rc, ro))
        for _ in range(warmup):
            al.score_pairs(qi, rj)
        al._b.history = []
        t0 = time.perf_counter()
        for _ in range(args.steps):
            got = al.score_pairs(qi, rj)
        wall = (time.perf_counter() - t0) / args.steps
        history, al._b.history = al._b.history, None
        timed.append((mode, scheme, al, got, wall, history, al._b.last_stats))
    out = []
    for mode, scheme, al, got, wall, history, st in timed:
        cells = int((al._queries.lengths[qi].astype(np.int64) * al._refs.lengths[rj].astype(np.int64)).sum())
        rng = np.random.default_rng(3)
        pick = rng.choice(qi.size, 20000, replace=False)
        want = native.score_encoded_pairs(al._queries, al._refs, qi[pick], rj[pick], scheme)
        mism = int((got[pick].astype(np.int64) != want).sum())
        sub = rng.choice(qi.size, 50000, replace=False)
        pb = gpusw_amd.backends.resolve_backend("hip-plain")
        ph, pdt, _ = pb.run_pairs(scheme, al._queries.max_len, "int16", 128, al._d_qbuf, al._d_qoff,
                                  al._d_rbuf, al._d_roff, qi[sub], rj[sub], sub.size)
        pcells = int((al._queries.lengths[qi[sub]].astype(np.int64) * al._refs.lengths[rj[sub]]).sum())
        value = cells / wall / 1e9
        out.append({
            "metric": METRIC, "value": round(value, 2), "unit": "GCUPS", "n_gpus": 1, "steps": args.steps,
            "warmup": warmup, "ms_per_step": round(wall * 1e3, 2), "higher_is_better": True,
            "scaling": "weak", "vs_baseline": None, "dtype": "int16", "data": "synthetic",
            "config": {"workload": f"read pairs: 1,000,000 (150-nt read, 200-400-nt window) pairs, dna "
                                   f"{mode}, int16, HIP-graph batches of 131,072", "cells": cells},
            "roofline": dict(_roofline(history, OPS_PER_CELL[mode], "readpairs"),
                             all_batches_ms=round(st.kernel_ms, 2), host_wall_ms=round(st.wall_ms, 2),
                             launches=int(st.launches)),
            "plain": {"gcups": round(pcells / pdt / 1e9, 2), "pairs": int(sub.size),
                      "speedup": round(value / (pcells / pdt / 1e9), 2),
                      "equal_to_batched": bool(np.array_equal(ph, got[sub]))},
            "checks": {"oracle_mismatches": mism, "oracle_sample": int(pick.size)},
        })
    return out


def run_serving(args):
    """Serving mix: Poisson arrivals of requests holding 1-512 queries (log-uniform count),
    70% DNA against the 16S-shape database and 30% protein against the protein database,
    top_k(5); continuous batching under a device-memory budget.  Reports per offered rate
    the achieved throughput and p50/p99 latency."""
    from gpusw_amd import Scheme, synth
    from gpusw_amd.serve import Server

    (dqc, dqo), (drc, dro) = synth.dna_16s_cross(seed=0)
    (pqc, pqo), (prc, pro) = synth.protein_topk(seed=0)
    dq = synth.unpack(dqc, dqo)
    pq = synth.unpack(pqc, pqo)
    rates = [float(x) for x in args.rates.split(",")]
    lines = []
    with Server(memory_budget=args.budget << 30, launch_cells=args.launch_cells) as srv:
        d_dna = srv.add_database((drc, dro), Scheme.preset("dna"))
        d_prot = srv.add_database((prc, pro), Scheme.preset("blosum62"))
        for rate in rates:
            rng = np.random.default_rng(int(rate * 1000) + 7)
            n = args.requests
            gaps = rng.exponential(1.0 / rate, n)
            sizes = np.floor(np.exp(rng.uniform(np.log(1), np.log(513), n))).astype(int)
            kinds = rng.random(n) < 0.7
            reqs = []
            for i in range(n):
                src = dq if kinds[i] else pq
                pick = rng.integers(0, len(src), sizes[i])
                seqs = [src[j] for j in pick]
                reqs.append((d_dna if kinds[i] else d_prot, synth.pack(seqs)))
            t0 = time.perf_counter()
            ids = []
            next_t = t0
            for i, (db, q) in enumerate(reqs):
                next_t += gaps[i]
                delay = next_t - time.perf_counter()
                if delay > 0:
                    time.sleep(delay)
                ids.append(srv.submit(db, q, k=5))
            offered_s = time.perf_counter() - t0
            res = [srv.result(r) for r in ids]
            wall = time.perf_counter() - t0
            lat = np.array([r.latency_ms for r in res])
            cells = sum(r.cells for r in res)
            lines.append({
                "metric": METRIC, "value": round(cells / wall / 1e9, 2), "unit": "GCUPS", "n_gpus": 1,
                "steps": 1, "warmup": 0, "ms_per_step": round(wall * 1e3, 2), "higher_is_better": True,
                "scaling": "weak", "vs_baseline": None, "dtype": "int16", "data": "synthetic",
                "config": {"workload": "serving mix: Poisson arrivals, 1-512 queries per request (log-uniform), "
                                       "70% DNA vs 16S-shape DB (20,000 refs), 30% protein vs 100,000-ref DB, top_k(5)",
                           "offered_rate_rps": rate, "requests": n,
                           "achieved_rate_rps": round(n / wall, 3), "offered_window_s": round(offered_s, 3),
                           "memory_budget_gib": args.budget, "launch_cells": args.launch_cells},
                "latency_ms": {"p50": round(float(np.percentile(lat, 50)), 2),
                               "p99": round(float(np.percentile(lat, 99)), 2),
                               "mean": round(float(lat.mean()), 2), "max": round(float(lat.max()), 2)},
            })
            log(f"serving rate {rate}: {lines[-1]['value']} GCUPS, p50 {lines[-1]['latency_ms']['p50']} ms, "
                f"p99 {lines[-1]['latency_ms']['p99']} ms, achieved {n / wall:.2f} req/s")
    return lines


def calibrate_copy(nbytes: int) -> dict:
    """One device-to-device copy of ``nbytes`` (the reference point that scales the
    FETCH_SIZE / WRITE_SIZE counters of the profiled kernels)."""
    from gpusw_amd import _hip

    src = _hip.DeviceBuffer(nbytes)
    dst = _hip.DeviceBuffer(nbytes)
    _hip.call("gsw_memset", src.ptr, 1, nbytes)
    _hip.call("gsw_synchronize")
    t0 = time.perf_counter()
    _hip.call("gsw_memcpy_d2d", dst.ptr, src.ptr, nbytes)
    _hip.call("gsw_synchronize")
    dt = time.perf_counter() - t0
    return {"calibrate_copy_bytes": nbytes, "seconds": dt, "GBps": nbytes / dt / 1e9}


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="dna_cross",
                    choices=["dna_cross", "plain_dna", "protein_topk", "read_pairs", "serving"])
    ap.add_argument("--rates", default="2,5,50,100,200,500,1000,2000",
                    help="serving: offered request rates (req/s); 50-2,000 is the configured sweep, "
                         "2 and 5 show the latency below saturation")
    ap.add_argument("--requests", type=int, default=40, help="serving: requests per rate")
    ap.add_argument("--budget", type=int, default=64, help="serving: device memory budget (GiB)")
    ap.add_argument("--launch-cells", type=float, default=2e11, help="serving: cells per launch")
    ap.add_argument("--plain-queries", type=int, default=32)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="BYTES",
                    help="only run one device copy of BYTES (profiler counter calibration)")
    args = ap.parse_args()
    if args.calibrate_copy:
        print(json.dumps(calibrate_copy(args.calibrate_copy)), flush=True)
        return 0
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU only: this build has no multi-GPU path "
              "(run with --gpus 1)", file=sys.stderr)
        return 2
    if args.workload == "dna_cross":
        lines = [run_dna_cross(args)]
    elif args.workload == "plain_dna":
        lines = [run_plain_dna(args)]
    elif args.workload == "protein_topk":
        lines = [run_protein_topk(args)]
    elif args.workload == "read_pairs":
        lines = run_read_pairs(args)
    else:
        lines = run_serving(args)
    for line in lines:
        print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
