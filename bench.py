#!/usr/bin/env python
"""Benchmark: bit-exact GCUPS of the batched HIP scoring path on one MI355X.

Default workload (one JSON line on stdout): the 16S-shape DNA cross -- 2,000 queries of
240-260 nt against 20,000 references of ~1,450 nt (synthetic, seed 0), local ``dna``
scheme, int16 with int32 fallback.  One step = the full (2000, 20000) score matrix
computed into HBM plus a device top_k(5) of every row.  Inputs are resident in HBM
before the timed region.  Also reported:

* ``roofline``: packed 16-bit VALU roofline of the cross kernels -- the VALU lane-slots the
  recurrence needs per cell on the tiles' element type (half2: 7.5 in local mode = 4 adds,
  3 three-input maxima and half a maximum for the running best, 5 otherwise in the
  diagonal-offset form, whose plain-form 7 is ``plain_form_frac``; int16x2: 9 and 6), times
  cells, over the device time of the call's batched launches, which overlap on
  two streams (HIP events from the first launch's start to the last one's end), against
  the gfx950 peak of 78.6 T lane-slots/s; ``recurrence_ops_rate`` counts add/sub/max
  operations instead (local 9, the clamp max(H, 0) folded; 10 with the clamp counted, in
  ``clamp_counted_ops_per_cell``), which a three-input max retires two at a time; the largest
  launch is reported beside it;
* ``plain``: the one-thread-per-pair comparator on the same GPU (a query subsample);
* ``cpu_baseline``: the oracle's C restatement on the host cores over a bounded sample
  of the same pairs (those pairs are also checked bit-exact against the GPU scores), and
  ``cpu_baseline.reference_python``: the reference's own pure-Python algorithm
  (``oracle.gotoh.score_codes``, a restatement of ``src/gpusw/reference.py:38-125``) on
  every host core through ``multiprocessing``, on a smaller sample;
* ``protein_topk``: the BLOSUM62 top-5 of 1,000 queries against 100,000 references --
  GCUPS, the call's roofline, the int16 plain comparator's speedup, oracle-checked hits;
* ``serving``: Poisson arrivals at one sub-saturation rate (>= 400 requests of 1-512
  queries, 70 % DNA / 30 % protein against both resident databases) through the
  continuous-batching server: p50/p99 latency, GCUPS, and every served top-k checked
  against the batched top-k of the same queries, plus an oracle sample of served hits;
* ``int16_tiles``: the headline step again with half2 off (``GSW_HALF=0``), every packed
  pair on the int16x2 tiles: its GCUPS, roofline and equality with the headline scores;
* ``read_pairs``: the 1,000,000 explicit read/window pairs of ``--workload read_pairs``,
  semiglobal and global, each with its value, per-batch roofline, plain speedup and
  oracle sample.

``--workload`` selects another configuration for extra lines: ``plain_dna`` (the 64 x 256
int32 case, oracle-checked in full), ``protein_topk`` (BLOSUM62
top-5 of 1,000 queries against 100,000 references), ``read_pairs`` (1,000,000 explicit
read/window pairs, semiglobal then global, graph-replayed batches; a step is one library
call with the scores left in HBM, the Python call that returns them is ``host_call``),
``serving`` (Poisson request mix through the continuous-batching server, p50/p99 latency).
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
PEAK_SLOTS = 78.6   # T packed-16-bit VALU lane-slots/s: 256 CU x 64 lanes x 2 halves x 2.4 GHz
# VALU lane-slots per cell update, by tile element kind and local mode.  half2 (binary16
# pairs): 4 adds + E, H, F each one v_pk_maximum3_f16 (local: clamps at 0 as third operand)
# + the running best over two columns per three-input max (local) = 7.5.  Global/semiglobal
# packed tiles run the diagonal-offset form (values carried as X + (i+j)*D, D = -gap_extend:
# extensions cost nothing), 2 adds + E, F and H maxima = 5.  int16x2: local 4 add/sub + 5
# max = 9 (E, F and H-gap kept >= 0 by saturating subtraction, so max(H, 0) costs nothing),
# offset form 2 adds + 4 max = 6.  int32 (plain form, one alignment per lane): twice 9 and 8.
SLOTS_PER_CELL = {("half2", True): 7.5, ("half2", False): 5.0, ("int16x2", True): 9.0,
                  ("int16x2", False): 6.0, ("int32", True): 18.0, ("int32", False): 16.0}
# the plain (not offset) form's global/semiglobal slots, reported beside the roofline
PLAIN_FORM_SLOTS = {"half2": 7.0, "int16x2": 8.0, "int32": 16.0}
# add/sub/max operations of the recurrence per cell (local: clamp folded; otherwise the
# offset form), and the plain Gotoh count (local: max(H, 0) as an operation of its own)
RECURRENCE_OPS = {True: 9, False: 6}
CLAMP_COUNTED_OPS = {True: 10, False: 8}
ROOF_UNIT = "T VALU lane-slots/s (packed 16-bit)"
# arithmetic type of a tile kind, and what the binary16 one means for the scores
DTYPE_OF = {"half2": "f16 (exact integers)", "int16x2": "int16", "int32": "int32"}
DTYPE_NOTE = ("half2 tiles carry two alignments per lane as binary16 values, which hold every integer "
              "in [-2048, 2048] exactly; a local result of 2048 or more is rescored in int32, and "
              "global/semiglobal pairs run in half2 only where every DP value stays within 2048, so "
              "every score is the exact int32 (checked against the oracle in 'checks')")


def _valu_roof(cells: float, seconds: float, elem: str, local: bool) -> dict:
    """VALU roofline fields of `cells` cell updates in `seconds` on tiles of kind `elem`."""
    slots = SLOTS_PER_CELL[(elem, local)]
    achieved = cells * slots / seconds / 1e12
    return {"bound": "valu", "achieved": round(achieved, 3), "peak": PEAK_SLOTS, "unit": ROOF_UNIT,
            "frac": round(achieved / PEAK_SLOTS, 4), "elem": elem, "slots_per_cell": slots,
            "recurrence_ops_per_cell": RECURRENCE_OPS[local], "clamp_counted_ops_per_cell": CLAMP_COUNTED_OPS[local],
            "recurrence_ops_rate": round(cells * RECURRENCE_OPS[local] / seconds / 1e12, 3),
            **({} if local else {"plain_form_slots_per_cell": PLAIN_FORM_SLOTS[elem],
                                 "plain_form_frac": round(cells * PLAIN_FORM_SLOTS[elem] / seconds / 1e12 / PEAK_SLOTS, 4)})}


def _dominant_kind(history) -> tuple[str, bool]:
    """Element kind and local mode of the most frequent dominant variant of the calls."""
    from gpusw_amd import _hip

    ids = [h.variant for h in history if h.variant >= 0]
    if not ids:
        return "int16x2", True
    _g, _r, elem, local = _hip.variant_info(max(set(ids), key=ids.count))
    return elem, local


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


# pairs scored by the pure-Python leg's worker processes (set before they are forked)
_PY_JOB = None


def _py_score(k: int) -> int:
    from oracle.gotoh import score_codes

    qc, qo, rc, ro, sub, mode, go, ge, fq, fr, qi, rj = _PY_JOB
    q, r = int(qi[k]), int(rj[k])
    return score_codes(qc[qo[q]:qo[q + 1]], rc[ro[r]:ro[r + 1]], sub, mode, go, ge, fq, fr)


def cpu_reference_python(q, r, scheme, rng, target_s: float = 10.0):
    """The reference's own CPU path -- its pure-Python double loop (``src/gpusw/reference.py
    :38-125``, restated as ``oracle.gotoh.score_codes``) -- on every host core: forked
    worker processes over a seeded sample of (query, reference) pairs, sized to about
    ``target_s`` seconds.  Runs before this process touches the GPU; returns the result line
    and the scored pairs, which are checked against the GPU scores afterwards."""
    import multiprocessing as mp

    global _PY_JOB
    (qc, qo), (rc, ro) = q, r
    cores = len(os.sched_getaffinity(0))
    qlen = np.diff(qo).astype(np.int64)
    rlen = np.diff(ro).astype(np.int64)
    n = 4 * cores + int(target_s * cores / 0.25)      # upper bound; the loop stops at target_s
    qi = rng.integers(0, qlen.size, n).astype(np.int32)
    rj = rng.integers(0, rlen.size, n).astype(np.int32)
    _PY_JOB = (qc, qo, rc, ro, scheme.substitution_table(), scheme.mode, scheme.gap_open, scheme.gap_extend,
               bool(scheme.free_end_gaps_query), bool(scheme.free_end_gaps_ref), qi, rj)
    done, scores = 0, []
    with mp.get_context("fork").Pool(cores) as pool:
        t0 = time.perf_counter()
        while done < n and time.perf_counter() - t0 < target_s:
            chunk = min(n - done, 2 * cores)
            scores += pool.map(_py_score, range(done, done + chunk), chunksize=1)
            done += chunk
        dt = time.perf_counter() - t0
    cells = float((qlen[qi[:done]] * rlen[rj[:done]]).sum())
    line = {"value": round(cells / dt / 1e9, 6), "unit": "GCUPS", "cores": cores, "kind": "reference",
            "sample": f"{done} random (query, ref) pairs of the same workload, {cells:.3e} cells, {dt:.1f} s, "
                      f"pure-Python Gotoh loop in {cores} worker processes"}
    return line, (qi[:done], rj[:done], np.array(scores, dtype=np.int64))


def _tile_name(variant: int) -> str | None:
    from gpusw_amd import _hip

    if variant < 0:
        return None
    g, r, kind, local = _hip.variant_info(variant)
    return f"G={g} lanes x R={r} columns ({kind}/{'local' if local else 'nonlocal'})"


# element type of a variant kind as rocprofv3 prints it in a kernel name
_ELEM_NAME = {"int16x2": "short __vector(2)", "half2": "_Float16 __vector(2)", "int32": "int"}


def _kernel_name(variant: int, striped: int) -> str | None:
    """The kernel's name as rocprofv3 prints it, for a variant id of the batched path."""
    from gpusw_amd import _hip

    if variant < 0:
        return None
    g, r, elem, local = _hip.variant_info(variant)
    kind, waves, extra = {0: ("cross", 8, ""), 1: ("stripe", 8, ""), 2: ("pairs", 8, ""), 3: ("pairs_tab", 4, ""),
                          4: ("pairs_tab", 4, ", true")}[int(striped)]
    return f"void gsw::{kind}_kernel<{_ELEM_NAME[elem]}, {g}, {r}, {str(local).lower()}, {waves}{extra}>(gsw::KArgs)"


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


def _roofline(history, mode: str, workload: str) -> dict:
    """Roofline of the dominant kernel: the batched launch with the most cells in each timed
    call (its cells x ops over its device time from HIP events on the stream it ran on)."""
    top = [h for h in history if h.top_variant >= 0 and h.top_ms > 0]
    if not top:
        return {"bound": "valu", "achieved": None, "peak": PEAK_SLOTS, "unit": ROOF_UNIT,
                "frac": None, "traffic": None}
    # the most frequent (variant, striped) among the calls' dominant launches (most cells)
    keys = [(h.top_variant, h.top_striped) for h in top]
    var, striped = max(set(keys), key=keys.count)
    sel = [h for h in top if (h.top_variant, h.top_striped) == (var, striped)]
    ms = float(np.mean([h.top_ms for h in sel]))
    cells = float(np.mean([h.top_cells for h in sel]))
    from gpusw_amd import _hip

    _g, _r, elem, local = _hip.variant_info(var)
    kernel = _kernel_name(var, striped)
    traffic, source = _measured_traffic(workload, kernel)
    out = _valu_roof(cells, ms * 1e-3, elem, local)
    out.update({"traffic": traffic, "kernel": kernel, "tile": _tile_name(var), "launch_ms": round(ms, 3),
                "launch_cells": int(cells), "launches_timed": len(sel)})
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


def _call_roofline(history, cells: int, kernel_s: float, steps: int, workload: str = "protein",
                   label: str = "all batched launches of a top_k call (cross_kernel + stripe_kernel tiles, two streams)"
                   ) -> dict:
    """Roofline of a cross-product call as a whole.  Its tile launches run on two streams
    and overlap (each launch's workgroups take CUs as the other stream's finish), so no
    single launch owns the GPU for its interval: the dominant unit is the call's set of
    batched launches, timed from the first launch's start to the last one's end (HIP events
    on the library stream, joined with the second stream).  cells and kernel_s are per step."""
    elem, local = _dominant_kind(history)
    calls = max(1, len(history) // max(steps, 1))
    traffic, src = _call_traffic(workload)
    out = _valu_roof(cells, kernel_s, elem, local)
    largest = _roofline(history, "local", workload)
    if largest.get("launch_ms") is not None:
        largest["note"] = ("launch_ms is the launch's own interval on its stream (it agrees with the "
                           "rocprofv3 kernel-trace average), which it shares with the other stream's "
                           "launches, so its frac is below the call's")
    out.update({"traffic": None if traffic is None else round(traffic * calls, 1), "kernel": label,
                "kernels_ms_per_step": round(kernel_s * 1e3, 2), "calls_per_step": calls,
                "largest_launch": largest})
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
    rng = np.random.default_rng(1)
    py_line = py_pairs = None
    if not args.no_cpu:   # before any GPU work: the worker processes are forked
        py_line, py_pairs = cpu_reference_python((qc, qo), (rc, ro), scheme, np.random.default_rng(11))
        log(f"cpu reference (pure Python): {py_line['value']} GCUPS on {py_line['cores']} processes")
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
    roof = _call_roofline(history, cells, kavg, args.steps, "dna16s",
                          "all batched launches of a cross call (cross_kernel tiles by query length, two streams)")
    roof["launches_per_step"] = int(stats.launches)

    # ---- correctness of the timed result (outside the timed region)
    host = scores.to_host()
    order = np.argsort(-host, axis=1, kind="stable")[:, :5]
    topk_ok = bool(np.array_equal(order, idx) and np.array_equal(np.take_along_axis(host, order, 1), sc))

    # ---- plain comparator on a query subsample
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

    # ---- the same step on the int16x2 tiles alone (half2 off through GSW_HALF, which the
    #      library reads per call), outside the headline: its rate, roofline and equality
    int16 = None
    if not args.no_int16:
        prev = os.environ.get("GSW_HALF")
        os.environ["GSW_HALF"] = "0"
        try:
            step()
            gpusw_amd._hip.call("gsw_synchronize")
            n16 = max(1, min(args.steps, 2))
            k16 = []
            t1 = time.perf_counter()
            for _ in range(n16):
                s16, dt16, _i16, _c16 = step()
                k16.append(dt16)
            gpusw_amd._hip.call("gsw_synchronize")
            w16 = (time.perf_counter() - t1) / n16
            elem16, _loc = _dominant_kind([b.last_stats])
        finally:
            if prev is None:
                os.environ.pop("GSW_HALF", None)
            else:
                os.environ["GSW_HALF"] = prev
        int16 = {"value": round(cells / w16 / 1e9, 2), "unit": "GCUPS", "ms_per_step": round(w16 * 1e3, 2),
                 "steps": n16, "dtype": DTYPE_OF[elem16],
                 "roofline": _valu_roof(cells, float(np.mean(k16)), elem16, True),
                 "scores_equal_half2": bool(np.array_equal(s16.to_host(), host)),
                 "what": "the headline step with GSW_HALF=0: every packed pair on the int16x2 tiles"}
        log(f"int16 tiles: {int16['value']} GCUPS, scores equal: {int16['scores_equal_half2']}")

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
        pqi, prj, pwant = py_pairs
        py_line["mismatches_vs_gpu"] = int((host[pqi, prj].astype(np.int64) != pwant).sum())
        cpu["reference_python"] = py_line

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
        "dtype": DTYPE_OF[roof["elem"]],
        "dtype_note": DTYPE_NOTE,
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
        "int16_tiles": int16,
        "cpu_baseline": cpu,
    }
    if not args.no_protein or not args.no_serving:
        prot, prot_pool = protein_leg(args)
        if not args.no_protein:
            line["protein_topk"] = prot
        if not args.no_serving:
            line["serving"] = serving_leg(args, ((qc, qo), (rc, ro), idx, sc), prot_pool)
    if not args.no_read_pairs:
        line["read_pairs"] = read_pairs_leg(args)
    return line


def read_pairs_leg(args) -> dict:
    """The default run's read-pairs key: the 1,000,000-pair workload of `--workload
    read_pairs` in both modes, each with its value, per-batch roofline, plain speedup and
    oracle check (the full lines come from that workload)."""
    out = {}
    for ln in run_read_pairs(args):
        mode = "semiglobal" if "semiglobal" in ln["config"]["workload"] else "global"
        r = ln["roofline"]
        out[mode] = {"value": ln["value"], "unit": "GCUPS", "ms_per_step": ln["ms_per_step"],
                     "steps": ln["steps"], "host_call": ln["host_call"], "dtype": ln["dtype"],
                     "config": ln["config"],
                     "roofline": {k: r.get(k) for k in ("bound", "achieved", "peak", "unit", "frac", "elem",
                                                        "slots_per_cell", "plain_form_slots_per_cell",
                                                        "plain_form_frac", "traffic", "kernel", "launch_ms",
                                                        "launch_cells", "all_batches_ms", "launches")},
                     "plain": ln["plain"], "checks": ln["checks"]}
    return out


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


def _plain_subsample_cross(al, scheme, qsel, dtype):
    """GCUPS of the one-thread-per-pair kernel on a query subsample (same GPU)."""
    import gpusw_amd

    pb = gpusw_amd.backends.resolve_backend("hip-plain")
    nr = al._refs.n
    sc, dt, _ = pb.run_cross(scheme, al._queries.max_len, dtype, 128, al._d_qbuf, al._d_qoff,
                             al._d_rbuf, al._d_roff, qsel, len(qsel), nr)
    cells = int(al._queries.lengths[qsel].astype(np.int64).sum()) * int(al._refs.lengths.astype(np.int64).sum())
    return sc.to_host(), cells / dt / 1e9


def protein_leg(args, steps: int | None = None, warmup: int | None = None):
    """Protein top-k: 1,000 lognormal queries vs 100,000 lognormal references, BLOSUM62 local
    -11/-1, device top_k(5) per query batch of 512 (the Aligner.top_k loop).  Returns the
    result dict and the workload with its top-5 (the serving leg's expected results)."""
    import gpusw_amd
    from gpusw_amd import Aligner, Scheme, synth
    from oracle import native

    steps = args.protein_steps if steps is None else steps
    warmup = 1 if warmup is None else warmup
    (qc, qo), (rc, ro) = synth.protein_topk(seed=0)
    scheme = Scheme.preset("blosum62")
    al = Aligner(scheme, backend="hip").index((rc, ro)).set_queries((qc, qo))
    nq, nr = al._queries.n, al._refs.n
    cells = int(al._queries.lengths.astype(np.int64).sum()) * int(al._refs.lengths.astype(np.int64).sum())
    log(f"protein top-k: {nq} x {nr}, {cells:.3e} cells")
    qb = args.protein_query_batch
    for _ in range(warmup):
        al.top_k(k=5, query_batch=qb)
    gpusw_amd._hip.call("gsw_synchronize")
    al._b.history = []          # the backend appends each call's stats
    t0 = time.perf_counter()
    for _ in range(steps):
        res = al.top_k(k=5, query_batch=qb)
    wall = (time.perf_counter() - t0) / steps
    history, al._b.history = al._b.history, None
    kernel_s = sum(h.kernel_ms for h in history) / 1e3 / steps
    value = cells / wall / 1e9
    log(f"protein top-k: {value:.1f} GCUPS ({wall * 1e3:.1f} ms per call)")
    # checks: top-k of a query subsample against the comparator's full rows (int16 stores,
    # as the reference's envelope resolves this workload), and the oracle on sampled hits
    rng = np.random.default_rng(2)
    dtype = scheme.resolve_dtype(int(al._queries.max_len), int(al._refs.max_len))
    qsel = np.sort(rng.choice(nq, 8, replace=False)).astype(np.int32)
    prow, plain_gcups = _plain_subsample_cross(al, scheme, qsel, dtype)
    order = np.argsort(-prow, axis=1, kind="stable")[:, :5]
    ok = bool(np.array_equal(order, res.topk_idx[qsel]) and
              np.array_equal(np.take_along_axis(prow, order, 1), res.topk_scores[qsel]))
    hq = rng.choice(nq, 200, replace=False)
    hqi = np.repeat(hq, 5).astype(np.int32)
    hrj = res.topk_idx[hq].reshape(-1).astype(np.int32)
    want = native.score_encoded_pairs(al._queries, al._refs, hqi, hrj, scheme)
    hit_mism = int((res.topk_scores[hq].reshape(-1).astype(np.int64) != want).sum())
    roof = _call_roofline(history, cells, kernel_s, steps)
    out = {
        "value": round(value, 2), "unit": "GCUPS", "ms_per_step": round(wall * 1e3, 2), "steps": steps,
        "dtype": DTYPE_OF[roof["elem"]], "dtype_note": DTYPE_NOTE, "rescored": int(history[-1].rescored),
        "warmup": warmup,
        "config": {"workload": "protein top-k: 1000 queries (lognormal median 350, 30-2000) x 100000 "
                               "references (lognormal median 300, 30-3000), BLOSUM62 local -11/-1, top_k(5)",
                   "cells": cells, "query_batch": qb or "Aligner default (2 GiB score block)"},
        "roofline": roof,
        "plain": {"gcups": round(plain_gcups, 2), "queries": int(len(qsel)), "store_dtype": dtype,
                  "speedup": round(value / plain_gcups, 2)},
        "checks": {"topk_equals_comparator_rows": ok, "oracle_hits_checked": int(hqi.size),
                   "oracle_hit_mismatches": hit_mism},
    }
    return out, ((qc, qo), (rc, ro), res.topk_idx, res.topk_scores)


def _serve(srv, dbs, pools, rate: float, n: int, seed: int):
    """Submit n Poisson-timed requests (1-512 queries, log-uniform; 70 % DNA, 30 % protein)
    and wait for all of them.  Returns per request (kind, query indices, ServedResult) and
    the submission window."""
    from gpusw_amd import synth

    rng = np.random.default_rng(seed)
    gaps = rng.exponential(1.0 / rate, n)
    sizes = np.floor(np.exp(rng.uniform(np.log(1), np.log(513), n))).astype(int)
    kinds = (rng.random(n) >= 0.7).astype(int)            # 0 DNA, 1 protein
    seqs = [synth.unpack(*pools[0][0]), synth.unpack(*pools[1][0])]
    reqs = []
    for i in range(n):
        pick = rng.integers(0, len(seqs[kinds[i]]), sizes[i])
        reqs.append((kinds[i], pick, synth.pack([seqs[kinds[i]][j] for j in pick])))
    t0 = time.perf_counter()
    ids = []
    next_t = t0
    for i, (kind, _pick, q) in enumerate(reqs):
        next_t += gaps[i]
        delay = next_t - time.perf_counter()
        if delay > 0:
            time.sleep(delay)
        ids.append(srv.submit(dbs[kind], q, k=5))
    offered_s = time.perf_counter() - t0
    res = [srv.result(r) for r in ids]
    wall = time.perf_counter() - t0
    return [(k, p, r) for (k, p, _q), r in zip(reqs, res)], offered_s, wall


def _check_served(served, pools, oracle_sample: int, seed: int):
    """Every served top-k row against the batched top-k of the same query (Aligner.top_k
    of the whole pool), and the oracle on a sample of served hits."""
    from gpusw_amd import Scheme
    from oracle import native

    bad_rows = 0
    hits = [[], []]
    for kind, pick, r in served:
        want_idx, want_sc = pools[kind][2][pick], pools[kind][3][pick]
        got_idx, got_sc = r.result.topk_idx, r.result.topk_scores
        bad_rows += int((~((got_idx == want_idx).all(1) & (got_sc == want_sc).all(1))).sum())
        for a in range(len(pick)):
            hits[kind].append((int(pick[a]), int(got_idx[a, 0]), int(got_sc[a, 0])))
    rng = np.random.default_rng(seed)
    mism, checked = 0, 0
    for kind, scheme in ((0, Scheme.preset("dna")), (1, Scheme.preset("blosum62"))):
        if not hits[kind]:
            continue
        h = np.array(hits[kind], dtype=np.int64)
        sel = h[rng.choice(len(h), min(oracle_sample, len(h)), replace=False)]
        (qc, qo), (rc, ro) = pools[kind][0], pools[kind][1]
        want = native.score_pairs(qc, qo, rc, ro, sel[:, 0].astype(np.int32), sel[:, 1].astype(np.int32),
                                  scheme.substitution_table(), scheme.mode, scheme.gap_open, scheme.gap_extend,
                                  scheme.free_end_gaps_query, scheme.free_end_gaps_ref)
        mism += int((want != sel[:, 2]).sum())
        checked += len(sel)
    return bad_rows, mism, checked


def _serving_line(served, offered_s, wall, rate, n):
    lat = np.array([r.latency_ms for _k, _p, r in served])
    cells = sum(r.cells for _k, _p, r in served)
    return {"offered_rate_rps": rate, "requests": n, "achieved_rate_rps": round(n / wall, 3),
            "offered_window_s": round(offered_s, 3), "gcups": round(cells / wall / 1e9, 2),
            "latency_ms": {"p50": round(float(np.percentile(lat, 50)), 2),
                           "p99": round(float(np.percentile(lat, 99)), 2),
                           "mean": round(float(lat.mean()), 2), "max": round(float(lat.max()), 2)}}


def serving_leg(args, dna_pool, prot_pool):
    """The serving mix at one sub-saturation rate, every served result checked."""
    from gpusw_amd import Scheme
    from gpusw_amd.serve import Server

    pools = (dna_pool, prot_pool)
    with Server(memory_budget=args.budget << 30, launch_cells=args.launch_cells) as srv:
        dbs = (srv.add_database(dna_pool[1], Scheme.preset("dna")),
               srv.add_database(prot_pool[1], Scheme.preset("blosum62")))
        served, offered_s, wall = _serve(srv, dbs, pools, args.serve_rate, args.serve_requests, seed=7007)
    line = _serving_line(served, offered_s, wall, args.serve_rate, args.serve_requests)
    bad, mism, checked = _check_served(served, pools, 400, seed=8)
    line["config"] = {"workload": "serving mix: Poisson arrivals, 1-512 queries per request (log-uniform), 70% DNA "
                                  "vs the 16S-shape DB (20,000 refs), 30% protein vs the 100,000-ref DB, top_k(5)",
                      "memory_budget_gib": args.budget, "launch_cells": args.launch_cells}
    line["checks"] = {"served_rows_checked": int(sum(len(p) for _k, p, _r in served)),
                      "rows_differing_from_batched_topk": bad, "oracle_hits_checked": checked,
                      "oracle_hit_mismatches": mism}
    log(f"serving at {args.serve_rate} req/s: {line['gcups']} GCUPS, p50 {line['latency_ms']['p50']} ms, "
        f"p99 {line['latency_ms']['p99']} ms, achieved {line['achieved_rate_rps']} req/s, "
        f"{bad} rows differ, {mism} oracle mismatches")
    return line


def run_protein_topk(args):
    prot, _pool = protein_leg(args, steps=args.steps, warmup=args.warmup)
    line = {"metric": METRIC, "value": prot["value"], "unit": "GCUPS", "n_gpus": 1, "steps": args.steps,
            "warmup": args.warmup, "ms_per_step": prot["ms_per_step"], "higher_is_better": True,
            "scaling": "weak", "vs_baseline": None, "data": "synthetic"}
    line.update({k: prot[k] for k in ("dtype", "dtype_note", "config", "roofline", "plain", "checks")})
    return line


def run_read_pairs(args):
    """1,000,000 explicit (read, window) pairs, dna scheme, semiglobal then global, int16."""
    import gpusw_amd
    from gpusw_amd import Aligner, Scheme, synth
    from oracle import native

    (rc, ro), (wc, wo), qi, rj = synth.read_pairs(seed=0)
    # both modes are timed before any check runs (the CPU oracle's threads and the comparator
    # stay out of the timed calls)
    timed = []
    # a call is ~6-10 ms: at least three untimed calls per mode (the first calls of a new
    # aligner still settle host buffers and planning threads) and at least 0.25 s of them
    # (a kernel trace showed a mode's first call with 1.5x longer batches than its later ones)
    warmup = max(args.warmup, 3)
    for mode in ("semiglobal", "global"):
        scheme = Scheme.preset("dna").replace(mode=mode, dtype="int16")
        al = Aligner(scheme, backend="hip").index((wc, wo)).set_queries((rc, ro))
        t_w, n_w = time.perf_counter(), 0
        while n_w < warmup or time.perf_counter() - t_w < 0.25:
            al.score_pairs(qi, rj)
            n_w += 1
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
        # value: the library call with the reference sequences resident and the scores left in
        # HBM (its wall time includes planning the batches on the host and copying their task
        # lists, overlapped with the kernels); the Python call that also returns the 4 MB of
        # scores over PCIe is reported beside it
        lib_s = float(np.mean([h.wall_ms for h in history])) / 1e3
        value = cells / lib_s / 1e9
        roof = dict(_roofline(history, mode, "readpairs"), all_batches_ms=round(st.kernel_ms, 2),
                    host_wall_ms=round(st.wall_ms, 2), launches=int(st.launches))
        out.append({
            "metric": METRIC, "value": round(value, 2), "unit": "GCUPS", "n_gpus": 1, "steps": args.steps,
            "warmup": warmup, "ms_per_step": round(lib_s * 1e3, 2), "higher_is_better": True,
            "host_call": {"gcups": round(cells / wall / 1e9, 2), "ms": round(wall * 1e3, 2),
                          "what": "Aligner.score_pairs in Python, scores returned to the host over PCIe"},
            "scaling": "weak", "vs_baseline": None, "dtype": DTYPE_OF.get(roof.get("elem"), "int16"),
            "dtype_note": DTYPE_NOTE, "data": "synthetic",
            "config": {"workload": f"read pairs: 1,000,000 (150-nt read, 200-400-nt window) pairs, dna "
                                   f"{mode}, int16 envelope, HIP-graph batches of 131,072; one step = one "
                                   f"gsw_score_pairs call, scores left in HBM", "cells": cells},
            "roofline": roof,
            "plain": {"gcups": round(pcells / pdt / 1e9, 2), "pairs": int(sub.size),
                      "speedup": round(value / (pcells / pdt / 1e9), 2),
                      "equal_to_batched": bool(np.array_equal(ph, got[sub]))},
            "checks": {"oracle_mismatches": mism, "oracle_sample": int(pick.size)},
        })
    return out


def run_serving(args):
    """The serving sweep: one line per offered rate (every served result checked)."""
    from gpusw_amd import Aligner, Scheme, synth
    from gpusw_amd.serve import Server

    (dqc, dqo), (drc, dro) = synth.dna_16s_cross(seed=0)
    dna = Aligner(Scheme.preset("dna"), backend="hip").index((drc, dro)).top_k((dqc, dqo), k=5)
    _prot, prot_pool = protein_leg(args, steps=1, warmup=0)
    pools = (((dqc, dqo), (drc, dro), dna.topk_idx, dna.topk_scores), prot_pool)
    rates = [float(x) for x in args.rates.split(",")]
    lines = []
    with Server(memory_budget=args.budget << 30, launch_cells=args.launch_cells) as srv:
        dbs = (srv.add_database(pools[0][1], Scheme.preset("dna")),
               srv.add_database(pools[1][1], Scheme.preset("blosum62")))
        for rate in rates:
            served, offered_s, wall = _serve(srv, dbs, pools, rate, args.requests, seed=int(rate * 1000) + 7)
            s = _serving_line(served, offered_s, wall, rate, args.requests)
            bad, mism, checked = _check_served(served, pools, 100, seed=int(rate))
            lines.append({
                "metric": METRIC, "value": s["gcups"], "unit": "GCUPS", "n_gpus": 1, "steps": 1, "warmup": 0,
                "ms_per_step": round(wall * 1e3, 2), "higher_is_better": True, "scaling": "weak",
                "vs_baseline": None, "dtype": "f16 (exact integers) / int16", "dtype_note": DTYPE_NOTE,
                "data": "synthetic",
                "config": {"workload": "serving mix: Poisson arrivals, 1-512 queries per request (log-uniform), "
                                       "70% DNA vs 16S-shape DB (20,000 refs), 30% protein vs 100,000-ref DB, top_k(5)",
                           "offered_rate_rps": rate, "requests": args.requests,
                           "achieved_rate_rps": s["achieved_rate_rps"], "offered_window_s": s["offered_window_s"],
                           "memory_budget_gib": args.budget, "launch_cells": args.launch_cells},
                "latency_ms": s["latency_ms"],
                "checks": {"rows_differing_from_batched_topk": bad, "oracle_hits_checked": checked,
                           "oracle_hit_mismatches": mism},
            })
            log(f"serving rate {rate}: {s['gcups']} GCUPS, p50 {s['latency_ms']['p50']} ms, "
                f"p99 {s['latency_ms']['p99']} ms, achieved {s['achieved_rate_rps']} req/s, {bad} rows differ")
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
    ap.add_argument("--requests", type=int, default=40, help="serving sweep: requests per rate")
    ap.add_argument("--serve-rate", type=float, default=5.0,
                    help="default run's serving leg: offered rate (req/s), below the ~8 req/s one GPU completes")
    ap.add_argument("--serve-requests", type=int, default=400, help="default run's serving leg: requests")
    ap.add_argument("--protein-steps", type=int, default=3, help="default run's protein leg: timed calls")
    ap.add_argument("--protein-query-batch", type=int, default=None,
                    help="protein top-k: queries per device batch (Aligner.top_k query_batch; default: the "
                         "Aligner's, a 2 GiB score block, here all 1,000; the reference uses 512)")
    ap.add_argument("--no-protein", action="store_true", help="default run: skip the protein top-k leg")
    ap.add_argument("--no-serving", action="store_true", help="default run: skip the serving leg")
    ap.add_argument("--no-read-pairs", action="store_true", help="default run: skip the read-pairs leg")
    ap.add_argument("--no-int16", action="store_true", help="default run: skip the int16-tiles leg")
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
