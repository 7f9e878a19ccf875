"""Measured tile autotuner for the batched wavefront kernels.

The library is compiled for a set of tiles: G lanes per reference (lanes-per-pair) and R
query columns per lane (query tile height W = G*R); some tiles also run wider queries as
stripes of W columns.  Its built-in cost model picks a tile (or stripes) per query length
from padding, per-step overhead and pipeline fill; this module replaces the model with
measurements.  For every query-length bin it times each tile
that covers the bin on a synthetic workload shaped like the caller's database (reference
lengths drawn from the real ones), for the packed element type the envelope picks (half2
or int16x2, dtype "auto"; rules marked ``packed``) and for int32, keeps the fastest, and
installs the rules with ``gsw_set_tile_table``.  Results are
cached as JSON keyed by (mode, element type, alphabet size, bin, device).
"""
from __future__ import annotations

import json
import os
from pathlib import Path

import numpy as np

from . import _hip
from .aligner import Aligner, _as_scheme

__all__ = ["tune", "install", "clear", "load", "default_cache"]


def default_cache() -> Path:
    return Path(os.environ.get("GSW_TILE_CACHE", Path.home() / ".cache" / "gpusw_amd" / "tiles.json"))


def _device_key() -> str:
    return _hip.device_name(_hip.get_device())


def install(rules: list[dict]) -> None:
    """Install rules ``{"lo", "hi", "packed", "local", "G", "R"}`` in the library."""
    flat = np.array([[r["lo"], r["hi"], int(r["packed"]), int(r["local"]), r["G"], r["R"]]
                     for r in rules], dtype=np.int32).reshape(-1)
    _hip.call("gsw_set_tile_table", _hip.ptr(flat) if flat.size else None, len(rules))


def clear() -> None:
    _hip.call("gsw_set_tile_table", None, 0)


def _time_tile(al: Aligner, G: int, R: int, reps: int) -> float:
    """Best kernel time of the bin's queries on tile (G, R); inf when the tile cannot run
    them (its profile does not fit LDS and the comparator would score them)."""
    _hip.force_variant(G, R)
    try:
        best = float("inf")
        for _ in range(reps):
            al.score_cross()
            if al._b.last_stats.plain_pairs:
                return float("inf")
            best = min(best, al._b.last_stats.kernel_ms)
        return best
    finally:
        _hip.force_variant(0, 0)


def tune(refs, scheme="dna", bins=None, n_queries: int = 128, n_refs: int = 4000,
         reps: int = 2, seed: int = 0, cache: Path | None = None, dtypes=("int16", "int32")) -> list[dict]:
    """Measure every covering tile per query-length bin and install the fastest.

    ``refs`` is the database (any funnel input); only its length distribution is used.
    ``bins`` is a list of inclusive ``(lo, hi)`` query-length ranges (default: 32-wide
    bins up to the widest tile).  Returns (and installs, and caches) the rules.
    """
    from .encode import funnel

    sch = _as_scheme(scheme)
    renc = funnel(refs, sch, id_prefix="r")
    rlen = renc.lengths[renc.lengths > 0]
    tiles = _hip.list_variants()
    stripe_tiles = set(_hip.list_stripe_variants())
    wmax = max(g * r for g, r in tiles)
    if bins is None:
        bins = [(lo, min(lo + 31, wmax)) for lo in range(1, wmax + 1, 32)]
    rng = np.random.default_rng(seed)
    k = len(sch.eff_alphabet)
    ref_l = rng.choice(rlen, size=n_refs)
    refs_syn = [rng.integers(0, k, int(L)).astype(np.uint8) for L in ref_l]
    from .synth import pack

    rules = []
    local = sch.mode == "local"
    for lo, hi in bins:
        qlens = rng.integers(lo, hi + 1, n_queries)
        queries = pack([rng.integers(0, k, int(L)).astype(np.uint8) for L in qlens])
        for dtype in dtypes:
            s2 = sch.replace(dtype="int32") if dtype == "int32" else sch.replace(dtype="auto")
            al = Aligner(s2, backend="hip").index(pack(refs_syn)).set_queries(queries)
            timings = []
            for G, R in tiles:
                if G * R < hi and (G, R) not in stripe_tiles:
                    continue   # neither covers the bin nor runs it as stripes
                ms = _time_tile(al, G, R, reps)
                if ms != float("inf"):
                    timings.append((ms, G, R))
            if not timings:
                continue
            ms, G, R = min(timings)
            # a rule names a tile; one narrower than the bin means stripes of that tile
            rules.append({"lo": int(lo), "hi": int(hi), "packed": dtype == "int16", "local": local,
                          "G": int(G), "R": int(R), "ms": round(ms, 3),
                          "candidates": [[g, r, round(t, 3)] for t, g, r in sorted(timings)]})
    install(rules)
    cache = Path(cache) if cache else default_cache()
    cache.parent.mkdir(parents=True, exist_ok=True)
    table = json.loads(cache.read_text()) if cache.exists() else {}
    key = f"{sch.mode}|ns={sch.ns}|{_device_key()}"
    table[key] = rules
    cache.write_text(json.dumps(table, indent=1))
    return rules


def load(scheme="dna", cache: Path | None = None) -> list[dict]:
    """Install the cached rules for ``scheme`` on this device (empty list if none)."""
    sch = _as_scheme(scheme)
    cache = Path(cache) if cache else default_cache()
    if not cache.exists():
        return []
    rules = json.loads(cache.read_text()).get(f"{sch.mode}|ns={sch.ns}|{_device_key()}", [])
    if rules:
        install(rules)
    return rules
