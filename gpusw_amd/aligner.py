"""The scoring engine: encode and upload references once, then score query sets
against them -- cross products, explicit pairs, or per-query top-k.

Methods, arguments and results follow the reference's ``src/gpusw/aligner.py:31-242``
(``index`` ``:69-77``, ``set_queries`` ``:79-87``, dtype scope ``:90-93``, GCUPS cells
``:107-109`` and ``:160-162``, ``score_cross`` ``:113-141``, ``score_pairs`` ``:143-167``,
``top_k`` ``:169-195``).  The device is reached only through the backend contract.
"""
from __future__ import annotations

import numpy as np

from .backends import resolve_backend
from .encode import Encoded, funnel
from .result import AlignResult
from .scheme import Scheme

__all__ = ["Aligner"]


def _as_scheme(scheme) -> Scheme:
    if isinstance(scheme, Scheme):
        return scheme
    if isinstance(scheme, str):
        return Scheme.preset(scheme)
    raise TypeError(f"scheme must be a Scheme or preset name, got {type(scheme).__name__}")


class Aligner:
    """Reusable scorer for one scheme.

    ``threads`` is the block size of the one-thread-per-pair comparator (``hip-plain``);
    ``device`` is the GPU ordinal; ``backend`` is ``"auto"``, ``"hip"`` or ``"hip-plain"``.
    Constructing an Aligner touches no device.
    """

    def __init__(self, scheme="dna", *, threads: int = 128, device: int = 0, backend: str = "auto"):
        self.scheme = _as_scheme(scheme)
        self.threads = int(threads)
        self.device = int(device)
        self._backend_name = backend
        self._backend = None
        self._refs: Encoded | None = None
        self._queries: Encoded | None = None
        self._d_rbuf = self._d_roff = None
        self._d_qbuf = self._d_qoff = None
        self._last_gcups = 0.0
        self._last_source = ""
        self._last_cells = 0
        self._last_seconds = 0.0

    @property
    def _b(self):
        if self._backend is None:
            self._backend = resolve_backend(self._backend_name)
        return self._backend

    # ------------------------------------------------------------------ inputs
    def index(self, refs) -> "Aligner":
        """Encode and upload the reference set once (returns ``self``)."""
        b = self._b
        with b.device_scope(self.device):
            enc = funnel(refs, self.scheme, id_prefix="r")
            self._refs = enc
            self._d_rbuf = b.upload(enc.codes)
            self._d_roff = b.upload(enc.offsets)
        return self

    def set_queries(self, queries) -> "Aligner":
        """Encode and upload a query set (returns ``self``)."""
        b = self._b
        with b.device_scope(self.device):
            enc = funnel(queries, self.scheme, id_prefix="q")
            self._queries = enc
            self._d_qbuf = b.upload(enc.codes)
            self._d_qoff = b.upload(enc.offsets)
        return self

    # --------------------------------------------------------------- internals
    def _resolved_dtype(self) -> str:
        maxq = self._queries.max_len if self._queries else 0
        maxr = self._refs.max_len if self._refs else 0
        return self.scheme.resolve_dtype(max(maxq, 1), max(maxr, 1))

    def _record(self, cells: int, seconds: float, source: str) -> None:
        self._last_cells = int(cells)
        self._last_seconds = float(seconds)
        self._last_gcups = cells / seconds / 1e9 if seconds > 0 else 0.0
        self._last_source = source

    def _launch_cross(self, qlist: np.ndarray):
        nr = self._refs.n
        scores, dt, src = self._b.run_cross(
            self.scheme, self._queries.max_len, self._resolved_dtype(), self.threads,
            self._d_qbuf, self._d_qoff, self._d_rbuf, self._d_roff, qlist, len(qlist), nr,
        )
        cells = int(self._queries.lengths[qlist].astype(np.int64).sum()) * int(
            self._refs.lengths.astype(np.int64).sum())
        return scores, dt, src, cells

    def _batches(self, nq: int, query_batch: int | None):
        nr = self._refs.n
        cap_rows = max(1, self._b.max_threads_per_launch // max(nr, 1))
        step = max(1, min(query_batch or nq or 1, cap_rows))
        for b0 in range(0, nq, step):
            yield np.arange(b0, min(b0 + step, nq), dtype=np.int32)

    def _need(self, queries, what: str):
        if self._refs is None:
            raise ValueError(f"call index(refs) before {what}()")
        if queries is not None:
            self.set_queries(queries)
        if self._queries is None:
            raise ValueError("no queries: pass queries= or call set_queries()")

    # ---------------------------------------------------------------- scoring
    def score_cross(self, queries=None, *, return_ids: bool = False, query_batch: int | None = None):
        """``(n_queries, n_refs)`` int32 scores of every query against every reference."""
        self._need(queries, "score_cross")
        nq, nr = self._queries.n, self._refs.n
        out = np.empty((nq, nr), dtype=np.int32)
        cells, secs, src = 0, 0.0, ""
        with self._b.device_scope(self.device):
            for qb in self._batches(nq, query_batch):
                scores, dt, src, c = self._launch_cross(qb)
                out[qb[0]:qb[0] + len(qb)] = self._b.as_host(scores)
                cells += c
                secs += dt
        self._record(cells, secs, src)
        if return_ids:
            return AlignResult(out, self._queries.ids, self._refs.ids, self.scheme)
        return out

    def score_pairs(self, qi, rj, *, return_ids: bool = False):
        """Scores of the explicit index pairs ``(qi[k], rj[k])`` as ``(n_pairs,)`` int32."""
        if self._refs is None or self._queries is None:
            raise ValueError("call index(refs) and set_queries(queries) first")
        qi = np.ascontiguousarray(qi, dtype=np.int32)
        rj = np.ascontiguousarray(rj, dtype=np.int32)
        if qi.shape != rj.shape:
            raise ValueError("qi and rj must have the same length")
        n = int(qi.shape[0])
        with self._b.device_scope(self.device):
            host, dt, src = self._b.run_pairs(
                self.scheme, self._queries.max_len, self._resolved_dtype(), self.threads,
                self._d_qbuf, self._d_qoff, self._d_rbuf, self._d_roff, qi, rj, n,
            )
        # sum of lq*lr (reference aligner.py:160-161); a backend that counted the cells
        # while planning the call hands them over once (last_cells)
        cells = getattr(self._b, "last_cells", None)
        if cells is not None:
            self._b.last_cells = None
        else:
            cells = int((self._queries.lengths[qi].astype(np.int64)
                         * self._refs.lengths[rj].astype(np.int64)).sum())
        self._record(cells, dt, src)
        host = np.asarray(host, dtype=np.int32)
        if return_ids:
            return AlignResult(host, [self._queries.ids[i] for i in qi],
                               [self._refs.ids[j] for j in rj], self.scheme)
        return host

    #: score-block size that sizes top_k's default query batch (device int32 scores)
    TOPK_BLOCK_BYTES = 2 << 30

    def top_k(self, queries=None, k: int = 5, *, query_batch: int | None = None) -> AlignResult:
        """Per-query best ``k`` references (score descending, ties to the lower index).

        Only the top-k of each batch leaves the device.  ``query_batch`` bounds the queries
        scored per device batch (their score block is ``query_batch x n_refs`` int32); the
        default (None) takes as many as keep that block within 2 GiB, a small share of an
        MI355X's 288 GB -- the reference's fixed 512 (``src/gpusw/aligner.py:169``) splits a
        1,000-query call in two.  Results do not depend on it.
        """
        self._need(queries, "top_k")
        nq, nr = self._queries.n, self._refs.n
        if query_batch is None:
            query_batch = max(1, self.TOPK_BLOCK_BYTES // (4 * max(nr, 1)))
        kk = min(k, nr)
        idx_out = np.empty((nq, kk), dtype=np.int32)
        sc_out = np.empty((nq, kk), dtype=np.int32)
        cells, secs, src = 0, 0.0, ""
        with self._b.device_scope(self.device):
            for qb in self._batches(nq, query_batch):
                scores, dt, src, c = self._launch_cross(qb)
                idx, sc = self._b.topk(scores, kk)
                idx_out[qb[0]:qb[0] + len(qb)] = idx
                sc_out[qb[0]:qb[0] + len(qb)] = sc
                cells += c
                secs += dt
        self._record(cells, secs, src)
        return AlignResult(None, self._queries.ids, self._refs.ids, self.scheme,
                           topk_idx=idx_out, topk_scores=sc_out)

    # ------------------------------------------------------------- traceback
    def align(self, qi, rj) -> list:
        """The alignments behind the scores of pairs ``(qi[k], rj[k])`` (GPU traceback,
        int32): a list of :class:`~gpusw_amd.result.Alignment`.  Each score equals the
        pair's ``score_pairs`` score.  No counterpart in the reference, whose results are
        score-only (its examples re-align top hits on the CPU)."""
        from .backends.hip import run_traceback
        from .result import Alignment, _cigar

        if self._refs is None or self._queries is None:
            raise ValueError("call index(refs) and set_queries(queries) first")
        qi = np.ascontiguousarray(qi, dtype=np.int32).reshape(-1)
        rj = np.ascontiguousarray(rj, dtype=np.int32).reshape(-1)
        if qi.shape != rj.shape:
            raise ValueError("qi and rj must have the same length")
        with self._b.device_scope(self.device):
            recs, ops, stats = run_traceback(self.scheme, self._d_qbuf, self._d_qoff, self._d_rbuf,
                                             self._d_roff, qi, rj)
        self._record(int(stats.cells), stats.kernel_ms / 1e3, "traceback (int32 fill, walk per pair)")
        qids, rids = self._queries.ids, self._refs.ids
        col = {f: recs[f].tolist() for f in recs.dtype.names}
        return [Alignment(qids[q], rids[r], sc, qb, qe, rb, re_, _cigar(ops, off, nops), m, x, ins, dl, gop)
                for q, r, sc, qb, qe, rb, re_, off, nops, m, x, ins, dl, gop in zip(
                    qi.tolist(), rj.tolist(), col["score"], col["q_begin"], col["q_end"], col["r_begin"],
                    col["r_end"], col["ops_offset"], col["n_ops"], col["matches"], col["mismatches"],
                    col["ins"], col["del_"], col["gap_opens"])]

    def top_k_alignments(self, queries=None, k: int = 5, *, query_batch: int | None = None) -> list:
        """``top_k`` followed by the traceback of every hit: ``[(query_id, [Alignment,
        ...]), ...]`` best first."""
        res = self.top_k(queries, k, query_batch=query_batch)
        nq, kk = res.topk_idx.shape
        qi = np.repeat(np.arange(nq, dtype=np.int32), kk)
        alns = self.align(qi, res.topk_idx.reshape(-1))
        return [(self._queries.ids[q], alns[q * kk:(q + 1) * kk]) for q in range(nq)]

    # ----------------------------------------------------------- introspection
    @property
    def backend(self) -> str:
        return self._b.name

    @property
    def query_ids(self) -> list:
        return list(self._queries.ids) if self._queries else []

    @property
    def reference_ids(self) -> list:
        return list(self._refs.ids) if self._refs else []

    @property
    def gcups(self) -> float:
        """Giga cell updates per second of the most recent scoring call (device time)."""
        return self._last_gcups

    def kernel_source(self) -> str:
        """Description of the kernels that ran last (or would run for this scheme)."""
        if self._last_source:
            return self._last_source
        maxq = self._queries.max_len if self._queries else 256
        dtype = self._resolved_dtype() if (self._queries and self._refs) else "int32"
        return self._b.render(self.scheme, maxq, dtype)
