"""HIP backends over ``libgsw_hip.so`` (AMD Instinct MI355X, gfx950).

:class:`HipBackend` implements the reference's ``Backend`` contract
(``src/gpusw/backends/__init__.py:31-87``) the way ``CudaBackend`` does
(``src/gpusw/backends/cuda.py:29-112``), but every call goes through the C ABI in
``include/gsw_hip.h``: uploads are ``gsw_malloc`` + copy, ``run_cross``/``run_pairs``
call the batched wavefront kernels (``gsw_score_cross``/``gsw_score_pairs``), ``topk``
is the device top-k (``gsw_topk_rows``).  ``dt`` is device time from HIP events.
:class:`HipPlainBackend` routes the same calls to the one-thread-per-pair comparator.
"""
from __future__ import annotations

import ctypes
import threading
from contextlib import contextmanager

import numpy as np

from .. import _hip
from . import Backend


class DeviceScores:
    """Native scores: a device int32 matrix ``(nq, nr)`` owned by Python."""

    def __init__(self, buf: _hip.DeviceBuffer, shape):
        self.buf = buf
        self.shape = tuple(shape)

    def to_host(self) -> np.ndarray:
        return self.buf.to_host().reshape(self.shape)


def _policy(scheme) -> int:
    """The library's element-type policy for a scheme's ``dtype``.

    ``"auto"`` applies the reference's int16 envelope (``src/gpusw/scheme.py:116-141``)
    to every pair with the pair's own lengths instead of the call's longest ones: a pair
    runs on packed 16-bit tiles only where no DP value can saturate int16, otherwise on the
    int32 tiles.  Every pair the call-wide envelope sends to int32 that could saturate is
    therefore on the int32 tiles, as in ``backends/cuda.py:68-85``.  The packed tiles are
    binary16 pairs (exact for integers up to 2048; local results from 2048 up are rescored
    in int32) or, for global/semiglobal pairs past that range, int16.  ``"int16"`` reaches
    the library only after ``resolve_dtype`` accepted it; local results at the top of the
    packed type would be rescored on the int32 tiles.  ``"int32"``: 32-bit tiles
    throughout.
    """
    if scheme.dtype == "int32":
        return _hip.DTYPE_INT32
    if scheme.dtype == "int16":
        return _hip.DTYPE_INT16
    return _hip.DTYPE_AUTO


def _describe(kind: str, scheme, stats: _hip.GswStats, dtype: str) -> str:
    return (
        f"// libgsw_hip.so ({_hip.load().gsw_version().decode()}) {kind}\n"
        f"// MODE={scheme.mode} GOPEN={scheme.gap_open} GEXT={scheme.gap_extend} "
        f"FREE_Q={int(scheme.free_end_gaps_query)} FREE_R={int(scheme.free_end_gaps_ref)} "
        f"NS={scheme.ns} dtype={dtype}\n"
        f"// launches={stats.launches} variant={stats.variant} rescored={stats.rescored} "
        f"plain_pairs={stats.plain_pairs} kernel_ms={stats.kernel_ms:.3f}\n"
    )


class HipBackend(Backend):
    """Batched wavefront kernels: packed-int16 lanes, int32 fallback, device top-k."""

    name = "hip"
    history = None   # set to a list to collect every call's GswStats (benchmarks)
    _scratch = threading.local()   # per thread and device: the pair scores' device buffer

    def _pair_out(self, n: int) -> _hip.DeviceBuffer:
        """A grow-only device buffer of at least ``n`` int32 for pair scores (one per
        thread and device, so concurrent callers never share one)."""
        bufs = getattr(self._scratch, "bufs", None)
        if bufs is None:
            bufs = self._scratch.bufs = {}
        dev = _hip.get_device()
        buf = bufs.get(dev)
        if buf is None or buf.nbytes < 4 * n:
            bufs[dev] = None   # free the old buffer first
            buf = bufs[dev] = _hip.DeviceBuffer(4 * max(n, 1), np.int32)
        return buf

    def available(self) -> bool:
        return _hip.gpu_available()

    def _note(self, stats) -> None:
        self.last_stats = stats
        if self.history is not None:
            self.history.append(stats)

    @contextmanager
    def device_scope(self, device: int):
        prev = _hip.get_device()
        if prev != device:
            _hip.set_device(device)
        try:
            yield
        finally:
            if prev != device:
                _hip.set_device(prev)

    def upload(self, arr):
        return _hip.DeviceBuffer.from_array(np.ascontiguousarray(arr))

    # ------------------------------------------------------------------ scoring
    def run_cross(self, scheme, maxq, dtype, threads, qbuf, qoff, rbuf, roff, qlist_host, nq, nr):
        qset = _hip.seqset_for(qbuf, qoff)
        rset = _hip.seqset_for(rbuf, roff)
        qlist = np.ascontiguousarray(qlist_host, dtype=np.int32)
        out = _hip.DeviceBuffer(4 * nq * nr, np.int32, (nq * nr,))
        s, table = scheme_struct_cached(scheme)
        stats = _hip.GswStats()
        _hip.call("gsw_score_cross", ctypes.byref(s), _policy(scheme), qset.handle, _hip.ptr(qlist),
                  int(nq), rset.handle, out.ptr, ctypes.byref(stats))
        self._note(stats)
        return DeviceScores(out, (nq, nr)), stats.kernel_ms / 1e3, _describe("cross", scheme, stats, dtype)

    def run_pairs(self, scheme, maxq, dtype, threads, qbuf, qoff, rbuf, roff, qi_host, rj_host, n):
        qset = _hip.seqset_for(qbuf, qoff)
        rset = _hip.seqset_for(rbuf, roff)
        qi = np.ascontiguousarray(qi_host, dtype=np.int32)
        rj = np.ascontiguousarray(rj_host, dtype=np.int32)
        s, table = scheme_struct_cached(scheme)
        host = np.empty(n, dtype=np.int32)
        # at most max_threads_per_launch pairs per library call (as the reference's Metal
        # backend chunks its pairs, src/gpusw/backends/metal.py:126); indices are checked by
        # the library while it plans (IndexError), not in extra passes
        cap = max(1, int(self.max_threads_per_launch))
        kernel_ms, cells, stats = 0.0, 0, _hip.GswStats()
        for k0 in range(0, max(n, 1), cap):
            m = min(cap, n - k0)
            buf = self._pair_out(m)
            stats = _hip.GswStats()
            _hip.call("gsw_score_pairs", ctypes.byref(s), _policy(scheme), qset.handle, rset.handle,
                      _hip.ptr(qi[k0:]), _hip.ptr(rj[k0:]), int(m), buf.ptr, ctypes.byref(stats))
            self._note(stats)
            kernel_ms += stats.kernel_ms
            cells += int(stats.cells)    # counted by the library while planning
            if m:
                _hip.call("gsw_memcpy_d2h", _hip.ptr(host[k0:]), buf.ptr, 4 * m)
        self.last_cells = cells
        return host, kernel_ms / 1e3, _describe("pairs", scheme, stats, dtype)

    def as_host(self, scores_native):
        return scores_native.to_host().astype(np.int32, copy=False)

    def topk(self, scores_native, kk):
        nq, nr = scores_native.shape
        idx = np.empty((nq, kk), dtype=np.int32)
        sc = np.empty((nq, kk), dtype=np.int32)
        stats = _hip.GswStats()
        _hip.call("gsw_topk_rows", scores_native.buf.ptr, int(nq), int(nr), int(kk), _hip.ptr(idx),
                  _hip.ptr(sc), ctypes.byref(stats))
        return idx, sc

    def render(self, scheme, maxq, dtype) -> str:
        return (
            "// libgsw_hip.so batched wavefront kernels (gfx950): cross_kernel / pairs_kernel\n"
            f"// MODE={scheme.mode} GOPEN={scheme.gap_open} GEXT={scheme.gap_extend} "
            f"FREE_Q={int(scheme.free_end_gaps_query)} FREE_R={int(scheme.free_end_gaps_ref)} "
            f"NS={scheme.ns} MAXQ={maxq} dtype={dtype}\n"
        )


class HipPlainBackend(HipBackend):
    """The comparator: one thread per pair (reference kernel design), same library."""

    name = "hip-plain"

    def run_cross(self, scheme, maxq, dtype, threads, qbuf, qoff, rbuf, roff, qlist_host, nq, nr):
        qlist = np.ascontiguousarray(qlist_host, dtype=np.int32)
        out = _hip.DeviceBuffer(4 * nq * nr, np.int32, (nq * nr,))
        s, table = scheme_struct_cached(scheme)
        stats = _hip.GswStats()
        store = _hip.DTYPE_INT16 if dtype == "int16" else _hip.DTYPE_INT32
        _hip.call("gsw_score_cross_plain", ctypes.byref(s), store, int(threads), qbuf.ptr, qoff.ptr,
                  rbuf.ptr, roff.ptr, _hip.ptr(qlist), int(nq), int(nr), max(int(maxq), 1), out.ptr,
                  ctypes.byref(stats))
        self.last_stats = stats
        return DeviceScores(out, (nq, nr)), stats.kernel_ms / 1e3, _describe("cross-plain", scheme, stats, dtype)

    def run_pairs(self, scheme, maxq, dtype, threads, qbuf, qoff, rbuf, roff, qi_host, rj_host, n):
        qi = np.ascontiguousarray(qi_host, dtype=np.int32)
        rj = np.ascontiguousarray(rj_host, dtype=np.int32)
        _check_pair_indices(qi, rj, qoff.shape[0] - 1, roff.shape[0] - 1)
        out = _hip.DeviceBuffer(4 * n, np.int32, (n,))
        s, table = scheme_struct_cached(scheme)
        stats = _hip.GswStats()
        store = _hip.DTYPE_INT16 if dtype == "int16" else _hip.DTYPE_INT32
        _hip.call("gsw_score_pairs_plain", ctypes.byref(s), store, int(threads), qbuf.ptr, qoff.ptr,
                  rbuf.ptr, roff.ptr, _hip.ptr(qi), _hip.ptr(rj), int(n), max(int(maxq), 1), out.ptr,
                  ctypes.byref(stats))
        self.last_stats = stats
        return out.to_host(), stats.kernel_ms / 1e3, _describe("pairs-plain", scheme, stats, dtype)


def run_traceback(scheme, qbuf, qoff, rbuf, roff, qi_host, rj_host):
    """GPU traceback of pairs (``gsw_traceback_pairs``): ``(records, ops, stats)`` with
    records a structured array of ``gsw_aln`` (``_hip.ALN_DTYPE``) and ops the
    concatenated CIGAR runs."""
    qset = _hip.seqset_for(qbuf, qoff)
    rset = _hip.seqset_for(rbuf, roff)
    qi = np.ascontiguousarray(qi_host, dtype=np.int32)
    rj = np.ascontiguousarray(rj_host, dtype=np.int32)
    _check_pair_indices(qi, rj, qset.n, rset.n)
    n = int(qi.size)
    qlen = np.diff(qoff.host.astype(np.int64))
    rlen = np.diff(roff.host.astype(np.int64))
    cap = int((qlen[qi] + rlen[rj] + 1).sum()) if n else 0
    recs = (_hip.GswAln * max(n, 1))()
    ops = np.zeros(max(cap, 1), dtype=np.uint32)
    s, _table = scheme_struct_cached(scheme)
    stats = _hip.GswStats()
    _hip.call("gsw_traceback_pairs", ctypes.byref(s), qset.handle, rset.handle, _hip.ptr(qi), _hip.ptr(rj),
              n, ctypes.cast(recs, ctypes.c_void_p), _hip.ptr(ops), cap, ctypes.byref(stats))
    return np.frombuffer(recs, dtype=_hip.ALN_DTYPE, count=n), ops, stats


def _check_pair_indices(qi, rj, nq, nr) -> None:
    if qi.size and (qi.min() < 0 or qi.max() >= nq):
        raise IndexError("query index out of range")
    if rj.size and (rj.min() < 0 or rj.max() >= nr):
        raise IndexError("reference index out of range")


_SCHEME_CACHE: dict = {}


def scheme_struct_cached(scheme):
    """C scheme struct for ``scheme`` (cached by value; the table stays alive)."""
    key = (scheme.mode, scheme.gap_open, scheme.gap_extend, bool(scheme.free_end_gaps_query),
           bool(scheme.free_end_gaps_ref), scheme.table_hash())
    hit = _SCHEME_CACHE.get(key)
    if hit is None:
        hit = _hip.scheme_struct(scheme)
        _SCHEME_CACHE[key] = hit
    return hit
