"""Serving: resident reference databases and continuously batched top-k requests.

:class:`Server` wraps the serving loop of ``libgsw_hip.so`` (``gsw_server_*`` in
``include/gsw_hip.h``).  Databases are encoded and uploaded once.  ``submit`` returns
immediately; requests are admitted in arrival order when their device memory (query
codes, score rows, top-k output) fits the budget, and the work of every admitted
request of a database is batched into shared persistent-kernel launches, so requests
that arrive while others run overlap with them on the GPU.  Each result is the same
top-k that ``Aligner.top_k`` returns for those queries alone (score descending, ties to
the lower reference index).  There is no counterpart in the reference package, whose
``Aligner.top_k`` (``src/gpusw/aligner.py:169-195``) is a synchronous batch loop.
"""
from __future__ import annotations

import ctypes
import threading
from dataclasses import dataclass

import numpy as np

from . import _hip
from .aligner import _as_scheme
from .backends.hip import scheme_struct_cached
from .encode import funnel
from .result import AlignResult

__all__ = ["Server", "ServedResult"]


@dataclass
class ServedResult:
    """A finished request: the top-k result and its timing (milliseconds since the
    server started; ``latency_ms`` = done - submit)."""

    result: AlignResult
    latency_ms: float
    submit_ms: float
    admit_ms: float
    done_ms: float
    cells: int
    launches: int


class Server:
    """Continuous-batching top-k server on one GPU.

    ``memory_budget`` bounds the device memory of admitted requests (bytes);
    ``launch_cells`` is the target number of cell updates per batched launch (smaller:
    lower latency, larger: fewer launches).
    """

    def __init__(self, memory_budget: int = 8 << 30, launch_cells: float = 2e11, device: int = 0):
        self._lock = threading.Lock()
        self._dbs: dict[int, tuple] = {}
        self._pending: dict[int, tuple] = {}
        self.device = int(device)
        _hip.set_device(self.device)
        self._h = ctypes.c_void_p()
        _hip.call("gsw_server_create", int(memory_budget), float(launch_cells), ctypes.byref(self._h))

    # ------------------------------------------------------------- databases
    def add_database(self, refs, scheme="dna") -> int:
        """Encode and upload a reference set once; returns its database id."""
        sch = _as_scheme(scheme)
        enc = funnel(refs, sch, id_prefix="r")
        codes = _hip.DeviceBuffer.from_array(enc.codes)
        offs = _hip.DeviceBuffer.from_array(enc.offsets)
        sset = _hip.SeqSet(codes, offs)
        s, _table = scheme_struct_cached(sch)
        db = ctypes.c_int32(0)
        _hip.call("gsw_server_add_db", self._h, ctypes.byref(s), sset.handle, ctypes.byref(db))
        with self._lock:
            self._dbs[db.value] = (sch, enc, codes, offs, sset)
        return int(db.value)

    # --------------------------------------------------------------- requests
    def submit(self, db: int, queries, k: int = 5) -> int:
        """Queue the top-``k`` of ``queries`` against database ``db``; returns a request id."""
        sch, renc = self._dbs[db][0], self._dbs[db][1]
        q = funnel(queries, sch, id_prefix="q")
        codes = np.ascontiguousarray(q.codes, dtype=np.uint8)
        offs = np.ascontiguousarray(q.offsets, dtype=np.int32)
        rid = ctypes.c_int64(0)
        _hip.call("gsw_server_submit", self._h, int(db), _hip.ptr(codes) if codes.size else None,
                  _hip.ptr(offs), int(q.n), int(k), ctypes.byref(rid))
        with self._lock:
            self._pending[rid.value] = (db, q.ids, min(k, renc.n))
        return int(rid.value)

    def result(self, request_id: int, timeout: float | None = None) -> ServedResult:
        """Wait for a request (``timeout`` in seconds, None = forever) and return it."""
        with self._lock:
            db, qids, kk = self._pending[request_id]
        sch, renc = self._dbs[db][0], self._dbs[db][1]
        nq = len(qids)
        idx = np.empty((nq, kk), dtype=np.int32)
        sc = np.empty((nq, kk), dtype=np.int32)
        info = _hip.GswRequestInfo()
        ms = -1 if timeout is None else int(max(timeout, 0) * 1000)
        try:
            _hip.call("gsw_server_wait", self._h, int(request_id), ms, _hip.ptr(idx), _hip.ptr(sc),
                      ctypes.byref(info))
        except TimeoutError:
            raise
        except Exception:
            with self._lock:
                self._pending.pop(request_id, None)
            raise
        with self._lock:
            self._pending.pop(request_id, None)
        res = AlignResult(None, qids, renc.ids, sch, topk_idx=idx, topk_scores=sc)
        return ServedResult(res, info.latency_ms, info.submit_ms, info.admit_ms, info.done_ms,
                            int(info.cells), int(info.launches))

    def stats(self) -> dict:
        st = _hip.GswServerStatus()
        _hip.call("gsw_server_stats", self._h, ctypes.byref(st))
        return st.as_dict()

    def close(self) -> None:
        """Finish every submitted request, then release the server."""
        if self._h is not None and self._h.value:
            _hip.call("gsw_server_destroy", self._h)
            self._h = ctypes.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
