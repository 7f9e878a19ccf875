"""ctypes boundary to ``libgsw_hip.so`` (declared in ``include/gsw_hip.h``).

Loads the in-tree library, declares every exported symbol, maps status codes to the
package's exception classes and wraps device allocations in :class:`DeviceBuffer`.
ctypes releases the GIL for the duration of each foreign call.  There is no fallback:
if the library is not built or no GPU is visible, :class:`NoGpuError` is raised.
"""
from __future__ import annotations

import ctypes
import os
import re
import threading
from pathlib import Path

import numpy as np

from .errors import GpuSWError, NoGpuError, OverflowRiskError

LIB_PATH = Path(os.environ.get("GSW_HIP_LIB", Path(__file__).resolve().parent / "libgsw_hip.so"))

(GSW_OK, GSW_ERR_NO_DEVICE, GSW_ERR_INVALID, GSW_ERR_HIP, GSW_ERR_LIMIT, GSW_ERR_OVERFLOW,
 GSW_ERR_BUSY, GSW_ERR_TIMEOUT) = range(8)
MODE = {"local": 0, "global": 1, "semiglobal": 2}
DTYPE_AUTO, DTYPE_INT16, DTYPE_INT32 = 0, 1, 2

_vp = ctypes.c_void_p
_i32 = ctypes.c_int32
_i64 = ctypes.c_int64
_sz = ctypes.c_size_t


class GswScheme(ctypes.Structure):
    _fields_ = [("mode", _i32), ("gap_open", _i32), ("gap_extend", _i32), ("free_q", _i32),
                ("free_r", _i32), ("ns", _i32), ("sub", _vp)]


class GswStats(ctypes.Structure):
    _fields_ = [("kernel_ms", ctypes.c_double), ("wall_ms", ctypes.c_double), ("cells", _i64),
                ("pairs", _i64), ("rescored", _i64), ("plain_pairs", _i64), ("launches", _i32),
                ("variant", _i32), ("top_ms", ctypes.c_double), ("top_cells", _i64),
                ("top_variant", _i32), ("top_striped", _i32)]

    def as_dict(self) -> dict:
        return {name: getattr(self, name) for name, _ in self._fields_}


class GswRequestInfo(ctypes.Structure):
    _fields_ = [("submit_ms", ctypes.c_double), ("admit_ms", ctypes.c_double),
                ("done_ms", ctypes.c_double), ("latency_ms", ctypes.c_double), ("cells", _i64),
                ("nq", _i32), ("k", _i32), ("launches", _i32), ("status", _i32)]

    def as_dict(self) -> dict:
        return {name: getattr(self, name) for name, _ in self._fields_}


class GswServerStatus(ctypes.Structure):
    _fields_ = [("submitted", _i64), ("completed", _i64), ("rejected", _i64), ("launches", _i64),
                ("cells", _i64), ("rescored", _i64), ("queued", _i32), ("active", _i32),
                ("memory_budget", _sz), ("memory_used", _sz), ("uptime_ms", ctypes.c_double)]

    def as_dict(self) -> dict:
        return {name: getattr(self, name) for name, _ in self._fields_}


class GswAln(ctypes.Structure):
    _fields_ = [("score", _i32), ("q_begin", _i32), ("q_end", _i32), ("r_begin", _i32), ("r_end", _i32),
                ("matches", _i32), ("mismatches", _i32), ("ins", _i32), ("del_", _i32), ("gap_opens", _i32),
                ("n_ops", _i32), ("ops_offset", _i64)]


#: numpy view of a gsw_aln array (same field offsets as the C struct)
ALN_DTYPE = np.dtype({"names": [f for f, _ in GswAln._fields_],
                      "formats": [np.int64 if t is _i64 else np.int32 for _, t in GswAln._fields_],
                      "offsets": [getattr(GswAln, f).offset for f, _ in GswAln._fields_],
                      "itemsize": ctypes.sizeof(GswAln)})


#: symbol -> argtypes (every function returns int status unless listed in _RESTYPE)
_SIGNATURES = {
    "gsw_last_error": [],
    "gsw_version": [],
    "gsw_device_count": [ctypes.POINTER(_i32)],
    "gsw_set_device": [_i32],
    "gsw_get_device": [ctypes.POINTER(_i32)],
    "gsw_device_name": [_i32, ctypes.c_char_p, _i32],
    "gsw_malloc": [ctypes.POINTER(_vp), _sz],
    "gsw_free": [_vp],
    "gsw_memcpy_h2d": [_vp, _vp, _sz],
    "gsw_memcpy_d2h": [_vp, _vp, _sz],
    "gsw_memset": [_vp, _i32, _sz],
    "gsw_synchronize": [],
    "gsw_score_cross_plain": [ctypes.POINTER(GswScheme), _i32, _i32, _vp, _vp, _vp, _vp, _vp, _i32,
                              _i32, _i32, _vp, ctypes.POINTER(GswStats)],
    "gsw_score_pairs_plain": [ctypes.POINTER(GswScheme), _i32, _i32, _vp, _vp, _vp, _vp, _vp, _vp,
                              _i64, _i32, _vp, ctypes.POINTER(GswStats)],
    "gsw_seqset_create": [_vp, _vp, _vp, _i32, ctypes.POINTER(_vp)],
    "gsw_seqset_destroy": [_vp],
    "gsw_score_cross": [ctypes.POINTER(GswScheme), _i32, _vp, _vp, _i32, _vp, _vp,
                        ctypes.POINTER(GswStats)],
    "gsw_score_pairs": [ctypes.POINTER(GswScheme), _i32, _vp, _vp, _vp, _vp, _i64, _vp,
                        ctypes.POINTER(GswStats)],
    "gsw_score_pairs_noalloc": [ctypes.POINTER(GswScheme), _i32, _vp, _vp, _vp, _vp, _i64, _vp,
                                ctypes.POINTER(GswStats)],
    "gsw_topk_rows": [_vp, _i32, _i32, _i32, _vp, _vp, ctypes.POINTER(GswStats)],
    "gsw_memcpy_d2d": [_vp, _vp, _sz],
    "gsw_server_create": [_sz, ctypes.c_double, ctypes.POINTER(_vp)],
    "gsw_server_add_db": [_vp, ctypes.POINTER(GswScheme), _vp, ctypes.POINTER(_i32)],
    "gsw_server_submit": [_vp, _i32, _vp, _vp, _i32, _i32, ctypes.POINTER(_i64)],
    "gsw_server_wait": [_vp, _i64, _i32, _vp, _vp, ctypes.POINTER(GswRequestInfo)],
    "gsw_server_stats": [_vp, ctypes.POINTER(GswServerStatus)],
    "gsw_server_destroy": [_vp],
    "gsw_set_tile_table": [_vp, _i32],
    "gsw_force_variant": [_i32, _i32],
    "gsw_list_variants": [_vp, _i32, ctypes.POINTER(_i32)],
    "gsw_variant_info": [_i32, _vp],
    "gsw_list_stripe_variants": [_vp, _i32, ctypes.POINTER(_i32)],
    "gsw_fasta_scan": [_vp, _i64, ctypes.POINTER(_i64), ctypes.POINTER(_i64)],
    "gsw_traceback_pairs": [ctypes.POINTER(GswScheme), _vp, _vp, _vp, _vp, _i64, _vp, _vp, _i64,
                            ctypes.POINTER(GswStats)],
    "gsw_fasta_encode": [_vp, _i64, _vp, _vp, _vp, _vp],
}
_RESTYPE = {"gsw_last_error": ctypes.c_char_p, "gsw_version": ctypes.c_char_p}

_lib = None
_lib_lock = threading.Lock()


def load():
    """Load and declare the library once; :class:`NoGpuError` if it is missing."""
    global _lib
    if _lib is not None:
        return _lib
    with _lib_lock:
        if _lib is None:
            if not LIB_PATH.exists():
                raise NoGpuError(
                    f"{LIB_PATH.name} is not built; run `python -m gpusw_amd.build` "
                    "(hipcc, ROCm) before using backend='hip'"
                )
            try:
                so = ctypes.CDLL(str(LIB_PATH))
            except OSError as exc:
                raise NoGpuError(f"cannot load {LIB_PATH}: {exc}") from exc
            for name, argtypes in _SIGNATURES.items():
                fn = getattr(so, name)
                fn.argtypes = argtypes
                fn.restype = _RESTYPE.get(name, ctypes.c_int)
            _lib = so
    return _lib


def last_error() -> str:
    return load().gsw_last_error().decode(errors="replace")


def check(rc: int) -> None:
    """Raise the exception matching a status code."""
    if rc == GSW_OK:
        return
    msg = last_error()
    if rc == GSW_ERR_NO_DEVICE:
        raise NoGpuError(msg)
    if rc == GSW_ERR_INVALID:
        # an index past the end of a sequence set, as numpy indexing would report it
        raise (IndexError if re.search(r"\bindex\b.*\bout of range\b", msg) else ValueError)(msg)
    if rc == GSW_ERR_OVERFLOW:
        raise OverflowRiskError(msg)
    if rc == GSW_ERR_TIMEOUT:
        raise TimeoutError(msg)
    raise GpuSWError(msg)


def call(name: str, *args) -> None:
    check(getattr(load(), name)(*args))


def device_count() -> int:
    n = _i32(0)
    call("gsw_device_count", ctypes.byref(n))
    return int(n.value)


def gpu_available() -> bool:
    """True iff the library loads and at least one AMD GPU is visible."""
    try:
        return device_count() > 0
    except NoGpuError:
        return False


def get_device() -> int:
    d = _i32(0)
    call("gsw_get_device", ctypes.byref(d))
    return int(d.value)


def set_device(dev: int) -> None:
    call("gsw_set_device", int(dev))


def device_name(dev: int = 0) -> str:
    buf = ctypes.create_string_buffer(256)
    call("gsw_device_name", int(dev), buf, 256)
    return buf.value.decode()


def ptr(a: np.ndarray):
    return a.ctypes.data_as(_vp)


class DeviceBuffer:
    """A device allocation owned by Python (freed with the object).

    ``host`` optionally mirrors the uploaded contents (kept for offsets, which the
    library needs on the host to index a sequence set).
    """

    __slots__ = ("ptr", "nbytes", "dtype", "shape", "host", "_cache", "__weakref__")

    def __init__(self, nbytes: int, dtype=np.uint8, shape=None):
        self.ptr = _vp()
        self.nbytes = int(nbytes)
        self.dtype = np.dtype(dtype)
        self.shape = tuple(shape) if shape is not None else (self.nbytes // self.dtype.itemsize,)
        self.host = None
        self._cache = {}
        call("gsw_malloc", ctypes.byref(self.ptr), max(self.nbytes, 1))

    @classmethod
    def from_array(cls, arr: np.ndarray, keep_host: bool = True) -> "DeviceBuffer":
        arr = np.ascontiguousarray(arr)
        buf = cls(arr.nbytes, arr.dtype, arr.shape)
        if arr.nbytes:
            call("gsw_memcpy_h2d", buf.ptr, ptr(arr), arr.nbytes)
        if keep_host:
            buf.host = arr
        return buf

    def to_host(self, dtype=None) -> np.ndarray:
        out = np.empty(self.shape, dtype=self.dtype)
        if out.nbytes:
            call("gsw_memcpy_d2h", ptr(out), self.ptr, out.nbytes)
        return out if dtype is None else out.astype(dtype, copy=False)

    def __del__(self):
        p = getattr(self, "ptr", None)
        if p is not None and p.value and _lib is not None:
            _lib.gsw_free(p)
            self.ptr = _vp()

    def __repr__(self) -> str:  # pragma: no cover
        return f"DeviceBuffer(shape={self.shape}, dtype={self.dtype})"


class SeqSet:
    """A ``gsw_seqset``: codes + offsets resident on the device, with a length index."""

    def __init__(self, codes: DeviceBuffer, offsets: DeviceBuffer):
        if offsets.host is None:
            raise ValueError("offsets buffer has no host mirror")
        self.codes = codes
        self.offsets = offsets
        off = np.ascontiguousarray(offsets.host, dtype=np.int32)
        self._off = off
        self.n = int(off.size - 1)
        self.handle = _vp()
        call("gsw_seqset_create", codes.ptr, offsets.ptr, ptr(off), self.n, ctypes.byref(self.handle))

    def __del__(self):
        h = getattr(self, "handle", None)
        if h is not None and h.value and _lib is not None:
            _lib.gsw_seqset_destroy(h)
            self.handle = _vp()


def seqset_for(codes: DeviceBuffer, offsets: DeviceBuffer) -> SeqSet:
    """Cached :class:`SeqSet` for an uploaded (codes, offsets) pair."""
    key = id(codes)
    s = offsets._cache.get(key)
    if s is None or s.codes is not codes:
        s = SeqSet(codes, offsets)
        offsets._cache[key] = s
    return s


def scheme_struct(scheme) -> tuple[GswScheme, np.ndarray]:
    """The C scheme for a :class:`~gpusw_amd.Scheme` (keep the returned table alive)."""
    table = np.ascontiguousarray(scheme.substitution_table(), dtype=np.int32)
    s = GswScheme(MODE[scheme.mode], scheme.gap_open, scheme.gap_extend,
                  int(bool(scheme.free_end_gaps_query)), int(bool(scheme.free_end_gaps_ref)),
                  table.shape[0], table.ctypes.data_as(_vp))
    return s, table


def list_variants() -> list[tuple[int, int]]:
    buf = np.zeros(2 * 256, dtype=np.int32)
    n = _i32(0)
    call("gsw_list_variants", ptr(buf), 256, ctypes.byref(n))
    return [(int(buf[2 * i]), int(buf[2 * i + 1])) for i in range(min(n.value, 256))]


KIND_NAMES = {0: "int32", 1: "int16x2", 2: "half2"}


def variant_info(variant: int) -> tuple[int, int, str, bool]:
    """(G, R, element kind, local mode) of a variant id as gsw_stats reports it."""
    info = np.zeros(4, dtype=np.int32)
    call("gsw_variant_info", int(variant), ptr(info))
    return int(info[0]), int(info[1]), KIND_NAMES[int(info[2])], bool(info[3])


def list_stripe_variants() -> list[tuple[int, int]]:
    buf = np.zeros(2 * 256, dtype=np.int32)
    n = _i32(0)
    call("gsw_list_stripe_variants", ptr(buf), 256, ctypes.byref(n))
    return [(int(buf[2 * i]), int(buf[2 * i + 1])) for i in range(min(n.value, 256))]


def force_variant(G: int = 0, R: int = 0) -> None:
    call("gsw_force_variant", int(G), int(R))
