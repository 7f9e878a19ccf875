"""ctypes wrapper of ``oracle/liboracle.so`` (test infrastructure: checker and CPU
baseline).  Same algorithm as ``oracle.gotoh.score_codes`` restated in C
(``oracle/gotoh.c``; reference ``src/gpusw/reference.py:38-125``)."""
from __future__ import annotations

import ctypes
import os
import subprocess
from pathlib import Path

import numpy as np

from .gotoh import MODE_CODE

_HERE = Path(__file__).resolve().parent
_LIB = None

_P = ctypes.c_void_p


def build() -> Path:
    subprocess.run(["make", "-s", "-C", str(_HERE), "liboracle.so"], check=True)
    return _HERE / "liboracle.so"


def lib():
    global _LIB
    if _LIB is None:
        path = _HERE / "liboracle.so"
        if not path.exists():
            build()
        so = ctypes.CDLL(str(path))
        i32, i64 = ctypes.c_int, ctypes.c_int64
        so.oracle_score_pairs.argtypes = [_P, _P, _P, _P, _P, _P, i64, _P, i32, i32, i64, i64,
                                          i32, i32, _P, i32]
        so.oracle_score_cross.argtypes = [_P, _P, _P, _P, _P, i32, i32, _P, i32, i32, i64, i64,
                                          i32, i32, _P, i32]
        _LIB = so
    return _LIB


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(_P)


def default_threads() -> int:
    return len(os.sched_getaffinity(0))


def score_pairs(qcodes, qoff, rcodes, roff, qi, rj, sub, mode, go, ge, free_q, free_r,
                nthreads: int | None = None) -> np.ndarray:
    qcodes = np.ascontiguousarray(qcodes, np.uint8)
    rcodes = np.ascontiguousarray(rcodes, np.uint8)
    qoff = np.ascontiguousarray(qoff, np.int32)
    roff = np.ascontiguousarray(roff, np.int32)
    qi = np.ascontiguousarray(qi, np.int32)
    rj = np.ascontiguousarray(rj, np.int32)
    sub = np.ascontiguousarray(sub, np.int64)
    out = np.empty(qi.size, np.int64)
    lib().oracle_score_pairs(_ptr(qcodes), _ptr(qoff), _ptr(rcodes), _ptr(roff), _ptr(qi),
                             _ptr(rj), qi.size, _ptr(sub), sub.shape[0], MODE_CODE[mode],
                             int(go), int(ge), int(free_q), int(free_r), _ptr(out),
                             nthreads or default_threads())
    return out


def score_cross(qcodes, qoff, rcodes, roff, qlist, sub, mode, go, ge, free_q, free_r,
                nthreads: int | None = None) -> np.ndarray:
    qcodes = np.ascontiguousarray(qcodes, np.uint8)
    rcodes = np.ascontiguousarray(rcodes, np.uint8)
    qoff = np.ascontiguousarray(qoff, np.int32)
    roff = np.ascontiguousarray(roff, np.int32)
    qlist = np.ascontiguousarray(qlist, np.int32)
    sub = np.ascontiguousarray(sub, np.int64)
    nr = roff.size - 1
    out = np.empty((qlist.size, nr), np.int64)
    lib().oracle_score_cross(_ptr(qcodes), _ptr(qoff), _ptr(rcodes), _ptr(roff), _ptr(qlist),
                             qlist.size, nr, _ptr(sub), sub.shape[0], MODE_CODE[mode], int(go),
                             int(ge), int(free_q), int(free_r), _ptr(out),
                             nthreads or default_threads())
    return out


def score_encoded_pairs(qenc, renc, qi, rj, scheme, nthreads=None) -> np.ndarray:
    """Convenience: score ``(qi[k], rj[k])`` of two ``Encoded`` sets under ``scheme``."""
    return score_pairs(qenc.codes, qenc.offsets, renc.codes, renc.offsets, qi, rj,
                       scheme.substitution_table(), scheme.mode, scheme.gap_open,
                       scheme.gap_extend, scheme.free_end_gaps_query, scheme.free_end_gaps_ref,
                       nthreads)


def score_encoded_cross(qenc, renc, scheme, qlist=None, nthreads=None) -> np.ndarray:
    if qlist is None:
        qlist = np.arange(qenc.n, dtype=np.int32)
    return score_cross(qenc.codes, qenc.offsets, renc.codes, renc.offsets, qlist,
                       scheme.substitution_table(), scheme.mode, scheme.gap_open,
                       scheme.gap_extend, scheme.free_end_gaps_query, scheme.free_end_gaps_ref,
                       nthreads)
