"""gpusw_amd -- affine-gap (Gotoh) alignment scoring on AMD Instinct MI355X (gfx950).

The public API keeps the names, arguments and results of ``gpusw`` (the reference's
``src/gpusw/__init__.py:43-189``): :func:`align_score`, :func:`align_pairs`,
:class:`Aligner`, :class:`AlignResult`, :class:`Scheme`, :class:`Matrix` and the
presets.  Scores come from the HIP library ``libgsw_hip.so`` through ctypes
(``backend="hip"``, or ``"hip-plain"`` for the one-thread-per-pair comparator).
"""
from __future__ import annotations

import numpy as np

from . import schemes
from ._hip import gpu_available
from ._version import __version__
from .aligner import Aligner
from .backends import available_backends
from .errors import EncodeError, GpuSWError, NoGpuError, OverflowRiskError, SchemeError
from .matrix import Matrix
from .matrix import available as available_matrices
from .result import AlignResult, Alignment
from .scheme import Scheme

__all__ = [
    "align_score",
    "align_pairs",
    "Aligner",
    "Scheme",
    "Matrix",
    "AlignResult",
    "Alignment",
    "schemes",
    "available_matrices",
    "gpu_available",
    "available_backends",
    "GpuSWError",
    "NoGpuError",
    "SchemeError",
    "OverflowRiskError",
    "EncodeError",
    "__version__",
]


def _build_scheme(scheme, *, mode=None, match=None, mismatch=None, gap_open=None, gap_extend=None,
                  matrix=None, alphabet=None, unknown=None, case_insensitive=None, dtype=None,
                  free_end_gaps_query=None, free_end_gaps_ref=None) -> Scheme:
    """A preset (or Scheme) with every non-None override applied
    (reference ``src/gpusw/__init__.py:66-104``)."""
    base = scheme if isinstance(scheme, Scheme) else Scheme.preset(scheme)
    if isinstance(matrix, str):
        matrix = Matrix.named(matrix)
    given = dict(mode=mode, match=match, mismatch=mismatch, gap_open=gap_open,
                 gap_extend=gap_extend, matrix=matrix, alphabet=alphabet, unknown=unknown,
                 case_insensitive=case_insensitive, dtype=dtype,
                 free_end_gaps_query=free_end_gaps_query, free_end_gaps_ref=free_end_gaps_ref)
    overrides = {k: v for k, v in given.items() if v is not None}
    return base.replace(**overrides) if overrides else base


def align_score(queries, refs, *, scheme="dna", mode=None, match=None, mismatch=None,
                gap_open=None, gap_extend=None, matrix=None, alphabet=None, unknown=None,
                dtype=None, threads=128, backend="auto", return_ids=True, query_batch=None):
    """Scores of every query against every reference (``(n_queries, n_refs)`` int32)."""
    sch = _build_scheme(scheme, mode=mode, match=match, mismatch=mismatch, gap_open=gap_open,
                        gap_extend=gap_extend, matrix=matrix, alphabet=alphabet, unknown=unknown,
                        dtype=dtype)
    al = Aligner(sch, threads=threads, backend=backend).index(refs).set_queries(queries)
    return al.score_cross(return_ids=return_ids, query_batch=query_batch)


def align_pairs(queries, refs, *, pairs=None, scheme="dna", mode=None, match=None, mismatch=None,
                gap_open=None, gap_extend=None, matrix=None, alphabet=None, unknown=None,
                dtype=None, threads=128, backend="auto", return_ids=True):
    """Scores of chosen pairs: a 1:1 zip when ``pairs`` is None (equal counts required),
    else ``pairs = [(query_index, reference_index), ...]``."""
    sch = _build_scheme(scheme, mode=mode, match=match, mismatch=mismatch, gap_open=gap_open,
                        gap_extend=gap_extend, matrix=matrix, alphabet=alphabet, unknown=unknown,
                        dtype=dtype)
    al = Aligner(sch, threads=threads, backend=backend).index(refs).set_queries(queries)
    nq, nr = al._queries.n, al._refs.n
    if pairs is None:
        if nq != nr:
            raise ValueError(
                f"pairs=None requires len(queries)==len(refs) for 1:1 scoring "
                f"(got {nq} vs {nr}); pass pairs=[(qi, rj), ...]"
            )
        qi = np.arange(nq, dtype=np.int32)
        rj = np.arange(nr, dtype=np.int32)
    else:
        qi = np.fromiter((p[0] for p in pairs), dtype=np.int32, count=len(pairs))
        rj = np.fromiter((p[1] for p in pairs), dtype=np.int32, count=len(pairs))
    return al.score_pairs(qi, rj, return_ids=return_ids)
