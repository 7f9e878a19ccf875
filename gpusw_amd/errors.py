"""Exception classes of the HIP scoring backend.

The names and the hierarchy are those of the reference package
(``src/gpusw/errors.py:5-31``) so that callers catching ``GpuSWError`` or one of its
subclasses keep working when they switch backends.
"""
from __future__ import annotations

__all__ = ["GpuSWError", "NoGpuError", "SchemeError", "OverflowRiskError", "EncodeError"]


class GpuSWError(Exception):
    """Root of every error this package raises."""


class NoGpuError(GpuSWError):
    """No usable AMD GPU, or the HIP library ``libgsw_hip.so`` is not built/loadable."""


class SchemeError(GpuSWError):
    """An invalid scoring scheme or substitution matrix."""


class OverflowRiskError(GpuSWError):
    """``dtype="int16"`` was forced but the worst-case score range does not fit."""


class EncodeError(GpuSWError):
    """Input sequences could not be turned into code/offset buffers."""
