"""The backend seam: the device operations the engine needs, and name -> backend.

The :class:`Backend` contract (members and signatures) is the reference's
``src/gpusw/backends/__init__.py:31-87``; :func:`resolve_backend` follows ``:90-121``
with the AMD names added:

* ``"hip"``        -- the batched wavefront kernels of ``libgsw_hip.so`` (gfx950);
* ``"hip-plain"``  -- the one-thread-per-pair comparator kernel of the same library;
* ``"auto"``       -- ``"hip"`` when an AMD GPU and the library are usable.

There is no CPU backend: without a GPU, resolving a backend raises :class:`NoGpuError`.
"""
from __future__ import annotations

from contextlib import contextmanager

from ..errors import NoGpuError

__all__ = ["Backend", "resolve_backend", "available_backends"]


class Backend:
    """Device operations behind :class:`~gpusw_amd.Aligner` (reference contract)."""

    name: str = "base"
    #: largest number of pairs one launch may cover; the engine chunks above it
    max_threads_per_launch: int = 1 << 62

    def available(self) -> bool:
        raise NotImplementedError

    @contextmanager
    def device_scope(self, device: int):
        yield

    def upload(self, arr):
        raise NotImplementedError

    def run_cross(self, scheme, maxq, dtype, threads, qbuf, qoff, rbuf, roff, qlist_host, nq, nr):
        """Return ``(scores_native (nq, nr), dt_seconds, description)``."""
        raise NotImplementedError

    def run_pairs(self, scheme, maxq, dtype, threads, qbuf, qoff, rbuf, roff, qi_host, rj_host, n):
        """Return ``(host int32 (n,), dt_seconds, description)``."""
        raise NotImplementedError

    def as_host(self, scores_native):
        raise NotImplementedError

    def topk(self, scores_native, kk):
        """Per-row best ``kk`` of native scores: ``(idx int32 (nq,kk), sc int32 (nq,kk))``."""
        raise NotImplementedError

    def render(self, scheme, maxq, dtype) -> str:
        raise NotImplementedError


def resolve_backend(name: str | None) -> Backend:
    """Backend for ``name`` (``"auto"`` | ``"hip"`` | ``"hip-plain"``)."""
    n = (name or "auto").lower()
    if n in ("hip", "hip-plain"):
        from .hip import HipBackend, HipPlainBackend

        return HipBackend() if n == "hip" else HipPlainBackend()
    if n == "auto":
        from .hip import HipBackend

        b = HipBackend()
        if b.available():
            return b
        raise NoGpuError(
            "no GPU backend available: libgsw_hip.so is not built or no AMD GPU is visible "
            "(build with `python -m gpusw_amd.build`)"
        )
    raise ValueError(f"unknown backend {name!r}; choose 'auto', 'hip' or 'hip-plain'")


def available_backends() -> list[str]:
    from .hip import HipBackend, HipPlainBackend

    return [b.name for b in (HipBackend(), HipPlainBackend()) if b.available()]
