"""The C ABI boundary (CPU): the built library exports every symbol that
include/gsw_hip.h declares, the ctypes declarations cover them, and backend resolution
behaves as the reference's (unknown name -> ValueError; no GPU -> NoGpuError)."""
import ctypes
import re
from pathlib import Path

import pytest

import gpusw_amd
from gpusw_amd import _hip
from gpusw_amd.backends import resolve_backend
from gpusw_amd.errors import GpuSWError, NoGpuError, OverflowRiskError

ROOT = Path(__file__).resolve().parent.parent
HEADER = (ROOT / "include" / "gsw_hip.h").read_text()
DECLARED = sorted(set(re.findall(r"^(?:int|const char \*)\s*(gsw_\w+)\(", HEADER, re.M)))


def test_header_declares_core_entry_points():
    for name in ("gsw_score_cross", "gsw_score_pairs", "gsw_score_cross_plain",
                 "gsw_score_pairs_plain", "gsw_topk_rows", "gsw_last_error", "gsw_seqset_create"):
        assert name in DECLARED


def test_library_exports_every_declared_symbol():
    if not _hip.LIB_PATH.exists():
        pytest.fail("libgsw_hip.so is not built (run python -m gpusw_amd.build)")
    so = ctypes.CDLL(str(_hip.LIB_PATH))
    missing = [n for n in DECLARED if not hasattr(so, n)]
    assert not missing, missing
    assert set(DECLARED) <= set(_hip._SIGNATURES)


def test_status_mapping():
    _hip.load()
    with pytest.raises(OverflowRiskError):
        _hip.check(_hip.GSW_ERR_OVERFLOW)
    with pytest.raises(ValueError):
        _hip.check(_hip.GSW_ERR_INVALID)
    with pytest.raises(NoGpuError):
        _hip.check(_hip.GSW_ERR_NO_DEVICE)
    with pytest.raises(GpuSWError):
        _hip.check(_hip.GSW_ERR_HIP)


def test_index_errors_map_to_index_error(monkeypatch):
    """The library's 'index out of range' failures surface as IndexError (what numpy
    indexing raises); other invalid-argument failures stay ValueError."""
    _hip.load()
    monkeypatch.setattr(_hip, "last_error", lambda: "pair 12 index out of range")
    with pytest.raises(IndexError):
        _hip.check(_hip.GSW_ERR_INVALID)
    monkeypatch.setattr(_hip, "last_error", lambda: "query index 7 out of range")
    with pytest.raises(IndexError):
        _hip.check(_hip.GSW_ERR_INVALID)
    monkeypatch.setattr(_hip, "last_error", lambda: "device 3 out of range (have 1)")
    with pytest.raises(ValueError) as ei:
        _hip.check(_hip.GSW_ERR_INVALID)
    assert not isinstance(ei.value, IndexError)


def test_bench_refuses_more_than_one_gpu():
    import subprocess
    import sys
    from pathlib import Path

    root = Path(__file__).resolve().parents[1]
    proc = subprocess.run([sys.executable, str(root / "bench.py"), "--gpus", "2"], capture_output=True,
                          text=True, timeout=120)
    assert proc.returncode != 0
    assert "one GPU" in (proc.stdout + proc.stderr)


def test_unknown_backend_name():
    with pytest.raises(ValueError):
        resolve_backend("cuda")


@pytest.mark.skipif(Path("/dev/kfd").exists(), reason="a GPU is present")
def test_no_gpu_raises_no_gpu_error():
    assert not gpusw_amd.gpu_available()
    assert gpusw_amd.available_backends() == []
    with pytest.raises(NoGpuError):
        resolve_backend("auto")
    with pytest.raises(NoGpuError):
        gpusw_amd.align_score(["ACGT"], ["ACGT"])


def test_aligner_is_gpu_free_until_scoring():
    al = gpusw_amd.Aligner("dna")
    assert al.gcups == 0.0 and al.query_ids == []


@pytest.mark.skipif(Path("/dev/kfd").exists(), reason="a GPU is present")
def test_server_without_gpu_raises():
    from gpusw_amd.serve import Server

    with pytest.raises(NoGpuError):
        Server()
