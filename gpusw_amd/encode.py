"""Sequence input funnel: every accepted input form becomes one ``uint8`` code buffer plus
an ``int32`` offsets array (``codes[offsets[i]:offsets[i+1]]`` is sequence ``i``).

Accepted forms and their rules follow the reference's ``src/gpusw/encode.py``:
LUT construction (``:30-44``), FASTA parsing (``:91-121``), input normalisation
(``:124-161``) and the funnel itself, including the pre-encoded fast path (``:164-220``).
Encoding is vectorised: ASCII inputs go through one LUT pass over the concatenated bytes.
"""
from __future__ import annotations

import gzip
import os
from dataclasses import dataclass

import numpy as np

from .errors import EncodeError

__all__ = ["Encoded", "funnel", "build_lut", "read_fasta", "encode_fasta", "save_db", "load_db"]

_PATH_SUFFIXES = (".fasta", ".fa", ".fna", ".faa", ".ffn", ".fastq", ".fq", ".txt", ".gz")


def build_lut(scheme) -> tuple[np.ndarray, int]:
    """256-entry byte -> code table for ``scheme.eff_alphabet``; other bytes map to the
    sentinel code ``K = len(alphabet)``.  Returns ``(lut, K)``."""
    alpha = scheme.eff_alphabet
    sentinel = len(alpha)
    lut = np.full(256, sentinel, dtype=np.uint8)
    for code, sym in enumerate(alpha):
        spellings = {sym}
        if scheme.case_insensitive:
            spellings |= {sym.lower(), sym.upper()}
        for s in spellings:
            lut[ord(s)] = code
    return lut, sentinel


@dataclass
class Encoded:
    """Concatenated codes, ``n+1`` offsets, ids, per-sequence lengths and the alphabet."""

    codes: np.ndarray
    offsets: np.ndarray
    ids: list
    lengths: np.ndarray
    alphabet: str

    @property
    def n(self) -> int:
        return len(self.ids)

    @property
    def max_len(self) -> int:
        return int(self.lengths.max()) if self.n else 0


def _is_path_like(s: str) -> bool:
    return "/" in s or os.sep in s or s.lower().endswith(_PATH_SUFFIXES)


def read_fasta(path_or_text) -> list[tuple[str, str]]:
    """``[(id, sequence), ...]`` from a FASTA path (``.gz`` allowed) or FASTA text.
    A path-like string naming no file raises :class:`FileNotFoundError`."""
    if isinstance(path_or_text, (str, os.PathLike)) and os.path.exists(path_or_text):
        path = str(path_or_text)
        opener = gzip.open if path.endswith(".gz") else open
        with opener(path, "rt") as fh:
            text = fh.read()
    else:
        text = str(path_or_text)
        if not text.lstrip().startswith(">") and _is_path_like(text):
            raise FileNotFoundError(f"FASTA file not found: {path_or_text}")
        text = path_or_text
    records: list[tuple[str, str]] = []
    name = None
    chunks: list[str] = []
    for raw in text.splitlines():
        line = raw.strip()
        if not line:
            continue
        if line.startswith(">"):
            if name is not None:
                records.append((name, "".join(chunks)))
            name = line[1:].split()[0] if len(line) > 1 else f"seq{len(records)}"
            chunks = []
        else:
            chunks.append(line)
    if name is not None:
        records.append((name, "".join(chunks)))
    return records


def _read_bytes(path: str) -> bytes:
    opener = gzip.open if str(path).endswith(".gz") else open
    with opener(path, "rb") as fh:
        return fh.read()


def encode_fasta(path, scheme) -> "Encoded | None":
    """FASTA file -> :class:`Encoded` in two native passes (``gsw_fasta_scan`` and
    ``gsw_fasta_encode``), with the records :func:`read_fasta` gives.  Returns None when
    the file is not ASCII or the native library is not built; the caller then takes the
    record-by-record path."""
    import ctypes

    from . import _hip
    from .errors import NoGpuError

    try:
        lib = _hip.load()
    except NoGpuError:
        return None
    raw = _read_bytes(str(path))
    buf = np.frombuffer(raw, dtype=np.uint8)
    nrec, nres = ctypes.c_int64(0), ctypes.c_int64(0)
    if lib.gsw_fasta_scan(_hip.ptr(buf), buf.size, ctypes.byref(nrec), ctypes.byref(nres)) != 0:
        return None
    lut, sentinel = build_lut(scheme)
    codes = np.empty(max(nres.value, 1), dtype=np.uint8)
    offsets = np.empty(nrec.value + 1, dtype=np.int64)
    spans = np.empty(2 * max(nrec.value, 1), dtype=np.int64)
    _hip.check(lib.gsw_fasta_encode(_hip.ptr(buf), buf.size, _hip.ptr(lut), _hip.ptr(codes),
                                    _hip.ptr(offsets), _hip.ptr(spans)))
    codes = codes[:nres.value]
    ids = [raw[a:a + m].decode("ascii") if m else f"seq{k}"
           for k, (a, m) in enumerate(spans[:2 * nrec.value].reshape(-1, 2).tolist())]
    if offsets[-1] > np.iinfo(np.int32).max:
        raise EncodeError("sequence set exceeds 2^31 residues")
    offsets = offsets.astype(np.int32)
    if scheme.unknown == "error" and codes.size:
        bad = np.flatnonzero(codes == sentinel)
        if bad.size:
            which = int(np.searchsorted(offsets, int(bad[0]), side="right")) - 1
            raise EncodeError(
                f"sequence {ids[which]!r} contains an out-of-alphabet symbol "
                f"(alphabet={scheme.eff_alphabet!r}); set unknown='mismatch' or 'zero'"
            )
    return Encoded(codes, offsets, ids, np.diff(offsets).astype(np.int32), scheme.eff_alphabet)


def save_db(path, enc: "Encoded") -> None:
    """Write an encoded sequence set as a directory: ``codes.npy`` (uint8),
    ``offsets.npy`` (int32, n+1), ``ids.txt`` (one id per line) and ``meta.json``."""
    import json

    d = os.fspath(path)
    os.makedirs(d, exist_ok=True)
    if any(("\n" in i or "\r" in i) for i in map(str, enc.ids)):
        raise EncodeError("sequence ids must not contain line breaks")
    np.save(os.path.join(d, "codes.npy"), np.ascontiguousarray(enc.codes, dtype=np.uint8))
    np.save(os.path.join(d, "offsets.npy"), np.ascontiguousarray(enc.offsets, dtype=np.int32))
    with open(os.path.join(d, "ids.txt"), "w", encoding="utf-8") as fh:
        fh.write("".join(f"{i}\n" for i in enc.ids))
    with open(os.path.join(d, "meta.json"), "w") as fh:
        json.dump({"format": "gsw-db-1", "alphabet": enc.alphabet, "n": enc.n,
                   "residues": int(enc.offsets[-1]) if enc.offsets.size else 0}, fh)


def load_db(path, scheme=None, mmap: bool = True) -> "Encoded":
    """Read a directory written by :func:`save_db` (arrays memory-mapped, nothing
    unpickled).  With ``scheme``, its alphabet must be the one the set was encoded for."""
    import json

    d = os.fspath(path)
    with open(os.path.join(d, "meta.json")) as fh:
        meta = json.load(fh)
    if meta.get("format") != "gsw-db-1":
        raise EncodeError(f"{d}: not a sequence database (format {meta.get('format')!r})")
    if scheme is not None and meta["alphabet"] != scheme.eff_alphabet:
        raise EncodeError(f"{d}: encoded for alphabet {meta['alphabet']!r}, scheme uses "
                          f"{scheme.eff_alphabet!r}")
    mode = "r" if mmap else None
    codes = np.load(os.path.join(d, "codes.npy"), mmap_mode=mode, allow_pickle=False)
    offsets = np.load(os.path.join(d, "offsets.npy"), mmap_mode=mode, allow_pickle=False)
    with open(os.path.join(d, "ids.txt"), encoding="utf-8") as fh:
        ids = fh.read().split("\n")[:-1]
    if len(ids) != offsets.size - 1 or meta["n"] != len(ids):
        raise EncodeError(f"{d}: ids, offsets and meta disagree on the sequence count")
    return Encoded(codes, offsets, ids, np.diff(offsets).astype(np.int32), meta["alphabet"])


def _is_db(obj) -> bool:
    return isinstance(obj, (str, os.PathLike)) and os.path.isfile(os.path.join(os.fspath(obj), "meta.json"))


def _records(obj, id_prefix: str) -> list[tuple[str, str]]:
    """Normalise a supported input object to ``[(id, sequence_text), ...]``."""
    if isinstance(obj, dict):
        return [(str(k), str(v)) for k, v in obj.items()]
    if isinstance(obj, (bytes, bytearray)):
        # latin-1 maps every byte to one code point, so no byte is lost
        return [(f"{id_prefix}0", bytes(obj).decode("latin-1"))]
    if isinstance(obj, str):
        if obj.lstrip().startswith(">") or os.path.exists(obj):
            return read_fasta(obj)
        if _is_path_like(obj):
            raise EncodeError(f"file not found: {obj!r}")
        return [(f"{id_prefix}0", obj)]
    if isinstance(obj, np.ndarray):
        obj = obj.tolist()
    try:
        items = list(obj)
    except TypeError as exc:
        raise EncodeError(f"cannot funnel object of type {type(obj).__name__}") from exc
    out: list[tuple[str, str]] = []
    for i, item in enumerate(items):
        if isinstance(item, str):
            out.append((f"{id_prefix}{i}", item))
        elif isinstance(item, (tuple, list)) and len(item) == 2:
            out.append((str(item[0]), str(item[1])))
        elif hasattr(item, "seq"):
            out.append((str(getattr(item, "id", None) or f"{id_prefix}{i}"), str(item.seq)))
        else:
            raise EncodeError(
                f"item {i} of type {type(item).__name__} is not a sequence, "
                "(id, seq) pair, or .seq-bearing record"
            )
    return out


def _from_preencoded(obj, scheme, id_prefix: str) -> Encoded:
    codes = np.ascontiguousarray(obj[0], dtype=np.uint8)
    offsets = np.ascontiguousarray(obj[1], dtype=np.int32)
    if offsets.size == 0:
        raise EncodeError("pre-encoded offsets must have length n+1 (>= 1)")
    if (offsets < 0).any():
        raise EncodeError("pre-encoded offsets must be non-negative")
    if (np.diff(offsets) < 0).any():
        raise EncodeError("pre-encoded offsets must be non-decreasing")
    if int(offsets[-1]) > codes.size:
        raise EncodeError(
            f"pre-encoded offsets[-1]={int(offsets[-1])} exceeds len(codes)={codes.size}; "
            "offsets must index within the codes buffer"
        )
    n = offsets.size - 1
    ids = list(obj[2]) if len(obj) == 3 else [f"{id_prefix}{i}" for i in range(n)]
    return Encoded(codes, offsets, ids, np.diff(offsets).astype(np.int32), scheme.eff_alphabet)


def _encode_texts(texts: list[str], scheme) -> tuple[np.ndarray, np.ndarray]:
    """Encode texts to ``(codes, lengths)``: case-fold when the scheme asks, encode each
    character as one latin-1 byte (unencodable -> ``?``), then map bytes through the LUT."""
    lut, _ = build_lut(scheme)
    if all(t.isascii() for t in texts):
        # ASCII: folding does not change lengths and the LUT already maps both cases.
        lengths = np.fromiter((len(t) for t in texts), dtype=np.int64, count=len(texts))
        joined = "".join(texts)
        if scheme.case_insensitive:
            joined = joined.upper()
        return np.frombuffer(joined.encode("ascii"), dtype=np.uint8), lengths
    folded = [t.upper() if scheme.case_insensitive else t for t in texts]
    lengths = np.fromiter((len(t) for t in folded), dtype=np.int64, count=len(folded))
    raw = np.frombuffer("".join(folded).encode("latin-1", "replace"), dtype=np.uint8)
    return raw, lengths


def funnel(obj, scheme, *, id_prefix: str = "q") -> Encoded:
    """Turn any supported input into an :class:`Encoded` for ``scheme``."""
    if isinstance(obj, Encoded):
        if obj.alphabet != scheme.eff_alphabet:
            raise EncodeError(
                f"Encoded alphabet {obj.alphabet!r} != scheme alphabet "
                f"{scheme.eff_alphabet!r}; re-encode for this scheme"
            )
        return obj
    if (
        isinstance(obj, tuple)
        and len(obj) in (2, 3)
        and isinstance(obj[0], np.ndarray)
        and isinstance(obj[1], np.ndarray)
        and obj[0].dtype == np.uint8
    ):
        return _from_preencoded(obj, scheme, id_prefix)

    if _is_db(obj):
        return load_db(obj, scheme)
    if isinstance(obj, (str, os.PathLike)) and not str(obj).lstrip().startswith(">") and os.path.isfile(obj):
        enc = encode_fasta(obj, scheme)   # None: not plain ASCII, take the record path
        if enc is not None:
            return enc
    records = _records(obj, id_prefix)
    ids = [r[0] for r in records]
    raw, lengths = _encode_texts([r[1] for r in records], scheme)
    lut, sentinel = build_lut(scheme)
    codes = lut[raw]
    offsets = np.zeros(len(records) + 1, dtype=np.int32)
    np.cumsum(lengths, out=offsets[1:])
    if scheme.unknown == "error" and codes.size:
        bad = np.flatnonzero(codes == sentinel)
        if bad.size:
            pos = int(bad[0])
            which = int(np.searchsorted(offsets, pos, side="right")) - 1
            sym = bytes(raw[pos:pos + 1]).decode("latin-1")
            raise EncodeError(
                f"sequence {ids[which]!r} contains out-of-alphabet symbol {sym!r} "
                f"(alphabet={scheme.eff_alphabet!r}); set unknown='mismatch' or 'zero'"
            )
    return Encoded(
        np.ascontiguousarray(codes, dtype=np.uint8), offsets, ids,
        lengths.astype(np.int32), scheme.eff_alphabet,
    )
