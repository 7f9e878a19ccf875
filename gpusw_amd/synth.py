"""Seeded synthetic workloads (sequences returned pre-encoded as ``(codes, offsets)``).

* ``plain_dna``      -- 64 queries x 256 references, uniform ACGT, lengths 100..300.
* ``dna_16s_cross``  -- 16S-shaped DNA: 200 templates ~N(1450, 60) nt clipped to
  1300..1600; 20,000 references copied from random templates with 5% substitution and
  1% indel events (insertion/deletion 50:50, length ~Geometric(0.5)); 2,000 queries of
  240..260 nt cut from random references with 2% extra substitution.
* ``protein_topk``   -- Robinson & Robinson (1991) background amino-acid frequencies;
  100,000 references of lognormal length (median 300, sigma 0.6, clipped 30..3000) and
  1,000 queries (median 350, sigma 0.6, clipped 30..2000).
* ``read_pairs``     -- 10,000 ACGT windows of 200..400 nt; 150-nt reads drawn from a
  random window with 1% substitution and 0.2% single-base indels; pair k = (read k,
  its window).

Codes index the scheme alphabet: ``ACGT`` -> 0..3 for the DNA schemes, and the first 20
letters of the BLOSUM62 alphabet (``ARNDCQEGHILKMFPSTWYV``) -> 0..19 for protein.
"""
from __future__ import annotations

import numpy as np

DNA = "ACGT"
PROT = "ARNDCQEGHILKMFPSTWYV"

# Robinson & Robinson (1991) amino-acid background frequencies, in PROT order.
RR_FREQ = np.array([0.07805, 0.05129, 0.04487, 0.05364, 0.01925, 0.04264, 0.06295, 0.07377,
                    0.02199, 0.05142, 0.09019, 0.05744, 0.02243, 0.03856, 0.05203, 0.07120,
                    0.05841, 0.01330, 0.03216, 0.06441])


def pack(seqs: list[np.ndarray]) -> tuple[np.ndarray, np.ndarray]:
    lengths = np.fromiter((len(s) for s in seqs), dtype=np.int64, count=len(seqs))
    offsets = np.zeros(len(seqs) + 1, dtype=np.int32)
    np.cumsum(lengths, out=offsets[1:])
    codes = np.concatenate(seqs).astype(np.uint8) if seqs else np.zeros(0, np.uint8)
    return codes, offsets


def unpack(codes: np.ndarray, offsets: np.ndarray) -> list[np.ndarray]:
    return [codes[offsets[i]:offsets[i + 1]] for i in range(len(offsets) - 1)]


def to_strings(codes: np.ndarray, offsets: np.ndarray, alphabet: str) -> list[str]:
    table = np.frombuffer(alphabet.encode(), dtype=np.uint8)
    raw = table[codes]
    return [raw[offsets[i]:offsets[i + 1]].tobytes().decode() for i in range(len(offsets) - 1)]


def _substitute(rng, seq: np.ndarray, rate: float, k: int) -> np.ndarray:
    """Replace each position with probability ``rate`` by a different symbol."""
    out = seq.copy()
    hit = rng.random(seq.size) < rate
    out[hit] = (out[hit] + rng.integers(1, k, int(hit.sum()))) % k
    return out


def _indels(rng, seq: np.ndarray, rate: float, k: int, geometric: bool) -> np.ndarray:
    """Insertion/deletion events at each position with probability ``rate`` (50:50)."""
    return _indels_at(rng, seq, np.flatnonzero(rng.random(seq.size) < rate), k, geometric)


def _indels_at(rng, seq: np.ndarray, pos: np.ndarray, k: int, geometric: bool) -> np.ndarray:
    """Apply one insertion-or-deletion event at each position in ``pos``."""
    if pos.size == 0:
        return seq
    pieces = []
    prev = 0
    for p in pos:
        if p < prev:
            continue
        length = int(rng.geometric(0.5)) if geometric else 1
        pieces.append(seq[prev:p])
        if rng.random() < 0.5:          # insertion before p
            pieces.append(rng.integers(0, k, length).astype(np.uint8))
            prev = p
        else:                           # deletion of [p, p+length)
            prev = p + length
    pieces.append(seq[prev:])
    return np.concatenate(pieces)


def plain_dna(seed: int = 0, nq: int = 64, nr: int = 256, lo: int = 100, hi: int = 300):
    """64 x 256 DNA, lengths ``integers(lo, hi+1)``; queries drawn first, then references."""
    rng = np.random.default_rng(seed)
    qs = [rng.integers(0, 4, int(rng.integers(lo, hi + 1))).astype(np.uint8) for _ in range(nq)]
    rs = [rng.integers(0, 4, int(rng.integers(lo, hi + 1))).astype(np.uint8) for _ in range(nr)]
    return pack(qs), pack(rs)


def dna_16s_cross(seed: int = 0, n_templates: int = 200, n_refs: int = 20000, n_queries: int = 2000):
    rng = np.random.default_rng(seed)
    tlen = np.clip(np.round(rng.normal(1450, 60, n_templates)), 1300, 1600).astype(int)
    templates = [rng.integers(0, 4, L).astype(np.uint8) for L in tlen]
    refs = []
    for _ in range(n_refs):
        t = templates[int(rng.integers(n_templates))]
        s = _substitute(rng, t, 0.05, 4)
        refs.append(_indels(rng, s, 0.01, 4, geometric=True))
    queries = []
    for _ in range(n_queries):
        r = refs[int(rng.integers(n_refs))]
        L = int(rng.integers(240, 261))
        L = min(L, r.size)
        start = int(rng.integers(0, r.size - L + 1))
        queries.append(_substitute(rng, r[start:start + L], 0.02, 4))
    return pack(queries), pack(refs)


def _lognormal_lengths(rng, n, median, sigma, lo, hi):
    return np.clip(np.round(rng.lognormal(np.log(median), sigma, n)), lo, hi).astype(int)


def protein_topk(seed: int = 0, n_refs: int = 100000, n_queries: int = 1000):
    rng = np.random.default_rng(seed)
    p = RR_FREQ / RR_FREQ.sum()
    rlen = _lognormal_lengths(rng, n_refs, 300, 0.6, 30, 3000)
    ref_codes = rng.choice(20, size=int(rlen.sum()), p=p).astype(np.uint8)
    roff = np.zeros(n_refs + 1, dtype=np.int32)
    np.cumsum(rlen, out=roff[1:])
    qlen = _lognormal_lengths(rng, n_queries, 350, 0.6, 30, 2000)
    q_codes = rng.choice(20, size=int(qlen.sum()), p=p).astype(np.uint8)
    qoff = np.zeros(n_queries + 1, dtype=np.int32)
    np.cumsum(qlen, out=qoff[1:])
    return (q_codes, qoff), (ref_codes, roff)


def read_pairs(seed: int = 0, n_windows: int = 10000, n_reads: int = 1000000, read_len: int = 150):
    """Returns ``((read_codes, read_off), (window_codes, window_off), qi, rj)``."""
    rng = np.random.default_rng(seed)
    wlen = rng.integers(200, 401, n_windows)
    wcodes = rng.integers(0, 4, int(wlen.sum())).astype(np.uint8)
    woff = np.zeros(n_windows + 1, dtype=np.int32)
    np.cumsum(wlen, out=woff[1:])
    w = rng.integers(0, n_windows, n_reads)
    take = read_len + 10                       # room for deletions, trimmed after
    start = (rng.random(n_reads) * (wlen[w] - take + 1)).astype(np.int64)
    idx = woff[w][:, None] + start[:, None] + np.arange(take)[None, :]
    reads = wcodes[idx]
    hit = rng.random(reads.shape) < 0.01
    reads[hit] = (reads[hit] + rng.integers(1, 4, int(hit.sum()))) % 4
    events = rng.random((n_reads, take)) < 0.002
    out = [r[:read_len] for r in reads]
    for i in np.flatnonzero(events.any(axis=1)):
        out[i] = _indels_at(rng, reads[i], np.flatnonzero(events[i]), 4, geometric=False)[:read_len]
    return pack(out), (wcodes, woff), np.arange(n_reads, dtype=np.int32), w.astype(np.int32)
