"""Shared helpers for the tests (seeded sequence generators, scheme table, C ABI calls
with an explicit dtype policy)."""
import ctypes

import numpy as np

from gpusw_amd import Matrix, Scheme

PROT = "ARNDCQEGHILKMFPSTWYV"


def rand_seqs(rng, n, lo, hi, alphabet):
    a = np.array(list(alphabet))
    return ["".join(rng.choice(a, int(rng.integers(lo, hi)))) for _ in range(n)]


def schemes():
    """The reference's GPU test scheme set (tests/test_gpu_bitexact.py:20-30), with the
    alphabet each one is exercised on."""
    b62 = Matrix.named("BLOSUM62")
    return {
        "dna_local": (Scheme(mode="local"), "ACGTN"),
        "dna_global": (Scheme(mode="global"), "ACGT"),
        "dna_semi_qr": (Scheme(mode="semiglobal"), "ACGT"),
        "dna_semi_q": (Scheme(mode="semiglobal", free_end_gaps_query=True, free_end_gaps_ref=False), "ACGT"),
        "dna_semi_r": (Scheme(mode="semiglobal", free_end_gaps_query=False, free_end_gaps_ref=True), "ACGT"),
        "dna_semi_none": (Scheme(mode="semiglobal", free_end_gaps_query=False, free_end_gaps_ref=False), "ACGT"),
        "ednafull_global": (Scheme(mode="global", matrix=Matrix.named("EDNAFULL"), gap_open=-10, gap_extend=-1), "ACGTN"),
        "blosum62_local": (Scheme(mode="local", matrix=b62, gap_open=-11, gap_extend=-1), PROT),
        "blosum62_global": (Scheme(mode="global", matrix=b62, gap_open=-11, gap_extend=-1), PROT + "X"),
        "protein_semiglobal": (Scheme(mode="semiglobal", matrix=b62, gap_open=-11, gap_extend=-1), PROT),
    }


def stable_seed(name: str) -> int:
    """Process-independent seed (the reference used hash(), which varies per process)."""
    return sum((i + 1) * ord(c) for i, c in enumerate(name)) % (2**32)


def call_pairs(al, qi, rj, policy):
    """gsw_score_pairs through the C ABI with an explicit dtype policy: (scores, stats)."""
    from gpusw_amd import _hip
    from gpusw_amd.backends.hip import scheme_struct_cached

    qset = _hip.seqset_for(al._d_qbuf, al._d_qoff)
    rset = _hip.seqset_for(al._d_rbuf, al._d_roff)
    s, _table = scheme_struct_cached(al.scheme)
    n = int(qi.size)
    out = _hip.DeviceBuffer(4 * n, np.int32, (n,))
    stats = _hip.GswStats()
    _hip.call("gsw_score_pairs", ctypes.byref(s), policy, qset.handle, rset.handle, _hip.ptr(qi),
              _hip.ptr(rj), n, out.ptr, ctypes.byref(stats))
    return out.to_host(), stats


def call_cross(al, policy):
    """gsw_score_cross of all queries against all references with an explicit dtype policy."""
    from gpusw_amd import _hip
    from gpusw_amd.backends.hip import scheme_struct_cached

    qset = _hip.seqset_for(al._d_qbuf, al._d_qoff)
    rset = _hip.seqset_for(al._d_rbuf, al._d_roff)
    s, _table = scheme_struct_cached(al.scheme)
    nq, nr = al._queries.n, al._refs.n
    qlist = np.arange(nq, dtype=np.int32)
    out = _hip.DeviceBuffer(4 * nq * nr, np.int32, (nq, nr))
    stats = _hip.GswStats()
    _hip.call("gsw_score_cross", ctypes.byref(s), policy, qset.handle, _hip.ptr(qlist), nq,
              rset.handle, out.ptr, ctypes.byref(stats))
    return out.to_host(), stats
