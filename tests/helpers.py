"""Shared helpers for the tests (seeded sequence generators, scheme table)."""
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
