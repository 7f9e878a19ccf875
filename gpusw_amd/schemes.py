"""Named scheme presets (the reference's ``src/gpusw/schemes.py:24-39``)."""
from __future__ import annotations

from .matrix import Matrix
from .scheme import Scheme

__all__ = [
    "DNA",
    "DNA_GLOBAL",
    "DNA_SEMIGLOBAL",
    "BLASTN",
    "EDNAFULL",
    "BLOSUM62",
    "BLOSUM62_GLOBAL",
    "PROTEIN_SEMIGLOBAL",
]

#: DNA local: match +2, mismatch -3, gap open -5, gap extend -2.
DNA = Scheme(mode="local", match=2, mismatch=-3, gap_open=-5, gap_extend=-2)
DNA_GLOBAL = DNA.replace(mode="global")
#: Overlap alignment: both sequences' end gaps are free.
DNA_SEMIGLOBAL = DNA.replace(mode="semiglobal")
#: blastn-style nucleotide defaults (same numbers as ``DNA``).
BLASTN = Scheme(mode="local", match=2, mismatch=-3, gap_open=-5, gap_extend=-2)
#: EMBOSS-style global nucleotide alignment with EDNAFULL, gaps -10/-1.
EDNAFULL = Scheme(mode="global", matrix=Matrix.named("EDNAFULL"), gap_open=-10, gap_extend=-1)
#: BLAST protein defaults: BLOSUM62, gaps -11/-1, local.
BLOSUM62 = Scheme(mode="local", matrix=Matrix.named("BLOSUM62"), gap_open=-11, gap_extend=-1)
BLOSUM62_GLOBAL = BLOSUM62.replace(mode="global")
PROTEIN_SEMIGLOBAL = BLOSUM62.replace(mode="semiglobal")
