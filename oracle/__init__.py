"""CPU oracle for the Gotoh scoring path -- TEST INFRASTRUCTURE, not product code.

Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s CPU-baseline leg may
import anything from this package, and only as the checker.  The product path
(``gpusw_amd``) never imports it and has no CPU fallback.

* ``oracle.gotoh``  -- pure-Python/NumPy restatement of the reference's
  ``src/gpusw/reference.py:38-151`` (int64, sentinel ``-(1<<30)``).
* ``oracle.native`` -- the same algorithm restated in C (``oracle/gotoh.c``), built by
  ``oracle/Makefile`` into ``oracle/liboracle.so``; used for larger parity cases and the
  timed CPU baseline.
* ``oracle.brute``  -- independent pin: explicit enumeration of every alignment path of
  tiny inputs (no dynamic programming).

Pinning (what checks each function):
* ``gotoh.score_codes``: reference known answers (tests/golden/known_answers.json),
  brute-force enumeration over all modes/flag combinations, closed forms (empty edges,
  identical sequences, all-gap ladders), symmetry invariants.
* ``native.score_*``: equality with ``gotoh.score_codes`` on seeded random inputs.
"""
