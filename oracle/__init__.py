"""FP64 CPU oracle for the batched MNDO/AM1/PM3 closed-shell SCF path.

TEST INFRASTRUCTURE ONLY.  Nothing in the product package imports this
directory; only ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` use it, and only as the checker or the
reported CPU baseline.

Every function cites the reference file:line it restates (paths relative to
the lanl/PYSEQM repository root).  Pinning status of each function is kept
in the module docstrings of ``oracle/seqm_ref.py``.
"""
