"""FP64 CPU oracle for the flux-reconstruction residual-and-update path.

TEST INFRASTRUCTURE ONLY.  Nothing in the shipped package imports this
directory; only ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` may import or run it, and there only as
the checker (or the timed CPU baseline), never as the thing measured.

Every function restates the reference algorithm of AutoFlowCFD's Numba/NumPy
CPU path (the reference's own FP64 semantics) and cites the reference
``file:line`` it follows.  Pins (closed forms, invariants, brute force) live
in ``tests/test_oracle_*.py``; functions no check pins are marked
"parity unpinned" in their docstring.
"""
