"""FP64 CPU oracle for the flux-reconstruction residual-and-update path.

TEST INFRASTRUCTURE ONLY.  Nothing in the shipped package imports this
directory; only ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` may import or run it, and there only as
the checker (or the timed CPU baseline), never as the thing measured.

Every function restates the reference algorithm of AutoFlowCFD's Numba/NumPy
CPU path (the reference's own FP64 semantics) and cites the reference
``file:line`` it follows.  Pins (closed forms, invariants, brute force, figures
the reference states in its source) live in ``tests/test_oracle_*.py``.

Parity unpinned (no reference fixture and no closed form pins them beyond special
states; each says so in its docstring): the values of ``residual.inviscid_residual``
and ``residual.viscous_residual`` on general states, ``residual.suppress_outliers``
decisions on general fields, ``sst.source_terms``, ``sst.transport`` and
``sst.turbulence_step`` values, ``time_step.local_time_step`` values and the
``time_step.ns_step`` trajectory.
"""
