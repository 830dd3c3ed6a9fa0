"""FP64 CPU oracle for the flux-reconstruction residual-and-update path.

TEST INFRASTRUCTURE ONLY.  Nothing in the shipped package imports this
directory; only ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` may import or run it, and there only as
the checker (or the timed CPU baseline), never as the thing measured.

Every function restates the reference algorithm of AutoFlowCFD's Numba/NumPy
CPU path (the reference's own FP64 semantics) and cites the reference
``file:line`` it follows.  Pins (closed forms, invariants, brute force, figures
the reference states in its source) live in ``tests/test_oracle_*.py``.

Pinned beyond the reference's own special states (tests/test_oracle_pins.py,
tests/test_oracle_convergence.py, tests/test_oracle_conditioning.py):
``residual.inviscid_residual`` and ``residual.viscous_residual`` on discontinuous states by
per-cell discrete conservation (prisms and tets, interior and boundary faces) and pointwise by the
one-dimensional FR reduction on layered prisms, and on smooth states by convergence to exact
solutions (isentropic vortex: rate P; manufactured viscous field: rate P - 1, the reference's
unlifted BR1 gradient) on prisms; each check rejects planted mistakes.
``continuation`` (tests/test_oracle_continuation.py): the order interpolation reproduces every
trilinear field exactly (inside and outside the old points) and its rows sum to one; the P0
finite-volume residual preserves the free stream and is discretely conservative against the
boundary flux with the per-cell ghost quirk; the P0 viscous residual equals the closed-form wall
penalty g'(+-1/2)(-eta |a| side) 2u / det J and vanishes without walls.
``topology`` (tests/test_oracle_topology.py): the face records close every cell (outward area
vectors sum to zero: a missing, duplicated or wrongly split triangle or a wrong orientation
fails), with 8 records per prism and 4 per tet; planted mistakes (the other quad diagonal, owner
and neighbour swapped, first-group-wins tagging) are caught against the product.

Parity unpinned (no reference fixture and no closed form pins them; each says so in its
docstring): the pointwise distribution of the tet interface corrections (the collapsed tet scheme
does not converge under refinement, so only its per-cell sums are pinned),
``residual.suppress_outliers`` decisions on general fields beyond the reference's own test cases,
``sst.source_terms``, ``sst.transport`` and ``sst.turbulence_step`` values,
``time_step.local_time_step`` values and the ``time_step.ns_step`` trajectory, and so the
trajectory of ``continuation.run`` (its P0 time step reuses the local time step formula).
"""
