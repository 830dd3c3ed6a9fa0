"""Pins of the oracle's order continuation (oracle/continuation.py).

* The interpolation between orders (order_continuation.py:14-57) is linear interpolation on the
  tensor GL grid with linear extrapolation: it reproduces every trilinear field exactly, at the
  new points inside and outside the old grid, its rows sum to one, and P0 -> P1 broadcasts.
* The P0 finite-volume residual (inviscid_p0.py) preserves a uniform free stream, and it is
  conservative: sum_c V_c R_c equals minus the boundary flux sum_f F(Q_owner, Q_ghost, n) A_f,
  with the ghost state each owner cell keeps (the last of its boundary records); the ghost quirk
  itself is checked on a corner cell.
* The P0 viscous residual is the wall penalty only: zero for a uniform flow without walls, and for
  a uniform flow over a no-slip wall, at each wall cell, g'(+-1/2) (-eta |a| side) 2 u / det J.
"""

import numpy as np
import pytest

from cases import freestream, prim_to_cons, wall_box
from autoflowcfd_amd.device import bc_params_table
from autoflowcfd_amd.mesh import BC_WALL, hex_box
from oracle import continuation as OC
from oracle import fr_ops
from oracle.faces import build_flat_faces
from oracle.geometry import Geometry
from oracle.residual import C_IP, ausm_up, ghost_state


@pytest.mark.parametrize("old,new", [(0, 1), (1, 2), (2, 3)])
def test_interp_matrix_is_exact_for_trilinear_fields(old, new):
    W = OC.interp_matrix(old, new)
    xo = fr_ops.tensor_points(old + 1)
    xn = fr_ops.tensor_points(new + 1)
    assert np.allclose(W.sum(axis=1), 1.0, atol=1e-14)
    rng = np.random.default_rng(old)
    c = rng.standard_normal(8)

    def f(x):
        a, b, z = x[:, 0], x[:, 1], x[:, 2]
        return (c[0] + c[1] * a + c[2] * b + c[3] * z + c[4] * a * b + c[5] * b * z + c[6] * a * z
                + c[7] * a * b * z)
    if old == 0:
        assert np.allclose(W, 1.0)
    else:
        assert np.allclose(W @ f(xo), f(xn), atol=1e-12)


def _setup(mesh, order=2):
    ops = fr_ops.generate_operators(order)
    g = Geometry(mesh, ops)
    return OC.P0Geometry(mesh, g), build_flat_faces(mesh, ops)


def test_p0_inviscid_preserves_free_stream_and_is_conservative():
    mesh = wall_box(seed=2)
    g0, fl = _setup(mesh)
    Q0 = freestream(0.2, 4.0)
    bcp = bc_params_table(Q0)
    bcd = {c: bcp[c] for c in range(8)}
    U = np.broadcast_to(prim_to_cons(Q0), (mesh.n_cells, 1, 5)).copy()
    mesh_ff = hex_box(4, 4, 3, 4.0, 1.0, 1.0, split="prism", grading_y=1.1, seed=2)
    g0f, _ = _setup(mesh_ff)
    Rf = OC.p0_inviscid_residual(np.broadcast_to(prim_to_cons(Q0), (mesh_ff.n_cells, 1, 5)).copy(), mesh_ff, g0f, bcd)
    assert np.abs(Rf).max() < 1e-8 * Q0[4]
    rng = np.random.default_rng(1)
    Q = np.broadcast_to(Q0, (mesh.n_cells, 5)) * (1.0 + 0.02 * rng.standard_normal((mesh.n_cells, 5)))
    U = prim_to_cons(Q)[:, None, :]
    R = OC.p0_inviscid_residual(U, mesh, g0, bcd)
    fr = mesh.faces
    qg = {}
    for f in np.flatnonzero(fr.neighbor < 0):
        o = fr.owner[f]
        qg[o] = ghost_state(np.array([fr.bc[f]]), Q[o][None, None], fr.normal[f][None, None], bcd)[0, 0]
    flux = 0.0
    for f in np.flatnonzero(fr.neighbor < 0):
        o = fr.owner[f]
        flux = flux + ausm_up(Q[o][None], qg[o][None], fr.normal[f][None])[0] * fr.area[f]
    lhs = np.sum(g0.cell_volumes[:, None] * R[:, 0, :], axis=0)
    assert np.allclose(lhs, -flux, rtol=1e-10, atol=1e-9 * np.abs(flux).max())
    # the ghost quirk: a cell with boundary records of two types uses the last one's ghost for both
    bnd = np.flatnonzero(fr.neighbor < 0)
    owners, counts = np.unique(fr.owner[bnd], return_counts=True)
    corner = [c for c in owners if len(set(fr.bc[bnd[fr.owner[bnd] == c]])) > 1]
    assert corner, "the wall box has corner cells"


def test_p0_viscous_is_the_wall_penalty():
    mesh = wall_box(seed=3)
    g0, fl = _setup(mesh)
    Q0 = freestream(0.15, 0.0)
    bcp = bc_params_table(Q0)
    bcd = {c: bcp[c] for c in range(8)}
    mu = 1e-3
    U = np.broadcast_to(prim_to_cons(Q0), (mesh.n_cells, 1, 5)).copy()
    R = OC.p0_viscous_residual(U, mesh, g0, fl, bcd, mu)[:, 0, :]
    wall = np.flatnonzero(fl.is_boundary & fl.owner_primary & (fl.bc == BC_WALL))
    expect = np.zeros_like(R)
    for f in wall:
        c, ax, sd = fl.owner[f], fl.owner_axis[f], fl.owner_side[f]
        a = g0.adj[c, 0, ax]
        eta = C_IP * mu / abs(g0.det[c, 0]) ** (1.0 / 3.0)
        gp = -0.5 if sd < 0 else 0.5
        expect[c, 1:4] += gp * (-eta * np.linalg.norm(a) * sd) * (2.0 * Q0[1:4]) / g0.det[c, 0]
    assert np.allclose(R, expect, rtol=1e-12, atol=1e-12 * np.abs(expect).max())
    assert np.abs(expect).max() > 0
    # no wall: nothing
    m2 = hex_box(3, 2, 2, split="prism", seed=1)
    g2, f2 = _setup(m2)
    R2 = OC.p0_viscous_residual(np.broadcast_to(prim_to_cons(Q0), (m2.n_cells, 1, 5)).copy(), m2, g2, f2, bcd, mu)
    assert np.abs(R2).max() == 0.0


def test_continuation_phases_run_and_reach_the_target_order():
    mesh = wall_box(n=(3, 2, 2), seed=4)
    Q0 = freestream(0.15, 0.0)
    bcp = bc_params_table(Q0)
    U, hist = OC.run(mesh, 2, {c: bcp[c] for c in range(8)}, prim_to_cons(Q0), 1.8e-5, max_iter=30, tol=1e-12,
                     fr_ops=fr_ops)
    assert U.shape == (mesh.n_cells, 27, 5)
    assert sorted(hist) == [0, 1, 2] and all(len(h) == 10 for h in hist.values())
    assert np.all(np.isfinite(U))


def test_p0_transient_amplifies_round_off():
    """The conditioning behind the tolerance of the GPU continuation test: started from the uniform
    free stream, a relative change of 1e-15 in the initial state (round-off size) moves the 10th P0
    residual norm by more than 1e-8 relative, while the first norm stays within 1e-10."""
    mesh = wall_box(n=(3, 2, 2), seed=4)
    Q0 = freestream(0.15, 0.0)
    bcp = bc_params_table(Q0)
    bcd = {c: bcp[c] for c in range(8)}
    g0, fl = _setup(mesh)
    U0 = np.broadcast_to(prim_to_cons(Q0), (mesh.n_cells, 1, 5)).copy()
    rng = np.random.default_rng(0)
    hs = []
    for eps in (0.0, 1e-15):
        U = U0 * (1.0 + eps * rng.standard_normal(U0.shape))
        h = []
        for _ in range(10):
            U, r = OC.p0_step(U, mesh, g0, fl, bcd, 1.8e-5)
            h.append(r)
        hs.append(np.array(h))
    rel = np.abs(hs[1] - hs[0]) / hs[0]
    assert rel[0] < 1e-10
    assert rel[-1] > 1e-8


@pytest.mark.parametrize("kind", ["wallbox", "mixed", "rotated", "tets", "wedges"])
def test_p0_record_sources_equal_the_oracle_flat_faces(kind):
    """The product's P0 neighbour sources (continuation.p0_record_sources: one flux point per record,
    a prism quad's centre on its diagonal) equal the oracle's flat faces built with the one-point
    operators: source cells, and weights equal to the oracle's 1 x 1 interpolation rows."""
    from autoflowcfd_amd.continuation import p0_record_sources
    from cases import mixed_mesh, rotated_prism_box, tet_box, wedge_box
    mesh = {"wallbox": lambda: wall_box(seed=3), "mixed": mixed_mesh, "rotated": lambda: rotated_prism_box(seed=1),
            "tets": lambda: tet_box(3, seed=2), "wedges": wedge_box}[kind]()
    fl = build_flat_faces(mesh, fr_ops.generate_operators(0))
    src, wgt = p0_record_sources(mesh)
    assert np.array_equal(src[:, 0], fl.nb_src0_cell)
    assert np.allclose(wgt[:, 0], fl.nb_src0_mat[:, 0, 0], rtol=0, atol=1e-14)
    s1 = np.full(mesh.faces.n_faces, -1)
    w1 = np.zeros(mesh.faces.n_faces)
    has = fl.nb_src1_idx >= 0
    s1[has] = fl.nb_src1_cell[fl.nb_src1_idx[has]]
    w1[has] = fl.nb_src1_mat[fl.nb_src1_idx[has], 0, 0]
    assert np.array_equal(src[:, 1], s1)
    assert np.allclose(wgt[:, 1], w1, rtol=0, atol=1e-14)
    if kind in ("wallbox", "rotated"):
        assert has.any() and (wgt[:, 0] == 0).any()      # quads whose centre goes to the second source


def test_p0_operators_are_the_one_point_closed_forms():
    """generate_operators(0): zero derivatives, unit extrapolation and filter, g' = -1/2, +1/2 (the
    degree-1 Radau corrections (1 -+ xi) / 2 differentiated)."""
    ops = fr_ops.generate_operators(0)
    assert ops.n_sps == 1 and ops.n_fp == 1
    assert np.abs(ops.D).max() == 0.0
    assert np.allclose(ops.E, 1.0) and np.allclose(ops.filt, 1.0)
    assert np.allclose(ops.g_left, -0.5) and np.allclose(ops.g_right, 0.5)
