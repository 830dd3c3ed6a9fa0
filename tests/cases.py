"""Seeded meshes and states shared by the CPU and GPU tests."""

import numpy as np

from autoflowcfd_amd.mesh import (BC_INLET, BC_OUTLET, BC_SYMMETRY, BC_WALL, hex_box,
                                  mesh_from_cells)

GAMMA = 1.4
RHO, P_INF, A_INF = 1.225, 101325.0, np.sqrt(1.4 * 101325.0 / 1.225)


def mixed_mesh():
    """2 tets sharing a face + 2 prisms sharing a quad (tests/unit/test_fr_residual_inviscid.py:42-70)."""
    nodes = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1], [1, 1, 1],
                      [10, 0, 0], [11, 0, 0], [10, 1, 0], [10, 0, 1], [11, 0, 1], [10, 1, 1],
                      [9, -1, 0], [9, -1, 1]], dtype=float)
    return mesh_from_cells(nodes, np.array([[5, 6, 7, 8, 9, 10], [5, 7, 11, 8, 10, 12]]),
                           np.array([[0, 1, 2, 3], [1, 2, 3, 4]]))


WALL_BOX_BC = {"y0": BC_WALL, "x0": BC_INLET, "x1": BC_OUTLET, "y1": BC_SYMMETRY,
               "z0": BC_SYMMETRY, "z1": BC_SYMMETRY}


def wall_box(n=(4, 4, 3), L=(4.0, 1.0, 1.0), grading=1.1, seed=0):
    """Small version of the graded wall-bounded prism box (y-extruded prisms, no-slip y=0)."""
    return hex_box(*n, *L, split="prism", grading_y=grading, seed=seed, bc_by_plane=WALL_BOX_BC)


def odd_prism_mesh(seed=5):
    """A graded wall box of 3 x 2 x 2 hex with its last prism removed: 23 prisms (an odd count,
    so kernels that take two prisms per wavefront meet a lone last prism).  Boundary records on
    y = 0 are WALL, every other boundary record (the removed prism's faces included) FARFIELD."""
    from autoflowcfd_amd.mesh import BC_FARFIELD, mesh_from_cells
    src = wall_box(n=(3, 2, 2), seed=seed)
    m = mesh_from_cells(src.nodes, src.prism_conn[:-1], default_bc=BC_FARFIELD)
    fr = m.faces
    on_wall = fr.is_boundary & (np.abs(fr.center[:, 1]) < 1e-12)
    fr.bc[on_wall] = BC_WALL
    m.wall_nodes = np.unique(fr.node_ids[on_wall])
    return m


def rotated_prism_box(seed=0):
    """A 4 x 3 x 3 hex wall box split into prisms whose vertex order is rotated at random within
    each prism: which quad is a prism's collapsed-direction (b) face then varies from cell to cell,
    so cells meet up to four neighbours through their b faces (up to four slots with 9-wide
    neighbour rows; the unrotated box has at most one).  All boundaries FARFIELD."""
    from autoflowcfd_amd.mesh import BC_FARFIELD, mesh_from_cells
    src = wall_box(n=(4, 3, 3), seed=3)
    rng = np.random.default_rng(seed)
    pc = src.prism_conn.copy()
    for i, r in enumerate(rng.integers(0, 3, len(pc))):
        pc[i] = np.concatenate([np.roll(pc[i, :3], r), np.roll(pc[i, 3:], r)])
    return mesh_from_cells(src.nodes, pc, default_bc=BC_FARFIELD)


def tet_box(n=3, seed=0, displacement=0.1, periodic=False):
    return hex_box(n, n, n, split="tet", displacement=displacement, seed=seed, periodic=periodic)


def freestream(mach=0.15, yaw_deg=0.0):
    u = mach * A_INF
    y = np.deg2rad(yaw_deg)
    return np.array([RHO, u * np.cos(y), 0.0, u * np.sin(y), P_INF])


def prim_to_cons(Q):
    rho, u, v, w, p = (Q[..., i] for i in range(5))
    E = p / ((GAMMA - 1.0) * rho) + 0.5 * (u * u + v * v + w * w)
    return np.stack([rho, rho * u, rho * v, rho * w, rho * E], axis=-1)


def smooth_state(xyz, Q0, amp=1e-2, seed=0):
    """Freestream plus a seeded smooth perturbation of relative amplitude ``amp`` (conserved)."""
    rng = np.random.default_rng(seed)
    L = np.ptp(xyz.reshape(-1, 3), axis=0) + 1e-30
    x = (xyz - xyz.reshape(-1, 3).min(axis=0)) / L
    Q = np.broadcast_to(Q0, xyz.shape[:-1] + (5,)).copy()
    scale = np.array([Q0[0], np.abs(Q0[1:4]).max() + 1.0, np.abs(Q0[1:4]).max() + 1.0,
                      np.abs(Q0[1:4]).max() + 1.0, Q0[4]])
    for v in range(5):
        k = rng.uniform(0.5, 2.0, size=3)
        ph = rng.uniform(0, 2 * np.pi)
        Q[..., v] += amp * scale[v] * np.sin(2 * np.pi * (x @ k) + ph)
    return prim_to_cons(Q)


def wedge_box(n=(4, 3, 3), seed=1):
    """A graded wall box whose first prism layer has one side edge collapsed to a point: the top
    node above an interior wall vertex is merged into it, so the prisms around that edge become
    wedges with v_i == w_i (the cells the reference's boundary-layer extrusion leaves where a layer
    stops growing) and the prisms above them are stretched down to the wall.  Records on y = 0 are
    WALL, every other boundary record FARFIELD."""
    from autoflowcfd_amd.mesh import BC_FARFIELD, mesh_from_cells
    src = wall_box(n=n, seed=seed)
    pc = src.prism_conn.copy()
    y = src.nodes[:, 1]
    lo, hi = src.nodes.min(axis=0), src.nodes.max(axis=0)
    edges = sorted({(int(v), int(w)) for c in pc for v, w in zip(c[:3], c[3:]) if abs(y[v]) < 1e-12 and y[w] > 0
                    and lo[0] < src.nodes[v, 0] < hi[0] and lo[2] < src.nodes[v, 2] < hi[2]})
    v, w = edges[len(edges) // 2]
    m = mesh_from_cells(src.nodes, np.where(pc == w, v, pc), default_bc=BC_FARFIELD)
    fr = m.faces
    on_wall = fr.is_boundary & (np.abs(fr.center[:, 1]) < 1e-12)
    fr.bc[on_wall] = BC_WALL
    m.wall_nodes = np.unique(fr.node_ids[on_wall])
    return m
