"""Mesh input of the CLI / run_steady: a .npz file or a built-in graded wall box.

.npz keys: nodes (N, 3) float, prism_conn (P, 6) and/or tet_conn (T, 4) int, and optionally
boundary_tris (B, 3) node ids with boundary_codes (B,) BC codes for the boundary triangles
(BC_WALL 0, FARFIELD 1, INLET 2, OUTLET 3, SYMMETRY 4, SLIP_WALL 5, EXTRAP 6); unlisted boundary
triangles get default_bc (EXTRAP).  "wallbox:N" builds the N^3-hex graded wall box split
into prisms (no-slip y = 0, inlet x = 0, outlet x = 4, symmetry elsewhere).
"""

import numpy as np

from .mesh import BC_EXTRAP, BC_INLET, BC_OUTLET, BC_SYMMETRY, BC_WALL, hex_box, mesh_from_cells

WALL_BOX_BC = {"y0": BC_WALL, "x0": BC_INLET, "x1": BC_OUTLET, "y1": BC_SYMMETRY, "z0": BC_SYMMETRY,
               "z1": BC_SYMMETRY}


def wall_box(n, length_unit=1.0, grading=1.1, seed=0):
    return hex_box(n, n, n, 4.0 * length_unit, length_unit, length_unit, split="prism", grading_y=grading,
                   seed=seed, bc_by_plane=WALL_BOX_BC)


def load_mesh(path, default_bc=BC_EXTRAP):
    if path.startswith("wallbox:"):
        return wall_box(int(path.split(":", 1)[1]))
    with np.load(path, allow_pickle=False) as f:
        nodes = f["nodes"]
        prisms = f["prism_conn"] if "prism_conn" in f else None
        tets = f["tet_conn"] if "tet_conn" in f else None
        m = mesh_from_cells(nodes, prisms, tets, default_bc=default_bc)
        if "boundary_tris" in f:
            tris = np.sort(np.asarray(f["boundary_tris"], dtype=np.int64), axis=1)
            codes = np.asarray(f["boundary_codes"], dtype=np.int64)
            fr = m.faces
            bnd = np.flatnonzero(fr.is_boundary)
            key = {tuple(t): c for t, c in zip(map(tuple, tris), codes)}
            for fi in bnd:
                c = key.get(tuple(fr.node_ids[fi]))
                if c is not None:
                    fr.bc[fi] = c
            wall = bnd[fr.bc[bnd] == BC_WALL]
            m.wall_nodes = np.unique(fr.node_ids[wall]) if len(wall) else np.zeros(0, np.int64)
    return m


def save_mesh(path, mesh):
    fr = mesh.faces
    bnd = np.flatnonzero(fr.is_boundary)
    np.savez(path, nodes=mesh.nodes, prism_conn=mesh.prism_conn, tet_conn=mesh.tet_conn,
             boundary_tris=fr.node_ids[bnd], boundary_codes=fr.bc[bnd])
