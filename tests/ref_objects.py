"""Stand-ins shaped like the reference's mesh and boundary objects, for the adapter tests.

They carry exactly the attributes the reference's own test meshes build
(tests/validation/_channel_mesh.py:69-229, tests/validation/_periodic_mesh.py:22-95): a
``VolumeMeshData``-like namespace (``nodes.get_coordinates()``, ``cells.connectivity``,
``prism_cells.connectivity``, ``boundaries.groups / bc_types``), a ``HighOrderMesh``-like namespace
(``_node_coords``, ``_fixed_prism_conn``, ``_fixed_tet_conn``, ``boundary_groups``,
``boundary_bc_types``, ``face_connectivity.face_node_ids``) and a ``BoundaryGhostStateProvider``-like
namespace (``group_code``, ``code_to_config``, ``default_config``).  Nothing of the reference is
imported.
"""

from types import SimpleNamespace

import numpy as np


class Nodes:
    def __init__(self, coords):
        self._coords = coords

    def get_coordinates(self):
        return self._coords


class Cells:
    def __init__(self, connectivity):
        self.connectivity = connectivity


class Boundaries:
    def __init__(self, groups, bc_types, parameters=None):
        self.groups = groups
        self.bc_types = bc_types
        self.parameters = parameters or {}

    def get_parameters(self, name):
        return self.parameters.get(name, {})


def _grid(nx, ny, nz, Lx, H, Lz):
    nx1, ny1, nz1 = nx + 1, ny + 1, nz + 1
    ii, jj, kk = np.meshgrid(np.arange(nx1), np.arange(ny1), np.arange(nz1), indexing="ij")
    order = np.argsort((ii + nx1 * jj + nx1 * ny1 * kk).ravel())
    nodes = np.column_stack([(ii.ravel() / nx * Lx)[order], (jj.ravel() / ny * H)[order],
                             (kk.ravel() / nz * Lz)[order]])
    return nodes, (lambda i, j, k: i + nx1 * j + nx1 * ny1 * k)


def _groups_on_planes(coords, Lx, H, Lz):
    tol = 1e-9 * max(Lx, H, Lz, 1.0)

    def on(c, v):
        return np.flatnonzero((np.abs(c - v) < tol).sum(axis=1) >= 3)
    return {"wall_bottom": on(coords[:, :, 1], 0.0), "wall_top": on(coords[:, :, 1], H),
            "z_min": on(coords[:, :, 2], 0.0), "z_max": on(coords[:, :, 2], Lz),
            "x_min": on(coords[:, :, 0], 0.0), "x_max": on(coords[:, :, 0], Lx)}


def channel_volume_mesh(nx, ny, nz, Lx, H, Lz, bc_types=None):
    """Prisms extruded along y over (x, z) triangles, cell-index boundary groups on the six planes
    (the layout of tests/validation/_channel_mesh.py:159-229)."""
    nodes, gid = _grid(nx, ny, nz, Lx, H, Lz)
    prisms = []
    for i in range(nx):
        for k in range(nz):
            for tri in ([(i, k), (i + 1, k), (i + 1, k + 1)], [(i, k), (i + 1, k + 1), (i, k + 1)]):
                for j in range(ny):
                    prisms.append([gid(a, j, b) for a, b in tri] + [gid(a, j + 1, b) for a, b in tri])
    conn = np.array(prisms, dtype=np.int32)
    groups = _groups_on_planes(nodes[conn], Lx, H, Lz)
    types = bc_types or {n: "WALL" for n in groups}
    return SimpleNamespace(cell_count=len(conn), node_count=len(nodes), nodes=Nodes(nodes),
                           cells=Cells(np.zeros((0, 4), dtype=np.int32)), prism_cells=Cells(conn),
                           boundaries=Boundaries(groups, types))


def periodic_channel_volume_mesh(nx, ny, nz, Lx, H, Lz, side="SYMMETRY"):
    """Prisms extruded along x, x_min / x_max a PERIODIC pair with translation (Lx, 0, 0)
    (the layout of tests/validation/_periodic_mesh.py:22-95)."""
    nodes, gid = _grid(nx, ny, nz, Lx, H, Lz)
    prisms = []
    for j in range(ny):
        for k in range(nz):
            for tri in ([(j, k), (j + 1, k), (j + 1, k + 1)], [(j, k), (j + 1, k + 1), (j, k + 1)]):
                for i in range(nx):
                    prisms.append([gid(i, a, b) for a, b in tri] + [gid(i + 1, a, b) for a, b in tri])
    conn = np.array(prisms, dtype=np.int32)
    groups = _groups_on_planes(nodes[conn], Lx, H, Lz)
    types = {"wall_bottom": side, "wall_top": side, "z_min": side, "z_max": side,
             "x_min": "PERIODIC", "x_max": "PERIODIC"}
    params = {"x_min": {"paired_with": "x_max", "translation": [Lx, 0.0, 0.0]}}
    return SimpleNamespace(cell_count=len(conn), node_count=len(nodes), nodes=Nodes(nodes),
                           cells=Cells(np.zeros((0, 4), dtype=np.int32)), prism_cells=Cells(conn),
                           boundaries=Boundaries(groups, types, params))


def face_exact_provider(faces, L, config_by_plane, default=None):
    """Provider whose group_code classifies every boundary record by the plane its centre lies on
    (the face-exact provider of tests/validation/_channel_mesh.py:19-66), over ``faces`` records
    (anything with n_faces, neighbor and center)."""
    Lx, H, Lz = L
    tol = 1e-6 * max(Lx, H, Lz, 1.0)
    code = np.full(faces.n_faces, -1, dtype=np.int32)
    name_to_code, code_to_config = {}, {}
    for f in np.flatnonzero(faces.neighbor < 0):
        c = faces.center[f]
        for name, ax, v in (("wall_bottom", 1, 0.0), ("wall_top", 1, H), ("z_min", 2, 0.0), ("z_max", 2, Lz),
                            ("x_min", 0, 0.0), ("x_max", 0, Lx)):
            if abs(c[ax] - v) < tol:
                break
        else:
            raise ValueError(f"boundary face centre {c} off every plane")
        if name not in name_to_code:
            name_to_code[name] = len(name_to_code)
            code_to_config[name_to_code[name]] = config_by_plane[name]
        code[f] = name_to_code[name]
    return SimpleNamespace(group_code=code, code_to_config=code_to_config,
                           default_config=default or {"type": "FARFIELD", "Q_free": [1.225, 0.0, 0.0, 0.0, 101325.0]})
