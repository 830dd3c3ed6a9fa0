"""Boundary groups, periodic pairing and the tabulated ghost-state provider, and adapters from the
reference's mesh objects.

The reference attaches boundary conditions to *groups*: ``BoundaryMap.groups`` maps a group name to
owner-cell indices (prisms first, then tets), ``bc_types`` maps it to a type name, and
``build_boundary_ghost_provider`` (src/autoflowcfd/core/fr_solver/boundary.py:52-145) turns them
into a ``BoundaryGhostStateProvider`` (src/autoflowcfd/boundary/fr_ghost_state.py:144-189): a
per-face ``group_code`` and ``code_to_config {code: {"type": ..., parameters}}``.  The provider is
a Python callable that cannot cross the C ABI; its tabulated form can.  Here every boundary record
gets a type code (autoflowcfd_amd.mesh.BC_*) and a row of an (8, 5) table of reference states,
one row per distinct (type, parameters), which is what the kernels read
(ghost_state in csrc/afr_physics.hpp).  Up to 8 distinct boundary states per mesh.

Adapters take the reference's objects by their attributes (no import of the reference):
``VolumeMeshData`` (``nodes.get_coordinates()``, ``prism_cells.connectivity``,
``cells.connectivity``, ``boundaries.groups / bc_types / parameters``) and ``HighOrderMesh``
(``_node_coords``, ``_fixed_prism_conn``, ``_fixed_tet_conn``, ``boundary_groups``,
``boundary_bc_types``, src/autoflowcfd/grid/high_order/high_order_mesh.py:86-93,154-175).
"""

from typing import Dict, Optional, Tuple

import numpy as np

from .mesh import (BC_FARFIELD, BC_INLET, BC_MOVING_WALL, BC_OUTLET, BC_SLIP_WALL, BC_SYMMETRY, BC_WALL,
                   N_BC_ROWS, FaceRecords, Mesh, is_wall, mesh_from_cells)

# boundary.py:73-79: the bc_types names the solver maps, and the provider config types
# (fr_ghost_state.py:165-189)
TYPE_MAP = {
    "WALL": ("WALL", {"is_no_slip": True}),
    "SLIP_WALL": ("WALL", {"is_no_slip": False}),
    "VELOCITY_INLET": ("INLET", {"Q_inlet": None}),
    "PRESSURE_OUTLET": ("OUTLET", {"p_outlet": None}),
    "SYMMETRY": ("SYMMETRY", {}),
}


def tag_boundary_groups(faces: FaceRecords, groups: Dict[str, np.ndarray]) -> Tuple[np.ndarray, Dict[str, int]]:
    """(group_code (F,), name_to_code) as src/autoflowcfd/grid/connectivity/face_connectivity.py:234-272:
    the groups are coded in their dict order, a boundary record takes the code of every group whose
    owner-cell list holds its owner (the last such group wins), unmatched and interior records -1."""
    code = np.full(faces.n_faces, -1, dtype=np.int32)
    name_to_code = {}
    bnd = np.flatnonzero(faces.neighbor < 0)
    owners = faces.owner[bnd]
    for c, (name, cells) in enumerate(groups.items()):
        name_to_code[name] = c
        code[bnd[np.isin(owners, np.asarray(cells))]] = c
    return code, name_to_code


def pair_periodic(faces: FaceRecords, groups, group_a: str, group_b: str, translation,
                  tol_scale: float = 1e-6) -> FaceRecords:
    """Merge the boundary records of ``group_a`` and ``group_b`` into interior records
    (src/autoflowcfd/grid/connectivity/face_connectivity.py:275-427): candidates are the boundary
    records owned by the group's cells whose normal is parallel to the translation (records of a
    corner cell's other faces are orthogonal and dropped; anything in between is an error); record a
    takes the matched b's owner as its neighbour and the translation, record b is removed."""
    t = np.asarray(translation, dtype=np.float64)
    tn = float(np.linalg.norm(t))
    if tn < 1e-300:
        raise ValueError("periodic translation must be nonzero")
    tdir = t / tn
    bnd = np.flatnonzero(faces.neighbor < 0)

    def on_plane(name):
        cand = bnd[np.isin(faces.owner[bnd], np.asarray(groups[name]))]
        align = np.abs(faces.normal[cand] @ tdir)
        par = align >= 1.0 - 1e-6
        if np.any(~par & (align > 0.05)):
            raise ValueError(f"periodic group {name!r}: a candidate face is neither parallel nor "
                             "orthogonal to the translation")
        return cand[par]

    ia, ib = on_plane(group_a), on_plane(group_b)
    if len(ia) != len(ib):
        raise ValueError(f"periodic pairing {group_a!r}<->{group_b!r}: {len(ia)} vs {len(ib)} faces")
    tol = tol_scale * max(tn, 1.0)
    if len(ia):
        from scipy.spatial import cKDTree
        dist, m = cKDTree(faces.center[ib]).query(faces.center[ia] + t[None, :], k=1)
        if len(set(m.tolist())) != len(ib) or np.any(dist > tol):
            raise ValueError(f"periodic pairing {group_a!r}<->{group_b!r}: faces do not match under the translation")
        fb = ib[m]
    else:
        fb = ib
    nb, nbc = faces.neighbor.copy(), faces.neighbor_code.copy()
    trans, bc = faces.translation.copy(), faces.bc.copy()
    nb[ia] = faces.owner[fb]
    nbc[ia] = faces.owner_code[fb]
    trans[ia] = t
    bc[ia] = -1
    keep = np.ones(faces.n_faces, dtype=bool)
    keep[fb] = False
    row = None if faces.bc_row is None else faces.bc_row.copy()
    if row is not None:
        row[ia] = -1
    return FaceRecords(owner=faces.owner[keep], neighbor=nb[keep], owner_code=faces.owner_code[keep],
                       neighbor_code=nbc[keep], normal=faces.normal[keep], area=faces.area[keep],
                       center=faces.center[keep], node_ids=faces.node_ids[keep], translation=trans[keep],
                       bc=bc[keep], bc_row=None if row is None else row[keep])


def _get_params(bmap_params, name):
    if bmap_params is None:
        return {}
    if callable(getattr(bmap_params, "get", None)):
        return bmap_params.get(name) or {}
    return {}


def apply_periodic_groups(faces: FaceRecords, groups, bc_types, parameters) -> FaceRecords:
    """Every PERIODIC group pair by its ``paired_with`` / ``translation`` parameters
    (face_connectivity.py:430-495, with the same consistency checks)."""
    names = [n for n, t in bc_types.items() if t == "PERIODIC"]
    done = set()
    for name in names:
        if name in done:
            continue
        p = _get_params(parameters, name)
        other, tr = p.get("paired_with"), p.get("translation")
        if other is None or tr is None:
            raise ValueError(f"periodic group {name!r} needs 'paired_with' and 'translation'")
        if bc_types.get(other) != "PERIODIC":
            raise ValueError(f"periodic group {name!r} is paired with {other!r}, which is not PERIODIC")
        q = _get_params(parameters, other)
        if q.get("paired_with") not in (None, name):
            raise ValueError(f"periodic pairing mismatch between {name!r} and {other!r}")
        if q.get("translation") is not None and not np.allclose(np.asarray(q["translation"], float), -np.asarray(tr, float)):
            raise ValueError(f"periodic translations of {name!r} and {other!r} are not opposite")
        faces = pair_periodic(faces, groups, name, other, tr)
        done.update((name, other))
    return faces


def configs_for_groups(name_to_code, bc_types, rho_inf, vel_inf, p_inf, bc_overrides=None):
    """(code_to_config, default_config) as src/autoflowcfd/core/fr_solver/boundary.py:52-145
    (without its synthetic-eddy inlet, which is not tabulatable): a group's override by name, else
    its bc_types entry through TYPE_MAP, else FARFIELD at the free stream."""
    Q_free = [rho_inf, vel_inf, 0.0, 0.0, p_inf]
    overrides = bc_overrides or {}
    code_to_config = {}
    for name, code in name_to_code.items():
        if name in overrides:
            code_to_config[code] = overrides[name]
            continue
        mapped, params = TYPE_MAP.get(bc_types.get(name, "FARFIELD"), ("FARFIELD", {"Q_free": Q_free}))
        params = dict(params)
        if mapped == "INLET":
            params["Q_inlet"] = Q_free
        elif mapped == "OUTLET":
            params["p_outlet"] = p_inf
        code_to_config[code] = {"type": mapped, **params}
    return code_to_config, {"type": "FARFIELD", "Q_free": Q_free}


def _type_and_state(cfg) -> Tuple[int, Tuple[float, ...]]:
    """(BC code, 5 reference values) of one provider config (fr_ghost_state.py:165-189)."""
    t = cfg["type"]
    zero = (0.0,) * 5
    if t == "WALL":
        if not cfg.get("is_no_slip", True):
            return BC_SLIP_WALL, zero
        vw = cfg.get("wall_velocity")
        if vw is None or not np.any(np.asarray(vw, dtype=np.float64)):
            return BC_WALL, zero
        v = np.asarray(vw, dtype=np.float64).ravel()
        return BC_MOVING_WALL, (0.0, float(v[0]), float(v[1]), float(v[2]), 0.0)
    if t == "FARFIELD":
        return BC_FARFIELD, tuple(float(x) for x in np.asarray(cfg["Q_free"], dtype=np.float64).ravel()[:5])
    if t == "INLET":
        if cfg.get("sem") is not None:
            raise NotImplementedError("synthetic-eddy inlets (InletSEMGhostState) are not tabulatable")
        q = np.asarray(cfg["Q_inlet"], dtype=np.float64)
        if q.ndim != 1:
            raise NotImplementedError("per-flux-point inlet states are not tabulatable")
        return BC_INLET, tuple(float(x) for x in q[:5])
    if t == "OUTLET":
        return BC_OUTLET, (0.0, 0.0, 0.0, 0.0, float(np.asarray(cfg["p_outlet"]).ravel()[-1]))
    if t == "SYMMETRY":
        return BC_SYMMETRY, zero
    raise ValueError(f"unknown boundary condition type {t!r}")


def tabulate_provider(faces: FaceRecords, group_code, code_to_config, default_config) -> np.ndarray:
    """Set faces.bc (type codes) and faces.bc_row on every boundary record from a provider's
    (group_code, code_to_config, default_config); returns the (8, 5) table of reference states."""
    group_code = np.asarray(group_code)
    if group_code.shape != (faces.n_faces,):
        raise ValueError("group_code must have one entry per face record")
    bnd = np.flatnonzero(faces.neighbor < 0)
    rows: Dict[Tuple[int, Tuple[float, ...]], int] = {}
    bc = faces.bc.copy()
    row = np.full(faces.n_faces, -1, dtype=np.int64)
    cache = {}
    for f in bnd:
        c = int(group_code[f])
        if c not in cache:
            cache[c] = _type_and_state(code_to_config.get(c, default_config))
        key = cache[c]
        if key not in rows:
            if len(rows) == N_BC_ROWS:
                raise ValueError(f"more than {N_BC_ROWS} distinct boundary states")
            rows[key] = len(rows)
        bc[f], row[f] = key[0], rows[key]
    table = np.zeros((N_BC_ROWS, 5))
    for (_, state), r in rows.items():
        table[r] = state
    faces.bc, faces.bc_row = bc, row
    return table


def remap_group_code(faces: FaceRecords, face_node_ids, group_code) -> np.ndarray:
    """A provider's group_code indexed by the reference's face order, re-indexed to these face
    records by their node triples (both sides key a triangle record by its sorted node ids)."""
    ref = np.sort(np.asarray(face_node_ids, dtype=np.int64)[:, :3], axis=1)
    lut = {tuple(k): int(c) for k, c in zip(ref.tolist(), np.asarray(group_code).tolist())}
    own = np.sort(faces.node_ids, axis=1)
    return np.array([lut.get(tuple(k), -1) for k in own.tolist()], dtype=np.int32)


def wall_nodes_of_groups(groups, bc_types, n_nodes, cell_conn) -> np.ndarray:
    """Wall nodes of the wall distance as src/autoflowcfd/cli/solve_wall_distance.py:30-60 takes
    them: for every WALL group, its indices are read as node ids, unless the largest is >= the node
    count, in which case they are cell ids and every node of those cells is taken."""
    out = set()
    for name, t in bc_types.items():
        if t != "WALL" or name not in groups:
            continue
        idx = np.asarray(groups[name], dtype=np.int64)
        if len(idx) == 0:
            continue
        if idx.max() >= n_nodes:
            for c in idx:
                if c < len(cell_conn):
                    out.update(int(n) for n in cell_conn[c] if 0 <= n < n_nodes)
        else:
            out.update(idx.tolist())
    return np.array(sorted(out), dtype=np.int64)


def _mesh_from_parts(nodes, prism_conn, tet_conn, groups, bc_types, parameters, topology="host"):
    nodes = np.asarray(nodes, dtype=np.float64)
    prisms = None if prism_conn is None or len(prism_conn) == 0 else np.asarray(prism_conn, dtype=np.int64)
    tets = None if tet_conn is None or len(tet_conn) == 0 else np.asarray(tet_conn, dtype=np.int64)
    mesh = mesh_from_cells(nodes, prisms, tets, default_bc=BC_FARFIELD, topology=topology)
    groups = {k: np.asarray(v) for k, v in (groups or {}).items()}
    bc_types = dict(bc_types or {})
    if groups:
        mesh.faces = apply_periodic_groups(mesh.faces, groups, bc_types, parameters)
    conn = []
    if prisms is not None:
        conn.extend(prisms.tolist())
    if tets is not None:
        conn.extend(tets.tolist())
    mesh.wall_nodes = wall_nodes_of_groups(groups, bc_types, len(nodes), conn)
    mesh.meta.update(boundary_groups=groups, boundary_bc_types=bc_types)
    # each record's type from its group (FARFIELD without one), as the solver maps them; the rows
    # stay the type codes until a solver tabulates the states with its free stream and overrides
    # (boundary_table)
    code, name_to_code = tag_boundary_groups(mesh.faces, groups)
    cfg, default = configs_for_groups(name_to_code, bc_types, 1.225, 0.0, 101325.0)
    tabulate_provider(mesh.faces, code, cfg, default)
    mesh.faces.bc_row = None
    return mesh


class _Params:
    def __init__(self, p):
        self.p = p or {}

    def get(self, name):
        return self.p.get(name)


def mesh_from_volume_mesh(volume_mesh, topology="host") -> Mesh:
    """Mesh of a reference ``VolumeMeshData`` (src/autoflowcfd/grid/high_order/high_order_mesh.py:111-140
    reads the same attributes), with its boundary groups, periodic pairs and wall nodes."""
    nodes = volume_mesh.nodes.get_coordinates()
    pc = getattr(volume_mesh, "prism_cells", None)
    prism = pc.connectivity if pc else None
    cells = getattr(volume_mesh, "cells", None)
    tets = cells.connectivity if cells is not None else None
    bm = getattr(volume_mesh, "boundaries", None)
    groups = getattr(bm, "groups", None) if bm is not None else None
    bc_types = getattr(bm, "bc_types", None) if bm is not None else None
    params = None
    if bm is not None:
        if callable(getattr(bm, "get_parameters", None)):
            params = _Params({n: bm.get_parameters(n) for n in (bc_types or {})})
        else:
            params = _Params(getattr(bm, "parameters", None))
    return _mesh_from_parts(nodes, prism, tets, groups, bc_types, params, topology)


def mesh_from_high_order_mesh(hom, parameters=None, topology="host") -> Mesh:
    """Mesh of a reference ``HighOrderMesh`` from its orientation-fixed connectivity and boundary
    groups (high_order_mesh.py:86-93,154-175); when it carries its face connectivity, the face node
    triples are kept (meta["reference_face_node_ids"]) so that a provider indexed by its faces maps."""
    mesh = _mesh_from_parts(hom._node_coords, hom._fixed_prism_conn, hom._fixed_tet_conn,
                            getattr(hom, "boundary_groups", None), getattr(hom, "boundary_bc_types", None),
                            _Params(parameters), topology)
    fc = getattr(hom, "face_connectivity", None)
    if fc is not None and getattr(fc, "face_node_ids", None) is not None:
        mesh.meta["reference_face_node_ids"] = np.asarray(fc.face_node_ids)
    return mesh


def as_mesh(obj, topology="host") -> Mesh:
    """This package's Mesh from a Mesh, a reference VolumeMeshData or a reference HighOrderMesh
    (face topology built by ``topology``: "host" or "gpu", the same records)."""
    if isinstance(obj, Mesh):
        return obj
    if hasattr(obj, "_node_coords"):
        return mesh_from_high_order_mesh(obj, topology=topology)
    if hasattr(obj, "nodes") and hasattr(getattr(obj, "nodes"), "get_coordinates"):
        return mesh_from_volume_mesh(obj, topology=topology)
    raise TypeError(f"cannot read a mesh from {type(obj).__name__}")


def boundary_table(mesh: Mesh, rho_inf, vel_inf, p_inf, bc_overrides=None, provider=None) -> Optional[np.ndarray]:
    """The (8, 5) reference-state table of a mesh whose boundary comes from groups or a provider,
    with its records' types and rows set; None for a mesh that carries plain type codes.

    provider: an object with the reference provider's ``group_code``, ``code_to_config`` and
    ``default_config`` (BoundaryGhostStateProvider); its group_code is indexed by the reference's
    face records, mapped here through meta["reference_face_node_ids"]."""
    if provider is not None and hasattr(provider, "group_code"):
        face_ids = mesh.meta.get("reference_face_node_ids")
        code = np.asarray(provider.group_code)
        if face_ids is not None:
            code = remap_group_code(mesh.faces, face_ids, code)
        elif len(code) != mesh.faces.n_faces:
            raise ValueError("provider.group_code does not index these face records; build the mesh with "
                             "mesh_from_high_order_mesh so that its face node triples are known")
        return tabulate_provider(mesh.faces, code, provider.code_to_config, provider.default_config)
    groups = mesh.meta.get("boundary_groups")
    if not groups:
        return None
    code, name_to_code = tag_boundary_groups(mesh.faces, groups)
    cfg, default = configs_for_groups(name_to_code, mesh.meta.get("boundary_bc_types", {}), rho_inf, vel_inf, p_inf,
                                      {k: v for k, v in (bc_overrides or {}).items() if isinstance(k, str)})
    return tabulate_provider(mesh.faces, code, cfg, default)


__all__ = ["tag_boundary_groups", "pair_periodic", "apply_periodic_groups", "configs_for_groups",
           "tabulate_provider", "remap_group_code", "wall_nodes_of_groups", "mesh_from_volume_mesh",
           "mesh_from_high_order_mesh", "as_mesh", "boundary_table", "is_wall", "BC_INLET", "BC_OUTLET"]
