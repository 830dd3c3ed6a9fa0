"""Synthetic hexahedral box meshes split into collapsed-coordinate prisms or tets,
and the face-record topology the FR residual consumes.

The FR element family of AutoFlowCFD is prisms and tets only
(src/autoflowcfd/grid/high_order/high_order_mesh.py:40,111), so a hex box is
split the way the reference's own validation meshes are:

  * 6 tets per hex on the v0-v6 diagonal   -- tests/validation/_channel_mesh.py:88-120
  * 2 prisms per hex, extruded along y     -- tests/validation/_channel_mesh.py:159-229

Topology follows FaceExtractor.extract_faces_mixed
(src/autoflowcfd/grid/mesh_gen/extraction/face_extractor.py:296-376): every
prism contributes 8 triangles (its three quads split on the global-node-id
diagonal rule, face_extraction_kernels.py:247-322), every tet its 4 faces;
records are deduplicated on the sorted node triple with a stable lexsort, the
first occurrence is the owner, and the record normal points away from the
owner centroid (face_geometry_finalize.py:92-160).  Cube-face codes per side
follow src/autoflowcfd/grid/connectivity/face_connectivity.py:53-231 with the
cube-face tables of src/autoflowcfd/grid/curved_mapping/curved_mapping.py:254-275.
Boundary faces are classified by the plane they lie on (face-exact, as
tests/validation/_channel_mesh.py:19-66 does) instead of by owner-cell group.
"""

from dataclasses import dataclass, field
from typing import Dict, Optional

import numpy as np

# cube-face code -> (axis, side); codes 0..5 = a=-1, a=+1, b=-1, b=+1, c=-1, c=+1
CUBE_FACE_AXIS = np.array([0, 0, 1, 1, 2, 2], dtype=np.int64)
CUBE_FACE_SIDE = np.array([-1.0, 1.0, -1.0, 1.0, -1.0, 1.0])
# local vertex indices of each real cube face (curved_mapping.py:254-275)
TET_CUBE_FACES = {0: (0, 2, 3), 1: (1, 2, 3), 2: (0, 1, 3), 4: (0, 1, 2)}
PRISM_CUBE_FACES = {4: (0, 1, 2), 5: (3, 4, 5), 0: (0, 2, 5, 3), 1: (1, 2, 5, 4), 2: (0, 1, 4, 3)}
PRISM_QUAD_CODES = (0, 1, 2)

# boundary-condition type codes (src/autoflowcfd/boundary/fr_ghost_state.py:144-189)
BC_INTERIOR = -1
BC_WALL = 0
BC_FARFIELD = 1
BC_INLET = 2
BC_OUTLET = 3
BC_SYMMETRY = 4
BC_SLIP_WALL = 5
BC_EXTRAP = 6   # ghost = owner trace (DefaultGhostProvider, src/autoflowcfd/core/fr_residual/inviscid.py:92-104)
BC_MOVING_WALL = 7  # no-slip wall with a wall velocity (fr_ghost_state.py:26-53, wall_velocity)
BC_NAMES = {BC_WALL: "WALL", BC_FARFIELD: "FARFIELD", BC_INLET: "INLET",
            BC_OUTLET: "OUTLET", BC_SYMMETRY: "SYMMETRY", BC_SLIP_WALL: "SLIP_WALL",
            BC_EXTRAP: "EXTRAPOLATE", BC_MOVING_WALL: "MOVING_WALL"}
N_BC_ROWS = 8   # rows of the per-case boundary-state table


def is_wall(bc):
    """No-slip wall records (resting or moving): force integral, k Dirichlet, wall distance."""
    bc = np.asarray(bc)
    return (bc == BC_WALL) | (bc == BC_MOVING_WALL)


@dataclass
class FaceRecords:
    """Triangle face records (one per unique node triple)."""
    owner: np.ndarray          # (F,) int64
    neighbor: np.ndarray       # (F,) int64, -1 on boundary
    owner_code: np.ndarray     # (F,) int64 cube-face code on the owner side
    neighbor_code: np.ndarray  # (F,) int64, -1 on boundary
    normal: np.ndarray         # (F,3) unit normal, owner -> neighbor / outward
    area: np.ndarray           # (F,)
    center: np.ndarray         # (F,3)
    node_ids: np.ndarray       # (F,3) sorted node triple
    translation: np.ndarray    # (F,3) periodic shift owner->neighbour image (0 if none)
    bc: np.ndarray             # (F,) int64 BC type code, BC_INTERIOR for interior records
    bc_row: Optional[np.ndarray] = None  # (F,) row of the boundary-state table (None: the type code)

    def rows(self) -> np.ndarray:
        """Row of the boundary-state table of every record (its type code unless set)."""
        return self.bc if self.bc_row is None else self.bc_row

    @property
    def n_faces(self) -> int:
        return int(self.owner.shape[0])

    @property
    def is_boundary(self) -> np.ndarray:
        return self.neighbor < 0


@dataclass
class Mesh:
    nodes: np.ndarray                 # (N,3) float64
    prism_conn: np.ndarray            # (n_prism,6) int64, orientation-fixed
    tet_conn: np.ndarray              # (n_tet,4) int64, orientation-fixed
    faces: Optional[FaceRecords] = None
    wall_nodes: Optional[np.ndarray] = None   # node ids on WALL planes
    meta: Dict = field(default_factory=dict)

    @property
    def n_prism(self) -> int:
        return int(self.prism_conn.shape[0])

    @property
    def n_tet(self) -> int:
        return int(self.tet_conn.shape[0])

    @property
    def n_cells(self) -> int:
        return self.n_prism + self.n_tet

    def cell_nodes(self, cell: int) -> np.ndarray:
        if cell < self.n_prism:
            return self.nodes[self.prism_conn[cell]]
        return self.nodes[self.tet_conn[cell - self.n_prism]]


# --------------------------------------------------------------------------
# orientation (src/autoflowcfd/grid/curved_mapping/curved_mapping_orientation.py:25-95)
# --------------------------------------------------------------------------

def _signed_tet_volume(p0, p1, p2, p3):
    return np.einsum("...i,...i->...", np.cross(p1 - p0, p2 - p0), p3 - p0) / 6.0


def fix_tet_orientation(conn: np.ndarray, nodes: np.ndarray) -> np.ndarray:
    conn = conn.copy()
    p = nodes[conn]
    neg = _signed_tet_volume(p[:, 0], p[:, 1], p[:, 2], p[:, 3]) < 0
    conn[neg, 1], conn[neg, 2] = conn[neg, 2].copy(), conn[neg, 1].copy()
    return conn


def fix_prism_orientation(conn: np.ndarray, nodes: np.ndarray) -> np.ndarray:
    conn = conn.copy()
    p = nodes[conn]
    v0, v1, v2, w0, w1, w2 = (p[:, i] for i in range(6))
    vol = (_signed_tet_volume(v0, v1, v2, w0) + _signed_tet_volume(v1, v2, w0, w1)
           + _signed_tet_volume(v2, w0, w1, w2))
    neg = vol < 0
    conn[neg, 1], conn[neg, 2] = conn[neg, 2].copy(), conn[neg, 1].copy()
    conn[neg, 4], conn[neg, 5] = conn[neg, 5].copy(), conn[neg, 4].copy()
    return conn


# --------------------------------------------------------------------------
# box generator
# --------------------------------------------------------------------------

def graded_coords(n: int, length: float, ratio: Optional[float]) -> np.ndarray:
    """n+1 node coordinates on [0,length], cell sizes growing by ``ratio`` away from 0."""
    if ratio is None or ratio == 1.0:
        return np.linspace(0.0, length, n + 1)
    j = np.arange(n + 1, dtype=np.float64)
    return length * (ratio ** j - 1.0) / (ratio ** n - 1.0)


def hex_box(nx, ny, nz, Lx=1.0, Ly=1.0, Lz=1.0, split="tet", grading_y=None,
            displacement=0.0, seed=0, permute=True, bc_by_plane=None,
            periodic=False, extrude="y", topology="host") -> Mesh:
    """Structured hex box split into tets or extruded prisms.

    Args:
        split: "tet" (6 per hex) or "prism" (2 per hex, extruded along ``extrude``).
        grading_y: geometric growth ratio of y spacing away from y=0 (wall-normal grading).
        displacement: relative amplitude of the seeded sinusoidal displacement of interior nodes
            (0.1 -> +-10% of the local spacing per axis).
        permute: apply a seeded random permutation to node ids and to cell order.
        bc_by_plane: {"x0","x1","y0","y1","z0","z1"} -> BC type code; default FARFIELD.
        periodic: True pairs the x0/x1, y0/y1 and z0/z1 planes into interior records (triply
            periodic box); a string of axis letters ("x", "xz") pairs only those planes, the others
            keep their ``bc_by_plane`` codes.
        extrude: prism extrusion axis, "y" (wall-normal, tests/validation/_channel_mesh.py:159-229)
            or "x" (along the periodic direction, tests/validation/_periodic_mesh.py:22-60: the
            x planes are then single triangle caps).
        topology: "host" (NumPy) or "gpu" (afr_face_topology): the same records bit for bit.
    """
    rng = np.random.default_rng(seed)
    xs = np.linspace(0.0, Lx, nx + 1)
    ys = graded_coords(ny, Ly, grading_y)
    zs = np.linspace(0.0, Lz, nz + 1)
    nx1, ny1, nz1 = nx + 1, ny + 1, nz + 1
    ii, jj, kk = np.meshgrid(np.arange(nx1), np.arange(ny1), np.arange(nz1), indexing="ij")
    gid_grid = ii + nx1 * jj + nx1 * ny1 * kk
    n_nodes = nx1 * ny1 * nz1
    nodes = np.zeros((n_nodes, 3))
    nodes[gid_grid.ravel(), 0] = xs[ii.ravel()]
    nodes[gid_grid.ravel(), 1] = ys[jj.ravel()]
    nodes[gid_grid.ravel(), 2] = zs[kk.ravel()]
    if displacement:
        hx = Lx / nx
        hy = np.diff(ys)
        hz = Lz / nz
        kvec = rng.uniform(0.5, 1.5, size=(3, 3))
        phase = rng.uniform(0.0, 2 * np.pi, size=3)
        interior = ((ii > 0) & (ii < nx) & (jj > 0) & (jj < ny) & (kk > 0) & (kk < nz)).ravel()
        g = gid_grid.ravel()[interior]
        x = nodes[g] / np.array([Lx, Ly, Lz])
        jloc = jj.ravel()[interior]
        hyl = np.minimum(hy[np.minimum(jloc, ny - 1)], hy[np.maximum(jloc - 1, 0)])
        h = np.stack([np.full_like(hyl, hx), hyl, np.full_like(hyl, hz)], axis=1)
        for ax in range(3):
            nodes[g, ax] += displacement * h[:, ax] * np.sin(2 * np.pi * (x @ kvec[ax]) + phase[ax])

    def gid(i, j, k):
        return i + nx1 * j + nx1 * ny1 * k

    I, J, K = np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij")
    I, J, K = I.ravel(), J.ravel(), K.ravel()
    prisms = np.zeros((0, 6), dtype=np.int64)
    tets = np.zeros((0, 4), dtype=np.int64)
    if split == "tet":
        v = [gid(I, J, K), gid(I + 1, J, K), gid(I + 1, J + 1, K), gid(I, J + 1, K),
             gid(I, J, K + 1), gid(I + 1, J, K + 1), gid(I + 1, J + 1, K + 1), gid(I, J + 1, K + 1)]
        tmpl = [(0, 1, 2, 6), (0, 2, 3, 6), (0, 3, 7, 6), (0, 7, 4, 6), (0, 4, 5, 6), (0, 5, 1, 6)]
        tets = np.stack([np.stack([v[a] for a in t], axis=1) for t in tmpl], axis=1).reshape(-1, 4)
    elif split == "prism":
        tri_a = [(0, 0), (1, 0), (1, 1)]
        tri_b = [(0, 0), (1, 1), (0, 1)]
        if extrude == "y":      # triangles in (x, z), layers in y
            corner = lambda dp, dq, de: gid(I + dp, J + de, K + dq)  # noqa: E731
        elif extrude == "x":    # triangles in (y, z), layers in x
            corner = lambda dp, dq, de: gid(I + de, J + dp, K + dq)  # noqa: E731
        else:
            raise ValueError(extrude)
        out = []
        for tri in (tri_a, tri_b):
            vv = [corner(dp, dq, 0) for dp, dq in tri]
            ww = [corner(dp, dq, 1) for dp, dq in tri]
            out.append(np.stack(vv + ww, axis=1))
        prisms = np.stack(out, axis=1).reshape(-1, 6)
    else:
        raise ValueError(split)

    if permute:
        perm_nodes = rng.permutation(n_nodes)  # old id -> new id
        new_nodes = np.empty_like(nodes)
        new_nodes[perm_nodes] = nodes
        nodes = new_nodes
        prisms = perm_nodes[prisms]
        tets = perm_nodes[tets]
        prisms = prisms[rng.permutation(len(prisms))]
        tets = tets[rng.permutation(len(tets))]

    prisms = fix_prism_orientation(prisms, nodes) if len(prisms) else prisms
    tets = fix_tet_orientation(tets, nodes) if len(tets) else tets
    mesh = Mesh(nodes=nodes, prism_conn=prisms.astype(np.int64), tet_conn=tets.astype(np.int64))
    mesh.meta.update(dict(nx=nx, ny=ny, nz=nz, L=(Lx, Ly, Lz), split=split, seed=seed,
                          grading_y=grading_y, displacement=displacement, extrude=extrude))
    build_faces(mesh, bc_by_plane or {}, (Lx, Ly, Lz), periodic=periodic, topology=topology)
    return mesh


# --------------------------------------------------------------------------
# face records
# --------------------------------------------------------------------------

def _prism_triangles(conn: np.ndarray) -> np.ndarray:
    """8 triangles per prism, quads split by the global-id diagonal rule.

    face_extraction_kernels.py:247-322: sort the bottom triangle by global id,
    carry the permutation to the top; quads (v0v1): (v0,v1,w1),(v0,w0,w1);
    (v1v2): (v1,v2,w2),(v1,w1,w2); (v0v2): (v0,v2,w2),(v0,w0,w2).
    """
    bottom = conn[:, 0:3]
    top = conn[:, 3:6]
    order = np.argsort(bottom, axis=1, kind="stable")
    r = np.arange(len(conn))[:, None]
    sb = bottom[r, order]
    st = top[r, order]
    v0, v1, v2 = sb[:, 0], sb[:, 1], sb[:, 2]
    w0, w1, w2 = st[:, 0], st[:, 1], st[:, 2]
    tris = np.stack([
        np.stack([v0, v1, v2], 1), np.stack([w0, w1, w2], 1),
        np.stack([v0, v1, w1], 1), np.stack([v0, w0, w1], 1),
        np.stack([v1, v2, w2], 1), np.stack([v1, w1, w2], 1),
        np.stack([v0, v2, w2], 1), np.stack([v0, w0, w2], 1),
    ], axis=1)
    return tris.reshape(-1, 3)


def _tet_triangles(conn: np.ndarray) -> np.ndarray:
    n0, n1, n2, n3 = conn[:, 0], conn[:, 1], conn[:, 2], conn[:, 3]
    tris = np.stack([np.stack([n0, n1, n2], 1), np.stack([n0, n1, n3], 1),
                     np.stack([n0, n2, n3], 1), np.stack([n1, n2, n3], 1)], axis=1)
    return tris.reshape(-1, 3)


def _resolve_codes(mesh: Mesh, cells: np.ndarray, tri: np.ndarray) -> np.ndarray:
    """Cube-face code of each (cell, triangle) pair (face_connectivity.py:53-76)."""
    codes = np.full(len(cells), -1, dtype=np.int64)
    tri_sorted = np.sort(tri, axis=1)
    is_prism = cells < mesh.n_prism
    for table, mask, conn, off in ((PRISM_CUBE_FACES, is_prism, mesh.prism_conn, 0),
                                   (TET_CUBE_FACES, ~is_prism, mesh.tet_conn, mesh.n_prism)):
        idx = np.flatnonzero(mask)
        if len(idx) == 0:
            continue
        cn = conn[cells[idx] - off]
        nmatch = np.zeros(len(idx), dtype=np.int64)
        for code, loc in table.items():
            corners = cn[:, list(loc)]
            if len(loc) == 3:
                hit = np.all(np.sort(corners, axis=1) == tri_sorted[idx], axis=1)
            else:
                hit = np.all((tri_sorted[idx][:, :, None] == corners[:, None, :]).any(axis=2), axis=1)
            codes[idx[hit]] = code
            nmatch += hit
        if np.any(nmatch != 1):
            raise RuntimeError("ambiguous or missing cube-face match (face_connectivity.py:53-76)")
    return codes


def collapsed_prisms(prism_conn: np.ndarray) -> np.ndarray:
    """(n_prism,) bool: prisms with one side edge collapsed to a point (v_i == w_i for exactly one
    i, every other vertex distinct): the wedges the reference's boundary-layer extrusion leaves
    where one bottom vertex stops growing (face_extraction_kernels.py:310-318)."""
    c = np.asarray(prism_conn).reshape(-1, 6)
    return np.sum(c[:, :3] == c[:, 3:], axis=1) == 1


def check_cells(mesh: Mesh) -> None:
    """Refuse cells without volume: a tet with a repeated vertex, a prism with a repeated vertex
    other than one collapsed side edge.  Collapsed-edge wedges are accepted: their two zero-area
    triangles are dropped from the face records as the reference drops them, so the collapsed quad
    keeps one triangle record and its slot one source."""
    for name, conn in (("prism", mesh.prism_conn), ("tet", mesh.tet_conn)):
        c = np.asarray(conn)
        if not c.size:
            continue
        c = c.reshape(len(c), -1)
        srt = np.sort(c, axis=1)
        n_rep = np.sum(srt[:, 1:] == srt[:, :-1], axis=1)
        ok = n_rep == 0
        if name == "prism":
            ok |= (n_rep == 1) & collapsed_prisms(c)
        bad = np.flatnonzero(~ok)
        if len(bad):
            raise ValueError(f"{len(bad)} {name} cell(s) with repeated vertices (first: cell {int(bad[0])}, "
                             f"nodes {c[bad[0]].tolist()}): only a prism with one collapsed side edge "
                             "(v_i == w_i) is accepted")


def topology_of(mesh: Mesh, topology="host") -> FaceRecords:
    """The face records of ``mesh``: "host" builds them with NumPy (build_topology), "gpu" on the
    device (topology.build_topology_gpu, the same records bit for bit)."""
    if topology == "gpu":
        from .topology import build_topology_gpu
        return build_topology_gpu(mesh)
    if topology != "host":
        raise ValueError(f"topology {topology!r}: 'host' or 'gpu'")
    return build_topology(mesh)


def mesh_from_cells(nodes, prism_conn=None, tet_conn=None, default_bc=BC_EXTRAP, topology="host") -> Mesh:
    """Mesh from arbitrary prism/tet connectivity; every boundary record gets ``default_bc``."""
    prisms = np.zeros((0, 6), np.int64) if prism_conn is None else np.asarray(prism_conn, np.int64)
    tets = np.zeros((0, 4), np.int64) if tet_conn is None else np.asarray(tet_conn, np.int64)
    nodes = np.asarray(nodes, dtype=np.float64)
    prisms = fix_prism_orientation(prisms, nodes) if len(prisms) else prisms
    tets = fix_tet_orientation(tets, nodes) if len(tets) else tets
    mesh = Mesh(nodes=nodes, prism_conn=prisms, tet_conn=tets)
    fr = topology_of(mesh, topology)
    fr.bc[fr.is_boundary] = default_bc
    mesh.faces = fr
    mesh.wall_nodes = np.zeros(0, np.int64)
    return mesh


def build_topology(mesh: Mesh) -> FaceRecords:
    """Deduplicated triangle records with owner/neighbour, cube-face codes and normals."""
    check_cells(mesh)
    tris = []
    owners = []
    if mesh.n_prism:
        pt = _prism_triangles(mesh.prism_conn)
        po = np.repeat(np.arange(mesh.n_prism), 8)
        # the zero-area triangles of collapsed-edge wedges are dropped (face_extraction_kernels.py:310-318)
        keep = (pt[:, 0] != pt[:, 1]) & (pt[:, 0] != pt[:, 2]) & (pt[:, 1] != pt[:, 2])
        tris.append(pt[keep])
        owners.append(po[keep])
    if mesh.n_tet:
        tris.append(_tet_triangles(mesh.tet_conn))
        owners.append(np.repeat(np.arange(mesh.n_tet) + mesh.n_prism, 4))
    tri = np.concatenate(tris)
    cell = np.concatenate(owners)
    keys = np.sort(tri, axis=1)
    order = np.lexsort((keys[:, 2], keys[:, 1], keys[:, 0]))  # stable
    k_sorted = keys[order]
    c_sorted = cell[order]
    t_sorted = tri[order]
    new = np.ones(len(order), dtype=bool)
    new[1:] = np.any(k_sorted[1:] != k_sorted[:-1], axis=1)
    uid = np.cumsum(new) - 1
    n_faces = int(uid[-1]) + 1
    first = np.flatnonzero(new)
    counts = np.bincount(uid, minlength=n_faces)
    if np.any(counts > 2):
        raise RuntimeError("face shared by more than two cells")
    owner = c_sorted[first]
    neighbor = np.full(n_faces, -1, dtype=np.int64)
    second = first[counts == 2] + 1
    neighbor[counts == 2] = c_sorted[second]
    node_ids = k_sorted[first]
    owner_tri = t_sorted[first]
    neigh_tri = np.zeros_like(owner_tri)
    neigh_tri[counts == 2] = t_sorted[second]

    p0, p1, p2 = (mesh.nodes[node_ids[:, i]] for i in range(3))
    center = (p0 + p1 + p2) / 3.0
    avec = 0.5 * np.cross(p1 - p0, p2 - p0)
    centroid = np.zeros((mesh.n_cells, 3))
    if mesh.n_prism:
        centroid[:mesh.n_prism] = mesh.nodes[mesh.prism_conn].mean(axis=1)
    if mesh.n_tet:
        centroid[mesh.n_prism:] = mesh.nodes[mesh.tet_conn].mean(axis=1)
    flip = np.einsum("ij,ij->i", avec, center - centroid[owner]) < 0
    avec[flip] *= -1.0
    area = np.linalg.norm(avec, axis=1)
    normal = avec / area[:, None]

    owner_code = _resolve_codes(mesh, owner, owner_tri)
    neighbor_code = np.full(n_faces, -1, dtype=np.int64)
    inner = neighbor >= 0
    neighbor_code[inner] = _resolve_codes(mesh, neighbor[inner], neigh_tri[inner])
    return FaceRecords(owner=owner, neighbor=neighbor, owner_code=owner_code, neighbor_code=neighbor_code,
                       normal=normal, area=area, center=center, node_ids=node_ids,
                       translation=np.zeros((n_faces, 3)), bc=np.full(n_faces, BC_INTERIOR, dtype=np.int64))


def build_faces(mesh: Mesh, bc_by_plane: Dict[str, int], L, periodic=False, tol_rel=1e-7, topology="host"):
    """Topology plus face-exact box-plane boundary classification and periodic pairing
    (``periodic``: True for all three axes, or a string of axis letters)."""
    fr = topology_of(mesh, topology)
    bnd = np.flatnonzero(fr.is_boundary)
    Lx, Ly, Lz = L
    tol = tol_rel * max(Lx, Ly, Lz, 1.0)
    plane = np.full(len(bnd), "", dtype=object)
    cb = fr.center[bnd]
    for name, ax, val in (("y0", 1, 0.0), ("y1", 1, Ly), ("z0", 2, 0.0), ("z1", 2, Lz),
                          ("x0", 0, 0.0), ("x1", 0, Lx)):
        on = (np.abs(cb[:, ax] - val) < tol) & (plane == "")
        plane[on] = name
    if np.any(plane == ""):
        raise RuntimeError("boundary face off every box plane")

    keep = np.ones(fr.n_faces, dtype=bool)
    axes = "xyz" if periodic is True else (periodic or "")
    paired = np.zeros(len(bnd), dtype=bool)
    for ax, (lo, hi), Lax in ((0, ("x0", "x1"), Lx), (1, ("y0", "y1"), Ly), (2, ("z0", "z1"), Lz)):
        if "xyz"[ax] not in axes:
            continue
        ia = bnd[plane == lo]
        ib = bnd[plane == hi]
        paired |= (plane == lo) | (plane == hi)
        shift = np.zeros(3)
        shift[ax] = Lax
        if len(ia) != len(ib):
            raise RuntimeError("periodic planes are not congruent")
        ka = np.round((fr.center[ia] + shift) / tol).astype(np.int64)
        kb = np.round(fr.center[ib] / tol).astype(np.int64)
        da = {tuple(k): i for k, i in zip(ka, ia)}
        for k, fb in zip(kb, ib):
            fa = da.get(tuple(k))
            if fa is None:
                raise RuntimeError("periodic pairing failed (face_connectivity.py:279-425)")
            fr.neighbor[fa] = fr.owner[fb]
            fr.neighbor_code[fa] = fr.owner_code[fb]
            fr.translation[fa] = shift
            keep[fb] = False
    rest = bnd[~paired]
    fr.bc[rest] = [bc_by_plane.get(p, BC_FARFIELD) for p in plane[~paired]]
    wall_planes = [p for p, t in bc_by_plane.items() if is_wall(t) and p[0] not in axes]
    wall_nodes = (np.unique(fr.node_ids[rest][np.isin(plane[~paired], wall_planes)])
                  if wall_planes else np.zeros(0, np.int64))
    mesh.wall_nodes = wall_nodes
    mesh.faces = FaceRecords(
        owner=fr.owner[keep], neighbor=fr.neighbor[keep], owner_code=fr.owner_code[keep],
        neighbor_code=fr.neighbor_code[keep], normal=fr.normal[keep], area=fr.area[keep],
        center=fr.center[keep], node_ids=fr.node_ids[keep], translation=fr.translation[keep], bc=fr.bc[keep],
        bc_row=None if fr.bc_row is None else fr.bc_row[keep])
    return mesh.faces


def concat_meshes(meshes):
    """Disjoint union of meshes as one mesh (the batched multi-case form).

    Cells are renumbered prisms first (case by case) then tets (case by case); face records
    are concatenated case by case with owner / neighbour remapped; node coordinates are kept
    exactly (cases may overlap in space: no geometric query mixes cases, the wall distance is
    taken case by case over mesh.meta["wall_off"]).  Returns (mesh, cell_case (C,), prism_off
    (n+1,), tet_off (n+1,), rec_off (n+1,)); the offsets are also in mesh.meta.
    """
    n = len(meshes)
    P = np.array([m.n_prism for m in meshes], dtype=np.int64)
    T = np.array([m.n_tet for m in meshes], dtype=np.int64)
    N = np.array([len(m.nodes) for m in meshes], dtype=np.int64)
    p_off = np.concatenate([[0], np.cumsum(P)])
    t_off = np.concatenate([[0], np.cumsum(T)])
    n_off = np.concatenate([[0], np.cumsum(N)])
    P_tot = int(p_off[-1])
    nodes, prisms, tets = [], [], []
    fields = {k: [] for k in ("owner", "neighbor", "owner_code", "neighbor_code", "normal", "area", "center",
                              "node_ids", "translation", "bc", "bc_row")}
    wall = []
    wall_off = [0]
    rec_off = [0]
    for i, m in enumerate(meshes):
        nodes.append(m.nodes)
        prisms.append(m.prism_conn + n_off[i])
        tets.append(m.tet_conn + n_off[i])

        def remap(c, i=i, m=m):
            c = np.asarray(c, dtype=np.int64)
            out = np.where(c < m.n_prism, p_off[i] + c, P_tot + t_off[i] + (c - m.n_prism))
            return np.where(c < 0, -1, out)

        fr = m.faces
        fields["owner"].append(remap(fr.owner))
        fields["neighbor"].append(remap(fr.neighbor))
        fields["owner_code"].append(fr.owner_code)
        fields["neighbor_code"].append(fr.neighbor_code)
        fields["normal"].append(fr.normal)
        fields["area"].append(fr.area)
        fields["center"].append(fr.center)
        fields["node_ids"].append(fr.node_ids + n_off[i])
        fields["translation"].append(fr.translation)
        fields["bc"].append(fr.bc)
        fields["bc_row"].append(fr.rows())
        rec_off.append(rec_off[-1] + fr.n_faces)
        if m.wall_nodes is not None and len(m.wall_nodes):
            wall.append(m.wall_nodes + n_off[i])
        wall_off.append(wall_off[-1] + (0 if m.wall_nodes is None else len(m.wall_nodes)))
    out = Mesh(nodes=np.concatenate(nodes), prism_conn=np.concatenate(prisms).reshape(-1, 6),
               tet_conn=np.concatenate(tets).reshape(-1, 4))
    out.faces = FaceRecords(**{k: np.concatenate(v) for k, v in fields.items()})
    out.wall_nodes = np.concatenate(wall) if wall else np.zeros(0, np.int64)
    cell_case = np.concatenate([np.repeat(np.arange(n), P), np.repeat(np.arange(n), T)]).astype(np.int32)
    rec_off = np.asarray(rec_off, dtype=np.int64)
    out.meta.update(n_cases=n, prism_off=p_off, tet_off=t_off, rec_off=rec_off,
                    wall_off=np.asarray(wall_off, dtype=np.int64))
    return out, cell_case, p_off, t_off, rec_off


def renumber_cells(mesh, prism_order=None, tet_order=None):
    """The same mesh with its cells stored in another order: new prism i is old prism
    prism_order[i], new tet i is old tet tet_order[i] (prisms stay first).  Face records keep
    their order with owner / neighbour renamed.  Returns (mesh, new_of_old (C,)).

    Storing cells in a locality order (e.g. the Morton order of their centroids) puts the
    neighbour data that the interface kernels gather into the same cache lines and pages."""
    P, T = mesh.n_prism, mesh.n_tet
    po = np.arange(P) if prism_order is None else np.asarray(prism_order, dtype=np.int64)
    to = np.arange(T) if tet_order is None else np.asarray(tet_order, dtype=np.int64)
    if sorted(po.tolist()) != list(range(P)) or sorted(to.tolist()) != list(range(T)):
        raise ValueError("renumber_cells needs permutations of the prisms and of the tets")
    new_of_old = np.empty(P + T, dtype=np.int64)
    new_of_old[po] = np.arange(P)
    new_of_old[P + to] = P + np.arange(T)
    out = Mesh(nodes=mesh.nodes, prism_conn=np.asarray(mesh.prism_conn).reshape(-1, 6)[po],
               tet_conn=np.asarray(mesh.tet_conn).reshape(-1, 4)[to])
    fr = mesh.faces
    nb = np.asarray(fr.neighbor)
    out.faces = FaceRecords(owner=new_of_old[np.asarray(fr.owner)], neighbor=np.where(nb >= 0, new_of_old[np.maximum(nb, 0)], -1),
                            owner_code=fr.owner_code, neighbor_code=fr.neighbor_code, normal=fr.normal, area=fr.area,
                            center=fr.center, node_ids=fr.node_ids, translation=fr.translation, bc=fr.bc,
                            bc_row=fr.bc_row)
    out.wall_nodes = mesh.wall_nodes
    out.meta = dict(mesh.meta)
    return out, new_of_old


def _spread3(x):
    """Insert two zero bits between the low 21 bits of x (uint64)."""
    x = x & np.uint64(0x1FFFFF)
    x = (x | (x << np.uint64(32))) & np.uint64(0x1F00000000FFFF)
    x = (x | (x << np.uint64(16))) & np.uint64(0x1F0000FF0000FF)
    x = (x | (x << np.uint64(8))) & np.uint64(0x100F00F00F00F00F)
    x = (x | (x << np.uint64(4))) & np.uint64(0x10C30C30C30C30C3)
    x = (x | (x << np.uint64(2))) & np.uint64(0x1249249249249249)
    return x


def morton_order(points):
    """Permutation of the points along a 3-D Morton (Z-order) curve of their bounding box."""
    if len(points) == 0:
        return np.zeros(0, dtype=np.int64)
    lo, hi = points.min(axis=0), points.max(axis=0)
    q = ((points - lo) / np.maximum(hi - lo, 1e-300) * (2 ** 21 - 1)).astype(np.uint64)
    key = _spread3(q[:, 0]) | (_spread3(q[:, 1]) << np.uint64(1)) | (_spread3(q[:, 2]) << np.uint64(2))
    return np.argsort(key, kind="stable")


def locality_renumbered(mesh):
    """The mesh with its prisms and its tets each stored in the Morton order of their centroids
    (mesh.renumber_cells); returns (mesh, new_of_old)."""
    P = mesh.n_prism
    po = morton_order(mesh.nodes[np.asarray(mesh.prism_conn).reshape(-1, 6)].mean(axis=1)) if P else None
    to = morton_order(mesh.nodes[np.asarray(mesh.tet_conn).reshape(-1, 4)].mean(axis=1)) if mesh.n_tet else None
    return renumber_cells(mesh, po, to)
