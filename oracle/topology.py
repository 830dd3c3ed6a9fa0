"""Face topology, boundary-group tags and periodic pairing, restated without the product.

Test infrastructure only: the product builds the same records vectorised on the host
(autoflowcfd_amd/mesh.py::build_topology) and by a device radix sort (afr_face_topology); this
module is the checker those are compared against, written as plain loops over Python dicts and
sets so that it shares no code, table or ordering trick with them.

  face_records        FaceExtractor.extract_faces_mixed
                      (src/autoflowcfd/grid/mesh_gen/extraction/face_extractor.py:296-376):
                      the 8 triangles of every prism (bottom vertices sorted, the top row carried
                      along; each quad split along the diagonal from its lower-index bottom
                      vertex, face_extraction_kernels.py:247-322, degenerate triangles of
                      collapsed wedges dropped, :310-318), the 4 of every tet (:325-342); one
                      record per distinct node triple in ascending (min, mid, max) order, owner
                      = first occurrence, neighbour = second (stable lexsort + scan, :344-386;
                      a third occurrence is an error, face_geometry_finalize.py:51-108);
                      centre = vertex mean, area vector = (p1 - p0) x (p2 - p0) / 2 of the sorted
                      triple, turned to point away from the owner's vertex-mean centre
                      (face_geometry_finalize.py:110-172), unit normal where area > 1e-12;
                      cube-face codes of both sides by node-set matching
                      (src/autoflowcfd/grid/connectivity/face_connectivity.py:53-76,127-231,
                      tables src/autoflowcfd/grid/curved_mapping/curved_mapping.py:227-242).
  tag_groups          tag_boundary_groups (face_connectivity.py:235-270): a boundary record takes
                      the code of every group holding its owner cell, the last such group
                      winning.
  pair_periodic       pair_periodic_boundary_faces (face_connectivity.py:273-425): candidates by
                      owner cell, kept when |n . t| >= 1 - 1e-6 (an error between 0.05 and that),
                      matched by centre + t within 1e-6 max(|t|, 1); the a-side record becomes
                      interior with the b-side owner and code and the translation, the b-side
                      record is dropped.  Matching here is by brute-force nearest distance.
"""

import numpy as np

# cube-face codes (face_connectivity.py:40) and local vertex sets (curved_mapping.py:227-242)
CODES = {"a=-1": 0, "a=+1": 1, "b=-1": 2, "b=+1": 3, "c=-1": 4, "c=+1": 5}
TET_FACES = {"a=-1": (0, 2, 3), "a=+1": (1, 2, 3), "b=-1": (0, 1, 3), "c=-1": (0, 1, 2)}
PRISM_FACES = {"c=-1": (0, 1, 2), "c=+1": (3, 4, 5), "a=-1": (0, 2, 5, 3), "a=+1": (1, 2, 5, 4),
               "b=-1": (0, 1, 4, 3)}


def _prism_triangles(cell):
    """The 8 boundary triangles of one prism (v0, v1, v2, w0, w1, w2), degenerate ones dropped."""
    bottom, top = [int(x) for x in cell[:3]], [int(x) for x in cell[3:]]
    order = sorted(range(3), key=lambda i: bottom[i])
    v0, v1, v2 = (bottom[i] for i in order)
    w0, w1, w2 = (top[i] for i in order)
    tris = [(v0, v1, v2), (w0, w1, w2), (v0, v1, w1), (v0, w0, w1), (v1, v2, w2), (v1, w1, w2),
            (v0, v2, w2), (v0, w0, w2)]
    return [t for t in tris if len(set(t)) == 3]


def _tet_triangles(cell):
    c = [int(x) for x in cell]
    return [(c[0], c[1], c[2]), (c[0], c[1], c[3]), (c[0], c[2], c[3]), (c[1], c[2], c[3])]


def _cube_code(cell_nodes, tri, table):
    face = set(tri)
    hits = []
    for key, local in table.items():
        cand = {int(cell_nodes[i]) for i in local}
        if (len(local) == 3 and cand == face) or (len(local) == 4 and face <= cand):
            hits.append(key)
    if len(hits) != 1:
        raise ValueError(f"ambiguous or missing cube-face match for face {tri}: {hits}")
    return CODES[hits[0]]


def face_records(nodes, prism_conn, tet_conn):
    """dict of owner, neighbor, owner_code, neighbor_code (int64), node_ids (F, 3) int64 sorted
    triples, normal (F, 3), area (F,), center (F, 3) FP64."""
    nodes = np.asarray(nodes, dtype=np.float64)
    prism_conn = np.asarray(prism_conn).reshape(-1, 6)
    tet_conn = np.asarray(tet_conn).reshape(-1, 4)
    n_prism = len(prism_conn)
    cells_of = {}
    for c, cell in enumerate(prism_conn):
        for t in _prism_triangles(cell):
            cells_of.setdefault(tuple(sorted(t)), []).append(c)
    for j, cell in enumerate(tet_conn):
        for t in _tet_triangles(cell):
            cells_of.setdefault(tuple(sorted(t)), []).append(n_prism + j)
    keys = sorted(cells_of)

    def conn_of(c):
        return (prism_conn[c], PRISM_FACES) if c < n_prism else (tet_conn[c - n_prism], TET_FACES)

    def centre_of(c):
        cn = conn_of(c)[0]
        return sum(nodes[int(i)] for i in cn) / float(len(cn))

    F = len(keys)
    out = {"owner": np.zeros(F, np.int64), "neighbor": np.full(F, -1, np.int64),
           "owner_code": np.zeros(F, np.int64), "neighbor_code": np.full(F, -1, np.int64),
           "node_ids": np.zeros((F, 3), np.int64), "normal": np.zeros((F, 3)), "area": np.zeros(F),
           "center": np.zeros((F, 3))}
    for f, key in enumerate(keys):
        cl = cells_of[key]
        if len(cl) > 2:
            raise ValueError(f"face {key} shared by {len(cl)} cells")
        owner = cl[0]
        out["owner"][f] = owner
        out["node_ids"][f] = key
        p0, p1, p2 = (nodes[k] for k in key)
        x = [(p0[d] + p1[d] + p2[d]) / 3.0 for d in range(3)]
        e1, e2 = p1 - p0, p2 - p0
        av = 0.5 * np.array([e1[1] * e2[2] - e1[2] * e2[1], e1[2] * e2[0] - e1[0] * e2[2],
                             e1[0] * e2[1] - e1[1] * e2[0]])
        if float(np.dot(av, np.array(x) - centre_of(owner))) < 0.0:
            av = -av
        a = float(np.sqrt((av[0] * av[0] + av[1] * av[1]) + av[2] * av[2]))   # norm along the row, in order
        out["center"][f] = x
        out["area"][f] = a
        if a > 1e-12:
            out["normal"][f] = av / a
        cn, table = conn_of(owner)
        out["owner_code"][f] = _cube_code(cn, key, table)
        if len(cl) == 2:
            out["neighbor"][f] = cl[1]
            cn, table = conn_of(cl[1])
            out["neighbor_code"][f] = _cube_code(cn, key, table)
    return out


def tag_groups(owner, neighbor, groups):
    """(F,) group code of every boundary record (-1: interior or in no group), and name -> code."""
    code = np.full(len(owner), -1, dtype=np.int64)
    names = {}
    for g, (name, cells) in enumerate(groups.items()):
        names[name] = g
        members = {int(c) for c in np.asarray(cells).ravel()}
        for f in range(len(owner)):
            if neighbor[f] < 0 and int(owner[f]) in members:
                code[f] = g
    return code, names


def pair_periodic(rec, groups, group_a, group_b, translation, tol_scale=1e-6):
    """Records (as face_records returns, plus "translation") with the group_a / group_b boundary
    records paired into interior records."""
    t = np.asarray(translation, dtype=np.float64)
    tn = float(np.sqrt(np.dot(t, t)))
    if tn < 1e-300:
        raise ValueError("zero periodic translation")
    d = t / tn
    F = len(rec["owner"])

    def candidates(name):
        members = {int(c) for c in np.asarray(groups[name]).ravel()}
        out = []
        for f in range(F):
            if rec["neighbor"][f] >= 0 or int(rec["owner"][f]) not in members:
                continue
            al = abs(float(np.dot(rec["normal"][f], d)))
            if al >= 1.0 - 1e-6:
                out.append(f)
            elif al > 0.05:
                raise ValueError(f"group {name}: record {f} neither parallel nor orthogonal to the translation")
        return out
    ia, ib = candidates(group_a), candidates(group_b)
    if len(ia) != len(ib):
        raise ValueError("periodic groups with different record counts")
    tol = tol_scale * max(tn, 1.0)
    match = {}
    for fa in ia:
        target = rec["center"][fa] + t
        dist = [float(np.sqrt(np.sum((rec["center"][fb] - target) ** 2))) for fb in ib]
        k = int(np.argmin(dist))
        if dist[k] > tol:
            raise ValueError(f"record {fa} has no periodic image within {tol:.3e}")
        match[fa] = ib[k]
    if len(set(match.values())) != len(ib):
        raise ValueError("periodic images matched twice")
    out = {k: np.array(v, copy=True) for k, v in rec.items()}
    if "translation" not in out:
        out["translation"] = np.zeros((F, 3))
    for fa, fb in match.items():
        out["neighbor"][fa] = rec["owner"][fb]
        out["neighbor_code"][fa] = rec["owner_code"][fb]
        out["translation"][fa] = t
    keep = np.ones(F, dtype=bool)
    keep[list(match.values())] = False
    return {k: v[keep] for k, v in out.items()}
