"""Independent geometry for the oracle pins (test helper).

Everything here is computed from the node coordinates and the cell connectivity alone:
faces of every cell with their outward area vectors, neighbours by matching node sets,
boundary planes of a box, and physical face quadrature.  None of it reads the face records of
``autoflowcfd_amd.mesh`` or the flux-point tables of the oracle, so an identity that holds
between these quantities and an oracle residual pins the oracle's assembly (record pairing,
sides, correction functions, metric scaling), not only its agreement with itself.
"""

import numpy as np

# local faces of the reference element vertex orders (cyclic vertex order on every face)
PRISM_FACES = ((0, 1, 2), (3, 4, 5), (0, 1, 4, 3), (1, 2, 5, 4), (2, 0, 3, 5))
TET_FACES = ((0, 1, 2), (0, 1, 3), (0, 2, 3), (1, 2, 3))


def cell_nodes(mesh, c):
    if c < mesh.n_prism:
        return mesh.prism_conn[c]
    return mesh.tet_conn[c - mesh.n_prism]


def cell_faces(mesh):
    """Per cell: list of dicts {key, verts (cyclic), area, normal (outward unit), centre}."""
    out = []
    for c in range(mesh.n_cells):
        cn = cell_nodes(mesh, c)
        tmpl = PRISM_FACES if c < mesh.n_prism else TET_FACES
        cc = mesh.nodes[cn].mean(axis=0)
        lst = []
        for loc in tmpl:
            verts = cn[list(loc)]
            p = mesh.nodes[verts]
            if len(loc) == 3:
                av = 0.5 * np.cross(p[1] - p[0], p[2] - p[0])
            else:
                av = 0.5 * np.cross(p[2] - p[0], p[3] - p[1])
            fc = p.mean(axis=0)
            if np.dot(av, fc - cc) < 0:
                av = -av
            A = float(np.linalg.norm(av))
            lst.append(dict(key=tuple(sorted(int(v) for v in verts)), verts=verts, area=A,
                            normal=av / A, centre=fc))
        out.append(lst)
    return out


def neighbours(faces):
    """Per cell and face: the other cell sharing the same node set, or -1 (boundary)."""
    users = {}
    for c, lst in enumerate(faces):
        for f in lst:
            users.setdefault(f["key"], []).append(c)
    nb = []
    for c, lst in enumerate(faces):
        row = []
        for f in lst:
            others = [o for o in users[f["key"]] if o != c]
            if len(others) > 1:
                raise ValueError("non-manifold face")
            row.append(others[0] if others else -1)
        nb.append(row)
    return nb


def box_plane(centre, L, tol=1e-9):
    """Name of the box plane ("x0", "x1", "y0", ...) a boundary face centre lies on."""
    for ax, name in enumerate("xyz"):
        if abs(centre[ax]) < tol:
            return name + "0"
        if abs(centre[ax] - L[ax]) < tol:
            return name + "1"
    raise ValueError(f"boundary face centre {centre} on no box plane")


def face_quadrature(p):
    """Physical quadrature of a planar face with cyclic vertices p (3|4, 3): points, weights that
    sum to 1 (multiply by the area).  Triangle: edge midpoints (degree 2); quadrilateral
    (parallelogram): 3x3 Gauss on the bilinear map (degree 5)."""
    if len(p) == 3:
        L = np.array([[0.5, 0.5, 0.0], [0.0, 0.5, 0.5], [0.5, 0.0, 0.5]])
        return L @ p, np.full(3, 1.0 / 3.0)
    x, wx = np.polynomial.legendre.leggauss(3)
    pts, ws = [], []
    for i, s in enumerate(x):
        for j, t in enumerate(x):
            N = np.array([(1 - s) * (1 - t), (1 + s) * (1 - t), (1 + s) * (1 + t), (1 - s) * (1 + t)]) / 4.0
            pts.append(N @ p)
            ws.append(wx[i] * wx[j] / 4.0)
    return np.array(pts), np.array(ws)


def radau_derivatives(order, x):
    """Derivatives of the left and right Radau correction polynomials of degree n = order + 1 at x:
    g_L = (-1)^n/2 (P_n - P_{n-1}) (g_L(-1) = 1, g_L(1) = 0), g_R = (P_n + P_{n-1})/2."""
    from numpy.polynomial import legendre as L
    n = order + 1
    Pn = np.zeros(n + 1); Pn[n] = 1.0
    Pm = np.zeros(n + 1); Pm[n - 1] = 1.0
    dPn = L.legval(x, L.legder(Pn))
    dPm = L.legval(x, L.legder(Pm))
    return (-1.0) ** n / 2.0 * (dPn - dPm), 0.5 * (dPn + dPm)
