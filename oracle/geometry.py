"""FP64 geometry of straight-sided collapsed-coordinate prisms and tets (oracle).

Restates:
  * Duffy maps and physical maps -- src/autoflowcfd/grid/curved_mapping/curved_mapping.py:133-241
  * exact Jacobians              -- src/autoflowcfd/grid/curved_mapping/curved_mapping_exact_jacobian.py:39-119
  * closed-form det / inverse    -- src/autoflowcfd/grid/curved_mapping/curved_mapping.py:39-111
  * adj(J) = det * J^{-1}        -- src/autoflowcfd/core/fr_operators/volume_contract.py:48-80
  * GL-quadrature cell volumes   -- src/autoflowcfd/grid/high_order/high_order_mesh.py:335-354
  * metric flux scale            -- src/autoflowcfd/core/fr_solver/solver_geometry.py:18-32

Cells are stored prisms first then tets (high_order_mesh.py:111); every array
here is indexed by that global cell index.
"""

import numpy as np


def map_points(is_prism: bool, ref: np.ndarray, nodes: np.ndarray) -> np.ndarray:
    """ref (..., n, 3) with nodes (..., 6|4, 3) -> physical (..., n, 3).

    tet: curved_mapping.py:191-214, prism: curved_mapping.py:217-241.
    """
    a, b, c = ref[..., 0], ref[..., 1], ref[..., 2]
    if is_prism:
        r = (1.0 + a) * (1.0 - b) / 2.0 - 1.0
        s = b
        l1 = -(r + s) / 2.0
        l2 = (1.0 + r) / 2.0
        l3 = (1.0 + s) / 2.0
        p = nodes
        bot = l1[..., None] * p[..., 0:1, :] + l2[..., None] * p[..., 1:2, :] + l3[..., None] * p[..., 2:3, :]
        top = l1[..., None] * p[..., 3:4, :] + l2[..., None] * p[..., 4:5, :] + l3[..., None] * p[..., 5:6, :]
        z = c[..., None]
        return 0.5 * (1.0 - z) * bot + 0.5 * (1.0 + z) * top
    t = c
    s = (1.0 + b) * (1.0 - c) / 2.0 - 1.0
    r = -(1.0 + a) * (s + t) / 2.0 - 1.0
    L1 = -(1.0 + r + s + t) / 2.0
    L2 = (1.0 + r) / 2.0
    L3 = (1.0 + s) / 2.0
    L4 = (1.0 + t) / 2.0
    p = nodes
    return (L1[..., None] * p[..., 0:1, :] + L2[..., None] * p[..., 1:2, :]
            + L3[..., None] * p[..., 2:3, :] + L4[..., None] * p[..., 3:4, :])


def exact_jacobian(is_prism: bool, ref: np.ndarray, nodes: np.ndarray) -> np.ndarray:
    """J[..., :, m] = d x / d xi_m for cells nodes (C, k, 3) at ref (n, 3) -> (C, n, 3, 3).

    curved_mapping_exact_jacobian.py:39-71 (tet) and :74-119 (prism).
    """
    a = ref[None, :, 0]
    b = ref[None, :, 1]
    c = ref[None, :, 2]
    P = nodes[:, None, :, :]  # (C,1,k,3)
    C = nodes.shape[0]
    n = ref.shape[0]
    jac = np.zeros((C, n, 3, 3))
    if not is_prism:
        p0, p1, p2, p3 = P[:, :, 0], P[:, :, 1], P[:, :, 2], P[:, :, 3]
        e1 = (p1 - p0) / 2.0
        e2 = (p2 - p0) / 2.0
        e3 = (p3 - p0) / 2.0
        s = (1.0 + b) * (1.0 - c) / 2.0 - 1.0
        t = c
        ca = -(s + t) / 2.0
        jac[..., 0] = ca[..., None] * e1
        cb1 = -(1.0 + a) * (1.0 - c) / 4.0
        cb2 = (1.0 - c) / 2.0
        jac[..., 1] = cb1[..., None] * e1 + cb2[..., None] * e2
        cc1 = -(1.0 + a) * (1.0 - b) / 4.0
        cc2 = -(1.0 + b) / 2.0
        jac[..., 2] = cc1[..., None] * e1 + cc2[..., None] * e2 + e3
        return jac
    p0, p1, p2, p3, p4, p5 = (P[:, :, i] for i in range(6))
    dbda = ((1.0 - b) / 4.0)[..., None] * (p1 - p0)
    dtda = ((1.0 - b) / 4.0)[..., None] * (p4 - p3)
    dbdb = (-(1.0 - a) / 4.0)[..., None] * p0 + (-(1.0 + a) / 4.0)[..., None] * p1 + 0.5 * p2
    dtdb = (-(1.0 - a) / 4.0)[..., None] * p3 + (-(1.0 + a) / 4.0)[..., None] * p4 + 0.5 * p5
    r = (1.0 + a) * (1.0 - b) / 2.0 - 1.0
    s = b
    l1 = -(r + s) / 2.0
    l2 = (1.0 + r) / 2.0
    l3 = (1.0 + s) / 2.0
    bot = l1[..., None] * p0 + l2[..., None] * p1 + l3[..., None] * p2
    top = l1[..., None] * p3 + l2[..., None] * p4 + l3[..., None] * p5
    jac[..., 0] = 0.5 * (1.0 - c)[..., None] * dbda + 0.5 * (1.0 + c)[..., None] * dtda
    jac[..., 1] = 0.5 * (1.0 - c)[..., None] * dbdb + 0.5 * (1.0 + c)[..., None] * dtdb
    jac[..., 2] = 0.5 * (top - bot)
    return jac


def det_inv(J: np.ndarray):
    """Closed-form cofactor det / inverse, curved_mapping.py:39-111."""
    m = J
    c00 = m[..., 1, 1] * m[..., 2, 2] - m[..., 1, 2] * m[..., 2, 1]
    c01 = -(m[..., 1, 0] * m[..., 2, 2] - m[..., 1, 2] * m[..., 2, 0])
    c02 = m[..., 1, 0] * m[..., 2, 1] - m[..., 1, 1] * m[..., 2, 0]
    c10 = -(m[..., 0, 1] * m[..., 2, 2] - m[..., 0, 2] * m[..., 2, 1])
    c11 = m[..., 0, 0] * m[..., 2, 2] - m[..., 0, 2] * m[..., 2, 0]
    c12 = -(m[..., 0, 0] * m[..., 2, 1] - m[..., 0, 1] * m[..., 2, 0])
    c20 = m[..., 0, 1] * m[..., 1, 2] - m[..., 0, 2] * m[..., 1, 1]
    c21 = -(m[..., 0, 0] * m[..., 1, 2] - m[..., 0, 2] * m[..., 1, 0])
    c22 = m[..., 0, 0] * m[..., 1, 1] - m[..., 0, 1] * m[..., 1, 0]
    d = m[..., 0, 0] * c00 + m[..., 0, 1] * c01 + m[..., 0, 2] * c02
    inv = np.empty_like(J)
    inv[..., 0, 0] = c00 / d
    inv[..., 0, 1] = c10 / d
    inv[..., 0, 2] = c20 / d
    inv[..., 1, 0] = c01 / d
    inv[..., 1, 1] = c11 / d
    inv[..., 1, 2] = c21 / d
    inv[..., 2, 0] = c02 / d
    inv[..., 2, 1] = c12 / d
    inv[..., 2, 2] = c22 / d
    return d, inv


class Geometry:
    """Per-cell metric terms at coarse and fine points (FP64).

    Attributes (C = n_cells):
      det (C, n_sps), inv (C, n_sps, 3, 3), adj (C, n_sps, 3, 3) with adj[...,m,i]
      det_f, adj_f on the fine (over-integration) points, sps_xyz (C, n_sps, 3),
      volume (C,), metric_flux_scale (C, n_sps).
    """

    def __init__(self, mesh, ops):
        np_ = mesh.n_prism
        self.n_prism = np_
        self.n_cells = mesh.n_cells
        det_c, inv_c, det_f, inv_f, xyz = [], [], [], [], []
        for is_prism, conn in ((True, mesh.prism_conn), (False, mesh.tet_conn)):
            if conn is None or len(conn) == 0:
                continue
            nodes = mesh.nodes[conn]
            Jc = exact_jacobian(is_prism, ops.ref_sps, nodes)
            Jf = exact_jacobian(is_prism, ops.ref_fine, nodes)
            d, i = det_inv(Jc)
            df, if_ = det_inv(Jf)
            det_c.append(d); inv_c.append(i); det_f.append(df); inv_f.append(if_)
            xyz.append(map_points(is_prism, ops.ref_sps[None], nodes))
        self.det = np.concatenate(det_c)
        self.inv = np.concatenate(inv_c)
        self.det_f = np.concatenate(det_f)
        inv_f = np.concatenate(inv_f)
        if np.any(self.det <= 0) or np.any(self.det_f <= 0):
            raise ValueError("non-positive Jacobian determinant (curved_mapping.py:324-334)")
        self.adj = self.det[..., None, None] * self.inv
        self.adj_f = self.det_f[..., None, None] * inv_f
        self.sps_xyz = np.concatenate(xyz)
        self.volume = np.sum(self.det * ops.w_sps[None, :], axis=1)
        self.metric_flux_scale = np.sum(np.linalg.norm(self.adj, axis=-1), axis=-1)
