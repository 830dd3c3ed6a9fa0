"""FP64 metric terms of straight-sided prisms/tets (host-side set-up, chunked).

Closed forms of src/autoflowcfd/grid/curved_mapping/curved_mapping_exact_jacobian.py:39-119
and the cofactor inverse of src/autoflowcfd/grid/curved_mapping/curved_mapping.py:39-111;
adj(J) = det(J) J^{-1} (src/autoflowcfd/core/fr_operators/volume_contract.py:48-80).
Metrics are computed in FP64 and rounded once to the device storage type.
"""

import numpy as np


def map_ref(is_prism, ref, nodes):
    """Physical points: ref (n,3) or (K,n,3), nodes (K,6|4,3) -> (K,n,3)."""
    if ref.ndim == 2:
        ref = np.broadcast_to(ref, (nodes.shape[0],) + ref.shape)
    a, b, c = ref[..., 0:1], ref[..., 1:2], ref[..., 2:3]
    P = nodes[:, None]
    if is_prism:
        r = (1.0 + a) * (1.0 - b) / 2.0 - 1.0
        s = b
        l1, l2, l3 = -(r + s) / 2.0, (1.0 + r) / 2.0, (1.0 + s) / 2.0
        bot = l1 * P[:, :, 0] + l2 * P[:, :, 1] + l3 * P[:, :, 2]
        top = l1 * P[:, :, 3] + l2 * P[:, :, 4] + l3 * P[:, :, 5]
        return 0.5 * (1.0 - c) * bot + 0.5 * (1.0 + c) * top
    t = c
    s = (1.0 + b) * (1.0 - c) / 2.0 - 1.0
    r = -(1.0 + a) * (s + t) / 2.0 - 1.0
    return (-(1.0 + r + s + t) / 2.0 * P[:, :, 0] + (1.0 + r) / 2.0 * P[:, :, 1]
            + (1.0 + s) / 2.0 * P[:, :, 2] + (1.0 + t) / 2.0 * P[:, :, 3])


def jacobian(is_prism, ref, nodes):
    """J[..., :, m] = dx/dxi_m; ref (n,3) or (K,n,3), nodes (K,6|4,3) -> (K,n,3,3)."""
    if ref.ndim == 2:
        ref = np.broadcast_to(ref, (nodes.shape[0],) + ref.shape)
    a, b, c = ref[..., 0:1], ref[..., 1:2], ref[..., 2:3]
    P = nodes[:, None]
    J = np.empty(ref.shape[:2] + (3, 3))
    if is_prism:
        p0, p1, p2, p3, p4, p5 = (P[:, :, i] for i in range(6))
        r = (1.0 + a) * (1.0 - b) / 2.0 - 1.0
        s = b
        l1, l2, l3 = -(r + s) / 2.0, (1.0 + r) / 2.0, (1.0 + s) / 2.0
        J[..., 0] = 0.5 * (1.0 - c) * ((1.0 - b) / 4.0 * (p1 - p0)) + 0.5 * (1.0 + c) * ((1.0 - b) / 4.0 * (p4 - p3))
        J[..., 1] = (0.5 * (1.0 - c) * (-(1.0 - a) / 4.0 * p0 - (1.0 + a) / 4.0 * p1 + 0.5 * p2)
                     + 0.5 * (1.0 + c) * (-(1.0 - a) / 4.0 * p3 - (1.0 + a) / 4.0 * p4 + 0.5 * p5))
        J[..., 2] = 0.5 * ((l1 * p3 + l2 * p4 + l3 * p5) - (l1 * p0 + l2 * p1 + l3 * p2))
        return J
    p0, p1, p2, p3 = (P[:, :, i] for i in range(4))
    e1, e2, e3 = (p1 - p0) / 2.0, (p2 - p0) / 2.0, (p3 - p0) / 2.0
    s = (1.0 + b) * (1.0 - c) / 2.0 - 1.0
    t = c
    J[..., 0] = -(s + t) / 2.0 * e1
    J[..., 1] = -(1.0 + a) * (1.0 - c) / 4.0 * e1 + (1.0 - c) / 2.0 * e2
    J[..., 2] = -(1.0 + a) * (1.0 - b) / 4.0 * e1 - (1.0 + b) / 2.0 * e2 + e3
    return J


def det_adj(J):
    """det(J) and adj(J) = cofactor^T, so that adj[..., m, i] = det * (J^{-1})[m, i]."""
    m = J
    adj = np.empty_like(J)
    adj[..., 0, 0] = m[..., 1, 1] * m[..., 2, 2] - m[..., 1, 2] * m[..., 2, 1]
    adj[..., 1, 0] = -(m[..., 1, 0] * m[..., 2, 2] - m[..., 1, 2] * m[..., 2, 0])
    adj[..., 2, 0] = m[..., 1, 0] * m[..., 2, 1] - m[..., 1, 1] * m[..., 2, 0]
    adj[..., 0, 1] = -(m[..., 0, 1] * m[..., 2, 2] - m[..., 0, 2] * m[..., 2, 1])
    adj[..., 1, 1] = m[..., 0, 0] * m[..., 2, 2] - m[..., 0, 2] * m[..., 2, 0]
    adj[..., 2, 1] = -(m[..., 0, 0] * m[..., 2, 1] - m[..., 0, 1] * m[..., 2, 0])
    adj[..., 0, 2] = m[..., 0, 1] * m[..., 1, 2] - m[..., 0, 2] * m[..., 1, 1]
    adj[..., 1, 2] = -(m[..., 0, 0] * m[..., 1, 2] - m[..., 0, 2] * m[..., 1, 0])
    adj[..., 2, 2] = m[..., 0, 0] * m[..., 1, 1] - m[..., 0, 1] * m[..., 1, 0]
    det = m[..., 0, 0] * adj[..., 0, 0] + m[..., 0, 1] * adj[..., 1, 0] + m[..., 0, 2] * adj[..., 2, 0]
    return det, adj


def cell_block_metrics(mesh, T, dtype=np.float32, chunk=65536):
    """det (C,ns), adj (C,ns,3,3), adj_fine (C,nf,3,3) and volumes, FP64-computed, stored ``dtype``."""
    C = mesh.n_cells
    det = np.empty((C, T.ns), dtype)
    adj = np.empty((C, T.ns, 3, 3), dtype)
    adjf = np.empty((C, T.nf, 3, 3), dtype)
    vol = np.empty(C, np.float64)
    det_mean = np.empty(C, np.float64)
    mfs = np.empty((C, T.ns), np.float64)
    gcl = np.empty((C, T.ns, 3), dtype)
    for is_prism, conn, off in ((True, mesh.prism_conn, 0), (False, mesh.tet_conn, mesh.n_prism)):
        for s0 in range(0, len(conn), chunk):
            sl = slice(s0, min(s0 + chunk, len(conn)))
            nodes = mesh.nodes[conn[sl]]
            d, a = det_adj(jacobian(is_prism, T.sps, nodes))
            df, af = det_adj(jacobian(is_prism, T.fine, nodes))
            if np.any(d <= 0) or np.any(df <= 0):
                raise ValueError("non-positive Jacobian determinant (curved_mapping.py:324-334)")
            g = slice(off + sl.start, off + sl.stop)
            det[g] = d
            adj[g] = a
            adjf[g] = af
            vol[g] = d @ T.w_sps
            det_mean[g] = d.mean(axis=1)
            mfs[g] = np.linalg.norm(a, axis=-1).sum(axis=-1)
            # constant-state divergence sum_{m,j} Op[s, m*nf+j] adj_fine[j, m, i] (FP64)
            ct = 0 if is_prism else 1
            Ak = np.transpose(af, (0, 2, 1, 3)).reshape(af.shape[0], 3 * T.nf, 3)
            gcl[g] = np.einsum("sk,cki->csi", T.Op[ct], Ak)
    return dict(det=det, adj=adj, adj_fine=adjf, volume=vol, det_mean=det_mean, metric_flux_scale=mfs, gcl=gcl)
