"""NumPy emulation of the device kernels' dataflow (slot form, folded operators).

Used by CPU tests to check the device-side set-up data (face slots, folded
operators, per-slot extrapolation tables) against the oracle without a GPU.
Mirrors autoflowcfd_amd/csrc/afr_residual.hip stage by stage.
"""

import numpy as np

from autoflowcfd_amd.face_slots import SLOT_CODES

GAMMA = 1.4


def prim(U):
    rho = np.maximum(U[..., 0], 1e-10)
    u, v, w, E = (U[..., i] / rho for i in (1, 2, 3, 4))
    return np.stack([rho, u, v, w, (GAMMA - 1) * rho * (E - 0.5 * (u * u + v * v + w * w))], -1)


def inviscid(mesh, T, geo, sl, U, bcp, flux, ausm, ghost, suppress):
    """flux/ausm/ghost: point functions (vectorised) from the oracle module."""
    C, ns, nfp = mesh.n_cells, T.ns, T.nfp
    Q = prim(U)
    R = np.zeros((C, ns, 5))
    for ct, (c0, c1) in enumerate(((0, mesh.n_prism), (mesh.n_prism, C))):
        if c1 <= c0:
            continue
        cells = np.arange(c0, c1)
        Qc = Q[cells]
        Qf = np.einsum("fs,csv->cfv", T.c2f[ct], Qc)
        Ft = np.matmul(geo["adj_fine"][cells], flux(Qf))  # (c,nf,3,5)
        Ftk = np.transpose(Ft, (0, 2, 1, 3)).reshape(len(cells), 3 * T.nf, 5)
        vol = np.einsum("sk,ckv->csv", T.Op[ct], Ftk)
        corr = np.zeros_like(vol)
        for si, code in enumerate(SLOT_CODES[ct]):
            ax, side = code // 2, (1.0 if code % 2 else -1.0)
            E = T.E[ct, ax, code % 2]
            qo = np.einsum("ps,csv->cpv", E, Qc)
            bc = sl.bc[cells, si]
            tn = sl.tn[cells, si]
            qn = np.zeros_like(qo)
            inner = bc == -1
            fp = np.arange(nfp)
            src = (sl.mask[cells, si][:, None] >> fp) & 1
            nbc = np.take_along_axis(sl.nb[cells, si], src, axis=1)
            blk = np.take_along_axis(sl.blk[cells, si], src, axis=1)
            rows = sl.rows[blk * nfp + fp]  # (c,nfp,ns)
            qn_int = np.einsum("cps,cpsv->cpv", rows, Q[np.maximum(nbc, 0)])
            qn[inner] = qn_int[inner]
            if np.any(~inner):
                qn[~inner] = ghost(bc[~inner], qo[~inner], np.repeat(tn[~inner][:, None], nfp, 1), bcp)
            ar = sl.adjrow[cells, si]
            mag = np.maximum(np.linalg.norm(ar, axis=-1), 1e-300)
            d = ar / mag[..., None] * side
            fb = ((sl.fallback[cells, si][:, None] >> fp) & 1).astype(bool)
            d = np.where(fb[..., None], tn[:, None, :], d)
            fc = ausm(qo, qn, d) * (mag * side)[..., None]
            jump = fc - np.einsum("cpi,cpiv->cpv", ar, flux(qo))
            g = (T.gR if side > 0 else T.gL)[T.coord_of_sp[ax]]
            corr += g[None, :, None] * jump[:, T.fp_of_sp[ax], :]
        det = geo["det"][cells][..., None]
        R[cells] = -vol / det - corr / det
    if suppress:
        med = np.median(np.abs(R), axis=1, keepdims=True)
        ref = np.maximum(np.maximum(med, 1e-9 * np.mean(np.abs(U), axis=1, keepdims=True)), 1e-300)
        R = np.where(np.abs(R) > 1e4 * ref, 0.0, R)
    return R


def inviscid_pert(mesh, T, geo, sl, U32, bcp, qref, gcl, flux, ausm, ghost, suppress):
    """Precision-faithful emulation of the perturbation form of the device kernel.

    FP32: delta states, all shared-operator products, metrics; FP64: point fluxes,
    AUSM+up, ghost states and the precomputed constant-state divergence ``gcl``.
    """
    f32, f64 = np.float32, np.float64
    C, ns, nfp = mesh.n_cells, T.ns, T.nfp
    Qd = (prim(U32.astype(f64)) - qref).astype(f32)          # delta primitive, FP32
    Fref = flux(qref[None])[0]                                 # (3,5) FP64
    R = np.zeros((C, ns, 5), f32)
    for ct, (c0, c1) in enumerate(((0, mesh.n_prism), (mesh.n_prism, C))):
        if c1 <= c0:
            continue
        cells = np.arange(c0, c1)
        dQ = Qd[cells]
        dqf = np.einsum("fs,csv->cfv", T.c2f[ct].astype(f32), dQ)                   # FP32
        dF = (flux(qref + dqf.astype(f64)) - Fref).astype(f32)                      # FP64 -> FP32
        Ft = np.matmul(geo["adj_fine"][cells].astype(f32), dF)
        Ftk = np.transpose(Ft, (0, 2, 1, 3)).reshape(len(cells), 3 * T.nf, 5)
        vol = np.einsum("sk,ckv->csv", T.Op[ct].astype(f32), Ftk)                  # FP32
        vol = vol + np.einsum("csi,iv->csv", gcl[cells].astype(f32), Fref.astype(f32))
        corr = np.zeros_like(vol)
        for si, code in enumerate(SLOT_CODES[ct]):
            ax, side = code // 2, (1.0 if code % 2 else -1.0)
            E = T.E[ct, ax, code % 2].astype(f32)
            dqo = np.einsum("ps,csv->cpv", E, dQ)
            bc = sl.bc[cells, si]
            tn = sl.tn[cells, si].astype(f32)
            qo = qref + dqo.astype(f64)
            fp = np.arange(nfp)
            src = (sl.mask[cells, si][:, None] >> fp) & 1
            nbc = np.take_along_axis(sl.nb[cells, si], src, axis=1)
            blk = np.take_along_axis(sl.blk[cells, si], src, axis=1)
            rows = sl.rows[blk * nfp + fp].astype(f32)
            dqn = np.einsum("cps,cpsv->cpv", rows, Qd[np.maximum(nbc, 0)])
            qn = qref + dqn.astype(f64)
            inner = bc == -1
            if np.any(~inner):
                qn[~inner] = ghost(bc[~inner], qo[~inner], np.repeat(tn[~inner][:, None].astype(f64), nfp, 1), bcp)
            ar = sl.adjrow[cells, si].astype(f32).astype(f64)
            mag = np.maximum(np.linalg.norm(ar, axis=-1), 1e-300)
            d = ar / mag[..., None] * side
            fb = ((sl.fallback[cells, si][:, None] >> fp) & 1).astype(bool)
            d = np.where(fb[..., None], tn[:, None, :].astype(f64), d)
            jump = (ausm(qo, qn, d) * (mag * side)[..., None] - np.einsum("cpi,cpiv->cpv", ar, flux(qo))).astype(f32)
            g = (T.gR if side > 0 else T.gL)[T.coord_of_sp[ax]].astype(f32)
            corr += g[None, :, None] * jump[:, T.fp_of_sp[ax], :]
        det = geo["det"][cells][..., None].astype(f32)
        R[cells] = -vol / det - corr / det
    if suppress:
        med = np.median(np.abs(R), axis=1, keepdims=True)
        ref = np.maximum(np.maximum(med, f32(1e-9) * np.mean(np.abs(U32), axis=1, keepdims=True)), f32(1.17549435e-38))
        R = np.where(np.abs(R) > f32(1e4) * ref, f32(0.0), R)
    return R


def gcl_table(T, geo, mesh):
    """Constant-state divergence per cell: G[c,s,i] = sum_{m,j} Op[s, m*nf+j] adj_fine[c,j,m,i] (FP64)."""
    C = mesh.n_cells
    G = np.zeros((C, T.ns, 3))
    for ct, (c0, c1) in enumerate(((0, mesh.n_prism), (mesh.n_prism, C))):
        if c1 <= c0:
            continue
        A = geo["adj_fine"][c0:c1]  # (c, nf, 3(m), 3(i))
        Ak = np.transpose(A, (0, 2, 1, 3)).reshape(c1 - c0, 3 * T.nf, 3)
        G[c0:c1] = np.einsum("sk,cki->csi", T.Op[ct], Ak)
    return G
