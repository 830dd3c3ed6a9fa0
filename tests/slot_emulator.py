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
