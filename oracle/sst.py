"""FP64 restatement of the reference SST k-omega path (test infrastructure).

  constants, F1, F2, nu_t, sources   src/autoflowcfd/core/turbulence/sst.py:21-330
  update_fields, positivity limiter  src/autoflowcfd/core/turbulence/sst.py:335-488
  turbulence source orchestration   src/autoflowcfd/core/fr_solver/turbulence.py:178-330
      (gradient of the *conserved* momentum for the strain rate, mu = 1.8e-5 in the
       sources, |grad k|, |grad omega| clipped at 1e6)
  transport residual                 src/autoflowcfd/core/turbulence/transport.py:122-520
  face extrapolation / distribution  src/autoflowcfd/core/turbulence/transport_kernel.py:103-327
      (record-major over every face record, no primary flags; Neumann ghost on boundary
       records, Dirichlet-zero mirror for k on WALL records; the owner-side correction is
       distributed with the owner's and with the neighbour's own flux-point maps)
  wall distance                      src/autoflowcfd/core/fr_solver/turbulence.py:112-121
      (nearest wall *node* of every solution point)

Only tests/, smoke() and bench.py's CPU baseline may use this module.
"""

import numpy as np

from .residual import physical_gradient, suppress_outliers
from .time_step import update_primitives

SIGMA_K1, SIGMA_K2, SIGMA_W1, SIGMA_W2 = 0.85, 1.0, 0.5, 0.856
BETA1, BETA2, A1, KAPPA, BETA_STAR = 0.075, 0.0828, 0.31, 0.41, 0.09
NUT_RATIO_MAX = 1.0e5
MU_SOURCE = 1.8e-5


def wall_distance(sps_xyz, wall_xyz):
    from scipy.spatial import cKDTree
    d, _ = cKDTree(wall_xyz).query(sps_xyz.reshape(-1, 3), k=1)
    return d.reshape(sps_xyz.shape[:-1])


def strain_rate(gv):
    S = 0.5 * (gv + np.swapaxes(gv, -1, -2))
    return np.sqrt(2.0 * np.einsum("...ij,...ij->...", S, S))


def blend_F1(k, w, d, nu, rho, CD):
    w = np.maximum(w, 1e-10)
    k = np.maximum(k, 1e-10)
    d = np.maximum(d, 1e-10)
    t1 = np.sqrt(k) / (BETA_STAR * w * d)
    t2 = 500.0 * nu / (d ** 2 * w)
    with np.errstate(over="ignore", invalid="ignore"):
        t3 = 4.0 * rho * SIGMA_W2 * k / (np.maximum(CD, 1e-10) * d ** 2)
    a = np.minimum(np.maximum(t1, t2), t3)
    a = np.where(np.isfinite(a), a, 1e75)
    a = np.minimum(a, 1e75)
    return np.tanh(a ** 4)


def blend_F2(k, w, d, nu):
    w = np.maximum(w, 1e-10)
    k = np.maximum(k, 1e-10)
    d = np.maximum(d, 1e-10)
    a = np.maximum(2.0 * np.sqrt(k) / (BETA_STAR * w * d), 500.0 * nu / (d ** 2 * w))
    a = np.where(np.isfinite(a), a, 1e150)
    a = np.minimum(a, 1e150)
    return np.tanh(a ** 2)


def eddy_viscosity(k, w, rho, S, F2, mu):
    k = np.maximum(k, 1e-10)
    w = np.maximum(w, 1e-10)
    S = np.maximum(S, 1e-10)
    with np.errstate(over="ignore", invalid="ignore"):
        nut = A1 * k / np.maximum(A1 * w, F2 * S)
    nut = np.minimum(nut, NUT_RATIO_MAX * mu / np.maximum(rho, 1e-10))
    return np.where(np.isfinite(nut), nut, 0.0)


def source_terms(Q, gv, d, mu, k_field, w_field, gk, gw):
    """(Sk, Somega, nu_t, beta_blend); sst.py:203-330.  Parity unpinned (no reference test pins
    the SST source values)."""
    rho = Q[..., 0]
    nu = mu / np.maximum(rho, 1e-10)
    k = np.minimum(k_field, 1e40)
    w = np.maximum(np.minimum(w_field, 1e40), 1e-10)
    S = strain_rate(gv)
    gdot = np.sum(gk * gw, axis=-1)
    CD = np.maximum(2.0 * rho * SIGMA_W2 / w * gdot, 1e-10)
    F1 = blend_F1(k, w, d, nu, rho, CD)
    F2 = blend_F2(k, w, d, nu)
    beta = F1 * BETA1 + (1.0 - F1) * BETA2
    nut = eddy_viscosity(k, w, rho, S, F2, mu)
    Pk = np.minimum(nut * rho * S ** 2, 10.0 * BETA_STAR * rho * k * w)
    Sk = Pk - rho * BETA_STAR * k * w
    g1 = BETA1 / BETA_STAR - SIGMA_W1 * KAPPA ** 2 / np.sqrt(BETA_STAR)
    g2 = BETA2 / BETA_STAR - SIGMA_W2 * KAPPA ** 2 / np.sqrt(BETA_STAR)
    gam = F1 * g1 + (1.0 - F1) * g2
    Sw = rho * gam * S ** 2 - rho * beta * w ** 2 + 2.0 * rho * (1.0 - F1) * SIGMA_W2 / w * gdot
    Sk = np.where(np.isfinite(Sk), Sk, 0.0)
    Sw = np.where(np.isfinite(Sw), Sw, 0.0)
    return Sk, Sw, nut, beta


def clipped_gradient(field, geom, ops, cap=1e6):
    g = physical_gradient(field[..., None], geom, ops)[:, :, 0, :]
    mag = np.linalg.norm(g, axis=-1)
    if np.any(mag > cap):
        g = g * np.clip(cap / np.maximum(mag, 1e-10), 0, 1)[..., None]
    return g


def _extrap_records(phi, ops, fl, wall_zero=None):
    """(phi_owner_fp, phi_neighbor_fp) (F, n_fp); transport_kernel.py:22-170."""
    F = fl.n_faces
    ct = (fl.owner >= fl.n_prism).astype(np.int64)
    E = ops.E[ct, fl.owner_axis, (fl.owner_side > 0).astype(np.int64)]
    po = np.einsum("fps,fs->fp", E, phi[fl.owner])
    pn = np.zeros((F, fl.n_fp))
    s0 = fl.nb_src0_cell >= 0
    pn[s0] = np.einsum("fps,fs->fp", fl.nb_src0_mat[s0], phi[fl.nb_src0_cell[s0]])
    s1 = np.flatnonzero(fl.nb_src1_idx >= 0)
    for f in s1:
        j = fl.nb_src1_idx[f]
        pn[f] += fl.nb_src1_mat[j] @ phi[fl.nb_src1_cell[j]]
    none = (fl.nb_src0_cell < 0) & (fl.nb_src1_idx < 0)
    wz = np.zeros(F, dtype=bool) if wall_zero is None else wall_zero
    pn[none & wz] = -po[none & wz]
    pn[none & ~wz] = po[none & ~wz]
    return po, pn


def _distribute_records(corr, ops, fl, det):
    """owner -= g' corr / det, neighbour += g' corr / det (owner fp order); transport_kernel.py:200-260."""
    C, ns = det.shape
    out = np.zeros((C, ns))
    for role, sgn in (("owner", -1.0), ("neighbor", 1.0)):
        cells = fl.owner if role == "owner" else fl.neighbor
        axis = fl.owner_axis if role == "owner" else fl.neighbor_axis
        side = fl.owner_side if role == "owner" else fl.neighbor_side
        sel = np.flatnonzero(cells >= 0)
        fp = ops.fp_of_sp[axis[sel]]
        co = ops.coord_of_sp[axis[sel]]
        g = np.where((side[sel] > 0)[:, None], ops.g_right[co], ops.g_left[co])
        contrib = g * np.take_along_axis(corr[sel], fp, axis=1)
        np.add.at(out, cells[sel], sgn * contrib / det[cells[sel]])
    return out


def _volume_div(Ft, geom, ops):
    """sum_m D_m Ft[..., m] per cell type, (C, ns)."""
    out = np.zeros(Ft.shape[:2])
    npr = geom.n_prism
    for ti, sl in ((0, slice(0, npr)), (1, slice(npr, geom.n_cells))):
        out[sl] = np.einsum("sjm,cjm->cs", ops.D[ti], Ft[sl])
    return out


def transport(k, w, nut, Q, geom, ops, fl, d, mu, wall_records):
    """(dk/dt, domega/dt) transport contributions; transport.py:122-520.  Pinned by uniform
    transport (zero); parity unpinned for general fields."""
    rho, vel = Q[..., 0], Q[..., 1:4]
    gk = clipped_gradient(k, geom, ops)
    gw = clipped_gradient(w, geom, ops)
    gdot = np.sum(gk * gw, axis=-1)
    ws = np.maximum(w, 1e-10)
    CD = np.maximum(2.0 * rho * SIGMA_W2 / ws * gdot, 1e-10)
    F1 = blend_F1(k, w, d, mu / np.maximum(rho, 1e-10), rho, CD)
    rnut = rho * nut
    gam_k = mu + (F1 * SIGMA_K1 + (1.0 - F1) * SIGMA_K2) * rnut
    gam_w = mu + (F1 * SIGMA_W1 + (1.0 - F1) * SIGMA_W2) * rnut
    n_true = fl.true_normal

    ro, _ = _extrap_records(rho, ops, fl)
    vo = np.stack([_extrap_records(vel[..., i], ops, fl)[0] for i in range(3)], axis=-1)
    mass = np.sum(ro[..., None] * vo * n_true, axis=-1)

    def one(phi, gam, wz):
        conv = -_volume_div(np.einsum("csmi,csi->csm", geom.adj, rho[..., None] * vel * phi[..., None]), geom, ops)
        gphi = physical_gradient(phi[..., None], geom, ops)[:, :, 0, :]
        diff = -_volume_div(np.einsum("csmi,csi->csm", geom.adj, gam[..., None] * gphi), geom, ops)
        po, pn = _extrap_records(phi, ops, fl, wz)
        go, gn = _extrap_records(gam, ops, fl)
        up = np.where(mass >= 0, po, pn)
        c_conv = mass * (up - po)
        c_diff = 0.5 * (go + gn) * (0.5 * (po + pn) - po)
        res = (conv / geom.det + _distribute_records(c_conv, ops, fl, geom.det) +
               diff / geom.det - _distribute_records(c_diff, ops, fl, geom.det))
        return res / np.maximum(rho, 1e-10)

    tk = one(k, gam_k, wall_records)
    tw = one(w, gam_w, None)
    tk = suppress_outliers(tk[..., None], k[..., None])[..., 0]
    tw = suppress_outliers(tw[..., None], w[..., None])[..., 0]
    return np.where(np.isfinite(tk), tk, 0.0), np.where(np.isfinite(tw), tw, 0.0)


def update_fields(k, w, dt, Sk_rho, Sw_rho, beta, tk, tw):
    """sst.py:382-479: point-implicit damping of the sources, explicit transport, floors."""
    ws = np.maximum(w, 1e-10)
    with np.errstate(over="ignore", invalid="ignore"):
        dk = Sk_rho / (1.0 + dt * BETA_STAR * ws)
        dw = Sw_rho / (1.0 + dt * beta * ws)
    dk = np.where(np.isfinite(dk), dk, 0.0) + tk
    dw = np.where(np.isfinite(dw), dw, 0.0) + tw
    dk = np.where(np.isfinite(dk), dk, 0.0)
    dw = np.where(np.isfinite(dw), dw, 0.0)
    k = k + dt * dk
    w = w + dt * dw
    k = np.where(np.isfinite(k), k, 1e-12)
    w = np.where(np.isfinite(w), w, 1e-12)
    return np.maximum(k, 1e-12), np.maximum(w, 1e-12)


def turbulence_step(U, k, w, geom, ops, fl, d, dt, mu_molecular, wall_records):
    """compute_turbulence_source (turbulence.py:178-330): returns (k_new, w_new, nu_t, Sk, Somega).
    Parity unpinned (no reference test pins the SST update values)."""
    Q = update_primitives(U)
    gU = physical_gradient(U[..., :5], geom, ops)
    gv = gU[:, :, 1:4, :]
    gk = clipped_gradient(k, geom, ops)
    gw = clipped_gradient(w, geom, ops)
    Sk, Sw, nut, beta = source_terms(Q, gv, d, MU_SOURCE, k, w, gk, gw)
    rho = Q[..., 0]
    tk, tw = transport(k, w, nut, Q, geom, ops, fl, d, mu_molecular, wall_records)
    kn, wn = update_fields(k, w, dt, Sk / np.maximum(rho, 1e-10), Sw / np.maximum(rho, 1e-10), beta, tk, tw)
    return kn, wn, nut, Sk, Sw
