"""FP64 restatement of the wall force integral (test infrastructure).

src/autoflowcfd/postprocess/fr_coefficients.py:27-152: over owner-primary WALL face
records, F = sum_fp p n w - sum_fp (tau . n) w with p, grad u and mu_t extrapolated
from the owner cell (state primitives with the floors of state.py:76-97), n the
record's unit normal repeated over the flux points and w = record area x normalised
tensor Gauss weights (face_flux_points_merge.py:60-64,182-184).  Cd = F_x / (q A),
Cl = F_z / (q A), Cs = F_y / (q A), q = rho_inf V_inf^2 / 2 (no yaw rotation).

Only tests/, smoke() and bench.py's CPU baseline may use this module.
"""

import numpy as np

from .residual import BC_MOVING_WALL, BC_WALL, physical_gradient
from .time_step import update_primitives


def wall_forces(U, geom, ops, fl, mu, mu_t=None, include_viscous=True):
    """(F_pressure (3,), F_viscous (3,))."""
    Q = update_primitives(U[..., :5])
    x, w = np.polynomial.legendre.leggauss(ops.n1d)
    rel = np.outer(w, w).ravel()
    rel = rel / rel.sum()
    gQ = physical_gradient(Q, geom, ops) if include_viscous else None
    fp_, fv_ = np.zeros(3), np.zeros(3)
    sel = np.flatnonzero(((fl.bc == BC_WALL) | (fl.bc == BC_MOVING_WALL)) & fl.owner_primary)
    for f in sel:
        c = fl.owner[f]
        ti = 0 if c < fl.n_prism else 1
        E = ops.E[ti, fl.owner_axis[f], int(fl.owner_side[f] > 0)]
        Qf = E @ Q[c]
        n = fl.true_normal[f]
        aw = fl.area[f] * rel
        fp_ += np.sum(Qf[:, 4:5] * n * aw[:, None], axis=0)
        if include_viscous:
            gv = np.einsum("ps,sij->pij", E, gQ[c, :, 1:4, :])
            mt = mu + (E @ mu_t[c] if mu_t is not None else 0.0)
            S = 0.5 * (gv + np.swapaxes(gv, -1, -2))
            div = gv[:, 0, 0] + gv[:, 1, 1] + gv[:, 2, 2]
            mt = np.broadcast_to(mt, div.shape)
            tau = 2.0 * mt[:, None, None] * S + (-2.0 / 3.0 * mt * div)[:, None, None] * np.eye(3)
            fv_ += -np.sum(np.einsum("pij,pj->pi", tau, n) * aw[:, None], axis=0)
    return fp_, fv_


def coefficients(F, rho_inf, vel_inf, area):
    q = 0.5 * rho_inf * vel_inf ** 2
    d = max(q * area, 1e-300)
    return {"Cd": F[0] / d, "Cl": F[2] / d, "Cs": F[1] / d}
