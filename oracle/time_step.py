"""FP64 restatement of one steady pseudo-step of the reference FR solver (test infrastructure).

  update_primitives     src/autoflowcfd/core/fr_solver/state.py:76-97 (rho >= 1e-10, p >= 1, k, w >= 1e-12)
  residual_norm         src/autoflowcfd/core/fr_solver/state.py:99-107 (RMS over all n_vars components)
  local_time_step       src/autoflowcfd/core/fr_solver/cfl.py:12-147
  enforce_positivity    src/autoflowcfd/core/time_integration/positivity.py:13-45
  apply_filter          src/autoflowcfd/core/fr_solver/filter.py:29-62 (first 5 variables)
  ssp_rk3               src/autoflowcfd/core/time_integration/base.py:51-57,250-335
  ns_step               src/autoflowcfd/core/fr_solver/step.py:15-221 (laminar / frozen mu_t path)

Only tests/, smoke() and bench.py's CPU baseline may use this module.
"""

import numpy as np

from .residual import inviscid_residual, viscous_residual

GAMMA = 1.4
CFL = 0.1
SSP_RK3 = {"alpha": [[1.0], [0.75, 0.25], [1.0 / 3.0, 0.0, 2.0 / 3.0]], "beta": [1.0, 0.25, 2.0 / 3.0]}


def update_primitives(U):
    rho = np.maximum(U[..., 0], 1e-10)
    u, v, w = U[..., 1] / rho, U[..., 2] / rho, U[..., 3] / rho
    ke = 0.5 * (u * u + v * v + w * w)
    p = np.maximum((U[..., 4] - rho * ke) * 0.4, 1.0)
    Q = [rho, u, v, w, p]
    if U.shape[-1] > 5:
        Q += [np.maximum(U[..., 5] / rho, 1e-12), np.maximum(U[..., 6] / rho, 1e-12)]
    return np.stack(Q, axis=-1)


def residual_norm(dUdt, n_vars=None):
    n_vars = dUdt.shape[-1] if n_vars is None else n_vars
    n_total = dUdt.shape[0] * dUdt.shape[1] * n_vars
    return float(np.sqrt(np.sum(dUdt.astype(np.float64) ** 2) / n_total)) if n_total else 0.0


def local_time_step(Q, geom, order, mu, mu_t=None):
    """cfl.py:12-147 (the three bounds of the local time step).  Parity unpinned for its values
    (no reference test pins them)."""
    rho, u, v, w, p = (Q[..., i] for i in range(5))
    a = np.sqrt(np.maximum(1.4 * p / np.maximum(rho, 1e-10), 1e-10))
    ws = np.maximum(np.sqrt(u * u + v * v + w * w) + a, 1e-10)
    h = np.power(np.abs(geom.volume), 1.0 / 3.0)[:, None]
    dt_adv = CFL / (2 * order + 1) * h / ws
    mu_eff = mu + (mu_t if mu_t is not None else 0.0)
    mu_eff = np.broadcast_to(mu_eff, rho.shape)
    dt_visc = 0.25 * CFL / (2 * order + 1) ** 2 * rho * h * h / np.maximum(mu_eff, 1e-30)
    dt_geo = CFL * np.abs(geom.det) / np.maximum(geom.metric_flux_scale * ws, 1e-300)
    return np.minimum(np.minimum(dt_adv, dt_visc), dt_geo)


def enforce_positivity(U, p_floor=1.0):
    """In place on (N, n_vars)."""
    rho = np.maximum(U[:, 0], 1e-6)
    U[:, 0] = rho
    vel = U[:, 1:4] / rho[:, None]
    vm = np.sqrt(np.sum(vel ** 2, axis=1))
    clip = vm > 1e4
    if np.any(clip):
        vel[clip] *= (1e4 / vm[clip])[:, None]
        U[clip, 1:4] = rho[clip, None] * vel[clip]
    ke = 0.5 * rho * np.sum(vel ** 2, axis=1)
    p = (GAMMA - 1.0) * (U[:, 4] - ke)
    low = p < p_floor
    if np.any(low):
        U[low, 4] = p_floor / (GAMMA - 1.0) + ke[low]
    if U.shape[1] > 5:
        U[:, 5] = np.maximum(U[:, 5], 0.0)
    if U.shape[1] > 6:
        U[:, 6] = np.maximum(U[:, 6], 1e-8)
    return U


def apply_filter(U, ops, n_prism):
    U = U.copy()
    U[:n_prism, :, :5] = np.einsum("sj,cjv->csv", ops.filt[0], U[:n_prism, :, :5])
    U[n_prism:, :, :5] = np.einsum("sj,cjv->csv", ops.filt[1], U[n_prism:, :, :5])
    return U


def ssp_rk3(U0, rhs, dt, ops, n_prism, R0=None):
    """rhs(U) returns dU/dt (C, ns, nv); dt (C, ns).  Positivity and filter after every stage."""
    C, ns, nv = U0.shape
    al, be = SSP_RK3["alpha"], SSP_RK3["beta"]
    L0 = rhs(U0) if R0 is None else R0
    d = dt[..., None]

    def post(X):
        X = enforce_positivity(X.reshape(C * ns, nv).copy()).reshape(C, ns, nv)
        return apply_filter(X, ops, n_prism)

    U1 = post(U0 + d * L0)
    L1 = rhs(U1)
    U2 = post(al[1][0] * U0 + al[1][1] * U1 + be[1] * d * L1)
    L2 = rhs(U2)
    return post(al[2][0] * U0 + al[2][1] * U1 + al[2][2] * U2 + be[2] * d * L2)


def ns_step(U, geom, ops, fl, bc_params, order, mu=1.8e-5, mu_t=None, Pr=0.72, Pr_t=0.9, viscous=True):
    """One pseudo-step: returns (U_new, RMS norm of dU/dt at U, dt_local).  The SSP-RK3 stages
    are pinned by the linear-ODE order test; the step trajectory on general states is parity
    unpinned."""
    Q = update_primitives(U)
    dt = local_time_step(Q, geom, order, mu, mu_t)

    def rhs(X):
        R = inviscid_residual(X, geom, ops, fl, bc_params)
        if viscous:
            R = R + viscous_residual(X, geom, ops, fl, bc_params, mu, mu_t=mu_t, Pr=Pr, Pr_t=Pr_t)
        return R

    R0 = rhs(U)
    U_new = ssp_rk3(U, rhs, dt, ops, geom.n_prism, R0=R0)
    return U_new, residual_norm(R0), dt
