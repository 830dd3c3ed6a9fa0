"""GPU: the reference's Taylor-Green vortex validation on the device, and P3 parity.

Restates tests/validation/test_tgv.py of the reference with its own acceptance criteria:
P2 on the 4^3-hex x 6-tet triply periodic box of side 2 pi (tests/validation/_tgv_mesh.py:38-74,
the same 6-tet template as hex_box(split="tet")), Re 20, the analytic TGV initial field
(test_tgv.py:96-112), 150 global steps with the test's global dt (test_tgv.py:70-92):
  * the state stays finite;
  * after the step-0 adaptation, the kinetic energy never rises by more than 0.1 % per step;
  * KE/KE0 after 150 steps lies in (0.5, 0.85) (the reference's calibrated range);
and the free-stream preservation of the same mesh (test_tgv.py:140-187): max|R_inv|/p_inf < 3e-5,
max|R_visc| < 3e-2.  The P3 tet path (operators too large for LDS, dense tet derivative rows) is
checked against the FP64 oracle: inviscid + viscous residual and one global-dt SSP-RK3 step.
"""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

L = 2.0 * np.pi
LC = 1.0
RHO_INF, P_INF, U0 = 1.225, 101325.0, 30.0
GAMMA = 1.4
RE = 20.0
MU = RHO_INF * U0 * LC / RE
N_STEPS = 150


def _mesh(n=4):
    from autoflowcfd_amd.mesh import hex_box
    return hex_box(n, n, n, L, L, L, split="tet", periodic=True, permute=False)


def _sps_xyz(mesh, order):
    from autoflowcfd_amd.fr_tables import tables
    from autoflowcfd_amd.geometry import map_ref
    return map_ref(False, tables(order).sps, mesh.nodes[mesh.tet_conn])


def _tgv_U(xyz):
    """test_tgv.py:96-112."""
    x, y, z = xyz[..., 0], xyz[..., 1], xyz[..., 2]
    u = U0 * np.sin(x / LC) * np.cos(y / LC) * np.cos(z / LC)
    v = -U0 * np.cos(x / LC) * np.sin(y / LC) * np.cos(z / LC)
    w = np.zeros_like(u)
    p = P_INF + (RHO_INF * U0 ** 2 / 16.0) * (np.cos(2 * x / LC) + np.cos(2 * y / LC)) * (np.cos(2 * z / LC) + 2.0)
    rho = np.full_like(u, RHO_INF)
    E = p / (GAMMA - 1.0) + 0.5 * rho * (u ** 2 + v ** 2 + w ** 2)
    return np.stack([rho, rho * u, rho * v, rho * w, E], axis=-1)


def _global_dt(solver, volumes):
    """test_tgv.py:70-92, restated as written (its pressure is formed from the primitive array)."""
    import torch
    Q = solver.get_state_cpu()["Q"]
    rho = Q[:, :, 0]
    u = Q[:, :, 1] / np.maximum(rho, 1e-10)
    v = Q[:, :, 2] / np.maximum(rho, 1e-10)
    w = Q[:, :, 3] / np.maximum(rho, 1e-10)
    p = (GAMMA - 1.0) * (Q[:, :, 4] - 0.5 * rho * (u ** 2 + v ** 2 + w ** 2))
    a = np.sqrt(np.maximum(GAMMA * p / np.maximum(rho, 1e-10), 1e-10))
    ws = np.maximum(np.sqrt(u ** 2 + v ** 2 + w ** 2) + a, 1e-10)
    h = np.power(np.abs(volumes), 1.0 / 3.0)[:, None]
    CFL = 0.1
    dt = np.minimum(CFL * h / ws, 0.25 * CFL * rho * h ** 2 / MU)
    local = solver._to_user(solver.dm.local_time_step(solver.U, None, mu=MU)).double().cpu().numpy()
    torch.cuda.synchronize()
    return float(np.minimum(dt, local).min())


def _kinetic_energy(solver):
    """test_tgv.py:115-119."""
    Q = solver.get_state_cpu()["Q"]
    return float((0.5 * Q[:, :, 0] * (Q[:, :, 1] ** 2 + Q[:, :, 2] ** 2 + Q[:, :, 3] ** 2)).mean())


def _solver(mesh, order, dtype):
    from autoflowcfd_amd.solver import GPUFRSolver
    return GPUFRSolver(mesh, order=order, rho_inf=RHO_INF, vel_inf=U0, p_inf=P_INF, mu_molecular=MU, dtype=dtype,
                       global_time_step=True)


@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_tgv_kinetic_energy_decays_monotonically(prec):
    import torch
    print(f"[gpu] TGV P2 {prec} 150 global steps", flush=True)
    dtype = torch.float64 if prec == "f64" else torch.float32
    mesh = _mesh()
    s = _solver(mesh, 2, dtype)
    s.set_state_from_cpu(_tgv_U(_sps_xyz(mesh, 2)))
    vol = s.cell_volumes()
    ke = [_kinetic_energy(s)]
    for i in range(N_STEPS):
        s.step(_global_dt(s, vol))
        U = s.get_state_cpu()["U"]
        assert np.all(np.isfinite(U)), f"diverged at step {i}"
        ke.append(_kinetic_energy(s))
    tail = ke[1:]
    increases = [tail[i] for i in range(1, len(tail)) if tail[i] > tail[i - 1] * 1.001]
    assert not increases, increases[:5]
    ratio = ke[-1] / ke[0]
    assert 0.5 < ratio < 0.85, ratio


def test_tgv_freestream_preservation():
    import torch
    print("[gpu] TGV mesh free-stream preservation P2", flush=True)
    mesh = _mesh()
    s = _solver(mesh, 2, torch.float64)
    C, ns = s.dm.n_cells, s.dm.ns
    U = np.zeros((C, ns, 5))
    U[..., 0] = RHO_INF
    U[..., 1] = RHO_INF * U0
    U[..., 4] = P_INF / (GAMMA - 1.0) + 0.5 * RHO_INF * U0 ** 2
    s.set_state_from_cpu(U)
    inv = s.compute_inviscid_residual_gpu().double().cpu().numpy()
    visc = s.compute_viscous_residual_gpu().double().cpu().numpy()
    assert np.max(np.abs(inv)) / P_INF < 3e-5
    assert np.max(np.abs(visc)) < 3e-2


def test_p3_requires_fp64():
    """P3 collapsed tet operators reach 6e7: FP32 rounding of the state is amplified into
    percent-level residual errors (measured 7e-2 relative on the TGV state), so the library
    refuses an FP32 P3 mesh."""
    import torch
    print("[gpu] P3 FP32 refused", flush=True)
    with pytest.raises(RuntimeError, match="FP64"):
        _solver(_mesh(2), 3, torch.float32)


@pytest.mark.parametrize("prec", ["f64"])
def test_p3_tet_residual_and_step_match_oracle(prec):
    """P3 on the periodic tet box (2^3 hex x 6 tets) vs the FP64 oracle: inviscid and viscous
    residual of the TGV state and one global-dt SSP-RK3 step.  The shipped P3 tables and the
    oracle's agree to 1e-6 of their largest entry (6e7; test_product_tables_equal_oracle), which
    bounds the agreement here (viscous residual measured 5.1e-5): tolerance 2e-4."""
    import torch
    from oracle import fr_ops
    from oracle.faces import build_flat_faces
    from oracle.geometry import Geometry
    from oracle.residual import inviscid_residual, viscous_residual
    from oracle.time_step import ssp_rk3
    from autoflowcfd_amd.device import bc_params_table
    print(f"[gpu] P3 tet parity {prec}", flush=True)
    dtype = torch.float64 if prec == "f64" else torch.float32
    mesh = _mesh(2)
    ops = fr_ops.generate_operators(3)
    g = Geometry(mesh, ops)
    fl = build_flat_faces(mesh, ops)
    bcp = bc_params_table(np.array([RHO_INF, U0, 0.0, 0.0, P_INF]))
    bcd = {c: bcp[c] for c in range(8)}
    U = _tgv_U(g.sps_xyz)
    if dtype == torch.float32:
        U = U.astype(np.float32).astype(np.float64)
    Ri = inviscid_residual(U, g, ops, fl, bcd)
    Rv = viscous_residual(U, g, ops, fl, bcd, MU)
    s = _solver(mesh, 3, dtype)
    s.set_state_from_cpu(U)
    ri = s.compute_inviscid_residual_gpu().double().cpu().numpy()
    rv = s.compute_viscous_residual_gpu().double().cpu().numpy()
    e = lambda a, b: np.linalg.norm(a - b) / np.linalg.norm(b)  # noqa: E731
    tol = 2e-4
    assert e(ri, Ri) < tol, f"inviscid {e(ri, Ri):.3e}"
    assert e(rv, Rv) < tol, f"viscous {e(rv, Rv):.3e}"
    # local time step at 64 stored points per cell (cfl.py:12-147 via the oracle): on straight-sided
    # tets the metric terms are constant per cell and do not pass through the P3 tables'
    # conditioning, so the P1 / P2 bound holds (measured 3.7e-16)
    from oracle.time_step import local_time_step, update_primitives
    dt_ref = local_time_step(update_primitives(U), g, 3, MU)
    dt_dev = s._to_user(s.dm.local_time_step(s.U, None, mu=MU)).double().cpu().numpy()
    e_dt = np.abs(dt_dev - dt_ref).max() / np.abs(dt_ref).max()
    print(f"[gpu] P3 local dt rel {e_dt:.3e}", flush=True)
    assert e_dt < 1e-12, f"local dt {e_dt:.3e}"
    dt = 1e-5
    U_ref = ssp_rk3(U, lambda X: inviscid_residual(X, g, ops, fl, bcd) + viscous_residual(X, g, ops, fl, bcd, MU),
                    np.full(U.shape[:2], dt), ops, 0)
    s.step(dt)
    U_new = s.get_state_cpu()["U"]
    inc = np.linalg.norm((U_new - U) - (U_ref - U)) / np.linalg.norm(U_ref - U)
    assert inc < 2e-4, f"step increment {inc:.3e}"
