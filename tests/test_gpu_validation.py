"""GPU: the reference's validation tests with their own acceptance criteria, through the
reference-shaped mesh objects, boundary groups and providers (autoflowcfd_amd/boundary.py).

* Isentropic vortex (tests/validation/test_isentropic_vortex.py:88-135): P2, 20 x 20 x 1 hex of
  x-extruded prisms, periodic in x (tests/validation/_periodic_mesh.py:22-95), FARFIELD on the
  y / z planes by a face-exact provider, the reference's closed form as it writes it
  (_isentropic_vortex.py:46-78), 150 steps with the global dt min(dt_local); the state stays finite
  and the mean errors against the closed form at the time reached stay below 2 % (density) and
  10 % (u, v); free-stream preservation max|R_inv| < 1e-3.
* Periodic BC (tests/validation/test_periodic_bc.py:57-123): free-stream preservation (inviscid
  < 1e-3, viscous < 1e-6) and, for a field periodic in x and linear in y, no cell beside the
  periodic faces whose residual exceeds 50 x the median.
* Couette flow (tests/validation/test_couette.py:78-226): y-extruded prism channel, no-slip wall
  at rest below, a wall moving at U_wall = 0.01 above (the moving-wall ghost state), SYMMETRY z,
  OUTLET x; stable over 800 steps from a wrong profile; from the exact profile the residual rises
  to a peak above twice its first value and falls below 0.96 x that peak within 1600 steps;
  free-stream preservation with FARFIELD everywhere (inviscid < 1e-3, viscous < 1e-6).

Couette runs in FP64: its 0.01 m/s velocities carry 6e-5 J/m^3 of kinetic energy beside 2.5e5
of internal energy, below the resolution of an FP32 energy.
"""

import numpy as np
import pytest

from ref_objects import channel_volume_mesh, face_exact_provider, periodic_channel_volume_mesh

pytestmark = pytest.mark.gpu

GAMMA = 1.4
RHO_INF, P_INF = 1.225, 101325.0


def _sps_xyz(mesh, order):
    from autoflowcfd_amd.fr_tables import tables
    from autoflowcfd_amd.geometry import map_ref
    return map_ref(True, tables(order).sps, mesh.nodes[mesh.prism_conn])


def _cons(rho, u, v, w, p):
    return np.stack([rho, rho * u, rho * v, rho * w, p / (GAMMA - 1.0) + 0.5 * rho * (u * u + v * v + w * w)], -1)


def _solver(mesh, provider, dtype, **kw):
    from autoflowcfd_amd.solver import GPUFRSolver
    return GPUFRSolver(mesh, order=2, rho_inf=RHO_INF, p_inf=P_INF, boundary_ghost_provider=provider,
                       dtype=dtype, storage="input", **kw)


# ------------------------------------------------------------------------------------------ vortex
VX_L, VX_H, VX_LZ, U_VX = 10.0, 10.0, 0.5, 100.0
BETA, RC = 5.0, 1.0


def vortex_primitive(x, y, t):
    """_isentropic_vortex.py:46-78 as the reference writes it (swirl in units of a_inf)."""
    a_inf = np.sqrt(GAMMA * P_INF / RHO_INF)
    x0, y0 = VX_L / 2, VX_H / 2
    xc = x0 + np.mod(U_VX * t, VX_L)
    dx = np.mod(x - xc + VX_L / 2, VX_L) - VX_L / 2
    dy = y - y0
    r2 = (dx / RC) ** 2 + (dy / RC) ** 2
    sw = BETA * a_inf / (2 * np.pi) * np.exp(0.5 * (1 - r2))
    th = 1 - (GAMMA - 1) * BETA ** 2 / (8 * GAMMA * np.pi ** 2) * np.exp(1 - r2)
    return (RHO_INF * th ** (1 / (GAMMA - 1)), U_VX - sw * dy / RC, sw * dx / RC, 0 * x,
            P_INF * th ** (GAMMA / (GAMMA - 1)))


def _vortex_setup(dtype):
    from autoflowcfd_amd import boundary as B
    vm = periodic_channel_volume_mesh(20, 20, 1, VX_L, VX_H, VX_LZ, side="FARFIELD")
    mesh = B.mesh_from_volume_mesh(vm)
    q = [RHO_INF, U_VX, 0.0, 0.0, P_INF]
    ff = {"type": "FARFIELD", "Q_free": q}
    prov = face_exact_provider(mesh.faces, (VX_L, VX_H, VX_LZ),
                               {n: ff for n in ("wall_bottom", "wall_top", "z_min", "z_max", "x_min", "x_max")})
    return mesh, _solver(mesh, prov, dtype, vel_inf=U_VX, global_time_step=True)


@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_isentropic_vortex_tracks_exact_solution(prec):
    import torch
    print(f"[gpu] isentropic vortex P2 {prec} 150 global steps", flush=True)
    mesh, s = _vortex_setup(torch.float64 if prec == "f64" else torch.float32)
    xyz = _sps_xyz(mesh, 2)
    x, y = xyz[..., 0], xyz[..., 1]
    s.set_state_from_cpu(_cons(*vortex_primitive(x, y, 0.0)))
    t = 0.0
    for i in range(150):
        dt = float(s.dm.local_time_step(s.U, None, mu=s.mu_molecular).min().item())
        s.step(dt)
        t += dt
        if i % 50 == 49:
            assert np.all(np.isfinite(s.get_state_cpu()["U"])), f"diverged by step {i}"
    Q = s.get_state_cpu()["Q"]
    assert np.all(np.isfinite(Q))
    rho_e, u_e, v_e, _, _ = vortex_primitive(x, y, t)
    e_rho = np.abs(Q[..., 0] - rho_e).mean() / RHO_INF
    e_u = np.abs(Q[..., 1] - u_e).mean() / U_VX
    e_v = np.abs(Q[..., 2] - v_e).mean() / U_VX
    print(f"vortex {prec}: t={t:.3e} mean errors rho {e_rho:.3e} u {e_u:.3e} v {e_v:.3e}")
    assert e_rho < 0.02 and e_u < 0.10 and e_v < 0.10


def test_isentropic_vortex_freestream_preservation():
    import torch
    _, s = _vortex_setup(torch.float64)
    C, ns = s.dm.n_cells, s.dm.ns
    one = np.ones((C, ns))
    s.set_state_from_cpu(_cons(RHO_INF * one, U_VX * one, 0 * one, 0 * one, P_INF * one))
    assert float(s.compute_inviscid_residual_gpu().abs().max().item()) < 1e-3


# ---------------------------------------------------------------------------------------- periodic
PB_L, PB_H, PB_LZ, U_PB = 1.0, 1.0, 0.3, 20.0


def _periodic_setup():
    import torch
    from autoflowcfd_amd import boundary as B
    mesh = B.mesh_from_volume_mesh(periodic_channel_volume_mesh(4, 3, 1, PB_L, PB_H, PB_LZ))
    sym = {"type": "SYMMETRY"}
    prov = face_exact_provider(mesh.faces, (PB_L, PB_H, PB_LZ), {n: sym for n in ("wall_bottom", "wall_top",
                                                                                  "z_min", "z_max")})
    return mesh, _solver(mesh, prov, torch.float64, vel_inf=U_PB)


def test_periodic_freestream_preservation():
    print("[gpu] periodic BC free stream", flush=True)
    _, s = _periodic_setup()
    C, ns = s.dm.n_cells, s.dm.ns
    one = np.ones((C, ns))
    s.set_state_from_cpu(_cons(RHO_INF * one, U_PB * one, 0 * one, 0 * one, P_INF * one))
    assert float(s.compute_inviscid_residual_gpu().abs().max().item()) < 1e-3
    assert float(s.compute_viscous_residual_gpu().abs().max().item()) < 1e-6


def test_periodic_nonuniform_field_no_outlier_at_periodic_face():
    print("[gpu] periodic BC nonuniform field", flush=True)
    mesh, s = _periodic_setup()
    xyz = _sps_xyz(mesh, 2)
    x, y = xyz[..., 0], xyz[..., 1]
    u = U_PB + 0.3 * U_PB * (y / PB_H) + 0.1 * U_PB * np.sin(2 * np.pi / PB_L * x)
    one = np.ones_like(u)
    rho = RHO_INF * one
    # test_periodic_bc.py:101-106 sets rho u and E = p/(gamma-1) + rho u^2 / 2 (its own energy form)
    U = np.stack([rho, rho * u, 0 * u, 0 * u, P_INF / (GAMMA - 1.0) + 0.5 * RHO_INF * u ** 2], -1)
    s.set_state_from_cpu(U)
    R = s.compute_inviscid_residual_gpu().cpu().numpy()
    fr = mesh.faces
    moved = np.any(fr.translation != 0, axis=1)
    cells = np.unique(fr.owner[moved])
    per_cell = np.abs(R[:, :, 1]).max(axis=1)
    assert np.all(np.isfinite(per_cell))
    ratio = per_cell[cells].max() / max(np.median(per_cell), 1e-300)
    assert ratio < 50.0, ratio


# ----------------------------------------------------------------------------------------- couette
CT_H, CT_NY = 1.0, 6
CT_S = CT_H / CT_NY
CT_LX, CT_LZ = 2.0 * CT_H, CT_S
CT_NX = round(CT_LX / CT_S)
U_WALL = 0.01


def _couette_setup(configs=None, vel_inf=U_WALL):
    import torch
    from autoflowcfd_amd import boundary as B
    mesh = B.mesh_from_volume_mesh(channel_volume_mesh(CT_NX, CT_NY, 1, CT_LX, CT_H, CT_LZ))
    configs = configs or {
        "wall_bottom": {"type": "WALL", "is_no_slip": True, "wall_velocity": [0.0, 0.0, 0.0]},
        "wall_top": {"type": "WALL", "is_no_slip": True, "wall_velocity": [U_WALL, 0.0, 0.0]},
        "z_min": {"type": "SYMMETRY"}, "z_max": {"type": "SYMMETRY"},
        "x_min": {"type": "OUTLET", "p_outlet": P_INF}, "x_max": {"type": "OUTLET", "p_outlet": P_INF}}
    prov = face_exact_provider(mesh.faces, (CT_LX, CT_H, CT_LZ), configs)
    return mesh, _solver(mesh, prov, torch.float64, vel_inf=vel_inf)


def _couette_state(s, mesh, frac):
    y = _sps_xyz(mesh, 2)[..., 1]
    U = s.get_state_cpu()["U"]
    u0 = frac * U_WALL * y / CT_H
    e0 = P_INF / ((GAMMA - 1.0) * RHO_INF) + 0.5 * u0 ** 2
    U[..., 1] = RHO_INF * u0
    U[..., 4] = RHO_INF * e0
    s.set_state_from_cpu(U)


def test_couette_stable_from_wrong_profile():
    print("[gpu] Couette 800 steps from the half-speed profile", flush=True)
    mesh, s = _couette_setup()
    _couette_state(s, mesh, 0.5)
    for i in range(800):
        res = s.step(1e-6)
        assert np.isfinite(res), f"diverged at step {i}"
    assert np.all(np.isfinite(s.get_state_cpu()["U"]))


def test_couette_residual_trend():
    print("[gpu] Couette 1600 steps from the exact profile", flush=True)
    mesh, s = _couette_setup()
    _couette_state(s, mesh, 1.0)
    hist = [s.step(1e-6) for _ in range(1600)]
    assert np.all(np.isfinite(hist))
    peak = max(hist)
    print(f"Couette residual: first {hist[0]:.4e} peak {peak:.4e} (step {hist.index(peak)}) final {hist[-1]:.4e}")
    assert peak > 2.0 * hist[0]
    assert hist[-1] < 0.96 * peak


def test_couette_freestream_preservation():
    print("[gpu] Couette mesh free stream", flush=True)
    u_inf = 30.0
    q = [RHO_INF, u_inf, 0.0, 0.0, P_INF]
    ff = {"type": "FARFIELD", "Q_free": q}
    mesh, s = _couette_setup({n: ff for n in ("wall_bottom", "wall_top", "z_min", "z_max", "x_min", "x_max")},
                             vel_inf=u_inf)
    C, ns = s.dm.n_cells, s.dm.ns
    one = np.ones((C, ns))
    s.set_state_from_cpu(_cons(RHO_INF * one, u_inf * one, 0 * one, 0 * one, P_INF * one))
    assert float(s.compute_inviscid_residual_gpu().abs().max().item()) < 1e-3
    assert float(s.compute_viscous_residual_gpu().abs().max().item()) < 1e-6


def test_moving_wall_group_table_matches_oracle():
    """The Couette boundary (two wall groups at different velocities, OUTLET, SYMMETRY) through the
    group table: device inviscid + viscous residual against the oracle with the same rows, FP64."""
    import torch
    from oracle import fr_ops
    from oracle.faces import build_flat_faces
    from oracle.geometry import Geometry
    from oracle.residual import inviscid_residual, viscous_residual
    print("[gpu] moving-wall group table vs oracle", flush=True)
    mesh, s = _couette_setup()
    rng = np.random.default_rng(4)
    y = _sps_xyz(mesh, 2)[..., 1]
    u = U_WALL * y / CT_H + 0.002 * rng.standard_normal(y.shape)
    one = np.ones_like(u)
    U = _cons(RHO_INF * (1 + 1e-3 * rng.standard_normal(u.shape)), u, 0.001 * rng.standard_normal(u.shape),
              0 * u, P_INF * (1 + 1e-4 * rng.standard_normal(u.shape)))
    s.set_state_from_cpu(U)
    table = s.dm.bc_params.cpu().numpy().reshape(8, 5)
    bcp = {r: table[r] for r in range(8)}
    ops = fr_ops.generate_operators(2)
    g = Geometry(mesh, ops)
    fl = build_flat_faces(mesh, ops)
    mu = s.mu_molecular
    ref = inviscid_residual(U, g, ops, fl, bcp) + viscous_residual(U, g, ops, fl, bcp, mu)
    Ud = torch.as_tensor(U, dtype=torch.float64, device="cuda").contiguous()
    dev = (s.compute_inviscid_residual_gpu(Ud) + s.compute_viscous_residual_gpu(Ud)).cpu().numpy()
    err = np.linalg.norm(dev - ref) / np.linalg.norm(ref)
    assert err < 1e-10, err
    assert np.any(fl.bc == 7) and len(np.unique(fl.bc_row[fl.is_boundary])) >= 4
    del one
