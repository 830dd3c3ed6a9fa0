"""GPU: the state stored as its offset from a per-case conserved reference (DeviceMesh state_ref,
AfrMeshDesc.state_ref / AfrP0Desc.state_ref), the storage of every FP32 solver.

* Equivalence in FP64: every kernel that loads the state (local time step, primitives and the
  outlier reference of the inviscid kernels, viscous volume, SST cell and update kernels, wall
  forces, SSP-RK3 stages with positivity and filter, the P0 phase) gives on the offset of a state
  what it gives on the state itself; the offset only moves FP64 rounding (1e-12 of the quantity;
  1e-7 of a pseudo-step's increment, which also carries the filter's row-sum rounding times the
  offset).  Points where the positivity projection acts (density below its floor, pressure below
  its floor, speed above 1e4) are included.
* FP32 against the FP64 oracle on the conserved state the storage represents (offset + FP32
  offset): one pseudo-step's norm 1e-5 and increment 1e-3 of its norm, as for absolute storage.
* The free-stream floor: at the exact free stream the FP32 residual of absolute storage carries
  the rounding of rho E ~ 2.5e5 (0.016) against the boundary states; the offset storage's equals
  the FP64 residual of the same state to the FP32 tolerance.
* The solver takes and returns conserved states (get / set, checkpoints) whatever it stores.
"""

import numpy as np
import pytest

from cases import freestream, mixed_mesh, prim_to_cons, smooth_state, tet_box, wall_box, wedge_box

pytestmark = pytest.mark.gpu

MU = 1.8e-5


def _rel(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / max(np.linalg.norm(np.asarray(b)), 1e-300))


def _meshes(mesh, order, dtype, turbulence=None):
    import torch  # noqa: F401
    from autoflowcfd_amd.device import DeviceMesh, bc_params_table
    Q0 = freestream(0.15, 4.0)
    bcp = bc_params_table(Q0)
    a = DeviceMesh(mesh, order, dtype=dtype, bc_params=bcp, turbulence=turbulence)
    b = DeviceMesh(mesh, order, dtype=dtype, bc_params=bcp, turbulence=turbulence, state_ref=prim_to_cons(Q0))
    return a, b, Q0


@pytest.mark.parametrize("order", [1, 2])
def test_fp64_kernels_equal_on_offset_storage(order):
    import torch
    from oracle import fr_ops
    from oracle.geometry import Geometry
    print(f"[gpu] offset storage FP64 equivalence P{order}", flush=True)
    mesh = wall_box(seed=7)
    a, b, Q0 = _meshes(mesh, order, torch.float64, turbulence="SST")
    g = Geometry(mesh, fr_ops.generate_operators(order))
    U = torch.as_tensor(smooth_state(g.sps_xyz, Q0, amp=2e-2, seed=3), device="cuda").contiguous()
    Ub = b.to_storage(U)
    assert float(torch.max(torch.abs(b.to_absolute(Ub) - U))) <= 1e-10
    mut = torch.as_tensor(np.random.default_rng(1).uniform(0, 20 * MU, U.shape[:2]), device="cuda").contiguous()
    # local time step, mean-flow residual, wall forces
    assert _rel(b.local_time_step(Ub, mut, mu=MU).cpu(), a.local_time_step(U, mut, mu=MU).cpu()) < 1e-12
    Ra = a.residual(U, mut, mu=MU).clone()
    Rb = b.residual(Ub, mut, mu=MU).clone()
    assert _rel(Rb.cpu(), Ra.cpu()) < 1e-12
    assert _rel(b.wall_forces(Ub, mut, mu=MU).cpu(), a.wall_forces(U, mut, mu=MU).cpu()) < 1e-12
    # SST: point physics, transport, update (rates) and nu_t / mu_t
    fa = a.sst_fields(1e-3, 2e3)
    fb = tuple(x.clone() for x in fa)
    dt = a.local_time_step_nut(U, fa[2], mu=MU).clone()
    ra = torch.empty((a.n_cells, a.ns, 2), dtype=torch.float64, device="cuda")
    rb = torch.empty_like(ra)
    a.sst_step(U, fa[0], fa[1], dt, fa[2], fa[3], mu=MU, rates=ra)
    b.sst_step(Ub, fb[0], fb[1], dt, fb[2], fb[3], mu=MU, rates=rb)
    assert _rel(rb.cpu(), ra.cpu()) < 1e-12
    assert _rel(fb[3].cpu(), fa[3].cpu()) < 1e-12
    # one SST pseudo-step
    UA, UB = U.clone(), Ub.clone()
    na = float(a.sst_rk3_step(UA, *fa, mu=MU).item())
    nb = float(b.sst_rk3_step(UB, *fb, mu=MU).item())
    print(f"[gpu] P{order} norms {na:.12e} {nb:.12e}", flush=True)
    assert abs(na - nb) <= 1e-12 * abs(na)
    inc_a = (UA - U).cpu().numpy()
    inc_b = (b.to_absolute(UB) - U).cpu().numpy()
    print(f"[gpu] P{order} increment rel {_rel(inc_b, inc_a):.3e}", flush=True)
    assert _rel(inc_b, inc_a) < 1e-7
    for x, y in zip(fb, fa):
        assert _rel(x.cpu(), y.cpu()) < 1e-9
    # one stage whose update drives three points through the positivity projection: p below its
    # floor, rho below its floor (|u| then above 1e4), |u| above 1e4
    dt = torch.full((a.n_cells, a.ns), 1e-3, dtype=torch.float64, device="cuda")
    R = torch.zeros_like(U)
    u0 = U.cpu().numpy()
    ke = 0.5 * float((u0[0, 0, 1:4] ** 2).sum() / u0[0, 0, 0])
    R[0, 0, 4] = (ke + 0.5 / 0.4 - u0[0, 0, 4]) / 1e-3
    R[1, 1, 0] = (1e-8 - u0[1, 1, 0]) / 1e-3
    R[2, 2, 1] = (2e4 * u0[2, 2, 0] - u0[2, 2, 1]) / 1e-3
    oa, ob = torch.empty_like(U), torch.empty_like(U)
    a.rk_stage(U, None, R, dt, oa, 1.0, 0.0, 1.0)
    b.rk_stage(Ub, None, R, dt, ob, 1.0, 0.0, 1.0)
    x = oa.cpu().numpy()
    assert x[0, :, 0].min() > 0 and np.isfinite(x).all()
    err = float(torch.max(torch.abs(b.to_absolute(ob) - oa) / (1.0 + torch.abs(oa))))
    print(f"[gpu] P{order} projected stage max rel {err:.3e}", flush=True)
    assert err < 1e-12


def test_p0_phase_equal_on_offset_storage():
    import torch
    from autoflowcfd_amd.continuation import P0Mesh
    from autoflowcfd_amd.device import bc_params_table
    from autoflowcfd_amd.geometry import cell_block_metrics
    from autoflowcfd_amd.fr_tables import tables
    print("[gpu] offset storage P0 phase FP64", flush=True)
    mesh = wall_box(seed=8)
    Q0 = freestream(0.2, 3.0)
    bcp = bc_params_table(Q0)
    vol = cell_block_metrics(mesh, tables(1), dtype=np.float64)["volume"]
    d = torch.as_tensor(np.random.default_rng(2).uniform(0.01, 1.0, mesh.n_cells), device="cuda")
    a = P0Mesh(mesh, vol, bcp, dtype=torch.float64, d_wall=d)
    b = P0Mesh(mesh, vol, bcp, dtype=torch.float64, d_wall=d, state_ref=prim_to_cons(Q0))
    rng = np.random.default_rng(4)
    Q = np.broadcast_to(Q0, (mesh.n_cells, 1, 5)).copy()
    Q[..., 1:4] += rng.normal(0, 5.0, Q[..., 1:4].shape)
    Q[..., 4] += rng.normal(0, 300.0, Q[..., 4].shape)
    U = torch.as_tensor(prim_to_cons(Q), device="cuda").contiguous()
    Ub = b.to_storage(U)
    mut = torch.full((mesh.n_cells,), 3 * MU, dtype=torch.float64, device="cuda")
    assert _rel(b.residual(Ub, mu=MU, mu_t=mut).cpu(), a.residual(U, mu=MU, mu_t=mut).cpu()) < 1e-12
    UA, UB = U.clone(), Ub.clone()
    fa = [torch.full((mesh.n_cells,), v, dtype=torch.float64, device="cuda") for v in (1e-3, 2e3, 0.0, 0.0)]
    fb = [x.clone() for x in fa]
    na = float(a.sst_step(UA, *fa, mu=MU).item())
    nb = float(b.sst_step(UB, *fb, mu=MU).item())
    assert abs(na - nb) <= 1e-12 * abs(na)
    assert _rel((b.to_absolute(UB) - U).cpu(), (UA - U).cpu()) < 1e-7
    for x, y in zip(fb, fa):
        assert _rel(x.cpu(), y.cpu()) < 1e-9


@pytest.mark.parametrize("kind,order", [("wallbox", 1), ("wallbox", 2), ("tets", 1), ("mixed", 2), ("wedges", 2)])
def test_fp32_offset_step_matches_oracle(kind, order):
    import torch
    from autoflowcfd_amd.device import DeviceMesh, bc_params_table
    from oracle import fr_ops
    from oracle.faces import build_flat_faces
    from oracle.geometry import Geometry
    from oracle.time_step import ns_step
    print(f"[gpu] offset storage FP32 pseudo-step vs oracle {kind} P{order}", flush=True)
    mesh = {"wallbox": lambda: wall_box(seed=2), "tets": lambda: tet_box(3, seed=1), "mixed": mixed_mesh,
            "wedges": wedge_box}[kind]()
    Q0 = freestream(0.15)
    bcp = bc_params_table(Q0)
    ops = fr_ops.generate_operators(order)
    g = Geometry(mesh, ops)
    fl = build_flat_faces(mesh, ops)
    dm = DeviceMesh(mesh, order, dtype=torch.float32, bc_params=bcp, state_ref=prim_to_cons(Q0))
    U = dm.to_storage(torch.as_tensor(smooth_state(g.sps_xyz, Q0, amp=1e-2, seed=0), device="cuda"))
    Urep = dm.to_absolute(U).cpu().numpy()            # the conserved state the storage represents
    U_ref, norm_ref, _ = ns_step(Urep, g, ops, fl, {c: bcp[c] for c in range(8)}, order, mu=MU)
    norm = float(dm.ssp_rk3_step(U, None, mu=MU).item())
    inc = dm.to_absolute(U).cpu().numpy() - Urep
    print(f"[gpu] P{order} norm rel {abs(norm - norm_ref) / norm_ref:.3e}, increment rel "
          f"{_rel(inc, U_ref - Urep):.3e}", flush=True)
    tol = 1e-4 if kind == "mixed" else 1e-5     # P2 tets: their operators amplify FP32 rounding
    assert abs(norm - norm_ref) <= tol * norm_ref
    assert _rel(inc, U_ref - Urep) < 1e-3


def test_free_stream_floor_removed():
    """A box without walls (inlet, outlet, symmetry planes) at its exact free stream: a steady
    solution, whose FP64 residual is round-off.  FP32 absolute storage rounds rho E, and the
    rounding meets the exact boundary states at the inlet and outlet; the offset storage holds the
    free stream exactly."""
    import torch
    from autoflowcfd_amd.device import DeviceMesh, bc_params_table
    from autoflowcfd_amd.mesh import BC_INLET, BC_OUTLET, BC_SYMMETRY, hex_box
    print("[gpu] FP32 residual at the exact free stream: absolute vs offset storage", flush=True)
    mesh = hex_box(6, 5, 4, 4.0, 1.0, 1.0, split="prism", grading_y=1.1, seed=1,
                   bc_by_plane={"x0": BC_INLET, "x1": BC_OUTLET, "y0": BC_SYMMETRY, "y1": BC_SYMMETRY,
                                "z0": BC_SYMMETRY, "z1": BC_SYMMETRY})
    Q0 = freestream(0.15)
    bcp = bc_params_table(Q0)
    res = {}
    for name, dt, off in (("f64", torch.float64, None), ("f32_abs", torch.float32, None),
                          ("f32_off", torch.float32, prim_to_cons(Q0))):
        dm = DeviceMesh(mesh, 2, dtype=dt, bc_params=bcp, state_ref=off)
        U = dm.to_storage(torch.as_tensor(np.broadcast_to(prim_to_cons(Q0), (mesh.n_cells, dm.ns, 5)).copy()))
        res[name] = np.abs(dm.residual(U, None, mu=MU).double().cpu().numpy()).max(axis=(0, 1))
    print(f"[gpu] free stream max |R| per variable: FP64 {res['f64']}, FP32 absolute storage "
          f"{res['f32_abs']}, FP32 offset storage {res['f32_off']}", flush=True)
    # measured: FP64 6.5e-6 and FP32 offset 6.6e-6 in rho E (round-off), FP32 absolute 133
    assert np.all(res["f32_off"] <= 2.0 * res["f64"] + 1e-9)
    assert res["f32_abs"][4] > 1.0


def test_fp32_solver_tracks_fp64_closer_with_offsets():
    import torch
    from autoflowcfd_amd.solver import GPUFRSolver
    print("[gpu] FP32 solver histories with and without offsets vs FP64", flush=True)
    mesh = wall_box(n=(6, 5, 4), seed=1)
    Q0 = freestream(0.15)
    kw = dict(order=2, rho_inf=Q0[0], vel_inf=Q0[1], p_inf=Q0[4])
    h = {}
    for name, dt, off in (("f64", torch.float64, False), ("f32_abs", torch.float32, False),
                          ("f32_off", torch.float32, True)):
        s = GPUFRSolver(mesh, dtype=dt, offset_storage=off, **kw)
        h[name] = np.array([s.step() for _ in range(200)])
    ea = np.abs(h["f32_abs"] / h["f64"] - 1.0)
    eb = np.abs(h["f32_off"] / h["f64"] - 1.0)
    print(f"[gpu] FP64 norms {h['f64'][0]:.3e} -> {h['f64'][-1]:.3e}; max rel deviation of the FP32 "
          f"histories: absolute storage {ea.max():.3e} (last {ea[-1]:.3e}), offsets {eb.max():.3e} "
          f"(last {eb[-1]:.3e})", flush=True)
    assert eb.max() < 1e-4
    assert ea.max() > 1e2 * eb.max()


def test_solver_state_round_trips_conserved(tmp_path):
    import torch
    from autoflowcfd_amd.solver import GPUFRSolver
    print("[gpu] FP32 offset solver: get / set state and checkpoint round trip", flush=True)
    mesh = wall_box(seed=3)
    Q0 = freestream(0.15)
    s = GPUFRSolver(mesh, order=2, rho_inf=Q0[0], vel_inf=Q0[1], p_inf=Q0[4], dtype=torch.float32)
    assert s.dm.state_ref is not None
    assert float(torch.abs(s.U).max()) == 0.0            # the initial free stream is stored exactly
    st0 = s.get_state_cpu()["U"]
    assert np.array_equal(st0, np.broadcast_to(prim_to_cons(Q0), st0.shape))
    from oracle import fr_ops
    from oracle.geometry import Geometry
    g = Geometry(mesh, fr_ops.generate_operators(2))
    U = smooth_state(g.sps_xyz, Q0, amp=1e-2, seed=6)
    s.set_state_from_cpu(U)
    got = s.get_state_cpu()["U"]
    off = prim_to_cons(Q0)
    assert np.all(np.abs(got - U) <= 6e-8 * np.abs(U - off) + 1e-9 * np.abs(off))
    s.step()
    st = s.get_state_cpu()["U"]
    p = tmp_path / "ck.npz"
    s.save_checkpoint(str(p))
    s2 = GPUFRSolver(mesh, order=2, rho_inf=Q0[0], vel_inf=Q0[1], p_inf=Q0[4], dtype=torch.float32)
    s2.load_checkpoint(str(p))
    # reloaded through (offset + stored) - offset in FP64: the same conserved state to FP64 rounding
    assert np.abs(s2.get_state_cpu()["U"] - st).max() <= 1e-9
