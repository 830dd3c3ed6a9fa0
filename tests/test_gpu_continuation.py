"""GPU: the P0 phase and the order continuation (continuation.py, afr_p0.hip) against the FP64
oracle (oracle/continuation.py).

* P0 residual (finite-volume AUSM+up with the per-cell ghost quirk, plus the boundary interior
  penalty) on a perturbed wall-box state: FP64 to 1e-10 relative, FP32 to 1e-5 relative L2.
* One P0 SSP-RK3 pseudo-step (dt, stages, positivity, norm): FP64 to 1e-10.
* The interpolation between orders on the device equals the oracle's matrix (scipy's
  RegularGridInterpolator) applied on the host, to round-off.
* The whole continuation P0 -> P1 -> P2 in FP64: per-phase residual histories (to 5e-6: round-off
  of the free-stream start, amplified by the transient) and the final state (1e-10) against the
  oracle's run, with the same phase rules (max_iter // 3 steps per phase).
"""

import numpy as np
import pytest

from cases import freestream, mixed_mesh, prim_to_cons, wall_box, wedge_box

pytestmark = pytest.mark.gpu


def _p0_setup(mesh, Q0, dtype):
    import torch
    from autoflowcfd_amd.continuation import P0Mesh
    from autoflowcfd_amd.device import bc_params_table
    from oracle import continuation as OC
    from oracle import fr_ops
    from oracle.faces import build_flat_faces
    from oracle.geometry import Geometry
    ops = fr_ops.generate_operators(2)
    g = Geometry(mesh, ops)
    g0 = OC.P0Geometry(mesh, g)
    fl = build_flat_faces(mesh, ops)
    bcp = bc_params_table(Q0)
    dev = P0Mesh(mesh, g.volume, bcp, dtype=dtype, device="cuda")
    return dev, g0, fl, {c: bcp[c] for c in range(8)}


def _state(mesh, Q0, seed=1):
    rng = np.random.default_rng(seed)
    Q = np.broadcast_to(Q0, (mesh.n_cells, 5)) * (1.0 + 0.02 * rng.standard_normal((mesh.n_cells, 5)))
    return prim_to_cons(Q)[:, None, :]


@pytest.mark.parametrize("kind", ["wallbox", "mixed", "wedges"])
@pytest.mark.parametrize("prec,tol", [("f64", 1e-10), ("f32", 1e-5)])
def test_p0_residual_matches_oracle(prec, tol, kind):
    import torch
    from oracle import continuation as OC
    dtype = torch.float64 if prec == "f64" else torch.float32
    print(f"[gpu] P0 residual {prec} {kind} vs oracle", flush=True)
    mesh = {"wallbox": lambda: wall_box(seed=5), "mixed": mixed_mesh, "wedges": wedge_box}[kind]()
    Q0 = freestream(0.2, 5.0)
    dev, g0, fl, bcd = _p0_setup(mesh, Q0, dtype)
    U = _state(mesh, Q0)
    mu = 1e-3
    ref = OC.p0_inviscid_residual(U, mesh, g0, bcd) + OC.p0_viscous_residual(U, mesh, g0, fl, bcd, mu)
    Ud = torch.as_tensor(U, dtype=dtype).cuda().contiguous()
    got = dev.residual(Ud, mu=mu).cpu().numpy().astype(np.float64)
    rel = np.linalg.norm(got - ref) / np.linalg.norm(ref)
    print(f"P0 residual {prec}: rel L2 {rel:.3e}")
    assert rel < tol
    if kind == "wallbox" and prec == "f64":
        # the wall penalty is small beside the perturbed inviscid residual: check it on its own as
        # the difference of the device residual with and without viscosity
        visc = OC.p0_viscous_residual(U, mesh, g0, fl, bcd, mu)
        d = got - dev.residual(Ud, mu=0.0).cpu().numpy()
        assert np.linalg.norm(visc) > 0
        assert np.linalg.norm(d - visc) < 1e-6 * np.linalg.norm(visc)


def test_p0_step_matches_oracle():
    import torch
    from oracle import continuation as OC
    print("[gpu] P0 SSP-RK3 step f64 vs oracle", flush=True)
    mesh = wall_box(seed=6)
    Q0 = freestream(0.15, 3.0)
    dev, g0, fl, bcd = _p0_setup(mesh, Q0, torch.float64)
    U = _state(mesh, Q0, seed=2)
    mu = 1.8e-5
    U_ref, res_ref = OC.p0_step(U, mesh, g0, fl, bcd, mu)
    Ud = torch.as_tensor(U, dtype=torch.float64).cuda().contiguous()
    res = float(dev.step(Ud, mu=mu).item())
    got = Ud.cpu().numpy()
    rel = np.linalg.norm(got - U_ref) / np.linalg.norm(U_ref - U)
    print(f"P0 step: norm {res:.12e} oracle {res_ref:.12e}; rel dU {rel:.3e}")
    assert abs(res - res_ref) <= 1e-10 * res_ref
    assert rel < 1e-9


@pytest.mark.parametrize("old,new", [(0, 1), (1, 2)])
def test_device_interpolation_equals_oracle_matrix(old, new):
    import torch
    from autoflowcfd_amd.continuation import interp_matrix, interpolate
    from oracle import continuation as OC
    print(f"[gpu] interpolation P{old} -> P{new}", flush=True)
    W = OC.interp_matrix(old, new)
    assert np.allclose(interp_matrix(old, new), W, atol=1e-14)
    rng = np.random.default_rng(old)
    X = rng.standard_normal((37, (old + 1) ** 3, 7))
    got = interpolate(torch.as_tensor(X).cuda(), old, new).cpu().numpy()
    assert np.allclose(got, np.einsum("ab,cbv->cav", W, X), rtol=0, atol=1e-13)


def test_order_continuation_matches_oracle_run():
    import torch
    from autoflowcfd_amd.continuation import solve_with_order_continuation
    from autoflowcfd_amd.solver import GPUFRSolver
    from oracle import continuation as OC
    from oracle import fr_ops
    from autoflowcfd_amd.device import bc_params_table
    print("[gpu] order continuation P0 -> P2 f64 vs oracle run", flush=True)
    mesh = wall_box(n=(3, 2, 2), seed=4)
    Q0 = freestream(0.15, 0.0)
    mu = 1.8e-5
    bcp = bc_params_table(Q0)
    U_ref, hist_ref = OC.run(mesh, 2, {c: bcp[c] for c in range(8)}, prim_to_cons(Q0), mu, max_iter=30, tol=1e-12,
                             fr_ops=fr_ops)

    def make(p):
        return GPUFRSolver(mesh, order=p, rho_inf=Q0[0], vel_inf=Q0[1], p_inf=Q0[4], mu_molecular=mu,
                           dtype=torch.float64)
    hist = {}
    s, out = solve_with_order_continuation(make, 2, max_iter=30, tol=1e-12, history=hist)
    assert sorted(hist) == sorted(hist_ref)
    for p in hist_ref:
        a, b = np.array(hist[p]), np.array(hist_ref[p])
        d = np.max(np.abs(a - b) / b) if len(a) == len(b) else np.inf
        print(f"P{p}: {len(a)} steps, first {a[0]:.10e} / {b[0]:.10e}, last {a[-1]:.10e} / {b[-1]:.10e}, "
              f"max rel {d:.2e}")
        assert len(a) == len(b)
        # the start is the uniform free stream: each residual is a sum of O(p A) face fluxes that cancel
        # to O(1e-2), so FP64 round-off alone leaves ~1e-11 relative per evaluation (the first P0 norm),
        # and the transient amplifies it: a 1e-15 relative change of the start moves the 10th P0 norm
        # of the oracle by 1.6e-7 (test_oracle_continuation.py).  5e-6 bounds that and still fails
        # for any wrong term (a dropped penalty or a wrong ghost changes the first norm by O(1))
        assert d < 5e-6
    got = s.get_state_cpu()["U"]
    rel = np.linalg.norm(got - U_ref) / np.linalg.norm(U_ref)
    print(f"final state rel {rel:.3e}; iterations {out['iterations']}")
    assert rel < 1e-10
    assert out["iterations"] == sum(len(h) for h in hist_ref.values())


def test_run_steady_runs_the_continuation_for_laminar_p2():
    """run_steady (the AutoFlowCFDAPI.run_steady surface) takes the continuation by default for a
    laminar P2 solve, as FRSolver.solve does (solver.py:236,292), and solves directly when it is
    switched off."""
    import torch
    from autoflowcfd_amd.api import run_steady
    from autoflowcfd_amd.continuation import solve_with_order_continuation
    from autoflowcfd_amd.solver import GPUFRSolver
    print("[gpu] run_steady laminar P2 with and without continuation", flush=True)
    mesh = wall_box(n=(3, 2, 2), seed=8)
    Q0 = freestream(0.15, 0.0)
    kw = dict(rho_inf=Q0[0], vel_inf=Q0[1], p_inf=Q0[4])
    res = run_steady(mesh, order=2, turbulence_model="none", max_iter=30, tol=1e-12, **kw)
    _, out = solve_with_order_continuation(lambda p: GPUFRSolver(mesh, order=p, **kw), 2, max_iter=30, tol=1e-12)
    assert res.iterations == out["iterations"] == 30
    assert res.final_residual == out["final_residual"]
    direct = run_steady(mesh, order=2, turbulence_model="none", max_iter=12, tol=1e-12, order_continuation=False,
                        **kw)
    s = GPUFRSolver(mesh, order=2, **kw)
    ref = s.solve(max_iter=12, tol=1e-12, output_interval=0)
    assert direct.iterations == 12 and direct.final_residual == ref["final_residual"]
    # SST at P2: the continuation by default too, as FRSolver.solve
    res_sst = run_steady(mesh, order=2, turbulence_model="sst", max_iter=6, tol=1e-12, **kw)
    _, out_sst = solve_with_order_continuation(lambda p: GPUFRSolver(mesh, order=p, turb_model="SST", **kw), 2,
                                               max_iter=6, tol=1e-12)
    assert res_sst.iterations == out_sst["iterations"] == 6
    assert res_sst.final_residual == out_sst["final_residual"]


@pytest.mark.parametrize("kind", ["wallbox", "wedges"])
@pytest.mark.parametrize("prec,tol", [("f64", 1e-10), ("f32", 1e-5)])
def test_p0_sst_step_matches_oracle(prec, tol, kind):
    """One P0 pseudo-step with SST (afr_p0_sst_step) against the oracle's general turbulence step
    with the one-point operators and the P0 stages: k, omega, nu_t, the state and the norm, on a
    perturbed wall-box state with non-uniform k, omega (so that convection, diffusion, the WALL
    mirror and the Neumann records all act)."""
    import torch
    from autoflowcfd_amd.continuation import P0Mesh
    from autoflowcfd_amd.device import bc_params_table
    from oracle import continuation as OC
    from oracle import fr_ops, sst
    from oracle.faces import build_flat_faces
    from oracle.geometry import Geometry
    print(f"[gpu] P0 SST step {prec} {kind} vs oracle", flush=True)
    dtype = torch.float64 if prec == "f64" else torch.float32
    mesh = wall_box(seed=7) if kind == "wallbox" else wedge_box()
    Q0 = freestream(0.2, 6.0)
    bcp = bc_params_table(Q0)
    bcd = {c: bcp[c] for c in range(8)}
    ops_t, ops0 = fr_ops.generate_operators(2), fr_ops.generate_operators(0)
    g_t = Geometry(mesh, ops_t)
    g0 = OC.P0Geometry(mesh, g_t)
    g0fr, fl0fr = Geometry(mesh, ops0), build_flat_faces(mesh, ops0)
    d = np.mean(sst.wall_distance(g_t.sps_xyz, mesh.nodes[mesh.wall_nodes]), axis=1, keepdims=True)
    U = _state(mesh, Q0, seed=4)
    rng = np.random.default_rng(5)
    k = 0.4 * (1.0 + 0.3 * rng.random((mesh.n_cells, 1)))
    w = 2600.0 * (1.0 + 0.3 * rng.random((mesh.n_cells, 1)))
    nut = 1e-5 * (1.0 + rng.random((mesh.n_cells, 1)))
    if dtype == torch.float32:
        U, k, w, nut = (x.astype(np.float32).astype(np.float64) for x in (U, k, w, nut))
    mu = 1.8e-5
    U_ref, k_ref, w_ref, nut_ref, res_ref = OC.p0_sst_step(U, k, w, nut, mesh, g0, g0fr, ops0, fl0fr, d, bcd, mu)
    dev = P0Mesh(mesh, g_t.volume, bcp, dtype=dtype, device="cuda", d_wall=d[:, 0])
    t = lambda x: torch.as_tensor(np.ascontiguousarray(x), dtype=dtype).cuda().contiguous()  # noqa: E731
    Ud, kd, wd, nutd = t(U), t(k[:, 0]), t(w[:, 0]), t(nut[:, 0])
    mutd = torch.zeros_like(kd)
    res = float(dev.sst_step(Ud, kd, wd, nutd, mutd, mu=mu).item())
    e = lambda a, b: np.linalg.norm(a - b) / np.linalg.norm(b)  # noqa: E731
    got = {n: x.double().cpu().numpy() for n, x in (("k", kd), ("w", wd), ("nut", nutd))}
    print(f"P0 SST {prec}: norm {res:.12e} / {res_ref:.12e}; k {e(got['k'], k_ref[:, 0]):.2e} "
          f"w {e(got['w'], w_ref[:, 0]):.2e} nut {e(got['nut'], nut_ref[:, 0]):.2e}")
    assert abs(res - res_ref) <= tol * res_ref
    assert e(got["nut"], nut_ref[:, 0]) < tol
    # the increments of k, omega and U (the fields themselves hide them)
    assert np.linalg.norm(got["k"] - k_ref[:, 0]) <= tol * np.linalg.norm(k_ref[:, 0] - k[:, 0]) * 10
    assert np.linalg.norm(got["w"] - w_ref[:, 0]) <= tol * np.linalg.norm(w_ref[:, 0] - w[:, 0]) * 10
    Un = Ud.double().cpu().numpy()
    assert np.linalg.norm(Un - U_ref) <= (1e-9 if prec == "f64" else 1e-3) * np.linalg.norm(U_ref - U)


def test_sst_order_continuation_matches_oracle_run():
    """The SST continuation P0 -> P1 -> P2 in FP64 against the oracle's run: per-phase histories,
    the final state and k, omega (reset fields at P0, cell-mean wall distance carried through)."""
    import torch
    from autoflowcfd_amd.continuation import solve_with_order_continuation
    from autoflowcfd_amd.device import bc_params_table
    from autoflowcfd_amd.solver import GPUFRSolver
    from oracle import continuation as OC
    from oracle import fr_ops
    print("[gpu] SST order continuation P0 -> P2 f64 vs oracle run", flush=True)
    mesh = wall_box(n=(3, 2, 2), seed=4)
    Q0 = freestream(0.15, 0.0)
    mu = 1.8e-5
    bcp = bc_params_table(Q0)
    U_ref, hist_ref, (k_ref, w_ref, nut_ref) = OC.run(mesh, 2, {c: bcp[c] for c in range(8)}, prim_to_cons(Q0), mu,
                                                      max_iter=30, tol=1e-12, fr_ops=fr_ops, turbulence=True)

    def make(p):
        return GPUFRSolver(mesh, order=p, rho_inf=Q0[0], vel_inf=Q0[1], p_inf=Q0[4], mu_molecular=mu,
                           dtype=torch.float64, turb_model="SST")
    hist = {}
    s, out = solve_with_order_continuation(make, 2, max_iter=30, tol=1e-12, history=hist)
    for p in hist_ref:
        a, b = np.array(hist[p]), np.array(hist_ref[p])
        d = np.max(np.abs(a - b) / b) if len(a) == len(b) else np.inf
        print(f"SST P{p}: {len(a)} steps, first {a[0]:.10e} / {b[0]:.10e}, last {a[-1]:.10e} / {b[-1]:.10e}, "
              f"max rel {d:.2e}")
        assert len(a) == len(b)
        assert d < 5e-6          # round-off of the free-stream start, amplified (as the laminar run)
    st = s.get_state_cpu()
    rel = np.linalg.norm(st["U"][..., :5] - U_ref) / np.linalg.norm(U_ref)
    kq = st["Q"][..., 5]
    rk = np.linalg.norm(kq - k_ref) / np.linalg.norm(k_ref)
    print(f"final state rel {rel:.3e}, k rel {rk:.3e}")
    assert rel < 1e-10 and rk < 1e-6
