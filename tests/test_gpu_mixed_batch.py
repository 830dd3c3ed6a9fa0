"""GPU: the mixed-order batch (P1 and P2 SST cases in one HIP graph per pseudo-step) against each case
alone, and its per-case Cd against the FP64 oracle.

* Bit identity: every case of a MixedBatch, after graph-replayed pseudo-steps, has exactly the
  norms and state of the same case stepped alone by GPUFRSolver with the batch's reference state.
* Cd against the oracle (the batched configuration's report): graded wall boxes of different size,
  Mach, yaw and order, SST, each started from its free stream with a developed wall layer (so that
  the skin friction is a real signal), a fixed budget of pseudo-steps; each case's Cd (P2 cases: at
  P1 the per-stage filter leaves no cell gradient, hence no skin friction) and Cl / Cs against the
  oracle's SST pseudo-steps (turbulence step with dt_local, then SSP-RK3 with mu_t frozen,
  src/autoflowcfd/core/fr_solver/step.py:15-221) and wall-force integral
  (src/autoflowcfd/postprocess/fr_coefficients.py:27-152): FP64 to 1e-6 (from a uniform start the
  interior residual is round-off of O(1e5) fluxes, which moves the trajectory by ~1e-7 between
  summation orders), FP32 to 0.1 %.
"""

import numpy as np
import pytest

from cases import freestream, wall_box
from test_gpu_batch import _alone

pytestmark = pytest.mark.gpu

MU = 1.8e-5


def _cases():
    from autoflowcfd_amd.batch import Case
    out, orders = [], []
    for j, (n, mach, yaw, p) in enumerate([((4, 3, 3), 0.15, 0.0, 2), ((3, 3, 2), 0.25, 7.0, 1),
                                           ((3, 2, 2), 0.08, 3.0, 2), ((4, 2, 2), 0.3, 10.0, 1)]):
        Q = freestream(mach, yaw)
        U = float(np.linalg.norm(Q[1:4]))
        k = 1.5 * (0.01 * U) ** 2
        out.append(Case(mesh=wall_box(n=n, seed=30 + j), Q_free=Q, k_inf=k, omega_inf=Q[0] * k / (10 * MU)))
        orders.append(p)
    return out, orders


def test_mixed_order_sst_batch_graph_equals_each_case_alone():
    import torch
    from autoflowcfd_amd.batch import MixedBatch
    print("[gpu] MixedBatch P1+P2 SST fp32 graph vs alone", flush=True)
    cases, orders = _cases()
    b = MixedBatch(cases, orders, dtype=torch.float32, turb_model="SST")
    b.capture_graph()
    for _ in range(3):
        b.step()
        torch.cuda.synchronize()
    hist = np.array(b.history)
    for i, c in enumerate(cases):
        s = _alone(c, orders[i], torch.float32, "SST", b.ref_state)
        ha = [s.step() for _ in range(4)]
        assert np.array_equal(hist[:, i], np.array(ha)), (i, hist[:, i], ha)
        st, alone = b.case_state(i), s.get_state_cpu()
        assert np.array_equal(st["U"], alone["U"][..., :5]), i
        assert np.array_equal(st["k"], alone["Q"][..., 5]) and np.array_equal(st["omega"], alone["Q"][..., 6]), i


def _wall_layer_state(case, order):
    """The case's free stream with a developed wall layer, velocity x (1 - exp(-y / (0.2 H))), on
    its solution points (case mesh order): the skin friction is then a real signal."""
    from autoflowcfd_amd.batch import cons
    from oracle import fr_ops
    from oracle.geometry import Geometry
    g = Geometry(case.mesh, fr_ops.generate_operators(order))
    H = np.ptp(case.mesh.nodes[:, 1])
    prof = 1.0 - np.exp(-g.sps_xyz[..., 1] / (0.2 * H))
    Q = np.broadcast_to(np.asarray(case.Q_free, dtype=np.float64), prof.shape + (5,)).copy()
    Q[..., 1:4] *= prof[..., None]
    return cons(Q)


def _oracle_sst_cd(case, order, n_steps):
    from autoflowcfd_amd.batch import cons
    from autoflowcfd_amd.device import bc_params_table
    from autoflowcfd_amd.mesh import BC_WALL
    from oracle import fr_ops, sst
    from oracle.faces import build_flat_faces
    from oracle.forces import coefficients, wall_forces
    from oracle.geometry import Geometry
    from oracle.residual import inviscid_residual, viscous_residual
    from oracle.time_step import local_time_step, ssp_rk3, update_primitives
    mesh, Q0 = case.mesh, np.asarray(case.Q_free, dtype=np.float64)
    ops = fr_ops.generate_operators(order)
    g, fl = Geometry(mesh, ops), build_flat_faces(mesh, ops)
    bcp = bc_params_table(Q0, Q0)
    bcd = {c: bcp[c] for c in range(8)}
    d = sst.wall_distance(g.sps_xyz, mesh.nodes[mesh.wall_nodes])
    U = _wall_layer_state(case, order)
    k = np.full((mesh.n_cells, ops.n_sps), case.k_inf)
    w = np.full_like(k, case.omega_inf)
    nut = np.zeros_like(k)
    for _ in range(n_steps):
        Q = update_primitives(U)
        # dt_local sees the previous step's eddy viscosity (step.py:69-115)
        dt = local_time_step(Q, g, order, MU, Q[..., 0] * nut)
        k, w, nut, _, _ = sst.turbulence_step(U, k, w, g, ops, fl, d, dt, MU, fl.bc == BC_WALL)
        mut = Q[..., 0] * nut

        def rhs(X, mut=mut):
            return inviscid_residual(X, g, ops, fl, bcd) + viscous_residual(X, g, ops, fl, bcd, MU, mu_t=mut)
        U = ssp_rk3(U, rhs, dt, ops, mesh.n_prism)
    # the eddy viscosity of the last step (rho at its start times nu_t), as the device keeps it
    fp, fv = wall_forces(U, g, ops, fl, MU, mu_t=mut)
    return coefficients(fp + fv, Q0[0], float(np.linalg.norm(Q0[1:4])), case.reference_area)


@pytest.mark.parametrize("prec,tol", [("f64", 1e-6), ("f32", 1e-3)])
def test_mixed_order_sst_batch_cd_matches_fp64_oracle(prec, tol):
    import torch
    from autoflowcfd_amd.batch import MixedBatch
    print(f"[gpu] MixedBatch P1+P2 SST {prec} per-case Cd vs oracle", flush=True)
    cases, orders = _cases()
    n_steps = 3
    dt = torch.float64 if prec == "f64" else torch.float32
    b = MixedBatch(cases, orders, dtype=dt, turb_model="SST")
    for i, c in enumerate(cases):
        U0 = _wall_layer_state(c, orders[i])
        b.set_case_state(i, U0 if prec == "f64" else U0.astype(np.float32).astype(np.float64))
    b.run(n_steps)
    got = b.coefficients()
    for i, c in enumerate(cases):
        ref = _oracle_sst_cd(c, orders[i], n_steps)
        scale = max(abs(ref["Cd"]), abs(ref["Cl"]), abs(ref["Cs"]))
        print(f"case {i} P{orders[i]}: Cd {got[i]['Cd']:.10e} oracle {ref['Cd']:.10e}, Cs {got[i]['Cs']:.8e} "
              f"oracle {ref['Cs']:.8e}")
        if orders[i] == 2:
            # at P1 the per-stage modal filter removes every cell gradient: no skin friction
            assert abs(got[i]["Cd"] - ref["Cd"]) <= tol * abs(ref["Cd"]), (i, got[i]["Cd"], ref["Cd"])
        for key in ("Cl", "Cs"):
            assert abs(got[i][key] - ref[key]) <= tol * scale, (i, key, got[i][key], ref[key])
