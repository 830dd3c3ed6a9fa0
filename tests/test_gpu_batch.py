"""GPU: the batched multi-case form (CaseBatch) against each case stepped alone.

Every kernel computes a cell from that cell's own data, the slot tables of a batch are each
case's own (concat_slots), the per-case norm reduces in the single-case order and the wall
distance is taken case by case; so with the batch's reference state, each case stepped alone
by GPUFRSolver must give the batched norms, state and wall forces BIT FOR BIT.
"""

import numpy as np
import pytest

from cases import freestream, mixed_mesh, tet_box, wall_box

pytestmark = pytest.mark.gpu


def _alone(case, order, dtype, turb, ref_state):
    from autoflowcfd_amd.batch import cons
    from autoflowcfd_amd.mesh import BC_FARFIELD, BC_INLET
    from autoflowcfd_amd.solver import GPUFRSolver
    Q = np.asarray(case.Q_free, dtype=np.float64)
    s = GPUFRSolver(case.mesh, order=order, rho_inf=Q[0], vel_inf=float(np.linalg.norm(Q[1:4])), p_inf=Q[4],
                    turb_model=turb, dtype=dtype, k_inf=case.k_inf, omega_inf=case.omega_inf,
                    bc_overrides={BC_FARFIELD: Q, BC_INLET: Q}, ref_state=ref_state)
    C, ns = s.dm.n_cells, s.dm.ns
    s.set_state_from_cpu(np.broadcast_to(cons(Q), (C, ns, 5)).copy())
    return s


def _run_pair(cases, order, dtype, turb, n_steps, graph):
    import torch
    from autoflowcfd_amd.batch import CaseBatch
    b = CaseBatch(cases, order=order, dtype=dtype, turb_model=turb)
    if graph:
        b.capture_graph()
        for _ in range(n_steps - 1):
            b.step()
            torch.cuda.synchronize()
    else:
        for _ in range(n_steps):
            b.step()
    hist_b = np.array(b.history)
    for i, c in enumerate(cases):
        s = _alone(c, order, dtype, turb, b.ref_state)
        hist_a = [s.step() for _ in range(n_steps)]
        assert np.array_equal(hist_b[:, i], np.array(hist_a)), (i, hist_b[:, i], hist_a)
        st = b.case_state(i)
        alone = s.get_state_cpu()
        assert np.array_equal(st["U"], alone["U"][..., :5]), i
        if turb == "SST":
            assert np.array_equal(st["k"], alone["Q"][..., 5]), i
            assert np.array_equal(st["omega"], alone["Q"][..., 6]), i
    return b


def test_sst_batch_equals_each_case_alone():
    """P2 SST FP32: three wall boxes of different sizes, Mach and yaw; eager steps."""
    import torch
    from autoflowcfd_amd.batch import Case
    print("[gpu] CaseBatch SST P2 fp32 vs alone", flush=True)
    cases = []
    for j, (n, mach, yaw) in enumerate([((4, 4, 3), 0.15, 0.0), ((3, 2, 2), 0.25, 7.0), ((5, 3, 2), 0.08, 3.0)]):
        Q = freestream(mach, yaw)
        U = float(np.linalg.norm(Q[1:4]))
        k = 1.5 * (0.01 * U) ** 2
        cases.append(Case(mesh=wall_box(n=n, seed=10 + j), Q_free=Q, k_inf=k, omega_inf=Q[0] * k / (10 * 1.8e-5)))
    b = _run_pair(cases, 2, torch.float32, "SST", 3, graph=False)
    coef = b.coefficients()
    assert len(coef) == 3 and all(np.isfinite(c["Cd"]) for c in coef)


def test_sst_batch_forces_equal_alone():
    import torch
    from autoflowcfd_amd.batch import Case, CaseBatch
    print("[gpu] CaseBatch SST P1 fp32 forces vs alone", flush=True)
    cases = [Case(mesh=wall_box(n=(3, 3, 2), seed=20 + j), Q_free=freestream(0.1 + 0.1 * j, 5.0 * j), k_inf=0.1,
                  omega_inf=500.0) for j in range(3)]
    b = CaseBatch(cases, order=1, dtype=torch.float32, turb_model="SST")
    for _ in range(2):
        b.step()
    F = b.wall_forces()
    for i, c in enumerate(cases):
        s = _alone(c, 1, torch.float32, "SST", b.ref_state)
        s.step()
        s.step()
        Fa = s.dm.wall_forces(s.U, s.mu_t, mu=s.mu_molecular).cpu().numpy()
        assert np.array_equal(F[i], Fa), (i, F[i], Fa)


def test_laminar_mixed_batch_graph_equals_alone():
    """P2 Navier-Stokes FP32 over prism, tet and mixed cases (tet offsets), graph replay."""
    import torch
    from autoflowcfd_amd.batch import Case
    print("[gpu] CaseBatch NS P2 fp32 mixed/tet/prism graph vs alone", flush=True)
    cases = [Case(mesh=mixed_mesh(), Q_free=freestream(0.2, 0.0)),
             Case(mesh=tet_box(3, seed=1), Q_free=freestream(0.1, 4.0)),
             Case(mesh=wall_box(n=(3, 2, 2), seed=3), Q_free=freestream(0.3, 9.0))]
    _run_pair(cases, 2, torch.float32, "NONE", 3, graph=True)


def test_batched_cd_matches_oracle_per_case():
    """The batched multi-case configuration's report: per-case Cd / Cl against the FP64 oracle.
    Three laminar P2 wall boxes of different sizes, Mach and yaw in one FP64 batch, each given its
    own smooth perturbed state (a skin friction and a pressure field to integrate): the batched
    per-case wall forces (afr_forces_cases_f64) against the oracle's wall-force integral of each
    case alone, to 1e-10 of the force scale.  (Pseudo-steps of the batch are tied to the oracle
    through the bitwise batch-equals-alone tests above and the single-case step tests.)"""
    import torch
    from autoflowcfd_amd.batch import Case, CaseBatch, free_stream_coefficients
    from cases import smooth_state
    from oracle import fr_ops
    from oracle.faces import build_flat_faces
    from oracle.forces import wall_forces
    from oracle.geometry import Geometry
    print("[gpu] CaseBatch P2 fp64 per-case Cd vs oracle", flush=True)
    spec = [((4, 3, 2), 0.15, 0.0), ((3, 2, 2), 0.25, 7.0), ((4, 2, 3), 0.08, 3.0)]
    cases = [Case(mesh=wall_box(n=n, seed=40 + j), Q_free=freestream(m, y)) for j, (n, m, y) in enumerate(spec)]
    b = CaseBatch(cases, order=2, dtype=torch.float64, turb_model="NONE")
    ops = fr_ops.generate_operators(2)
    states = []
    for i, c in enumerate(cases):
        g = Geometry(c.mesh, ops)
        U = smooth_state(g.sps_xyz, c.Q_free, amp=2e-2, seed=50 + i)
        states.append((g, build_flat_faces(c.mesh, ops), U))
        idx = torch.as_tensor(b.case_cells(i), device=b.U.device)
        b.U[idx] = torch.as_tensor(U, dtype=torch.float64, device=b.U.device)
    coef = b.coefficients()
    for i, (c, (g, fl, U)) in enumerate(zip(cases, states)):
        Fp, Fv = wall_forces(U, g, ops, fl, 1.8e-5)
        ref = free_stream_coefficients(np.concatenate([Fp, Fv]), c.Q_free, c.reference_area)
        scale = max(abs(ref["Cd"]), abs(ref["Cl"]), abs(ref["Cs"]))
        assert scale > 0
        for key in ("Cd", "Cl", "Cs"):
            assert abs(coef[i][key] - ref[key]) <= 1e-10 * scale, (i, key, coef[i][key], ref[key])
