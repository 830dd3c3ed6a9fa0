"""GPU: the accuracy criterion of the FP32 path, Cd within 0.1 % of the FP64 oracle after a fixed
number of pseudo-steps.

A laminar graded wall box (no-slip wall at y = 0, INLET / OUTLET / SYMMETRY elsewhere, as the C2
workload), Mach 0.15 with 5 degrees of yaw, a seeded smooth 1 % perturbation of the free stream and
a developed wall layer u = u_inf (1 - exp(-y / 0.2)) so that the skin friction, the whole of Cd on
the flat wall, is a real signal; the FP32 GPUFRSolver takes 20 SSP-RK3 pseudo-steps with the local time step
(src/autoflowcfd/core/fr_solver/step.py:15-221) and the FP64 oracle the same 20 from the same
FP32-rounded start; the wall-force coefficients (src/autoflowcfd/postprocess/fr_coefficients.py:27-152,
fixed axes) must agree to 0.1 % of Cd (and of |F|/q A for Cl and Cs, whose reference the
pressure on the wall sets).  P2: at P1 the reference's per-stage modal filter removes every
non-constant mode (sigma = exp(-alpha) = 2.2e-16 at eta = 1, src/autoflowcfd/fr/modal_filter.py:47-128),
so the cells' own velocity gradients, and with them the skin friction, vanish after the first
stage (tests/test_oracle_operators.py pins that filter)."""

import numpy as np
import pytest

from cases import smooth_state, wall_box

pytestmark = pytest.mark.gpu

MU = 1.8e-5
N_STEPS = 20


@pytest.mark.parametrize("order", [2])
def test_fp32_cd_after_fixed_steps_matches_fp64_oracle(order):
    import torch
    from autoflowcfd_amd.device import bc_params_table
    from autoflowcfd_amd.solver import GPUFRSolver
    from oracle import fr_ops
    from oracle.faces import build_flat_faces
    from oracle.forces import coefficients, wall_forces
    from oracle.geometry import Geometry
    from oracle.time_step import ns_step
    print(f"[gpu] FP32 Cd after {N_STEPS} steps vs FP64 oracle, P{order}", flush=True)
    mesh = wall_box(n=(4, 4, 3), seed=11)
    u = 0.15 * np.sqrt(1.4 * 101325.0 / 1.225)
    yaw = np.deg2rad(5.0)
    Q0 = np.array([1.225, u * np.cos(yaw), 0.0, u * np.sin(yaw), 101325.0])
    ops = fr_ops.generate_operators(order)
    g = Geometry(mesh, ops)
    fl = build_flat_faces(mesh, ops)
    # a developed wall layer: the velocity of the free stream times 1 - exp(-y / 0.2), so that the
    # skin friction (the whole of Cd on the flat y = 0 wall) is a real signal, plus the seeded
    # perturbation
    prof = 1.0 - np.exp(-g.sps_xyz[..., 1] / 0.2)
    U0 = smooth_state(g.sps_xyz, Q0, amp=1e-2, seed=7)
    for v in (1, 2, 3):
        U0[..., 4] -= 0.5 * U0[..., v] ** 2 / U0[..., 0] * (1.0 - prof ** 2)
        U0[..., v] *= prof
    U0 = U0.astype(np.float32).astype(np.float64)
    bcp = bc_params_table(Q0)
    s = GPUFRSolver(mesh, order=order, rho_inf=1.225, vel_inf=u, p_inf=101325.0, mu_molecular=MU,
                    bc_overrides={1: Q0, 2: Q0}, dtype=torch.float32)
    s.set_state_from_cpu(U0)
    for _ in range(N_STEPS):
        s.step()
    c_dev = s.compute_aerodynamic_coefficients(reference_area=1.0)
    U = U0
    for _ in range(N_STEPS):
        U, _, _ = ns_step(U, g, ops, fl, {c: bcp[c] for c in range(8)}, order, mu=MU)
    fp, fv = wall_forces(U, g, ops, fl, MU)
    c_ref = coefficients(fp + fv, 1.225, u, 1.0)
    scale = max(abs(c_ref["Cd"]), abs(c_ref["Cl"]), abs(c_ref["Cs"]))
    print(f"P{order}: Cd dev {c_dev['Cd']:.8e} oracle {c_ref['Cd']:.8e}; Cs dev {c_dev['Cs']:.6e} oracle {c_ref['Cs']:.6e}")
    assert abs(c_dev["Cd"] - c_ref["Cd"]) <= 1e-3 * abs(c_ref["Cd"])
    for k in ("Cl", "Cs"):
        assert abs(c_dev[k] - c_ref[k]) <= 1e-3 * scale
