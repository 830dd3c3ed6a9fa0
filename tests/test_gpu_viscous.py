"""GPU parity: viscous BR1 residual and full Navier-Stokes residual vs the FP64 oracle.

Tolerances: FP64 device path rel. L2 <= 1e-10 (1e-7 for meshes with P2 tets,
whose collapsed-basis derivative operators reach 4e5); FP32 device path
rel. L2 <= 1e-5 against the oracle evaluated on the same FP32-rounded state.
"""

import numpy as np
import pytest

from cases import freestream, mixed_mesh, rotated_prism_box, smooth_state, tet_box, wall_box, wedge_box

pytestmark = pytest.mark.gpu

MU = 1.8e-5


def _oracle(mesh, order, U, bcp, mu_t, mu):
    from oracle import fr_ops
    from oracle.faces import build_flat_faces
    from oracle.geometry import Geometry
    from oracle.residual import inviscid_residual, viscous_residual
    ops = fr_ops.generate_operators(order)
    g = Geometry(mesh, ops)
    fl = build_flat_faces(mesh, ops)
    Rv = viscous_residual(U, g, ops, fl, bcp, mu, mu_t=mu_t)
    Ri = inviscid_residual(U, g, ops, fl, bcp)
    return Rv, Ri


def _run(mesh, order, dtype, amp=1e-2, seed=0, mu=MU, with_mut=True):
    import torch
    from autoflowcfd_amd.device import DeviceMesh, bc_params_table
    from oracle import fr_ops
    from oracle.geometry import Geometry
    Q0 = freestream(0.15)
    bcp = bc_params_table(Q0)
    g = Geometry(mesh, fr_ops.generate_operators(order))
    U = smooth_state(g.sps_xyz, Q0, amp=amp, seed=seed)
    rng = np.random.default_rng(seed + 7)
    mut = (mu * rng.uniform(0.0, 50.0, size=U.shape[:2])) if with_mut else None
    if dtype == torch.float32:
        U = U.astype(np.float32).astype(np.float64)
        mut = None if mut is None else mut.astype(np.float32).astype(np.float64)
    Rv_ref, Ri_ref = _oracle(mesh, order, U, {c: bcp[c] for c in range(8)}, mut, mu)
    dm = DeviceMesh(mesh, order, dtype=dtype, bc_params=bcp)
    Ud = torch.as_tensor(U, dtype=dtype, device="cuda").contiguous()
    mtd = None if mut is None else torch.as_tensor(mut, dtype=dtype, device="cuda").contiguous()
    Rv = dm.viscous_residual(Ud, mtd, mu=mu).double().cpu().numpy()
    Rf = dm.residual(Ud, mtd, mu=mu).double().cpu().numpy()
    torch.cuda.synchronize()
    ev = np.linalg.norm(Rv - Rv_ref) / np.linalg.norm(Rv_ref)
    R_ref = Ri_ref + Rv_ref
    ef = np.linalg.norm(Rf - R_ref) / np.linalg.norm(R_ref)
    return ev, ef


CASES = [
    ("wallbox", lambda: wall_box(seed=2), 1), ("wallbox", lambda: wall_box(seed=2), 2),
    ("mixed", mixed_mesh, 1), ("mixed", mixed_mesh, 2),
    ("tetbox", lambda: tet_box(3, seed=1), 1), ("tetbox", lambda: tet_box(3, seed=1), 2),
    ("rotprisms", rotated_prism_box, 1), ("rotprisms", rotated_prism_box, 2),
    ("wedges", wedge_box, 1), ("wedges", wedge_box, 2),
]


@pytest.mark.parametrize("name,make,order", CASES)
def test_viscous_fp64_matches_oracle(name, make, order):
    import torch
    print(f"[gpu] viscous fp64 {name} P{order}", flush=True)
    ev, ef = _run(make(), order, torch.float64)
    tol = 1e-7 if name in ("tetbox", "mixed") and order == 2 else 1e-10
    assert ev < tol, f"{name} P{order}: viscous rel L2 {ev:.3e}"
    assert ef < tol, f"{name} P{order}: full residual rel L2 {ef:.3e}"


@pytest.mark.parametrize("name,make,order", [c for c in CASES if c[0] in ("wallbox", "rotprisms", "wedges")] + [CASES[2]])
def test_viscous_fp32_matches_oracle(name, make, order):
    import torch
    print(f"[gpu] viscous fp32 {name} P{order}", flush=True)
    ev, ef = _run(make(), order, torch.float32)
    assert ev < 1e-5, f"{name} P{order}: viscous rel L2 {ev:.3e}"
    assert ef < 1e-5, f"{name} P{order}: full residual rel L2 {ef:.3e}"
