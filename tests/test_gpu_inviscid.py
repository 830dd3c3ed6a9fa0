"""GPU parity: inviscid FR residual of the HIP path vs the FP64 oracle.

Tolerances: FP64 device path rel. L2 <= 1e-10 (1e-7 with P2 tets, whose
collapsed-basis operators reach 4e5); FP32 device path rel. L2 <= 1e-5 against
the oracle evaluated on the same FP32 state.
"""

import numpy as np
import pytest

from cases import freestream, mixed_mesh, rotated_prism_box, smooth_state, tet_box, wall_box, wedge_box

pytestmark = pytest.mark.gpu


def _oracle(mesh, order, U, bcp):
    from oracle import fr_ops
    from oracle.faces import build_flat_faces
    from oracle.geometry import Geometry
    from oracle.residual import inviscid_residual
    ops = fr_ops.generate_operators(order)
    g = Geometry(mesh, ops)
    fl = build_flat_faces(mesh, ops)
    return inviscid_residual(U, g, ops, fl, bcp), g


def _run(mesh, order, dtype, amp=1e-2, seed=0, mach=0.15):
    import torch
    from autoflowcfd_amd.device import DeviceMesh, bc_params_table
    from oracle import fr_ops
    from oracle.geometry import Geometry
    Q0 = freestream(mach)
    bcp = bc_params_table(Q0)
    g = Geometry(mesh, fr_ops.generate_operators(order))
    U = smooth_state(g.sps_xyz, Q0, amp=amp, seed=seed)
    if dtype == torch.float32:
        U = U.astype(np.float32).astype(np.float64)  # the oracle sees the same (FP32) state
    R_ref, _ = _oracle(mesh, order, U, {c: bcp[c] for c in range(8)})
    dm = DeviceMesh(mesh, order, dtype=dtype, bc_params=bcp)
    Ud = torch.as_tensor(U, dtype=dtype, device="cuda").contiguous()
    R = dm.inviscid_residual(Ud).double().cpu().numpy()
    torch.cuda.synchronize()
    err = np.linalg.norm(R - R_ref) / np.linalg.norm(R_ref)
    return err, R, R_ref


CASES = [
    ("mixed", mixed_mesh, 1), ("mixed", mixed_mesh, 2),
    ("tetbox", lambda: tet_box(3, seed=1), 1), ("tetbox", lambda: tet_box(3, seed=1), 2),
    ("wallbox", lambda: wall_box(seed=2), 1), ("wallbox", lambda: wall_box(seed=2), 2),
    ("periodic", lambda: tet_box(3, seed=3, displacement=0.0, periodic=True), 2),
    ("rotprisms", rotated_prism_box, 2),
    ("wedges", wedge_box, 1), ("wedges", wedge_box, 2),
]


@pytest.mark.parametrize("name,make,order", CASES)
def test_inviscid_fp64_matches_oracle(name, make, order):
    import torch
    print(f"[gpu] inviscid fp64 {name} P{order}", flush=True)
    err, R, R_ref = _run(make(), order, torch.float64)
    tol = 1e-7 if name in ("tetbox", "periodic", "mixed") and order == 2 else 1e-10  # P2 tet operators reach 4e5
    assert err < tol, f"{name} P{order}: rel L2 {err:.3e}"


@pytest.mark.parametrize("name,make,order", [c for c in CASES if c[0] in ("wallbox", "mixed", "rotprisms", "wedges") and not (c[0] == "mixed" and c[2] == 2)])
def test_inviscid_fp32_matches_oracle(name, make, order):
    import torch
    print(f"[gpu] inviscid fp32 {name} P{order}", flush=True)
    err, R, R_ref = _run(make(), order, torch.float32)
    assert err < 1e-5, f"{name} P{order}: rel L2 {err:.3e}"
