"""GPU: resident solves stepped as one batched mesh assembled on the device (DeviceMesh.concat +
ResidentBatch), the serving path's continuous batching.

Each solve keeps its own device mesh; the batch concatenates their arrays on the device.  With a
common reference state every member must follow its trajectory alone bit for bit, through graph
replays and through a membership change (a member leaves, the others continue in a new batch)."""

import numpy as np
import pytest

from cases import freestream, wall_box

pytestmark = pytest.mark.gpu


def _requests(turb):
    from autoflowcfd_amd.batch import Case
    from autoflowcfd_amd.serve import SolveRequest
    out = []
    for j, (n, mach, yaw) in enumerate([((4, 3, 3), 0.15, 0.0), ((3, 3, 2), 0.25, 7.0), ((5, 2, 2), 0.08, 3.0)]):
        Q = freestream(mach, yaw)
        U = float(np.linalg.norm(Q[1:4]))
        k = 1.5 * (0.01 * U) ** 2
        out.append(SolveRequest(rid=j, case=Case(mesh=wall_box(n=n, seed=40 + j), Q_free=Q, k_inf=k,
                                                 omega_inf=Q[0] * k / (10 * 1.8e-5)), order=2))
    return out


@pytest.mark.parametrize("turb", ["NONE", "SST"])
def test_resident_batch_equals_each_solve_alone_through_a_membership_change(turb):
    import torch
    from autoflowcfd_amd.batch import ResidentBatch, static_ref_state
    from autoflowcfd_amd.serve import build_solver
    print(f"[gpu] ResidentBatch {turb} fp32 vs alone", flush=True)
    reqs = _requests(turb)
    ref = static_ref_state([r.case for r in reqs])
    alone = [build_solver(r, torch.float32, turb, ref_state=ref) for r in reqs]
    for s in alone:
        for _ in range(5):
            s.step()
    mem = [build_solver(r, torch.float32, turb, ref_state=ref) for r in reqs]
    for s in mem:
        s.step()                                     # each solve's first step alone, as the server does
    b = ResidentBatch(mem, ref)
    b.capture_graph()
    hist = []
    for _ in range(2):
        b.launch()
        hist.append(b.norms.cpu().numpy().copy())
    for i, s in enumerate(alone):
        assert [h[i] for h in hist] == s.residual_history[1:3], (i, [h[i] for h in hist], s.residual_history[1:3])
    b.write_back()
    del b
    b2 = ResidentBatch([mem[0], mem[2]], ref)        # solve 1 retired
    b2.capture_graph()
    for _ in range(2):
        b2.launch()
        n = b2.norms.cpu().numpy()
        for j, i in enumerate((0, 2)):
            mem[i].residual_history.append(float(n[j]))
    b2.write_back()
    for i in (0, 2):
        assert mem[i].residual_history[-2:] == alone[i].residual_history[3:5], i
        a, m = alone[i].get_state_cpu(), mem[i].get_state_cpu()
        assert np.array_equal(a["U"], m["U"]), i
