"""GPU: GPUFRSolver surface (step / solve / state I/O / checkpoints / coefficients) and the CLI.

The pseudo-step arithmetic itself is pinned by test_gpu_step.py and test_gpu_sst.py; here
the solver's first step is compared with one oracle pseudo-step from the uniform initial
state (FP64, 1e-10), and the I/O paths are checked for exact round trips.
"""

import os
import subprocess
import sys

import numpy as np
import pytest

from cases import smooth_state, wall_box

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_solver_first_step_matches_oracle_and_io_round_trips(tmp_path):
    import torch
    from autoflowcfd_amd.solver import GPUFRSolver
    from oracle import fr_ops
    from oracle.faces import build_flat_faces
    from oracle.geometry import Geometry
    from oracle.time_step import ns_step
    from autoflowcfd_amd.device import bc_params_table
    print("[gpu] GPUFRSolver NS first step + I/O", flush=True)
    mesh = wall_box(seed=4)
    s = GPUFRSolver(mesh, order=2, rho_inf=1.225, vel_inf=51.04, p_inf=101325.0, dtype=torch.float64)
    ops = fr_ops.generate_operators(2)
    g = Geometry(mesh, ops)
    # a perturbed start: from the uniform freestream the first increment is ~1e-11 of |U|, at
    # the FP64 rounding level of the state itself, and compares nothing
    U0 = smooth_state(g.sps_xyz, np.array([1.225, 51.04, 0.0, 0.0, 101325.0]), amp=1e-2, seed=5)
    s.set_state_from_cpu(U0)
    res = s.step()
    fl = build_flat_faces(mesh, ops)
    bcp = bc_params_table(np.array([1.225, 51.04, 0.0, 0.0, 101325.0]))
    U_ref, norm_ref, _ = ns_step(U0, g, ops, fl, {c: bcp[c] for c in range(8)}, 2, mu=1.8e-5)
    U1 = s.get_state_cpu()["U"]
    inc = np.linalg.norm((U1 - U0) - (U_ref - U0)) / max(np.linalg.norm(U_ref - U0), 1e-300)
    assert inc < 1e-8, inc
    assert abs(res - norm_ref) <= 1e-10 * max(norm_ref, 1e-300) + 1e-12
    out = s.solve(max_iter=3, tol=0.0, output_interval=0)
    assert out["iterations"] == 4 and len(out["residual_history"]) == 4
    path = str(tmp_path / "ck.npz")
    s.save_checkpoint(path)
    U_save = s.get_state_cpu()["U"]
    s.step()
    s.load_checkpoint(path)
    assert np.array_equal(s.get_state_cpu()["U"], U_save) and s.iteration == 4
    s.set_state_from_cpu(U0)
    assert np.array_equal(s.get_state_cpu()["U"], U0)
    c = s.compute_aerodynamic_coefficients(reference_area=1.0)
    assert np.isfinite(c["Cd"]) and np.isfinite(c["Cl"])


def test_solver_sst_runs_and_cli():
    import torch
    from autoflowcfd_amd.solver import GPUFRSolver
    print("[gpu] GPUFRSolver SST + CLI", flush=True)
    s = GPUFRSolver(wall_box(seed=5), order=1, vel_inf=51.04, turb_model="SST", k_inf=0.3908, omega_inf=2660.0)
    out = s.solve(max_iter=2, tol=0.0, output_interval=0)
    assert np.isfinite(out["final_residual"]) and s.get_state_cpu()["U"].shape[-1] == 7
    r = subprocess.run([sys.executable, "-m", "autoflowcfd_amd", "solve", "steady", "wallbox:3", "--order", "1",
                        "--max-iter", "2", "--turbulence-model", "none", "--output", "/tmp/afr_cli_out",
                        "--reference-area", "1.0"], cwd=ROOT, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr[-2000:]
    assert '"Cd"' in r.stdout


def test_graph_replay_matches_eager():
    """A captured pseudo-step, replayed several times with eager work and syncs between, gives
    the eager trajectory bit for bit."""
    import torch
    from autoflowcfd_amd.solver import GPUFRSolver
    print("[gpu] hipGraph pseudo-step replay", flush=True)
    mk = lambda: GPUFRSolver(wall_box(seed=6), order=2, vel_inf=51.04, turb_model="SST",  # noqa: E731
                             k_inf=0.3908, omega_inf=2660.0)
    a, b = mk(), mk()
    eager = [a.step() for _ in range(4)]
    b.capture_graph()            # includes one eager warm-up step
    replay = [b.residual_history[-1]]
    for _ in range(3):
        replay.append(b.step())
        torch.cuda.synchronize()
        _ = b.get_state_cpu()    # eager work between replays
    assert replay == eager, (replay, eager)
    assert np.array_equal(a.get_state_cpu()["U"], b.get_state_cpu()["U"])
