"""Writes the FP64 oracle result of the full P1 Euler configuration as a small fixture (run on the
CPU; the GPU test compares against it because the oracle takes minutes at this size).

  config_c1_p1_euler_16.npz   P1 inviscid Euler, 16^3 hex x 6 tets (24,576 cells), seeded random
                              cell / node permutation and +-10 % sinusoidal node displacement, all
                              FARFIELD, Mach 0.2 (rho = 1, p = 1/gamma) plus a Gaussian density /
                              pressure pulse of amplitude 1e-2 at the centre; 10 local-time-step
                              SSP-RK3 steps.  Stored: the residual RMS before each step, the
                              global increment norm, per-variable checksums of the final state and
                              the final state at 2000 seeded solution points.

Run:  python tests/golden/make_config_goldens.py
"""

import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

GAMMA = 1.4
SAMPLE_SEED = 0  # seed of the 2000 stored sample points (PCG64)


def c1_setup():
    from autoflowcfd_amd.mesh import hex_box
    mesh = hex_box(16, 16, 16, split="tet", displacement=0.1, seed=11)
    Q0 = np.array([1.0, 0.2 * np.sqrt(GAMMA * (1 / GAMMA) / 1.0), 0.0, 0.0, 1 / GAMMA])
    return mesh, Q0


def c1_state(xyz, Q0):
    r2 = np.sum((xyz - 0.5) ** 2, axis=-1)
    pulse = 1e-2 * np.exp(-r2 / (2 * 0.1 ** 2))
    rho = Q0[0] * (1 + pulse)
    p = Q0[4] * (1 + pulse)
    u = Q0[1]
    E = p / (GAMMA - 1) + 0.5 * rho * u * u
    return np.stack([rho, rho * u, 0 * rho, 0 * rho, E], axis=-1)


def sample_points(n_cells, ns, k=2000):
    rng = np.random.default_rng(SAMPLE_SEED)
    return rng.integers(0, n_cells, k), rng.integers(0, ns, k)


def make_c1():
    from oracle import fr_ops
    from oracle.faces import build_flat_faces
    from oracle.geometry import Geometry
    from oracle.time_step import ns_step
    from autoflowcfd_amd.device import bc_params_table
    t0 = time.time()
    mesh, Q0 = c1_setup()
    ops = fr_ops.generate_operators(1)
    g = Geometry(mesh, ops)
    fl = build_flat_faces(mesh, ops)
    bcp = bc_params_table(Q0)
    bcd = {c: bcp[c] for c in range(8)}
    U0 = c1_state(g.sps_xyz, Q0)
    U = U0.copy()
    norms = []
    for i in range(10):
        U, nr, _ = ns_step(U, g, ops, fl, bcd, 1, mu=0.0, viscous=False)
        norms.append(nr)
        print(f"c1 step {i + 1}: norm {nr:.12e} ({time.time() - t0:.0f} s)", flush=True)
    ci, si = sample_points(U.shape[0], U.shape[1])
    np.savez_compressed(os.path.join(HERE, "config_c1_p1_euler_16.npz"), norms=np.array(norms),
                        inc_norm=np.linalg.norm(U - U0), checksum=U.sum(axis=(0, 1)),
                        checksum2=(U ** 2).sum(axis=(0, 1)), sample_cell=ci, sample_sp=si, sample_U=U[ci, si],
                        n_cells=mesh.n_tet)


if __name__ == "__main__":
    make_c1()
