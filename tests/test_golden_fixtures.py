"""The stored oracle fixtures still match the oracle (first step recomputed on the CPU)."""

import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))


def test_c1_fixture_first_step_matches_oracle():
    from oracle import fr_ops
    from oracle.faces import build_flat_faces
    from oracle.geometry import Geometry
    from oracle.time_step import ns_step
    from autoflowcfd_amd.device import bc_params_table
    sys.path.insert(0, os.path.join(HERE, "golden"))
    import make_config_goldens as gold
    ref = np.load(os.path.join(HERE, "golden", "config_c1_p1_euler_16.npz"))
    mesh, Q0 = gold.c1_setup()
    assert mesh.n_tet == int(ref["n_cells"]) == 24576
    ops = fr_ops.generate_operators(1)
    g = Geometry(mesh, ops)
    fl = build_flat_faces(mesh, ops)
    bcp = bc_params_table(Q0)
    U0 = gold.c1_state(g.sps_xyz, Q0)
    _, nr, _ = ns_step(U0, g, ops, fl, {c: bcp[c] for c in range(8)}, 1, mu=0.0, viscous=False)
    assert abs(nr - float(ref["norms"][0])) <= 1e-12 * abs(nr)
    assert ref["sample_U"].shape == (2000, 5) and np.all(np.isfinite(ref["sample_U"]))


def test_c1_fixture_sample_points_are_reproducible():
    """The recipe's sample_points (seeded) regenerates the stored sample indices."""
    sys.path.insert(0, os.path.join(HERE, "golden"))
    import make_config_goldens as gold
    ref = np.load(os.path.join(HERE, "golden", "config_c1_p1_euler_16.npz"))
    ci, si = gold.sample_points(int(ref["n_cells"]), 8)
    assert np.array_equal(ci, ref["sample_cell"]) and np.array_equal(si, ref["sample_sp"])
