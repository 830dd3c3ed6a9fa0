"""Oracle wall-force integral pinned by closed forms.

* uniform pressure on a closed all-WALL box integrates to zero;
* a linear shear u = a y over the no-slip plane y = 0 gives the skin-friction force
  F_x = mu a S and the pressure force -p S along y (outward normal -y), S the wall
  area (the prisms are extruded along y, so the wall is tiled by triangle records;
  each record is integrated with its own area, fr_coefficients.py:108-118);
* Cd/Cl use F_x and F_z with q = rho V^2 / 2 (fr_coefficients.py:140-150).
"""

import numpy as np

from cases import prim_to_cons
from autoflowcfd_amd.mesh import BC_SYMMETRY, BC_WALL, hex_box
from oracle import fr_ops
from oracle.faces import build_flat_faces
from oracle.forces import coefficients, wall_forces
from oracle.geometry import Geometry


def _setup(bc, split="prism", order=2, n=(3, 2, 2), L=(2.0, 1.0, 1.0)):
    m = hex_box(*n, *L, split=split, seed=1, bc_by_plane=bc)
    ops = fr_ops.generate_operators(order)
    return m, ops, Geometry(m, ops), build_flat_faces(m, ops)


def test_uniform_pressure_closed_box_is_zero():
    bc = {k: BC_WALL for k in ("x0", "x1", "y0", "y1", "z0", "z1")}
    for split in ("prism", "tet"):
        m, ops, g, fl = _setup(bc, split=split)
        Q = np.broadcast_to([1.2, 0.0, 0.0, 0.0, 1e5], (m.n_cells, ops.n_sps, 5))
        fp, fv = wall_forces(prim_to_cons(Q), g, ops, fl, 1.8e-5)
        assert np.abs(fp).max() < 1e-9 * 1e5
        assert np.abs(fv).max() < 1e-12


def test_linear_shear_skin_friction():
    bc = {"y0": BC_WALL, "x0": BC_SYMMETRY, "x1": BC_SYMMETRY, "y1": BC_SYMMETRY, "z0": BC_SYMMETRY,
          "z1": BC_SYMMETRY}
    a, mu = 30.0, 1.8e-5
    for split in ("prism", "tet"):
        m, ops, g, fl = _setup(bc, split=split)
        y = g.sps_xyz[..., 1]
        Q = np.zeros(y.shape + (5,))
        Q[..., 0], Q[..., 1], Q[..., 4] = 1.2, a * y, 1e5
        fp, fv = wall_forces(prim_to_cons(Q), g, ops, fl, mu)
        wall = (fl.bc == BC_WALL) & fl.owner_primary
        counted = fl.area[wall].sum()
        assert np.isclose(counted, 2.0)
        assert np.isclose(fv[0], mu * a * counted, rtol=1e-9)
        assert abs(fv[1]) < 1e-9 and abs(fv[2]) < 1e-9
        assert np.isclose(fp[1], -1e5 * counted, rtol=1e-12)  # outward normal -y
        c = coefficients(fp + fv, 1.2, 50.0, 2.0)
        assert np.isclose(c["Cd"], (fp + fv)[0] / (0.5 * 1.2 * 2500.0 * 2.0))
