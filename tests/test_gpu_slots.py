"""GPU: the face-slot tables built on the device (afr_face_slots + the host's location of the
block representatives) equal the host builder's (face_slots.build_slots) array for array and bit
for bit: sources, row blocks, split masks, boundary codes and rows, role-signed normals and the
interpolation rows themselves, at P1 and P2.

Meshes: graded prism wall boxes (permuted), displaced tet boxes, periodic tet and x-extruded prism
boxes (corner correspondence by position), the mixed prism / tet mesh, rotated prisms, the odd
23-prism mesh, a prism quad split between two tets (the split-slot path), a mesh with boundary-group
rows, collapsed-edge wedges (a quad face held by one triangle record), and a 30^3-hex wall box."""

import numpy as np
import pytest

from cases import mixed_mesh, odd_prism_mesh, rotated_prism_box, tet_box, wall_box, wedge_box

pytestmark = pytest.mark.gpu


def _split_quad_mesh():
    from autoflowcfd_amd.mesh import mesh_from_cells
    nodes = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1], [1, 0, 1], [0, 1, 1], [0.5, -1, 0.5]], float)
    return mesh_from_cells(nodes, np.array([[0, 1, 2, 3, 4, 5]]), np.array([[0, 1, 4, 6], [0, 3, 4, 6]]))


def _mesh(kind):
    from autoflowcfd_amd.mesh import hex_box
    if kind == "wallbox":
        return wall_box(n=(5, 4, 3), seed=2)
    if kind == "tets":
        return tet_box(3, seed=1, displacement=0.12)
    if kind == "tets_periodic":
        return tet_box(3, seed=2, displacement=0.0, periodic=True)
    if kind == "prisms_periodic_x":
        return hex_box(5, 3, 2, 2.0, 1.0, 1.0, split="prism", periodic="x", extrude="x", seed=3,
                       bc_by_plane={"y0": 0, "y1": 1, "z0": 1, "z1": 1})
    if kind == "mixed":
        return mixed_mesh()
    if kind == "rotated":
        return rotated_prism_box(seed=4)
    if kind == "odd":
        return odd_prism_mesh()
    if kind == "split_quad":
        return _split_quad_mesh()
    if kind == "wedges":
        return wedge_box()
    if kind == "group_rows":
        m = wall_box(n=(3, 3, 2), seed=5)
        m.faces.bc_row = np.where(m.faces.neighbor < 0, np.arange(m.faces.n_faces) % 8, -1)
        return m
    return hex_box(30, 30, 30, 4.0, 1.0, 1.0, split="prism", grading_y=1.1, seed=7,
                   bc_by_plane={"y0": 0, "x0": 2, "x1": 3, "y1": 4, "z0": 4, "z1": 4})


@pytest.mark.parametrize("order", [1, 2])
@pytest.mark.parametrize("kind", ["wallbox", "tets", "tets_periodic", "prisms_periodic_x", "mixed", "rotated",
                                  "odd", "split_quad", "group_rows", "wallbox30", "wedges"])
def test_device_slots_equal_host(kind, order):
    from autoflowcfd_amd.face_slots import build_slots
    from autoflowcfd_amd.fr_tables import tables
    from autoflowcfd_amd.slots_gpu import build_slots_gpu
    if kind == "wallbox30" and order == 1:
        pytest.skip("one order suffices at this size")
    print(f"[gpu] face slots {kind} P{order}", flush=True)
    mesh = _mesh(kind)
    T = tables(order)
    a = build_slots_gpu(mesh, T)
    b = build_slots(mesh, T, with_geometry=False)
    for k in ("nb", "blk", "mask", "bc", "tn", "rows"):
        x, y = getattr(a, k), getattr(b, k)
        assert x.shape == y.shape and x.dtype == y.dtype, (k, x.shape, y.shape, x.dtype, y.dtype)
        assert np.array_equal(x, y), k
    assert (a.bc_row is None) == (b.bc_row is None)
    if a.bc_row is not None:
        assert np.array_equal(a.bc_row, b.bc_row)
    assert (a.n_blocks, a.keyed_slots, a.split_slots) == (b.n_blocks, b.keyed_slots, b.split_slots)
    if kind == "split_quad":
        assert a.split_slots == 1
