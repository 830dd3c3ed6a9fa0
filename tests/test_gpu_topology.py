"""GPU: the face topology built on the device (afr_face_topology) equals the host build
(mesh.build_topology) record for record and bit for bit: owners, neighbours, cube-face codes and
node triples exactly, the FP64 normals, areas and centres as the same correctly rounded values.

Meshes: permuted graded prism boxes, displaced tet boxes, the mixed prism/tet mesh of the
reference's residual tests, a periodic box through hex_box (boundary classification and pairing
on top of either topology), node ids above 2^21 (the two-pass sort), collapsed-edge wedges (their
zero-area triangles dropped, also with ids above 2^21), and a face shared by three cells
(refused)."""

import numpy as np
import pytest

from cases import mixed_mesh, tet_box, wall_box, wedge_box

pytestmark = pytest.mark.gpu

FIELDS_INT = ("owner", "neighbor", "owner_code", "neighbor_code", "node_ids")
FIELDS_F64 = ("normal", "area", "center")


def _same(a, b):
    assert a.n_faces == b.n_faces
    for k in FIELDS_INT:
        assert np.array_equal(getattr(a, k), getattr(b, k)), k
    for k in FIELDS_F64:
        x, y = getattr(a, k), getattr(b, k)
        assert np.array_equal(x, y), (k, float(np.max(np.abs(x - y))))


def _shifted(mesh, shift):
    """The same cells with every node id raised by ``shift`` (unused nodes in front)."""
    from autoflowcfd_amd.mesh import Mesh
    nodes = np.concatenate([np.zeros((shift, 3)), mesh.nodes])
    return Mesh(nodes=nodes, prism_conn=mesh.prism_conn + shift, tet_conn=mesh.tet_conn + shift)


@pytest.mark.parametrize("kind", ["wallbox", "tets", "mixed", "big_ids", "wallbox30", "wedges", "wedges_big_ids"])
def test_device_topology_equals_host(kind):
    from autoflowcfd_amd.mesh import build_topology, hex_box
    from autoflowcfd_amd.topology import build_topology_gpu
    print(f"[gpu] face topology {kind}", flush=True)
    if kind == "wallbox":
        mesh = wall_box(n=(6, 5, 4), seed=3)
    elif kind == "tets":
        mesh = tet_box(4, seed=2, displacement=0.15)
    elif kind == "mixed":
        mesh = mixed_mesh()
    elif kind == "big_ids":
        mesh = _shifted(wall_box(n=(3, 3, 2), seed=1), 2 ** 21 + 5)
    elif kind == "wedges":
        mesh = wedge_box()
    elif kind == "wedges_big_ids":
        mesh = _shifted(wedge_box(n=(3, 3, 2)), 2 ** 21 + 5)
    else:
        mesh = hex_box(30, 30, 30, 4.0, 1.0, 1.0, split="prism", grading_y=1.1, seed=7)
    _same(build_topology_gpu(mesh), build_topology(mesh))


def test_hex_box_with_device_topology_equals_host_build():
    from autoflowcfd_amd.mesh import hex_box
    print("[gpu] hex_box periodic, topology gpu vs host", flush=True)
    kw = dict(split="prism", grading_y=1.05, seed=4, periodic="x", extrude="x",
              bc_by_plane={"y0": 0, "y1": 4, "z0": 4, "z1": 4})
    a = hex_box(5, 4, 3, 2.0, 1.0, 1.0, topology="gpu", **kw)
    b = hex_box(5, 4, 3, 2.0, 1.0, 1.0, topology="host", **kw)
    _same(a.faces, b.faces)
    assert np.array_equal(a.faces.bc, b.faces.bc) and np.array_equal(a.faces.translation, b.faces.translation)
    assert np.array_equal(a.wall_nodes, b.wall_nodes)


def test_face_of_three_cells_is_refused():
    from autoflowcfd_amd.mesh import Mesh
    from autoflowcfd_amd.topology import build_topology_gpu
    print("[gpu] face topology: a triple shared by three cells", flush=True)
    nodes = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1], [0, 0, -1], [1, 1, 1]], dtype=float)
    tets = np.array([[0, 1, 2, 3], [0, 1, 2, 4], [0, 1, 2, 5]])
    with pytest.raises(RuntimeError, match="more than two cells"):
        build_topology_gpu(Mesh(nodes=nodes, prism_conn=np.zeros((0, 6), np.int64), tet_conn=tets))
