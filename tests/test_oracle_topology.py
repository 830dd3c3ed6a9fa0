"""CPU: the product's face topology, boundary-group tags and periodic pairing against the oracle's
own restatement (oracle/topology.py: plain loops over dicts and sets, sharing no code or table
with autoflowcfd_amd.mesh / autoflowcfd_amd.boundary).

Every record's owner, neighbour, cube-face codes and node triple are compared exactly, and its
FP64 normal, area and centre bit for bit (both follow the reference's operation order).  The GPU
topology is compared with the product's host build bit for bit in tests/test_gpu_topology.py, so
it is tied to the oracle through this test.  A planted mistake in the oracle (the other diagonal
of the prism quads, owner and neighbour swapped, the last group not winning) is caught."""

import numpy as np
import pytest

from cases import mixed_mesh, odd_prism_mesh, rotated_prism_box, tet_box, wall_box, wedge_box


def _same(prod, ref):
    for k in ("owner", "neighbor", "owner_code", "neighbor_code", "node_ids"):
        assert np.array_equal(np.asarray(getattr(prod, k), dtype=np.int64), ref[k]), k
    for k in ("normal", "area", "center"):
        assert np.array_equal(getattr(prod, k), ref[k]), (k, float(np.max(np.abs(getattr(prod, k) - ref[k]))))


MESHES = {
    "wallbox": lambda: wall_box(n=(4, 3, 3), seed=2),
    "tets": lambda: tet_box(3, seed=1, displacement=0.15),
    "mixed": mixed_mesh,
    "rotated": lambda: rotated_prism_box(seed=4),
    "odd": odd_prism_mesh,
    "wedges": wedge_box,
}


@pytest.mark.parametrize("kind", sorted(MESHES))
def test_host_topology_equals_oracle(kind):
    from autoflowcfd_amd.mesh import build_topology
    from oracle.topology import face_records
    mesh = MESHES[kind]()
    _same(build_topology(mesh), face_records(mesh.nodes, mesh.prism_conn, mesh.tet_conn))


def _groups_box():
    """A graded prism box, x-extruded so that its x = 0 and x = 2 planes carry congruent
    triangles, with boundary groups given as owner-cell lists (the reference's BoundaryMap
    convention), corner cells in two groups."""
    from autoflowcfd_amd.mesh import build_topology, hex_box
    m = hex_box(4, 3, 3, 2.0, 1.0, 1.0, split="prism", grading_y=1.1, seed=3, periodic="x", extrude="x")
    fr = build_topology(m)
    bnd = fr.neighbor < 0
    x, y = fr.center[:, 0], fr.center[:, 1]
    groups = {
        "inlet": np.unique(fr.owner[bnd & (np.abs(x) < 1e-12)]),
        "outlet": np.unique(fr.owner[bnd & (np.abs(x - 2.0) < 1e-12)]),
        "wall": np.unique(fr.owner[bnd & (np.abs(y) < 1e-12)]),
        "far": np.unique(fr.owner[bnd & (np.abs(y - 1.0) < 1e-12)]),
    }
    return m, fr, groups


def test_group_tags_equal_oracle():
    from autoflowcfd_amd.boundary import tag_boundary_groups
    from oracle.topology import tag_groups
    m, fr, groups = _groups_box()
    code, names = tag_boundary_groups(fr, groups)
    ref, ref_names = tag_groups(fr.owner, fr.neighbor, groups)
    assert names == ref_names
    assert np.array_equal(code.astype(np.int64), ref)
    # corner cells sit in two groups: the later group's code wins on all their boundary records
    assert np.any(np.isin(groups["inlet"], groups["wall"]))


def test_periodic_pairing_equals_oracle():
    from autoflowcfd_amd.boundary import pair_periodic
    from oracle.topology import face_records, pair_periodic as pair_ref
    m, fr, groups = _groups_box()
    t = np.array([2.0, 0.0, 0.0])
    got = pair_periodic(fr, groups, "inlet", "outlet", t)
    rec = face_records(m.nodes, m.prism_conn, m.tet_conn)
    ref = pair_ref(rec, groups, "inlet", "outlet", t)
    _same(got, ref)
    assert np.array_equal(got.translation, ref["translation"])
    assert got.n_faces == fr.n_faces - int(np.sum((fr.neighbor < 0) & (np.abs(fr.center[:, 0]) < 1e-12)))


def test_planted_oracle_mistakes_are_caught(monkeypatch):
    from autoflowcfd_amd.boundary import tag_boundary_groups
    from autoflowcfd_amd.mesh import build_topology
    import oracle.topology as top
    mesh = wall_box(n=(3, 2, 2), seed=1)
    prod = build_topology(mesh)

    def other_diagonal(cell):
        b, t = [int(x) for x in cell[:3]], [int(x) for x in cell[3:]]
        o = sorted(range(3), key=lambda i: b[i])
        v0, v1, v2 = (b[i] for i in o)
        w0, w1, w2 = (t[i] for i in o)
        return [(v0, v1, v2), (w0, w1, w2), (v0, v1, w0), (v1, w0, w1), (v1, v2, w1), (v2, w1, w2),
                (v0, v2, w0), (v2, w0, w2)]
    with monkeypatch.context() as mp:
        mp.setattr(top, "_prism_triangles", other_diagonal)
        with pytest.raises(AssertionError):
            _same(prod, top.face_records(mesh.nodes, mesh.prism_conn, mesh.tet_conn))
    ref = top.face_records(mesh.nodes, mesh.prism_conn, mesh.tet_conn)
    swapped = dict(ref)
    inner = np.flatnonzero(ref["neighbor"] >= 0)[0]
    swapped["owner"] = ref["owner"].copy()
    swapped["neighbor"] = ref["neighbor"].copy()
    swapped["owner"][inner], swapped["neighbor"][inner] = ref["neighbor"][inner], ref["owner"][inner]
    with pytest.raises(AssertionError):
        _same(prod, swapped)
    # a tagging where the first group holding the owner wins differs on the corner records
    m, fr, groups = _groups_box()
    code, _ = tag_boundary_groups(fr, groups)
    first = np.full(fr.n_faces, -1, dtype=np.int64)
    for g, cells in enumerate(groups.values()):
        first[(fr.neighbor < 0) & np.isin(fr.owner, cells) & (first < 0)] = g
    assert not np.array_equal(code.astype(np.int64), first)


def _closure(rec, n_cells):
    """Per-cell sum of the outward area vectors of its records (owner +, neighbour -)."""
    s = np.zeros((n_cells, 3))
    av = rec["normal"] * rec["area"][:, None]
    np.add.at(s, rec["owner"], av)
    inner = rec["neighbor"] >= 0
    np.add.at(s, rec["neighbor"][inner], -av[inner])
    return s


@pytest.mark.parametrize("kind", sorted(MESHES))
def test_oracle_records_close_every_cell(kind):
    """The divergence theorem pins the oracle's records without the product: every cell is closed
    by its records (outward area vectors sum to zero), so a missing or duplicated triangle, a wrong
    quad split or a wrong orientation fails; codes name distinct cube faces of each cell."""
    from oracle.topology import face_records
    mesh = MESHES[kind]()
    rec = face_records(mesh.nodes, mesh.prism_conn, mesh.tet_conn)
    scale = float(rec["area"].max())
    assert np.abs(_closure(rec, mesh.n_cells)).max() <= 1e-13 * scale
    # a prism has 8 triangle records (2 per quad; 6 for a collapsed-edge wedge), a tet 4
    count = np.bincount(rec["owner"], minlength=mesh.n_cells) + np.bincount(
        rec["neighbor"][rec["neighbor"] >= 0], minlength=mesh.n_cells)
    wedge = np.sum(mesh.prism_conn[:, :3] == mesh.prism_conn[:, 3:], axis=1) == 1
    assert np.all(count[:mesh.n_prism] == np.where(wedge, 6, 8)) and np.all(count[mesh.n_prism:] == 4)
    # planted: one record's normal turned round breaks the closure of its cells
    bad = dict(rec)
    bad["normal"] = rec["normal"].copy()
    bad["normal"][0] *= -1.0
    assert np.abs(_closure(bad, mesh.n_cells)).max() > 1e-3 * scale


def test_collapsed_wedge_records():
    """A prism whose side edge v2-w2 collapses to a point: the reference drops its two zero-area
    triangles (6 records), the oracle and the product do the same; other repeated vertices (a
    degenerate cap, two collapsed edges, a tet with a repeated vertex) are refused."""
    from autoflowcfd_amd.mesh import Mesh, build_topology
    from oracle.topology import face_records
    nodes = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1], [1, 0, 1], [0.3, 0.3, 1.0]], dtype=float)
    none = np.zeros((0, 4), np.int64)
    conn = np.array([[0, 1, 2, 3, 4, 2]])
    rec = face_records(nodes, conn, none)
    assert len(rec["owner"]) == 6 and np.all(rec["area"] > 0)
    assert np.abs(_closure(rec, 1)).max() <= 1e-15
    _same(build_topology(Mesh(nodes=nodes, prism_conn=conn, tet_conn=none)), rec)
    for bad_p, bad_t in (([[0, 1, 1, 3, 4, 5]], None), ([[0, 1, 2, 0, 4, 2]], None), (None, [[0, 1, 2, 2]])):
        m = Mesh(nodes=nodes, prism_conn=np.array(bad_p if bad_p else np.zeros((0, 6), np.int64)),
                 tet_conn=np.array(bad_t if bad_t else none))
        with pytest.raises(ValueError, match="repeated vertices"):
            build_topology(m)
