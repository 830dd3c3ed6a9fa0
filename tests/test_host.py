"""Host-side logic: meshes, face slots, sparse rows, locality order, C ABI.

The face-slot tables that the HIP kernels gather through are checked by a NumPy
emulation of the device algorithm (tests/slot_emulator.py) against the FP64
oracle's face-major inviscid residual on mixed, prism, tet and periodic meshes.
"""

import ctypes
import os
import re
import subprocess

import numpy as np
import pytest

import slot_emulator as se
from cases import freestream, mixed_mesh, smooth_state, tet_box, wall_box
from autoflowcfd_amd.device import bc_params_table, morton_order, sparse_rows
from autoflowcfd_amd.face_slots import build_slots
from autoflowcfd_amd.fr_tables import tables
from autoflowcfd_amd.geometry import cell_block_metrics
from autoflowcfd_amd.mesh import hex_box
from oracle import fr_ops
from oracle import residual as R
from oracle.faces import build_flat_faces
from oracle.geometry import Geometry

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

MESHES = [
    ("mixed", mixed_mesh, 1), ("mixed", mixed_mesh, 2),
    ("tetbox", lambda: tet_box(3, seed=1), 2),
    ("wallbox", lambda: wall_box(seed=2), 1), ("wallbox", lambda: wall_box(seed=2), 2),
    ("periodic", lambda: tet_box(3, seed=3, displacement=0.0, periodic=True), 2),
]


@pytest.mark.parametrize("name,make,order", MESHES)
def test_slot_tables_reproduce_oracle_residual(name, make, order):
    m = make()
    T = tables(order)
    ops = fr_ops.generate_operators(order)
    g = Geometry(m, ops)
    fl = build_flat_faces(m, ops)
    Q0 = freestream(0.15)
    bcp = bc_params_table(Q0)
    bcd = {c: bcp[c] for c in range(8)}
    U = smooth_state(g.sps_xyz, Q0, seed=4)
    ref = R.inviscid_residual(U, g, ops, fl, bcd)
    geo = cell_block_metrics(m, T, np.float64)
    sl = build_slots(m, T)
    em = se.inviscid(m, T, geo, sl, U, bcd, R.euler_flux, R.ausm_up, R.ghost_state, True)
    tol = 1e-7 if (order == 2 and m.n_tet) else 1e-11
    assert np.linalg.norm(em - ref) / np.linalg.norm(ref) < tol


def test_hex_box_prism_split_counts_and_orientation():
    m = hex_box(3, 2, 2, 1.0, 1.0, 1.0, split="prism", seed=0)
    assert m.n_prism == 2 * 3 * 2 * 2 and m.n_tet == 0
    T = tables(1)
    geo = cell_block_metrics(m, T, np.float64)
    assert np.all(geo["det"] > 0)
    assert np.isclose(geo["volume"].sum(), 1.0, rtol=1e-12)
    m = hex_box(2, 2, 2, 1.0, 2.0, 1.0, split="tet", seed=1, displacement=0.05)
    geo = cell_block_metrics(m, tables(2), np.float64)
    assert m.n_tet == 6 * 8 and np.all(geo["det"] > 0)
    assert np.isclose(geo["volume"].sum(), 2.0, rtol=1e-10)


def test_slots_are_a_partition_of_faces():
    """Every interior face appears in exactly two slots (one per side), boundary faces in one."""
    m = wall_box(seed=5)
    sl = build_slots(m, tables(2), with_geometry=False)
    interior = sl.bc == -1
    assert np.all(sl.nb[..., 0][interior] >= 0)
    # neighbour relation is symmetric
    pairs = set()
    for c in range(m.n_cells):
        for s in np.flatnonzero(interior[c]):
            pairs.add((c, int(sl.nb[c, s, 0])))
    assert all((b, a) in pairs for a, b in pairs)


def test_sparse_rows_round_trip():
    rng = np.random.default_rng(0)
    rows = np.zeros((20, 27))
    for r in range(20):
        k = rng.integers(1, 10)
        rows[r, rng.choice(27, size=k, replace=False)] = rng.normal(size=k)
    rows[3, 5] = 1e-14  # round-off of a structural zero is dropped
    nnz, idx, w = sparse_rows(rows)
    dense = np.zeros_like(rows)
    for r in range(20):
        dense[r, idx[r, :nnz[r]]] = w[r, :nnz[r]]
        assert np.all(w[r, nnz[r]:] == 0)
    rows[3, 5] = 0.0
    assert np.array_equal(dense, rows)


def test_morton_order_is_a_locality_permutation():
    rng = np.random.default_rng(1)
    pts = rng.uniform(size=(5000, 3))
    perm = morton_order(pts)
    assert np.array_equal(np.sort(perm), np.arange(5000))
    step = np.linalg.norm(np.diff(pts[perm], axis=0), axis=1).mean()
    rand = np.linalg.norm(np.diff(pts, axis=0), axis=1).mean()
    assert step < 0.25 * rand


# ------------------------------------------------------------------ C ABI

def _header_functions():
    txt = open(os.path.join(ROOT, "include", "afr.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\bint\s+(afr_[a-z0-9_]+)\s*\(", txt)))


def test_library_exports_every_header_function():
    from autoflowcfd_amd import _lib
    lib = _lib.load()
    names = _header_functions()
    assert names, "no functions parsed from include/afr.h"
    for n in names:
        assert hasattr(lib, n), n
        assert n in _lib.SIGNATURES, n
    assert set(_lib.SIGNATURES) == set(names)
    assert lib.afr_version() == _lib.ABI_VERSION


def test_abi_struct_layouts_match_ctypes(tmp_path):
    from autoflowcfd_amd import _lib
    src = tmp_path / "layout.c"
    fields_m = [f[0] for f in _lib.AfrMeshDesc._fields_]
    fields_g = [f[0] for f in _lib.AfrGeomDesc._fields_]
    fields_p = [f[0] for f in _lib.AfrP0Desc._fields_]
    fields_t = [f[0] for f in _lib.AfrTopoDesc._fields_]
    fields_s = [f[0] for f in _lib.AfrSlotDesc._fields_]
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "afr.h"', 'int main(void) {',
             'printf("%zu %zu %zu %zu %zu\\n", sizeof(AfrMeshDesc), sizeof(AfrGeomDesc), sizeof(AfrP0Desc), '
             'sizeof(AfrTopoDesc), sizeof(AfrSlotDesc));']
    lines += [f'printf("%zu\\n", offsetof(AfrMeshDesc, {f}));' for f in fields_m]
    lines += [f'printf("%zu\\n", offsetof(AfrGeomDesc, {f}));' for f in fields_g]
    lines += [f'printf("%zu\\n", offsetof(AfrP0Desc, {f}));' for f in fields_p]
    lines += [f'printf("%zu\\n", offsetof(AfrTopoDesc, {f}));' for f in fields_t]
    lines += [f'printf("%zu\\n", offsetof(AfrSlotDesc, {f}));' for f in fields_s]
    lines += ['return 0; }']
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-std=c99", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split()
    vals = [int(x) for x in out]
    assert vals[0] == ctypes.sizeof(_lib.AfrMeshDesc) and vals[1] == ctypes.sizeof(_lib.AfrGeomDesc)
    assert vals[2] == ctypes.sizeof(_lib.AfrP0Desc) and vals[3] == ctypes.sizeof(_lib.AfrTopoDesc)
    assert vals[4] == ctypes.sizeof(_lib.AfrSlotDesc)
    off_m = [getattr(_lib.AfrMeshDesc, f).offset for f in fields_m]
    off_g = [getattr(_lib.AfrGeomDesc, f).offset for f in fields_g]
    off_p = [getattr(_lib.AfrP0Desc, f).offset for f in fields_p]
    off_t = [getattr(_lib.AfrTopoDesc, f).offset for f in fields_t]
    off_s = [getattr(_lib.AfrSlotDesc, f).offset for f in fields_s]
    offs = off_m + off_g + off_p + off_t + off_s
    assert vals[5:] == offs


def test_error_paths_without_gpu():
    from autoflowcfd_amd import _lib
    lib = _lib.load()
    h = ctypes.c_void_p()
    assert lib.afr_mesh_create(None, ctypes.byref(h)) != 0
    buf = ctypes.create_string_buffer(256)
    lib.afr_last_error(buf, 256)
    assert b"null" in buf.value
    d = _lib.AfrMeshDesc()
    d.abi_version = _lib.ABI_VERSION + 100
    assert lib.afr_mesh_create(ctypes.byref(d), ctypes.byref(h)) != 0
    lib.afr_last_error(buf, 256)
    assert b"ABI" in buf.value


def test_meshio_round_trip(tmp_path):
    from autoflowcfd_amd.meshio import load_mesh, save_mesh
    m = wall_box(seed=3)
    p = str(tmp_path / "m.npz")
    save_mesh(p, m)
    m2 = load_mesh(p)
    assert m2.n_cells == m.n_cells and m2.faces.n_faces == m.faces.n_faces
    key = lambda fr: {tuple(t): c for t, c in zip(map(tuple, fr.node_ids), fr.bc)}  # noqa: E731
    assert key(m2.faces) == key(m.faces)
    assert np.array_equal(np.sort(m2.wall_nodes), np.sort(m.wall_nodes))


def test_concat_meshes_is_a_disjoint_union():
    """Batched multi-case mesh: every case's slot tables, rows and records stay inside the case,
    and the face-major oracle residual of the union, cut back into cases, is each case's own."""
    from autoflowcfd_amd.mesh import concat_meshes
    parts = [mixed_mesh(), wall_box(n=(3, 2, 2), seed=7), tet_box(2, seed=5), wall_box(seed=2)]
    m, cc, po, to, ro = concat_meshes(parts)
    assert m.n_prism == sum(p.n_prism for p in parts) and m.n_tet == sum(p.n_tet for p in parts)
    assert list(po) == list(np.concatenate([[0], np.cumsum([p.n_prism for p in parts])]))
    assert list(ro) == list(np.concatenate([[0], np.cumsum([p.faces.n_faces for p in parts])]))
    assert np.array_equal(m.meta["wall_off"], np.concatenate([[0], np.cumsum(
        [0 if p.wall_nodes is None else len(p.wall_nodes) for p in parts])]))
    T = tables(2)
    sl = build_slots(m, T)
    src = sl.nb                                   # (C, 5, 2) source cells, -1 where none
    own = np.broadcast_to(cc[:, None, None], src.shape)
    assert np.all((src < 0) | (cc[np.maximum(src, 0)] == own))
    fr = m.faces
    inner = fr.neighbor >= 0
    assert np.all(cc[fr.owner[inner]] == cc[fr.neighbor[inner]])
    # per-case slot tables packed by concat_slots: the same structure as the union's own, and
    # bit for bit each case's rows at every used flux point
    from autoflowcfd_amd.face_slots import concat_slots
    per = [build_slots(p, T) for p in parts]
    sc = concat_slots(per, [p.n_prism for p in parts], [p.n_tet for p in parts], T.nfp)
    for k in ("nb", "mask", "bc", "fallback"):
        assert np.array_equal(getattr(sc, k), getattr(sl, k)), k
    for i, p in enumerate(parts):
        cells = np.concatenate([np.arange(po[i], po[i + 1]), m.n_prism + np.arange(to[i], to[i + 1])])
        for c_loc, c in enumerate(cells):
            for s in range(5):
                if sc.bc[c, s] != -1:
                    continue
                for fp in range(T.nfp):
                    src = (sc.mask[c, s] >> fp) & 1
                    assert np.array_equal(sc.rows[sc.blk[c, s, src] * T.nfp + fp],
                                          per[i].rows[per[i].blk[c_loc, s, src] * T.nfp + fp])
        assert np.array_equal(sc.adjrow[cells], per[i].adjrow)
    # residual of the union == per-case residuals (one free stream for every case here)
    ops = fr_ops.generate_operators(2)
    Q0 = freestream(0.15)
    bcp = bc_params_table(Q0)
    bcd = {c: bcp[c] for c in range(8)}
    g = Geometry(m, ops)
    U = smooth_state(g.sps_xyz, Q0, seed=4)
    Rall = R.inviscid_residual(U, g, ops, build_flat_faces(m, ops), bcd)
    for i, p in enumerate(parts):
        cells = np.concatenate([np.arange(po[i], po[i + 1]), m.n_prism + np.arange(to[i], to[i + 1])])
        gi = Geometry(p, ops)
        Ri = R.inviscid_residual(U[cells], gi, ops, build_flat_faces(p, ops), bcd)
        # FP64 noise (point-location Newton, summation order): 1.5e-12 of max|R| measured
        assert np.abs(Rall[cells] - Ri).max() < 1e-10 * np.abs(Ri).max(), i


def test_every_module_imports():
    """Every product module imports on a machine without a GPU (syntax and import-time errors
    show here, not first on the GPU)."""
    import importlib
    for name in ("autoflowcfd_amd.solver", "autoflowcfd_amd.batch", "autoflowcfd_amd.serve",
                 "autoflowcfd_amd.autotune", "autoflowcfd_amd.api", "autoflowcfd_amd.meshio",
                 "autoflowcfd_amd.device", "autoflowcfd_amd.face_slots", "autoflowcfd_amd.mesh",
                 "autoflowcfd_amd.fr_tables", "autoflowcfd_amd.geometry", "autoflowcfd_amd.build"):
        importlib.import_module(name)


def test_renumbered_mesh_keeps_the_residual():
    """A mesh stored in another cell order gives the same oracle residual, cell for cell."""
    from autoflowcfd_amd.mesh import locality_renumbered
    for make, order in ((lambda: wall_box(seed=2), 2), (mixed_mesh, 1)):
        m = make()
        m2, new_of_old = locality_renumbered(m)
        ops = fr_ops.generate_operators(order)
        Q0 = freestream(0.15)
        bcp = bc_params_table(Q0)
        bcd = {c: bcp[c] for c in range(8)}
        g = Geometry(m, ops)
        U = smooth_state(g.sps_xyz, Q0, seed=4)
        R1 = R.inviscid_residual(U, g, ops, build_flat_faces(m, ops), bcd)
        U2 = np.empty_like(U)
        U2[new_of_old] = U
        R2 = R.inviscid_residual(U2, Geometry(m2, ops), ops, build_flat_faces(m2, ops), bcd)
        assert np.abs(R2[new_of_old] - R1).max() <= 1e-12 * np.abs(R1).max()


def test_graph_entry_points_reject_bad_arguments_without_a_gpu():
    """The C-ABI graph calls validate their arguments before any HIP call: a null stream, output
    or graph is an error with a message (CPU-checkable), and destroying a null graph is a no-op."""
    import ctypes
    import pytest
    from autoflowcfd_amd import _lib
    lib = _lib.load()
    with pytest.raises(RuntimeError, match="null stream"):
        _lib.check(lib.afr_graph_begin(None), "afr_graph_begin")
    with pytest.raises(RuntimeError, match="null stream or output"):
        _lib.check(lib.afr_graph_end(None, None), "afr_graph_end")
    with pytest.raises(RuntimeError, match="null graph"):
        _lib.check(lib.afr_graph_launch(None, None), "afr_graph_launch")
    assert lib.afr_graph_destroy(None) == 0
    from autoflowcfd_amd.device import LibGraph
    with pytest.raises(RuntimeError, match="nothing captured"):
        LibGraph(lib).launch(None)


def test_slot_mask_wide_flag():
    """Bit 30 of the slot mask marks slots whose first-source row block has a row wider than
    order + 1; the per-flux-point source bits are kept; slots without a source stay unflagged."""
    import numpy as np
    from autoflowcfd_amd.device import slot_mask_with_wide_flag
    nfp, order = 9, 2
    nnz = np.array([3] * 9 + [3] * 8 + [9] + [0])          # block 0 narrow, block 1 one wide row, + zero row
    mask = np.array([[0, 5, 0], [1, 0, 0]])
    blk0 = np.array([[0, 1, -1], [1, 0, -1]])
    out = slot_mask_with_wide_flag(mask, blk0, nnz, 2, nfp, order)
    assert out.tolist() == [[0, 5 | (1 << 30), 0], [1 | (1 << 30), 0, 0]]
    assert ((out & ((1 << 30) - 1)) == mask).all()
    m = wall_box(seed=1)
    from autoflowcfd_amd import face_slots
    from autoflowcfd_amd.device import sparse_rows
    sl = face_slots.slots_for(m, 2)
    nz, _, _ = sparse_rows(np.concatenate([sl.rows, np.zeros((1, sl.rows.shape[1]))]))
    flags = (slot_mask_with_wide_flag(sl.mask, np.asarray(sl.blk)[..., 0], nz, sl.n_blocks, 9, 2) >> 30) & 1
    assert flags.sum(axis=1).max() <= 1 and flags.any()     # a graded prism box: at most one wide slot per cell


def test_bench_refuses_more_than_one_gpu():
    """bench.py runs on one GPU: --gpus N with N > 1 exits (status 2) with a message, before
    touching a device."""
    import subprocess
    import sys
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    r = subprocess.run([sys.executable, os.path.join(root, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU only" in r.stderr and r.stdout == ""


def test_bench_algorithmic_byte_models():
    """The bench's algorithmic-byte models at the C2 sizes (the figures quoted for the kernel
    rooflines): inviscid 4.196 GB, viscous volume 5.27 GB, viscous interface 4.87 GB, and the
    full-residual B_res 5.863 GB."""
    import importlib.util
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    spec = importlib.util.spec_from_file_location("bench_models", os.path.join(root, "bench.py"))
    b = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(b)
    C, ns, nfine, nfp, nfaces = 1102736, 27, 64, 9, 2783736
    n_sp, n_fine = C * ns, C * nfine
    assert abs(b.algorithmic_bytes_inviscid(n_sp, n_fine, nfaces, nfp) / 1e9 - 4.196) < 0.01
    assert abs(b.algorithmic_bytes_visc_volume(n_sp, C) / 1e9 - 5.27) < 0.01
    assert abs(b.algorithmic_bytes_visc_iface(n_sp, C, nfp) / 1e9 - 4.87) < 0.01
    assert abs(b.algorithmic_bytes_residual(n_sp, n_fine, nfaces, nfp) / 1e9 - 5.863) < 0.01
