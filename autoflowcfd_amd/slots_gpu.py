"""Face-slot tables built on the GPU (libafr.so: afr_face_slots).

The same ``Slots`` as face_slots.build_slots(mesh, T, with_geometry=False), bit for bit: the
device assigns every real cube face its primary record, writes the per-slot sources, boundary
codes and rows, role-signed normals, and keys every single-source slot by its face
correspondence; a device radix sort numbers the distinct keys (the row blocks, in ascending key
order) and picks each block's first slot.  The host then locates the flux points of those few
representatives (face_flux_points.py:164-236), checks that the key determines them on up to
``check_samples`` members per block, and builds the rows of the rare prism quads split between two
cells (face_flux_points_data.py:145-217).
"""

import ctypes

import numpy as np
import torch

from . import _lib
from .face_slots import NSLOT, SLOT_CODES, Slots, _own_fp_phys, _quad_lower_mask, _rows_for
from .fr_tables import HOST_LOCK


def build_slots_gpu(mesh, T, check_samples=4, device=None) -> Slots:
    lib = _lib.load()
    if not torch.cuda.is_available():
        raise RuntimeError("build_slots_gpu needs a visible GPU (HIP device)")
    dv = torch.device("cuda", torch.cuda.current_device()) if device is None else torch.device(device)
    fr = mesh.faces
    C, P, F, nfp = mesh.n_cells, mesh.n_prism, fr.n_faces, T.nfp
    n_nodes = mesh.nodes.shape[0]
    # inputs the kernels index with: validated here so that every device access stays in bounds
    for conn in (mesh.prism_conn, mesh.tet_conn):
        if conn.size and (conn.min() < 0 or conn.max() >= n_nodes):
            raise ValueError("cell connectivity names a node outside the node array")
    if F < 1 or fr.owner.min() < 0 or fr.owner.max() >= C or fr.neighbor.min() < -1 or fr.neighbor.max() >= C:
        raise ValueError("face records name cells outside the mesh")
    inner = fr.neighbor >= 0
    if (fr.owner_code.min() < 0 or fr.owner_code.max() > 5 or np.any(fr.neighbor_code[inner] < 0)
            or np.any(fr.neighbor_code[inner] > 5)):
        raise ValueError("face records with cube-face codes outside 0..5")
    n = C * NSLOT

    def up(a, dt):
        return torch.as_tensor(np.ascontiguousarray(a, dtype=dt)).to(dv)
    nodes = up(mesh.nodes, np.float64)
    pc, tc = up(mesh.prism_conn.reshape(-1, 6), np.int32), up(mesh.tet_conn.reshape(-1, 4), np.int32)
    rec_in = {"owner": up(fr.owner, np.int32), "neighbor": up(fr.neighbor, np.int32),
              "owner_code": up(fr.owner_code, np.int32), "neighbor_code": up(fr.neighbor_code, np.int32),
              "bc": up(fr.bc, np.int32), "normal": up(fr.normal, np.float64),
              "translation": up(fr.translation, np.float64)}
    bc_row_in = None if fr.bc_row is None else up(fr.bc_row, np.int32)
    nbytes = ctypes.c_size_t(0)
    _lib.check(lib.afr_face_slots_workspace(C, ctypes.byref(nbytes)), "afr_face_slots_workspace")
    work = torch.empty((max(int(nbytes.value), 1),), dtype=torch.uint8, device=dv)
    i32 = lambda *s: torch.empty(s, dtype=torch.int32, device=dv)      # noqa: E731
    f64 = lambda *s: torch.empty(s, dtype=torch.float64, device=dv)    # noqa: E731
    out = {"nb": i32(C, NSLOT, 2), "blk": i32(C, NSLOT, 2), "slot_bc": i32(C, NSLOT),
           "tn": f64(C, NSLOT, 3), "trans": f64(C, NSLOT, 3), "other": i32(C, NSLOT, 4), "rec": i32(n),
           "slot_flags": i32(C, NSLOT), "rep": i32(n), "run_start": i32(n), "sorted_slot": i32(n),
           "counts": i32(2), "status": i32(1)}
    out["slot_bc_row"] = i32(C, NSLOT) if bc_row_in is not None else None
    p = lambda t: ctypes.c_void_p(None if t is None else t.data_ptr())  # noqa: E731
    d = _lib.AfrSlotDesc()
    d.abi_version, d.n_prism, d.n_tet, d.n_records = _lib.ABI_VERSION, P, mesh.n_tet, F
    d.nodes, d.prism_conn, d.tet_conn = p(nodes), p(pc), p(tc)
    for k, t in rec_in.items():
        setattr(d, k, p(t))
    d.bc_row = p(bc_row_in)
    d.work, d.work_bytes = p(work), int(work.numel())
    for k, t in out.items():
        setattr(d, k, p(t))
    with torch.cuda.device(dv):
        stream = ctypes.c_void_p(torch.cuda.current_stream(dv).cuda_stream)
        _lib.check(lib.afr_face_slots(ctypes.byref(d), stream), "afr_face_slots")
        torch.cuda.current_stream(dv).synchronize()
    status = int(out["status"].item())
    if status & 1:
        raise RuntimeError("every real cube face must be the primary side of exactly one record")
    if status & 2:
        raise RuntimeError("face corners of the two sides do not coincide")
    n_blocks, n_keyed = (int(x) for x in out["counts"].cpu().numpy())
    nb = out["nb"].cpu().numpy()
    blk = out["blk"].cpu().numpy()
    bc = out["slot_bc"].cpu().numpy()
    tn = out["tn"].cpu().numpy()
    bc_row = None if out["slot_bc_row"] is None else out["slot_bc_row"].cpu().numpy()
    mask = np.zeros((C, NSLOT), dtype=np.int32)
    own_code = np.full((C, NSLOT), -1, dtype=np.int64)
    own_code[:P] = np.array(SLOT_CODES[0])
    own_code[P:, :4] = np.array(SLOT_CODES[1])
    with HOST_LOCK:
        rows, n_blocks, n_split = _host_rows(mesh, T, out, n, n_blocks, n_keyed, own_code, blk, mask, check_samples,
                                             dv)
    return Slots(nb=nb, blk=blk, mask=mask, bc=bc, tn=tn, adjrow=None, fallback=None, rows=rows,
                 n_blocks=n_blocks, keyed_slots=n_keyed, split_slots=n_split, bc_row=bc_row)


def _host_rows(mesh, T, out, n, n_blocks, n_keyed, own_code, blk, mask, check_samples, dv):
    """Interpolation rows of the block representatives and of the split slots (host, LAPACK)."""
    fr, nfp = mesh.faces, T.nfp
    flags = out["slot_flags"]
    rows_list = []
    if n_keyed:
        rep = out["rep"][:n_blocks].long()
        other = out["other"].view(n, 4)
        trans = out["trans"].view(n, 3)
        rep_h = rep.cpu().numpy()
        oth_h = other[rep].cpu().numpy()
        rc, rs = rep_h // NSLOT, rep_h % NSLOT
        tg = _own_fp_phys(mesh, T, rc, own_code[rc, rs]) + trans[rep].cpu().numpy()[:, None, :]
        rows = _rows_for(mesh, T, oth_h[:, 0].astype(np.int64), oth_h[:, 1].astype(np.int64), tg)
        rows_list.append(rows.reshape(-1, T.ns))
        # the key determines the located points: check the first members of every block
        starts = out["run_start"][:n_blocks].cpu().numpy().astype(np.int64)
        ends = np.append(starts[1:], n_keyed)
        pos = np.concatenate([np.arange(a, min(b, a + check_samples)) for a, b in zip(starts, ends)])
        which = np.repeat(np.arange(n_blocks), np.minimum(ends - starts, check_samples))
        chk = out["sorted_slot"][torch.as_tensor(pos, device=dv)].long()
        ch = chk.cpu().numpy()
        cc, cs = ch // NSLOT, ch % NSLOT
        o2 = other[chk].cpu().numpy()
        tg = _own_fp_phys(mesh, T, cc, own_code[cc, cs]) + trans[chk].cpu().numpy()[:, None, :]
        rows_chk = _rows_for(mesh, T, o2[:, 0].astype(np.int64), o2[:, 1].astype(np.int64), tg)
        ref_rows = rows.reshape(n_blocks, nfp, T.ns)[which]
        if np.max(np.abs(rows_chk - ref_rows)) > 1e-7 * max(1.0, np.abs(ref_rows).max()):
            raise RuntimeError("cross-interpolation rows are not determined by the face correspondence")
    # split prism quad faces (two different source cells): rows per slot
    split = torch.nonzero((flags.view(-1) & 2) != 0).view(-1)
    n_split = int(split.numel())
    if n_split:
        sp = split.cpu().numpy()
        oth = out["other"].view(n, 4)[split].cpu().numpy().astype(np.int64)
        recs = out["rec"][split].cpu().numpy().astype(np.int64)
        trs = out["trans"].view(n, 3)[split].cpu().numpy()
        nb_blocks = n_blocks
        for j, flat in enumerate(sp):
            c, s = int(flat // NSLOT), int(flat % NSLOT)
            m0 = _quad_lower_mask(mesh, c, own_code[c, s], fr.node_ids[recs[j]], T.n1)
            tg = _own_fp_phys(mesh, T, np.array([c]), np.array([own_code[c, s]])) + trs[j]
            r0 = _rows_for(mesh, T, np.array([oth[j, 0]]), np.array([oth[j, 1]]), tg)
            r1 = _rows_for(mesh, T, np.array([oth[j, 2]]), np.array([oth[j, 3]]), tg)
            rows_list.append(r0[0])
            rows_list.append(r1[0])
            blk[c, s] = (nb_blocks, nb_blocks + 1)
            nb_blocks += 2
            mask[c, s] = int(np.sum((~m0).astype(np.int64) << np.arange(nfp)))
        n_blocks = nb_blocks
    rows = np.concatenate(rows_list, axis=0) if rows_list else np.zeros((nfp, T.ns))
    return rows, n_blocks, n_split
