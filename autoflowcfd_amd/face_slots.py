"""Cell-centric face "slots": the device form of the reference's flat face geometry.

The reference walks face records and scatter-adds each primary side's FR
correction into its cell (src/autoflowcfd/core/fr_residual/inviscid_kernel.py:87-284,
colored sweeps at src/autoflowcfd/core/fr_residual/inviscid.py:299-320).  Every
real cube face of every cell is the primary side of exactly one record
(src/autoflowcfd/fr/face_flux_points_merge.py:83-124), so the same sum can be
computed by the cell itself, one slot per real cube face, with no scatter,
coloring or atomics: slot = (cell, cube face) -> (other cell(s), cross
interpolation rows, true normal with the role's sign, boundary code).

Cross interpolation rows (face_flux_points.py:164-236: the other cell's
collapsed basis evaluated at the located flux points, times V_sps^{-1}) depend
only on the two cell types, the two cube faces and the corner correspondence
of the shared face: exactly for triangles (affine) and for quad/quad
(the same bilinear patch).  Slots are keyed on that and one representative per
key is located physically; every row block is stored once (rows table) and
slots index it.  A prism quad face split between two different cells
(face_flux_points_data.py:145-217) gets its own per-slot blocks.
"""

from dataclasses import dataclass

from typing import Optional

import numpy as np

from .fr_tables import face_pts, gl
from .geometry import det_adj, jacobian, map_ref
from .mesh import CUBE_FACE_AXIS, CUBE_FACE_SIDE, PRISM_CUBE_FACES, PRISM_QUAD_CODES, TET_CUBE_FACES

SLOT_CODES = ((0, 1, 2, 4, 5), (0, 1, 2, 4))   # real cube faces: prism, tet
NSLOT = 5
BC_NONE = -2            # unused slot (tets have 4 real faces)
_CODE_TO_SLOT = (np.array([0, 1, 2, -1, 3, 4]), np.array([0, 1, 2, -1, 3, -1]))
_TET_FACE = {0: (0, 2, 3), 1: (1, 2, 3), 2: (0, 1, 3), 4: (0, 1, 2)}


@dataclass
class Slots:
    nb: np.ndarray        # (C,5,2) int32 source cells (src1 = -1 unless split quad)
    blk: np.ndarray       # (C,5,2) int32 row-block ids
    mask: np.ndarray      # (C,5) int32 bit fp -> flux point fp takes source 1
    bc: np.ndarray        # (C,5) int32 BC code, -1 interior, -2 unused
    tn: np.ndarray        # (C,5,3) float64 true normal, role-signed
    adjrow: np.ndarray    # (C,5,nfp,3) float64 E @ adj[:, axis, :]
    fallback: np.ndarray  # (C,5) int32 bit fp -> direction falls back to tn
    rows: np.ndarray      # (n_blk*nfp, ns) float64
    n_blocks: int
    keyed_slots: int
    split_slots: int
    bc_row: Optional[np.ndarray] = None  # (C,5) int32 boundary-state table row (None: the BC code)

    def rows_of_bc(self) -> np.ndarray:
        """Boundary-state table row of every slot (its BC code unless set; < 0 off the boundary)."""
        return self.bc if self.bc_row is None else self.bc_row


def _cells_nodes(mesh, cells):
    if len(cells) == 0:
        return np.zeros((0, 6, 3))
    if cells[0] < mesh.n_prism:
        return mesh.nodes[mesh.prism_conn[cells]]
    return mesh.nodes[mesh.tet_conn[cells - mesh.n_prism]]


def locate(mesh, cells, codes, targets):
    """Free coordinates (K,n,2) on face ``codes`` of homogeneous ``cells`` hitting ``targets``.

    tet: closed-form barycentric (face_flux_points_locate.py:53-111);
    prism: Newton least squares on the bilinear face (face_flux_points_locate.py:114-246).
    """
    K, n, _ = targets.shape
    free = np.zeros((K, n, 2))
    if K == 0:
        return free
    nodes = _cells_nodes(mesh, cells)
    tet = cells[0] >= mesh.n_prism
    for code in np.unique(codes):
        sel = np.flatnonzero(codes == code)
        ax, sd = int(CUBE_FACE_AXIS[code]), float(CUBE_FACE_SIDE[code])
        other = [d for d in range(3) if d != ax]
        nd, tg = nodes[sel], targets[sel]
        if tet:
            vi, vj, vk = _TET_FACE[int(code)]
            Pi = nd[:, vi]
            e1, e2 = nd[:, vj] - Pi, nd[:, vk] - Pi
            G = np.stack([np.einsum("ki,ki->k", e1, e1), np.einsum("ki,ki->k", e1, e2),
                          np.einsum("ki,ki->k", e2, e2)], 1)
            rr = tg - Pi[:, None]
            b1 = np.einsum("kni,ki->kn", rr, e1)
            b2 = np.einsum("kni,ki->kn", rr, e2)
            det = (G[:, 0] * G[:, 2] - G[:, 1] ** 2)[:, None]
            al = (b1 * G[:, 2:3] - b2 * G[:, 1:2]) / det
            be = (G[:, 0:1] * b2 - G[:, 1:2] * b1) / det
            L = np.zeros(al.shape + (4,))
            L[..., vi] = 1.0 - al - be
            L[..., vj] = al
            L[..., vk] = be
            r, s, t = 2.0 * L[..., 1] - 1.0, 2.0 * L[..., 2] - 1.0, 2.0 * L[..., 3] - 1.0
            c = t
            b = 2.0 * (s + 1.0) / np.where(np.abs(1.0 - c) > 1e-300, 1.0 - c, 1e-300) - 1.0
            a = -2.0 * (r + 1.0) / np.where(np.abs(s + t) > 1e-300, s + t, 1e-300) - 1.0
            free[sel] = np.stack([a, b, c], -1)[..., other]
            continue
        x = np.zeros((len(sel), n, 2))
        polish = False
        for _ in range(60):
            ref = np.empty((len(sel), n, 3))
            ref[..., ax] = sd
            ref[..., other[0]] = x[..., 0]
            ref[..., other[1]] = x[..., 1]
            res = map_ref(True, ref, nd) - tg
            Jf = jacobian(True, ref, nd)[..., other]
            JtJ = np.einsum("knia,knib->knab", Jf, Jf)
            Jtr = np.einsum("knia,kni->kna", Jf, res)
            dx = -np.linalg.solve(JtJ, Jtr[..., None])[..., 0]
            x += dx
            # quadratic convergence: once the step is below 1e-12 one more step reaches the
            # round-off floor (a 1e-15 test on the batch maximum can wait on round-off noise)
            if polish:
                break
            polish = np.max(np.abs(dx)) < 1e-12
        free[sel] = x
    return free


def _face_corners(mesh, cells, codes):
    """Global corner node ids of cube face ``codes`` of ``cells`` (K, 3|4), padded with -1."""
    out = np.full((len(cells), 4), -1, dtype=np.int64)
    for code in np.unique(codes):
        sel = np.flatnonzero(codes == code)
        cs = cells[sel]
        pr = cs < mesh.n_prism
        if np.any(pr):
            loc = PRISM_CUBE_FACES[int(code)]
            out[sel[pr], :len(loc)] = mesh.prism_conn[cs[pr]][:, list(loc)]
        if np.any(~pr):
            loc = TET_CUBE_FACES[int(code)]
            out[sel[~pr], :3] = mesh.tet_conn[cs[~pr] - mesh.n_prism][:, list(loc)]
    return out


def _quad_lower_mask(mesh, cell, code, tri_nodes, n1):
    """FP mask of the diagonal half the record ``tri_nodes`` covers (face_flux_points_data.py:145-217)."""
    cn = mesh.prism_conn[cell]
    quad = PRISM_CUBE_FACES[int(code)]
    order = sorted(range(3), key=lambda k: int(cn[k]))
    vs = [order[k] for k in range(3)]
    ws = [order[k] + 3 for k in range(3)]
    edge = {quad[0], quad[1]}
    if edge == {vs[0], vs[1]}:
        diag = {vs[0], ws[1]}
    elif edge == {vs[1], vs[2]}:
        diag = {vs[1], ws[2]}
    else:
        diag = {vs[0], ws[2]}
    i0, i1, i2, i3 = quad
    if diag == {i0, i2}:
        std, lower = True, (i0, i1, i2)
    else:
        std, lower = False, (i0, i1, i3)
    x, _ = gl(n1)
    g1, g2 = np.meshgrid(x, x, indexing="ij")
    low = (g1.ravel() >= g2.ravel()) if std else ((g1.ravel() + g2.ravel()) <= 0.0)
    is_lower = set(int(v) for v in tri_nodes) == set(int(cn[i]) for i in lower)
    return low if is_lower else ~low


def _own_fp_phys(mesh, T, cells, codes):
    """Physical positions (K, nfp, 3) of the flux points of cube face ``codes`` of ``cells``."""
    out = np.zeros((len(cells), T.nfp, 3))
    for code in np.unique(codes):
        sel = np.flatnonzero(codes == code)
        ref = face_pts(T.n1, int(CUBE_FACE_AXIS[code]), float(CUBE_FACE_SIDE[code]))
        for is_prism in (True, False):
            ss = sel[(cells[sel] < mesh.n_prism) == is_prism]
            if len(ss):
                out[ss] = map_ref(is_prism, ref, _cells_nodes(mesh, cells[ss]))
    return out


def _rows_for(mesh, T, oth_cells, oth_codes, targets):
    """Interpolation rows (K, nfp, ns) of the other cells at the located targets
    (face_flux_points.py:164-236: the other cell's basis at the located points, times V^-1)."""
    nfp = T.nfp
    out = np.zeros((len(oth_cells), nfp, T.ns))
    for is_prism in (True, False):
        ss = np.flatnonzero((oth_cells < mesh.n_prism) == is_prism)
        if len(ss) == 0:
            continue
        free = locate(mesh, oth_cells[ss], oth_codes[ss], targets[ss])
        for code in np.unique(oth_codes[ss]):
            s2 = ss[oth_codes[ss] == code]
            ax = int(CUBE_FACE_AXIS[code])
            other = [d for d in range(3) if d != ax]
            abc = np.empty((len(s2), nfp, 3))
            abc[..., ax] = CUBE_FACE_SIDE[code]
            f2 = free[np.searchsorted(ss, s2)]
            abc[..., other[0]] = f2[..., 0]
            abc[..., other[1]] = f2[..., 1]
            out[s2] = T.interp[0 if is_prism else 1].at(abc.reshape(-1, 3)).reshape(len(s2), nfp, T.ns)
    return out


def build_slots(mesh, T, check_samples=4, rng_seed=0, with_geometry=True) -> Slots:
    """with_geometry=False leaves adjrow/fallback to the device metric precompute."""
    fr = mesh.faces
    F, C, n1, nfp = fr.n_faces, mesh.n_cells, T.n1, T.nfp
    bnd = fr.is_boundary
    prism_owner_quad = (fr.owner < mesh.n_prism) & np.isin(fr.owner_code, PRISM_QUAD_CODES)
    prism_neigh_quad = (~bnd) & (fr.neighbor < mesh.n_prism) & np.isin(fr.neighbor_code, PRISM_QUAD_CODES)

    def primary(cells, codes, grouped):
        prim = np.ones(F, dtype=bool)
        idx = np.flatnonzero(grouped)
        if len(idx):
            o = np.lexsort((idx, codes[idx], cells[idx]))
            s = idx[o]
            same = np.zeros(len(s), dtype=bool)
            same[1:] = (cells[s[1:]] == cells[s[:-1]]) & (codes[s[1:]] == codes[s[:-1]])
            prim[s[same]] = False
        return prim

    op = primary(fr.owner, fr.owner_code, prism_owner_quad)          # interior and boundary groups
    npri = primary(fr.neighbor, fr.neighbor_code, prism_neigh_quad)

    # second record of each prism-quad group (for split sources)
    def partner(cells, codes, grouped):
        part = np.full(F, -1, dtype=np.int64)
        idx = np.flatnonzero(grouped)
        if len(idx):
            o = np.lexsort((idx, codes[idx], cells[idx]))
            s = idx[o]
            same = np.zeros(len(s), dtype=bool)
            same[1:] = (cells[s[1:]] == cells[s[:-1]]) & (codes[s[1:]] == codes[s[:-1]])
            part[s[:-1][same[1:]]] = s[1:][same[1:]]
        return part

    op_part = partner(fr.owner, fr.owner_code, prism_owner_quad & ~bnd)
    np_part = partner(fr.neighbor, fr.neighbor_code, prism_neigh_quad)

    # slot assignment: (cell, slot) <- (record, role)
    rec = np.full((C, NSLOT), -1, dtype=np.int64)
    role = np.zeros((C, NSLOT), dtype=np.int64)   # 0 owner, 1 neighbour
    o_f = np.flatnonzero(op)
    ct_o = (fr.owner[o_f] >= mesh.n_prism).astype(np.int64)
    sl_o = np.where(ct_o == 0, _CODE_TO_SLOT[0][fr.owner_code[o_f]], _CODE_TO_SLOT[1][fr.owner_code[o_f]])
    n_f = np.flatnonzero(~bnd & npri)
    ct_n = (fr.neighbor[n_f] >= mesh.n_prism).astype(np.int64)
    sl_n = np.where(ct_n == 0, _CODE_TO_SLOT[0][fr.neighbor_code[n_f]], _CODE_TO_SLOT[1][fr.neighbor_code[n_f]])
    if np.any(sl_o < 0) or np.any(sl_n < 0):
        raise RuntimeError("record on a degenerate cube face")
    cnt = np.zeros((C, NSLOT), dtype=np.int64)
    np.add.at(cnt, (fr.owner[o_f], sl_o), 1)
    np.add.at(cnt, (fr.neighbor[n_f], sl_n), 1)
    rec[fr.owner[o_f], sl_o] = o_f
    rec[fr.neighbor[n_f], sl_n] = n_f
    role[fr.neighbor[n_f], sl_n] = 1
    expect = np.zeros((C, NSLOT), dtype=np.int64)
    expect[:mesh.n_prism, :] = 1
    expect[mesh.n_prism:, :4] = 1
    if np.any(cnt != expect):
        raise RuntimeError("every real cube face must be the primary side of exactly one record")

    cell_ids = np.repeat(np.arange(C), NSLOT).reshape(C, NSLOT)
    ctype = (cell_ids >= mesh.n_prism).astype(np.int64)
    slot_idx = np.tile(np.arange(NSLOT), (C, 1))
    own_code = np.full((C, NSLOT), -1, dtype=np.int64)
    own_code[:mesh.n_prism] = np.array(SLOT_CODES[0])
    own_code[mesh.n_prism:, :4] = np.array(SLOT_CODES[1])
    valid = rec >= 0
    r = np.where(valid, rec, 0)
    is_own = role == 0
    bc = np.where(valid, np.where(is_own, fr.bc[r], -1), BC_NONE).astype(np.int32)
    bc[valid & bnd[r]] = fr.bc[r][valid & bnd[r]]
    bc_row = None
    if fr.bc_row is not None:
        bc_row = np.where(valid & bnd[r], fr.bc_row[r], -1).astype(np.int32)
    sgn = np.where(is_own, 1.0, -1.0)
    tn = fr.normal[r] * sgn[..., None]
    trans = fr.translation[r] * sgn[..., None]      # search point = own FP + trans
    other0 = np.where(is_own, fr.neighbor[r], fr.owner[r])
    ocode0 = np.where(is_own, fr.neighbor_code[r], fr.owner_code[r])
    part = np.where(is_own, op_part[r], np_part[r])
    other1 = np.where(part >= 0, np.where(is_own, fr.neighbor[np.maximum(part, 0)], fr.owner[np.maximum(part, 0)]), -1)
    ocode1 = np.where(part >= 0, np.where(is_own, fr.neighbor_code[np.maximum(part, 0)],
                                          fr.owner_code[np.maximum(part, 0)]), -1)
    interior = valid & ~bnd[r]
    split = interior & (other1 >= 0) & (other1 != other0)

    nb = np.full((C, NSLOT, 2), -1, dtype=np.int32)
    nb[..., 0] = np.where(interior, other0, -1)
    nb[..., 1] = np.where(split, other1, -1)
    blk = np.zeros((C, NSLOT, 2), dtype=np.int32)
    mask = np.zeros((C, NSLOT), dtype=np.int32)

    def own_fp_phys(cells, codes):
        return _own_fp_phys(mesh, T, cells, codes)

    rows_list = []
    n_blocks = 0

    def rows_for(oth_cells, oth_codes, targets):
        return _rows_for(mesh, T, oth_cells, oth_codes, targets), None

    # ---- keyed (single-source) slots
    keyed = interior & ~split
    kc, ks = np.nonzero(keyed)
    keyed_slots = len(kc)
    if keyed_slots:
        own_corners = _face_corners(mesh, kc, own_code[kc, ks])
        oth_corners = _face_corners(mesh, other0[kc, ks], ocode0[kc, ks])
        ncorner = np.where(own_corners[:, 3] >= 0, 4, 3)
        # corner correspondence: shared node ids, or positions for periodic images
        hit = own_corners[:, :, None] == oth_corners[:, None, :]
        hit &= (own_corners[:, :, None] >= 0)
        pi = np.argmax(hit, axis=2)
        ok = np.all(hit.any(axis=2) | (own_corners < 0), axis=1)
        per = np.flatnonzero(~ok)
        if len(per):
            Xo = mesh.nodes[np.maximum(own_corners[per], 0)] + trans[kc[per], ks[per]][:, None, :]
            Xt = mesh.nodes[np.maximum(oth_corners[per], 0)]
            d2 = np.zeros((len(per), 4, 4))
            for i in range(4):
                for j in range(4):
                    dd = Xo[:, i] - Xt[:, j]
                    d2[:, i, j] = np.einsum("ki,ki->k", dd, dd)
            d2[:, :, 3][oth_corners[per, 3] < 0] = np.inf
            pi[per] = np.argmin(d2, axis=2)
            scale = np.einsum("ki,ki->k", Xo[:, 1] - Xo[:, 0], Xo[:, 1] - Xo[:, 0])
            if np.any(np.min(d2, axis=2)[:, :3].max(axis=1) > 1e-12 * scale + 1e-300):
                raise RuntimeError("face corners of the two sides do not coincide")
        pi[ncorner == 3, 3] = 0
        key = (((ctype[kc, ks] * 6 + own_code[kc, ks]) * 2 + (other0[kc, ks] >= mesh.n_prism)) * 6
               + ocode0[kc, ks]) * 256 + pi[:, 0] + 4 * pi[:, 1] + 16 * pi[:, 2] + 64 * pi[:, 3]
        ukeys, first, inv = np.unique(key, return_index=True, return_inverse=True)
        rep_c, rep_s = kc[first], ks[first]
        tg = own_fp_phys(rep_c, own_code[rep_c, rep_s]) + trans[rep_c, rep_s][:, None, :]
        rows, _ = rows_for(other0[rep_c, rep_s], ocode0[rep_c, rep_s], tg)
        rows_list.append(rows.reshape(-1, T.ns))
        blk[kc, ks, 0] = inv
        n_blocks = len(ukeys)
        # verify the key determines the located points on a few slots per key
        rng = np.random.default_rng(rng_seed)
        chk = []
        for u in range(len(ukeys)):
            members = np.flatnonzero(inv == u)
            chk.extend(rng.choice(members, size=min(check_samples, len(members)), replace=False).tolist())
        chk = np.array(chk)
        cc, cs = kc[chk], ks[chk]
        tg = own_fp_phys(cc, own_code[cc, cs]) + trans[cc, cs][:, None, :]
        rows_chk, _ = rows_for(other0[cc, cs], ocode0[cc, cs], tg)
        ref_rows = rows.reshape(len(ukeys), nfp, T.ns)[inv[chk]]
        if np.max(np.abs(rows_chk - ref_rows)) > 1e-7 * max(1.0, np.abs(ref_rows).max()):
            raise RuntimeError("cross-interpolation rows are not determined by the face correspondence")

    # ---- split prism quad faces (two different source cells)
    sc, ss_ = np.nonzero(split)
    for c, s in zip(sc, ss_):
        f0 = rec[c, s]
        m0 = _quad_lower_mask(mesh, c, own_code[c, s], fr.node_ids[f0], n1)
        tg = own_fp_phys(np.array([c]), np.array([own_code[c, s]])) + trans[c, s]
        r0, _ = rows_for(np.array([other0[c, s]]), np.array([ocode0[c, s]]), tg)
        r1, _ = rows_for(np.array([other1[c, s]]), np.array([ocode1[c, s]]), tg)
        rows_list.append(r0[0])
        rows_list.append(r1[0])
        blk[c, s] = (n_blocks, n_blocks + 1)
        n_blocks += 2
        mask[c, s] = int(np.sum((~m0).astype(np.int64) << np.arange(nfp)))

    rows = np.concatenate(rows_list, axis=0) if rows_list else np.zeros((nfp, T.ns))

    # ---- extrapolated adjugate rows and direction fallback
    adjrow = np.zeros((C, NSLOT, nfp, 3)) if with_geometry else None
    fallback = np.zeros((C, NSLOT), dtype=np.int32) if with_geometry else None
    chunk = 65536
    for is_prism, conn, off in ((True, mesh.prism_conn, 0), (False, mesh.tet_conn, mesh.n_prism)):
        ct = 0 if is_prism else 1
        for s0 in range(0, len(conn) if with_geometry else 0, chunk):
            sl = slice(s0, min(s0 + chunk, len(conn)))
            _, adj = det_adj(jacobian(is_prism, T.sps, mesh.nodes[conn[sl]]))
            g = np.arange(off + sl.start, off + sl.stop)
            for si, code in enumerate(SLOT_CODES[ct]):
                ax, sd = int(CUBE_FACE_AXIS[code]), float(CUBE_FACE_SIDE[code])
                E = T.E[ct, ax, int(sd > 0)]
                ar = np.einsum("ps,ksi->kpi", E, adj[:, :, ax, :])
                adjrow[g, si] = ar
                mag = np.maximum(np.linalg.norm(ar, axis=-1), 1e-300)
                d = ar / mag[..., None] * sd
                al = np.einsum("kpi,ki->kp", d, tn[g, si])
                fb = (al < 0.5)
                fallback[g, si] = np.sum(fb.astype(np.int64) << np.arange(nfp), axis=1).astype(np.int32)
    return Slots(nb=nb, blk=blk, mask=mask, bc=bc, tn=tn, adjrow=adjrow, fallback=fallback, rows=rows,
                 n_blocks=n_blocks, keyed_slots=keyed_slots, split_slots=len(sc), bc_row=bc_row)


def concat_slots(slot_list, n_prism_list, n_tet_list, nfp):
    """Slot tables of a disjoint union of meshes (mesh.concat_meshes numbering: all prisms case
    by case, then all tets case by case) from each case's own tables.

    Built per case, every row, normal and flag of a cell is the one its case alone has, bit
    for bit (a union built in one pass would locate flux points in batches of another
    composition, which moves interpolation weights by round-off)."""
    P = np.asarray(n_prism_list, dtype=np.int64)
    Tn = np.asarray(n_tet_list, dtype=np.int64)
    p_off = np.concatenate([[0], np.cumsum(P)])
    t_off = np.concatenate([[0], np.cumsum(Tn)])
    P_tot = int(p_off[-1])
    C = P_tot + int(t_off[-1])
    s0 = slot_list[0]

    def alloc(a, fill):
        return None if a is None else np.full((C,) + a.shape[1:], fill, dtype=a.dtype)

    out = {k: alloc(getattr(s0, k), f) for k, f in (("nb", -1), ("blk", 0), ("mask", 0), ("bc", -2),
                                                     ("tn", 0.0), ("fallback", 0), ("adjrow", 0.0))}
    any_rows = any(sl.bc_row is not None for sl in slot_list)
    out["bc_row"] = np.full(out["bc"].shape, -1, dtype=np.int32) if any_rows else None
    rows, blk_off = [], 0
    for i, sl in enumerate(slot_list):
        Pi = int(P[i])
        g = np.concatenate([p_off[i] + np.arange(Pi), P_tot + t_off[i] + np.arange(int(Tn[i]))])
        if len(g) != sl.nb.shape[0]:
            raise ValueError(f"case {i}: slot tables do not match its cell counts")
        if sl.rows.shape[0] % nfp:
            raise ValueError(f"case {i}: rows are not whole blocks of {nfp}")
        nb = sl.nb.astype(np.int64)
        nb_g = np.where(nb < Pi, p_off[i] + nb, P_tot + t_off[i] + (nb - Pi))
        out["nb"][g] = np.where(nb < 0, -1, nb_g).astype(np.int32)
        out["blk"][g] = sl.blk + blk_off
        for k in ("mask", "bc", "tn", "fallback", "adjrow"):
            if out[k] is not None:
                out[k][g] = getattr(sl, k)
        if any_rows:
            out["bc_row"][g] = np.where(sl.bc >= 0, sl.rows_of_bc(), -1)
        rows.append(sl.rows)
        blk_off += sl.rows.shape[0] // nfp
    return Slots(nb=out["nb"], blk=out["blk"], mask=out["mask"], bc=out["bc"], tn=out["tn"], adjrow=out["adjrow"],
                 fallback=out["fallback"], rows=np.concatenate(rows), n_blocks=blk_off,
                 keyed_slots=sum(s.keyed_slots for s in slot_list), split_slots=sum(s.split_slots for s in slot_list),
                 bc_row=out["bc_row"])


def slots_for(mesh, order, storage="locality"):
    """Slot tables of a mesh at an order without geometry (the device metric precompute fills
    the rest): the host set-up of one solve, runnable in a worker process.  storage="locality"
    builds them for the mesh renumbered as GPUFRSolver stores it (mesh.locality_renumbered)."""
    from .fr_tables import tables
    from .mesh import locality_renumbered
    if storage == "locality":
        mesh, _ = locality_renumbered(mesh)
    return build_slots(mesh, tables(order), with_geometry=False)
