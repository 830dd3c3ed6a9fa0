"""Device-resident mesh/operator data and the residual entry points.

PyTorch-ROCm tensors own every device buffer; the C ABI (libafr.so) receives
raw device pointers.  Host-side set-up (operator tables, metrics, face slots)
is FP64 and rounded once to the device precision.
"""

import ctypes

import numpy as np
import torch

from . import _lib
from .face_slots import NSLOT, SLOT_CODES, build_slots
from .fr_tables import prism_factors, tables
from .geometry import cell_block_metrics
from .mesh import BC_FARFIELD, BC_INLET, BC_OUTLET, locality_renumbered, morton_order  # noqa: F401

GAMMA = 1.4
ROW_DROP = 1e-11   # cross-interpolation weights below this are round-off of structural zeros


def sparse_rows(rows, drop=ROW_DROP):
    """Dense (n_rows, ns) interpolation rows -> (nnz (n_rows,), idx (n_rows, k), w (n_rows, k))."""
    keep = np.abs(rows) > drop
    lost = np.abs(np.where(keep, 0.0, rows)).sum(axis=1)
    if lost.size and lost.max() > 1e-9:
        raise ValueError(f"sparse rows drop weight {lost.max():.2e}")
    nnz = keep.sum(axis=1).astype(np.int32)
    k = max(int(nnz.max()) if nnz.size else 1, 1)
    order = np.argsort(~keep, axis=1, kind="stable")[:, :k]
    idx = np.where(np.arange(k)[None, :] < nnz[:, None], order, 0).astype(np.int32)
    w = np.where(np.arange(k)[None, :] < nnz[:, None], np.take_along_axis(rows, order, axis=1), 0.0)
    return nnz, idx, w


def slot_mask_with_wide_flag(mask, blk0, rows_nnz, n_blocks, nfp, order):
    """Slot masks (bit fp: flux point fp takes source 1) with bit 30 set on the slots whose
    first-source row block (blk0, -1 for none) has a row with more than order + 1 non-zeros
    (rows_nnz: non-zeros of every row, blocks of nfp rows)."""
    if nfp > 30:
        raise ValueError("slot masks hold at most 30 flux points")
    mask = np.asarray(mask, dtype=np.int64)
    blk0 = np.asarray(blk0, dtype=np.int64)
    if n_blocks <= 0:
        return mask.copy()
    blk_wide = np.asarray(rows_nnz)[:n_blocks * nfp].reshape(n_blocks, nfp).max(axis=1) > order + 1
    wide = (blk0 >= 0) & blk_wide[np.clip(blk0, 0, n_blocks - 1)]
    return mask | (wide.astype(np.int64) << 30)


def bc_params_table(Q_free, Q_inlet=None, p_outlet=None):
    """(8,5) per-BC-code reference states (src/autoflowcfd/core/fr_solver/boundary.py:52-145)."""
    P = np.zeros((8, 5))
    Q_free = np.asarray(Q_free, dtype=np.float64)
    P[BC_FARFIELD] = Q_free
    P[BC_INLET] = Q_free if Q_inlet is None else np.asarray(Q_inlet, dtype=np.float64)
    P[BC_OUTLET, 4] = Q_free[4] if p_outlet is None else p_outlet
    return P


def sparse_fixed(rows, widths, drop=ROW_DROP):
    """Dense (..., n) rows -> (w, idx) of shape (..., W) with the non-zeros of each row first and zero
    padding; widths broadcast against the leading shape and bound the non-zeros per row (checked)."""
    rows = np.asarray(rows, dtype=np.float64)
    lead, n = rows.shape[:-1], rows.shape[-1]
    W = int(np.max(widths))
    if W >= n and np.all(np.broadcast_to(widths, lead) >= n):
        # full width: the rows stay dense (nothing dropped)
        idx = np.broadcast_to(np.arange(n, dtype=np.int32), rows.shape).copy()
        return rows.copy(), idx
    keep = np.abs(rows) > drop * max(np.abs(rows).max(), 1.0)
    lost = np.abs(np.where(keep, 0.0, rows)).max() if rows.size else 0.0
    if lost > 1e-9:
        raise ValueError(f"sparse rows drop weight {lost:.2e}")
    cnt = keep.sum(axis=-1)
    if np.any(cnt > np.broadcast_to(widths, lead)):
        raise ValueError("sparse row wider than its kernel width")
    order = np.argsort(~keep, axis=-1, kind="stable")[..., :W]
    valid = np.arange(W) < cnt[..., None]
    idx = np.where(valid, order, 0).astype(np.int32)
    w = np.where(valid, np.take_along_axis(rows, order, axis=-1), 0.0)
    return w, idx


def suppress_residual_outliers(R, U):
    """Zero the outliers of the residual field R (C, ns, nv) in place on the device, with U (C, ns,
    >= nv) the reference field, as suppress_residual_outliers
    (src/autoflowcfd/core/fr_operators/troubled_cell.py:382-408); returns R."""
    if R.device.type != "cuda" or U.device != R.device or R.dtype != U.dtype:
        raise ValueError("R and U must be device tensors of one precision")
    if R.dim() != 3 or U.dim() != 3 or tuple(U.shape[:2]) != tuple(R.shape[:2]) or U.shape[2] < R.shape[2]:
        raise ValueError(f"shapes R {tuple(R.shape)} and U {tuple(U.shape)} do not match")
    if not (R.is_contiguous() and U.is_contiguous()) or R.dtype not in (torch.float32, torch.float64):
        raise ValueError("R and U must be contiguous FP32 or FP64 tensors")
    lib = _lib.load()
    st = ctypes.c_void_p(torch.cuda.current_stream(R.device).cuda_stream)
    _lib.check(lib.afr_suppress_residual_outliers(ctypes.c_void_p(R.data_ptr()), ctypes.c_void_p(U.data_ptr()),
                                                  int(R.shape[0]), int(R.shape[1]), int(R.shape[2]),
                                                  int(U.shape[2]), 0 if R.dtype == torch.float32 else 1, st),
               "afr_suppress_residual_outliers")
    return R


class LibGraph:
    """A HIP graph of afr_* calls captured and replayed through the C ABI (afr_graph_begin /
    afr_graph_end / afr_graph_launch), for callers that do not use torch's graphs:

        g = LibGraph(dm.lib)
        with torch.cuda.stream(s):
            g.begin(s); dm.ssp_rk3_step(U); g.end(s)     # captured, not run
        g.launch(s)                                       # one pseudo-step per launch

    Nothing may allocate on the stream while it is captured (the DeviceMesh step calls do not)."""

    def __init__(self, lib):
        self.lib, self.handle = lib, None

    def begin(self, stream):
        _lib.check(self.lib.afr_graph_begin(ctypes.c_void_p(stream.cuda_stream)), "afr_graph_begin")

    def end(self, stream):
        h = ctypes.c_void_p()
        _lib.check(self.lib.afr_graph_end(ctypes.c_void_p(stream.cuda_stream), ctypes.byref(h)), "afr_graph_end")
        self.handle = h

    def launch(self, stream):
        if self.handle is None:
            raise RuntimeError("LibGraph.launch: nothing captured")
        _lib.check(self.lib.afr_graph_launch(self.handle, ctypes.c_void_p(stream.cuda_stream)), "afr_graph_launch")

    def close(self):
        if self.handle is not None:
            _lib.check(self.lib.afr_graph_destroy(self.handle), "afr_graph_destroy")
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def record_incidence(C, fr, width=8):
    """(C, width) int64 incidence of every cell: record * 16 + role * 8 + the cell's face code
    (role 1: the cell is the neighbour), the owner entries in record order, then the neighbour
    entries in record order, -1 padded."""
    F = fr.n_faces
    rid = np.arange(F, dtype=np.int64)
    nb = np.flatnonzero(fr.neighbor >= 0)
    cells = np.concatenate([fr.owner, fr.neighbor[nb]]).astype(np.int64)
    codes = np.concatenate([rid * 16 + fr.owner_code, nb * 16 + 8 + fr.neighbor_code[nb]])
    order = np.argsort(cells, kind="stable")
    cs = cells[order]
    count = np.bincount(cs, minlength=C)
    if count.size and count.max() > width:
        raise ValueError(f"more than {width} face records on one cell")
    start = np.concatenate([[0], np.cumsum(count)[:-1]])
    slot = np.arange(len(cs)) - start[cs]
    inc = np.full((C, width), -1, dtype=np.int64)
    inc[cs, slot] = codes[order]
    return inc


class DeviceMesh:
    """One mesh at one order on one GPU."""

    def __init__(self, mesh, order, dtype=torch.float32, device="cuda", bc_params=None, ref_state=None,
                 slots=None, metrics="gpu", turbulence=None, cell_case=None, state_ref=None):
        """metrics: "gpu" (FP64 precompute on the device) or "host" (NumPy FP64, small meshes).
        turbulence: None or "SST" (face-record tables, wall distance and SST work buffers).
        cell_case: None, or the case index of every cell of a disjoint union of cases
        (mesh.concat_meshes); bc_params is then (n_cases, 8, 5) and the wall distance is taken
        case by case (mesh.meta["wall_off"] splits mesh.wall_nodes).
        state_ref: None (every state this mesh takes holds the conserved state itself) or the
        conserved offset of each case, (5,) or (n_cases, 5) FP64: every state this mesh takes or
        writes (U, stage states) then holds U - state_ref[case of the cell], and the kernels add
        the offset back in FP64 when they load a point (``to_storage`` / ``to_absolute``
        convert).  With each case's free stream as the offset, FP32 storage holds the flow's
        variation about the free stream instead of rounding rho E ~ 2.5e5 to 0.016."""
        self.lib = _lib.load()
        self.turbulence = turbulence
        self.sps_xyz = None
        self.mesh, self.order, self.dtype, self.device = mesh, order, dtype, torch.device(device)
        T = tables(order)
        self.T = T
        C = mesh.n_cells
        self.n_cells, self.ns, self.nfp, self.nf = C, T.ns, T.nfp, T.nf
        on_gpu = metrics == "gpu"
        if slots is not None:
            sl = slots
        elif on_gpu:
            from .slots_gpu import build_slots_gpu
            sl = build_slots_gpu(mesh, T, device=self.device)
        else:
            sl = build_slots(mesh, T, with_geometry=True)
        self.slots = sl
        self._validate(sl, C, T, mesh.n_prism)

        def dev(x, dt=None):
            return torch.as_tensor(np.ascontiguousarray(x), dtype=dt or dtype).to(self.device)

        def devi(x):
            return torch.as_tensor(np.ascontiguousarray(x, dtype=np.int32)).to(self.device)

        self.slot_nb, self.slot_blk = devi(sl.nb), devi(sl.blk)
        # boundary slots: type | (row << 4), the row of the boundary-state table holding the slot's
        # reference state (the type code itself unless boundary groups set their own rows)
        self.slot_bc = devi(np.where(sl.bc >= 0, sl.bc | (np.maximum(sl.rows_of_bc(), 0) << 4), sl.bc))
        self.slot_tn = dev(sl.tn)
        self.rows = dev(sl.rows)
        # sparse rows plus one trailing all-zero row (index n_rows) for lanes without a source
        nnz, idx, w = sparse_rows(np.concatenate([sl.rows, np.zeros((1, sl.rows.shape[1]))]))
        # slot mask: bit fp = flux point fp takes source 1; bit 30 = the slot's first-source rows
        # are wider than order + 1 (the viscous interface kernel's helper lanes need no row load
        # to find them)
        self.slot_mask = devi(slot_mask_with_wide_flag(sl.mask, np.asarray(sl.blk)[..., 0], nnz, int(sl.n_blocks),
                                                       T.nfp, order))
        # fixed row width: (P+1)^2 when every row fits (prism sources), else (P+1)^3, zero-padded,
        # so the kernels read a compile-time number of entries per row without a bound check
        width = (order + 1) ** 2 if idx.shape[1] <= (order + 1) ** 2 else T.ns
        if idx.shape[1] > width:
            raise ValueError("interpolation row wider than the solution points")
        pad = width - idx.shape[1]
        idx = np.pad(idx, ((0, 0), (0, pad)))
        w = np.pad(w, ((0, 0), (0, pad)))
        self.rows_nnz, self.rows_idx, self.rows_w = devi(nnz), devi(idx), dev(w)
        self.rows_knz = int(width)
        self.prism_tab = dev(prism_factors(order).packed())
        # prisms are processed in Morton order of their centroids (L2 reuse of source cells)
        cen = mesh.nodes[mesh.prism_conn].mean(axis=1) if mesh.n_prism else np.zeros((0, 3))
        if cell_case is not None and "prism_off" in mesh.meta and mesh.n_prism:
            # a batch walks case by case (each case in its own Morton order): the cases' boxes overlap
            # in space, and one Morton order over all of them would interleave the cases' cells
            po = np.asarray(mesh.meta["prism_off"], dtype=np.int64)
            sched = np.concatenate([po[i] + morton_order(cen[po[i]:po[i + 1]]) for i in range(len(po) - 1)])
        else:
            sched = morton_order(cen)
        self.prism_schedule = devi(sched)
        self.E = []
        self.E64 = []
        for ct in (0, 1):
            E = np.zeros((NSLOT, T.nfp, T.ns))
            for si, code in enumerate(SLOT_CODES[ct]):
                E[si] = T.E[ct, code // 2, code % 2]
            self.E.append(dev(E))
            self.E64.append(E)
        if on_gpu:
            self._gpu_metrics(sl)
        else:
            geo = metrics if isinstance(metrics, dict) else cell_block_metrics(mesh, T, dtype=np.float64)
            self.det = dev(geo["det"])
            self.adj = dev(geo["adj"])
            self.adj_fine = dev(geo["adj_fine"])
            self.gcl = dev(geo["gcl"])
            self.metric_flux_scale = dev(geo["metric_flux_scale"])
            self.volume = np.asarray(geo["volume"], dtype=np.float64)
            self.det_mean = np.asarray(geo["det_mean"], dtype=np.float64)
            self.slot_fb = devi(sl.fallback)
            self.slot_adjrow = dev(sl.adjrow)
            if turbulence is not None:
                from .geometry import map_ref
                xyz = np.zeros((C, T.ns, 3))
                if mesh.n_prism:
                    xyz[:mesh.n_prism] = map_ref(True, T.sps, mesh.nodes[mesh.prism_conn])
                if mesh.n_tet:
                    xyz[mesh.n_prism:] = map_ref(False, T.sps, mesh.nodes[mesh.tet_conn])
                self.sps_xyz = torch.as_tensor(xyz).to(self.device)
        # sparse derivative / face-trace rows of the viscous kernels (widths: d/da N1, d/db N1^2,
        # d/dc N1 (prism) or N1^3 (tet); face rows N1^2 (prism) or N1^3 (tet))
        n1 = order + 1
        self.dsp_w, self.dsp_i, self.esp_w, self.esp_i = [], [], [], []
        for ct in (0, 1):
            wd = np.array([n1, n1 * n1, n1 if ct == 0 else n1 ** 3])
            if ct == 1 and order >= 3:
                wd = np.array([T.ns, T.ns, T.ns])   # P3 tet rows stay dense (DW<P,CT> in the kernel)
            Dm = np.transpose(T.D[ct], (2, 0, 1))  # (3, ns, ns)
            w_, i_ = sparse_fixed(Dm, wd[:, None])
            self.dsp_w.append(dev(w_))
            self.dsp_i.append(devi(i_))
            ke = n1 * n1 if ct == 0 else n1 ** 3
            w_, i_ = sparse_fixed(self.E64[ct], ke)
            self.esp_w.append(dev(w_))
            self.esp_i.append(devi(i_))
        self.c2f = [dev(T.c2f[ct]) for ct in (0, 1)]
        self.op = [dev(T.Op[ct]) for ct in (0, 1)]
        self.D = [dev(np.transpose(T.D[ct], (2, 0, 1))) for ct in (0, 1)]
        self.filt = [dev(T.filt[ct]) for ct in (0, 1)]
        self.gL, self.gR = dev(T.gL), dev(T.gR)
        self.fp_of_sp, self.coord_of_sp = devi(T.fp_of_sp), devi(T.coord_of_sp)
        bcp = np.zeros((8, 5)) if bc_params is None else np.asarray(bc_params, dtype=np.float64)
        self.cell_case = None
        self.n_cases = 1
        if cell_case is not None:
            cc = np.asarray(cell_case)
            self.n_cases = int(bcp.shape[0]) if bcp.ndim == 3 else 1
            if bcp.ndim != 3 or bcp.shape[1:] != (8, 5):
                raise ValueError("batched mesh needs bc_params of shape (n_cases, 8, 5)")
            if cc.shape != (C,) or (C and (cc.min() < 0 or cc.max() >= self.n_cases)):
                raise ValueError("cell_case must give a case in [0, n_cases) for every cell")
            self.cell_case = devi(cc)
        elif bcp.shape != (8, 5):
            raise ValueError("bc_params must be (8, 5)")
        self.bc_params = dev(bcp, torch.float64)
        self._set_state_ref(state_ref)
        # reference state of the perturbation form: the freestream by default
        if ref_state is None and bcp.ndim == 3:
            raise ValueError("a batched mesh needs an explicit common ref_state")
        self.ref_state = np.asarray(ref_state if ref_state is not None else bcp[BC_FARFIELD], dtype=np.float64)
        self.work_q = torch.empty((C, T.ns, 5), dtype=dtype, device=self.device)
        self.work_grad = torch.empty((C, T.ns, 12), dtype=dtype, device=self.device)
        self.work_umean = torch.empty((C, 5), dtype=dtype, device=self.device)
        self.work_rvol = torch.empty((C, T.ns, 5), dtype=dtype, device=self.device)
        self.work_vpack = torch.empty((C, T.ns, 20), dtype=dtype, device=self.device)
        self.work_h = torch.empty((C,), dtype=dtype, device=self.device)
        self.volume_dev = torch.as_tensor(self.volume, dtype=dtype).to(self.device)
        self.work_r = torch.empty((C, T.ns, 5), dtype=dtype, device=self.device)
        self.work_s1 = torch.empty((C, T.ns, 5), dtype=dtype, device=self.device)
        self.work_s2 = torch.empty((C, T.ns, 5), dtype=dtype, device=self.device)
        self.work_dt = torch.empty((C, T.ns), dtype=dtype, device=self.device)
        self.work_part = torch.zeros((4096,), dtype=torch.float64, device=self.device)
        self.norm_dev = torch.zeros((self.n_cases,), dtype=torch.float64, device=self.device)

        self._build_wall_records(dev, devi)
        if turbulence is not None:
            if turbulence != "SST":
                raise ValueError(f"unsupported turbulence model {turbulence!r}")
            self._build_sst(dev, devi)
        self._create(int(sl.rows.shape[0]), mesh.n_prism,
                     mesh.meta["prism_off"] if self.cell_case is not None else None,
                     mesh.meta["tet_off"] if self.cell_case is not None else None)

    def _set_state_ref(self, state_ref):
        if state_ref is None:
            self.state_ref, self.state_ref_dev = None, None
            return
        sr = np.asarray(state_ref, dtype=np.float64).reshape(-1, 5)
        if sr.shape[0] != self.n_cases or not np.all(np.isfinite(sr)):
            raise ValueError(f"state_ref must hold {self.n_cases} finite conserved states of 5 values")
        self.state_ref = sr
        self.state_ref_dev = torch.as_tensor(np.ascontiguousarray(sr)).to(self.device)

    def cell_offset(self):
        """(C, 1, 5) FP64 device tensor: the conserved offset of every cell's stored state (zeros
        without a state_ref)."""
        if self.state_ref is None:
            return torch.zeros((self.n_cells, 1, 5), dtype=torch.float64, device=self.device)
        if self.cell_case is None:
            return self.state_ref_dev.view(1, 1, 5).expand(self.n_cells, 1, 5)
        return self.state_ref_dev[self.cell_case.long()].view(self.n_cells, 1, 5)

    def to_absolute(self, U):
        """The conserved state (FP64, device) of a state U stored by this mesh."""
        X = U[..., :5].double()
        return X if self.state_ref is None else X + self.cell_offset()

    def to_storage(self, U_abs):
        """The stored form (this mesh's precision, device) of a conserved state (FP64 preferred:
        an FP32 input has already lost what the offset keeps)."""
        X = torch.as_tensor(U_abs, device=self.device)[..., :5].double()
        if self.state_ref is not None:
            X = X - self.cell_offset()
        return X.to(self.dtype).contiguous()

    def _create(self, n_rows, n_prism, prism_off=None, tet_off=None):
        """Fill the AfrMeshDesc from this object's device arrays and create the library's mesh."""
        order, dtype, C = self.order, self.dtype, self.n_cells
        turbulence = self.turbulence
        d = _lib.AfrMeshDesc()
        d.abi_version = _lib.ABI_VERSION
        d.order = order
        d.dtype = 0 if dtype == torch.float32 else 1
        d.n_rows = int(n_rows)
        d.n_cells = C
        d.n_prism = int(n_prism)
        self.n_prism, self.n_rows = int(n_prism), int(n_rows)
        p = lambda t: ctypes.c_void_p(t.data_ptr())  # noqa: E731
        d.det, d.adj, d.adj_fine = p(self.det), p(self.adj), p(self.adj_fine)
        d.slot_nb, d.slot_blk, d.slot_mask = p(self.slot_nb), p(self.slot_blk), p(self.slot_mask)
        d.slot_bc, d.slot_fallback = p(self.slot_bc), p(self.slot_fb)
        d.slot_tn, d.slot_adjrow, d.rows = p(self.slot_tn), p(self.slot_adjrow), p(self.rows)
        d.rows_nnz, d.rows_idx, d.rows_w = p(self.rows_nnz), p(self.rows_idx), p(self.rows_w)
        d.rows_knz = self.rows_knz
        d.prism_tab = p(self.prism_tab)
        d.prism_schedule = p(self.prism_schedule) if n_prism else None
        for ct in (0, 1):
            d.c2f[ct], d.op[ct], d.E[ct] = p(self.c2f[ct]), p(self.op[ct]), p(self.E[ct])
            d.D[ct], d.filt[ct] = p(self.D[ct]), p(self.filt[ct])
        d.gL, d.gR = p(self.gL), p(self.gR)
        d.fp_of_sp, d.coord_of_sp = p(self.fp_of_sp), p(self.coord_of_sp)
        d.gcl, d.bc_params = p(self.gcl), p(self.bc_params)
        if self.cell_case is not None:
            # host copies of the case offsets: prisms, tets (from the mesh) and wall records
            po = np.ascontiguousarray(prism_off, dtype=np.int64)
            to = np.ascontiguousarray(tet_off, dtype=np.int64)
            self.case_prism_off, self.case_tet_off = po, to
            d.cell_case, d.n_cases = p(self.cell_case), self.n_cases
            d.case_prism_off = po.ctypes.data_as(ctypes.c_void_p)
            d.case_tet_off = to.ctypes.data_as(ctypes.c_void_p)
            if self.n_wall_rec:
                d.case_wall_off = self.case_wall_off.ctypes.data_as(ctypes.c_void_p)
        for v in range(5):
            d.ref_state[v] = float(self.ref_state[v])
        d.state_ref = p(self.state_ref_dev) if self.state_ref is not None else None
        d.work_q, d.work_grad, d.work_umean = p(self.work_q), p(self.work_grad), p(self.work_umean)
        d.work_rvol, d.work_h = p(self.work_rvol), p(self.work_h)
        d.work_vpack = p(self.work_vpack)
        for ct in (0, 1):
            d.dsp_w[ct], d.dsp_i[ct] = p(self.dsp_w[ct]), p(self.dsp_i[ct])
            d.esp_w[ct], d.esp_i[ct] = p(self.esp_w[ct]), p(self.esp_i[ct])
        d.metric_flux_scale, d.volume = p(self.metric_flux_scale), p(self.volume_dev)
        d.work_r, d.work_s1, d.work_s2 = p(self.work_r), p(self.work_s1), p(self.work_s2)
        d.work_dt, d.work_part = p(self.work_dt), p(self.work_part)
        d.n_wall_rec = self.n_wall_rec
        if self.n_wall_rec:
            d.wrec_owner, d.wrec_oslot = p(self.wrec_owner), p(self.wrec_oslot)
            d.wrec_geom, d.w1d, d.work_fpart = p(self.wrec_geom), p(self.w1d), p(self.work_fpart)
        if turbulence is not None:
            d.n_records = self.n_records
            d.rec_owner, d.rec_oslot, d.rec_flags = p(self.rec_owner), p(self.rec_oslot), p(self.rec_flags)
            d.rec_tn, d.cell_inc = p(self.rec_tn), p(self.cell_inc)
            d.work_pack, d.work_tvol = p(self.work_pack), p(self.work_tvol)
            d.work_nbp = p(self.work_nbp)
            d.work_tsrc, d.work_crec = p(self.work_tsrc), p(self.work_crec)
        self._desc = d
        handle = ctypes.c_void_p()
        _lib.check(self.lib.afr_mesh_create(ctypes.byref(d), ctypes.byref(handle)), "afr_mesh_create")
        self.handle = handle


    @classmethod
    def concat(cls, parts, bc_params, ref_state):
        """The batched mesh of independent cases from their resident DeviceMeshes, built on the
        device: every per-cell array is concatenated (all prisms case by case, then all tets case
        by case, as mesh.concat_meshes numbers them), cell, row-block and face-record indices are
        offset, nothing is recomputed.  Every cell keeps exactly its case's data, so a case in the
        batch follows its trajectory alone bit for bit (same ``ref_state`` in every part).
        bc_params: (n_cases, 8, 5) boundary states; the parts must share order, precision and
        turbulence model."""
        if not parts:
            raise ValueError("no parts to concatenate")
        a = parts[0]
        for q in parts[1:]:
            if (q.order, q.dtype, q.turbulence, q.device) != (a.order, a.dtype, a.turbulence, a.device):
                raise ValueError("parts differ in order, precision, turbulence model or device")
        ref = np.asarray(ref_state, dtype=np.float64)
        for q in parts:
            if not np.array_equal(np.asarray(q.ref_state, dtype=np.float64), ref):
                raise ValueError("every part must have been built with the batch's ref_state")
        bcp = np.asarray(bc_params, dtype=np.float64)
        if bcp.shape != (len(parts), 8, 5):
            raise ValueError("bc_params must be (n_cases, 8, 5)")
        offs = [q.state_ref for q in parts]
        if any(o is None for o in offs) and not all(o is None for o in offs):
            raise ValueError("parts must all store offsets (state_ref) or all store the state itself")
        for q in parts:
            if q.state_ref is not None and q.n_cases != 1:
                raise ValueError("a part must be a single case")
        self = cls.__new__(cls)
        dv = a.device
        self.lib, self.turbulence, self.mesh, self.slots = a.lib, a.turbulence, None, None
        self.order, self.dtype, self.device, self.T = a.order, a.dtype, dv, a.T
        self.ns, self.nfp, self.nf = a.ns, a.nfp, a.nf
        n = len(parts)
        P = np.array([q.n_prism for q in parts], dtype=np.int64)
        Cs = np.array([q.n_cells for q in parts], dtype=np.int64)
        p_off = np.concatenate([[0], np.cumsum(P)])
        t_off = np.concatenate([[0], np.cumsum(Cs - P)])
        P_tot, C = int(p_off[-1]), int(Cs.sum())
        self.n_cells, self.n_cases = C, n

        def cells(get):
            xs = [get(q) for q in parts]
            return torch.cat([x[:q.n_prism] for x, q in zip(xs, parts)] +
                             [x[q.n_prism:] for x, q in zip(xs, parts)]).contiguous()

        def remap(x, i):
            """Cell indices of part i (-1 kept) in the batch numbering."""
            q = parts[i]
            x = x.long()
            out = torch.where(x < q.n_prism, x + int(p_off[i]), x - q.n_prism + P_tot + int(t_off[i]))
            return torch.where(x < 0, x, out).to(torch.int32)

        def cells_i(get):
            """Per-cell arrays holding cell indices, remapped part by part."""
            xs = [remap(get(q), i) for i, q in enumerate(parts)]
            return torch.cat([x[:q.n_prism] for x, q in zip(xs, parts)] +
                             [x[q.n_prism:] for x, q in zip(xs, parts)]).contiguous()

        self.slot_nb = cells_i(lambda q: q.slot_nb)
        blk_off = np.concatenate([[0], np.cumsum([q.n_rows // q.nfp for q in parts])])
        blks = [q.slot_blk + int(blk_off[i]) for i, q in enumerate(parts)]
        self.slot_blk = torch.cat([b[:q.n_prism] for b, q in zip(blks, parts)] +
                                  [b[q.n_prism:] for b, q in zip(blks, parts)]).contiguous()
        for name in ("slot_mask", "slot_bc", "slot_tn", "slot_fb", "slot_adjrow", "det", "adj", "adj_fine", "gcl",
                     "metric_flux_scale", "volume_dev"):
            setattr(self, name, cells(lambda q, name=name: getattr(q, name)))
        self.volume = np.concatenate([q.volume[:q.n_prism] for q in parts] + [q.volume[q.n_prism:] for q in parts])
        self.det_mean = np.concatenate([q.det_mean[:q.n_prism] for q in parts] +
                                       [q.det_mean[q.n_prism:] for q in parts])
        # interpolation rows: each part's rows without its trailing zero row, then one zero row
        knz = max(q.rows_knz for q in parts)

        def padw(x):
            return torch.nn.functional.pad(x, (0, knz - x.shape[1]))
        self.rows = torch.cat([q.rows for q in parts]).contiguous()
        self.rows_knz = knz
        n_rows = int(self.rows.shape[0])
        self.rows_nnz = torch.cat([q.rows_nnz[:q.n_rows] for q in parts] + [a.rows_nnz[-1:]]).contiguous()
        self.rows_idx = torch.cat([padw(q.rows_idx[:q.n_rows]) for q in parts] + [padw(a.rows_idx[-1:])]).contiguous()
        self.rows_w = torch.cat([padw(q.rows_w[:q.n_rows]) for q in parts] + [padw(a.rows_w[-1:])]).contiguous()
        self.prism_schedule = torch.cat([q.prism_schedule + int(p_off[i]) for i, q in enumerate(parts)]).to(
            torch.int32).contiguous()
        for name in ("prism_tab", "E", "E64", "dsp_w", "dsp_i", "esp_w", "esp_i", "c2f", "op", "D", "filt", "gL", "gR",
                     "fp_of_sp", "coord_of_sp"):
            setattr(self, name, getattr(a, name))
        cc = np.concatenate([np.repeat(np.arange(n), P), np.repeat(np.arange(n), Cs - P)]).astype(np.int32)
        self.cell_case = torch.as_tensor(cc).to(dv)
        self.bc_params = torch.as_tensor(bcp).to(dv)
        self.ref_state = ref
        self._set_state_ref(None if offs[0] is None else np.concatenate(offs, axis=0))
        dt = self.dtype
        ns = self.ns
        self.work_q = torch.empty((C, ns, 5), dtype=dt, device=dv)
        self.work_grad = torch.empty((C, ns, 12), dtype=dt, device=dv)
        self.work_umean = torch.empty((C, 5), dtype=dt, device=dv)
        self.work_rvol = torch.empty((C, ns, 5), dtype=dt, device=dv)
        self.work_vpack = torch.empty((C, ns, 20), dtype=dt, device=dv)
        self.work_h = torch.empty((C,), dtype=dt, device=dv)
        self.work_r = torch.empty((C, ns, 5), dtype=dt, device=dv)
        self.work_s1 = torch.empty((C, ns, 5), dtype=dt, device=dv)
        self.work_s2 = torch.empty((C, ns, 5), dtype=dt, device=dv)
        self.work_dt = torch.empty((C, ns), dtype=dt, device=dv)
        self.work_part = torch.zeros((4096,), dtype=torch.float64, device=dv)
        self.norm_dev = torch.zeros((n,), dtype=torch.float64, device=dv)
        # wall records, grouped case by case (each part's records in its own order)
        self.n_wall_rec = int(sum(q.n_wall_rec for q in parts))
        if self.n_wall_rec:
            have = [i for i, q in enumerate(parts) if q.n_wall_rec]
            self.wrec_owner = torch.cat([remap(parts[i].wrec_owner, i) for i in have]).contiguous()
            self.wrec_oslot = torch.cat([parts[i].wrec_oslot for i in have]).contiguous()
            self.wrec_geom = torch.cat([parts[i].wrec_geom for i in have]).contiguous()
            self.w1d = a.w1d if a.n_wall_rec else parts[have[0]].w1d
            self.case_wall_off = np.concatenate([[0], np.cumsum([q.n_wall_rec for q in parts])]).astype(np.int64)
            self.work_fpart = torch.zeros((self.n_wall_rec, 6), dtype=torch.float64, device=dv)
            self.forces_dev = torch.zeros((6,), dtype=torch.float64, device=dv)
        if self.turbulence is not None:
            rec_off = np.concatenate([[0], np.cumsum([q.n_records for q in parts])])
            self.n_records = int(rec_off[-1])
            self.rec_owner = torch.cat([remap(q.rec_owner, i) for i, q in enumerate(parts)]).contiguous()
            self.rec_oslot = torch.cat([q.rec_oslot for q in parts]).contiguous()
            self.rec_flags = torch.cat([q.rec_flags for q in parts]).contiguous()
            self.rec_tn = torch.cat([q.rec_tn for q in parts]).contiguous()
            incs = [torch.where(q.cell_inc >= 0, q.cell_inc + 16 * int(rec_off[i]), q.cell_inc)
                    for i, q in enumerate(parts)]
            self.cell_inc = torch.cat([x[:q.n_prism] for x, q in zip(incs, parts)] +
                                      [x[q.n_prism:] for x, q in zip(incs, parts)]).contiguous()
            if self.n_records * 16 >= 2 ** 31:
                raise ValueError("too many face records for the int32 incidence encoding")
            self.d_wall = cells(lambda q: q.d_wall)
            self.sps_xyz = cells(lambda q: q.sps_xyz) if all(q.sps_xyz is not None for q in parts) else None
            nfp = self.nfp
            self.work_pack = torch.empty((C, 5, nfp, 8), dtype=dt, device=dv)
            self.work_nbp = torch.empty((C, ns, 4), dtype=dt, device=dv)
            self.work_tvol = torch.empty((C, ns, 2), dtype=dt, device=dv)
            self.work_tsrc = torch.empty((C, ns, 3), dtype=dt, device=dv)
            self.work_crec = torch.empty((self.n_records, nfp, 2), dtype=dt, device=dv)
        else:
            self.sps_xyz = None
        self.part_prism_off, self.part_tet_off = p_off, t_off
        self._create(n_rows, P_tot, p_off, t_off)
        return self

    def part_cells(self, i):
        """Device cell indices of part i of a concatenated mesh, in that part's own cell order."""
        p0, p1 = int(self.part_prism_off[i]), int(self.part_prism_off[i + 1])
        t0, t1 = int(self.part_tet_off[i]), int(self.part_tet_off[i + 1])
        return torch.cat([torch.arange(p0, p1, device=self.device),
                          self.n_prism + torch.arange(t0, t1, device=self.device)])

    @staticmethod
    def owner_primary(mesh):
        """First record (lowest index) of each prism quad on its owner side; every triangle face is primary
        (src/autoflowcfd/fr/face_flux_points_merge.py:101-110)."""
        fr = mesh.faces
        F = fr.n_faces
        quad = (fr.owner < mesh.n_prism) & np.isin(fr.owner_code, (0, 1, 2))
        primary = np.ones(F, dtype=bool)
        qi = np.flatnonzero(quad)
        if len(qi):
            key = fr.owner[qi] * 8 + fr.owner_code[qi]
            order = np.lexsort((qi, key))
            ks = key[order]
            first = np.ones(len(ks), dtype=bool)
            first[1:] = ks[1:] != ks[:-1]
            primary[qi[order[~first]]] = False
        return primary

    def _primary(self):
        if getattr(self, "_primary_cache", None) is None:
            self._primary_cache = self.owner_primary(self.mesh)
        return self._primary_cache

    def _build_wall_records(self, dev, devi):
        """Owner-primary WALL records of the force integral (fr_coefficients.py:108-112)."""
        from .mesh import is_wall
        fr = self.mesh.faces
        sel = np.flatnonzero(is_wall(fr.bc) & self._primary())
        self.n_wall_rec = len(sel)
        if not len(sel):
            return
        code_to_slot = np.full(6, -1, dtype=np.int64)
        for si, code in enumerate(SLOT_CODES[0]):
            code_to_slot[code] = si
        if self.cell_case is not None:
            cc = self.cell_case.cpu().numpy()
            cnt = np.bincount(cc[fr.owner[sel]], minlength=self.n_cases)
            if np.any(np.diff(cc[fr.owner[sel]]) < 0):
                raise ValueError("wall records of a batched mesh must be grouped case by case")
            self.case_wall_off = np.concatenate([[0], np.cumsum(cnt)]).astype(np.int64)
        self.wrec_owner = devi(fr.owner[sel])
        self.wrec_oslot = devi(code_to_slot[fr.owner_code[sel]])
        geom = np.concatenate([fr.normal[sel], fr.area[sel, None]], axis=1)
        self.wrec_geom = torch.as_tensor(np.ascontiguousarray(geom, dtype=np.float64)).to(self.device)
        self.w1d = torch.as_tensor(np.polynomial.legendre.leggauss(self.order + 1)[1]).to(self.device)
        self.work_fpart = torch.zeros((len(sel), 6), dtype=torch.float64, device=self.device)
        self.forces_dev = torch.zeros((6,), dtype=torch.float64, device=self.device)

    def wall_forces(self, U, mu_t=None, mu=1.8e-5):
        """(6,) FP64 device tensor: pressure force (3), viscous force (3) on the WALL boundary."""
        if not self.n_wall_rec:
            raise ValueError("mesh has no WALL boundary records")
        self._check_state(U)
        _lib.check(self.lib.afr_forces_f64(self.handle, ctypes.c_void_p(U.data_ptr()), self._check_mut(mu_t),
                                           float(mu), ctypes.c_void_p(self.forces_dev.data_ptr()), self._stream()),
                   "afr_forces_f64")
        return self.forces_dev

    def wall_forces_cases(self, U, mu_t=None, mu=1.8e-5, out=None):
        """(n_cases, 6) FP64 device tensor of a batched mesh: per-case pressure and viscous force."""
        if self.cell_case is None:
            raise ValueError("wall_forces_cases needs a batched mesh (cell_case)")
        if not self.n_wall_rec:
            raise ValueError("mesh has no WALL boundary records")
        self._check_state(U)
        out = torch.zeros((self.n_cases, 6), dtype=torch.float64, device=self.device) if out is None else out
        if out.shape != (self.n_cases, 6) or out.dtype != torch.float64 or not out.is_contiguous():
            raise ValueError("out must be a contiguous (n_cases, 6) float64 tensor")
        _lib.check(self.lib.afr_forces_cases_f64(self.handle, ctypes.c_void_p(U.data_ptr()), self._check_mut(mu_t),
                                                 float(mu), ctypes.c_void_p(out.data_ptr()), self._stream()),
                   "afr_forces_cases_f64")
        return out

    def _build_sst(self, dev, devi):
        """Face-record tables of the SST transport and the wall distance of every solution point.

        Records follow the mesh's triangle face records; the owner-primary flag marks the
        first record (lowest index) of each prism quad on its owner side, the records whose
        neighbour values the reference assembles from neighbour sources
        (src/autoflowcfd/fr/face_flux_points_merge.py:101-110); every other record takes a
        Neumann ghost in the transport (src/autoflowcfd/core/turbulence/transport_kernel.py:103-170).
        """
        from .mesh import is_wall
        mesh, T, C = self.mesh, self.T, self.n_cells
        fr = mesh.faces
        F = fr.n_faces
        bnd = fr.neighbor < 0
        primary = self._primary()
        flags = ((~bnd & primary).astype(np.int32)) | (is_wall(fr.bc).astype(np.int32) << 1)
        code_to_slot = np.full(6, -1, dtype=np.int64)
        for si, code in enumerate(SLOT_CODES[0]):
            code_to_slot[code] = si
        oslot = code_to_slot[fr.owner_code]
        if np.any(oslot < 0):
            raise ValueError("face record with an unknown owner face code")
        inc = record_incidence(C, fr)
        if F * 16 >= 2 ** 31:
            raise ValueError("too many face records for the int32 incidence encoding")
        # records renumbered in the locality order of their owners (prisms: the Morton schedule,
        # then tets): the face kernel then walks neighbouring cells together; each cell keeps
        # its incidence list in the same order, only the record ids are renamed
        rank = np.arange(C, dtype=np.int64)
        if mesh.n_prism:
            sched = self.prism_schedule.cpu().numpy().astype(np.int64)
            rank[sched] = np.arange(mesh.n_prism)
        perm = np.argsort(rank[fr.owner], kind="stable")
        new_id = np.empty(F, dtype=np.int64)
        new_id[perm] = np.arange(F)
        inc = np.where(inc >= 0, new_id[np.maximum(inc, 0) >> 4] * 16 + (inc & 15), -1)
        self.n_records = F
        self.rec_perm = perm
        self.rec_owner, self.rec_oslot, self.rec_flags = devi(fr.owner[perm]), devi(oslot[perm]), devi(flags[perm])
        self.rec_tn = dev(fr.normal[perm])
        self.cell_inc = devi(inc)
        dt = self.dtype
        self.work_pack = torch.empty((C, 5, T.nfp, 8), dtype=dt, device=self.device)
        self.work_nbp = torch.empty((C, T.ns, 4), dtype=dt, device=self.device)
        self.work_tvol = torch.empty((C, T.ns, 2), dtype=dt, device=self.device)
        self.work_tsrc = torch.empty((C, T.ns, 3), dtype=dt, device=self.device)
        self.work_crec = torch.empty((F, T.nfp, 2), dtype=dt, device=self.device)
        # wall distance: nearest WALL node of every solution point
        self.d_wall = torch.empty((C, T.ns), dtype=dt, device=self.device)
        wall = np.ascontiguousarray(mesh.nodes[mesh.wall_nodes], dtype=np.float64) if mesh.wall_nodes is not None \
            else np.zeros((0, 3))
        if len(wall) == 0:
            raise ValueError("SST needs WALL boundary nodes for the wall distance")
        wall_d = torch.as_tensor(wall).to(self.device)
        xyz = self.sps_xyz.contiguous()
        P = mesh.n_prism
        if self.cell_case is None:
            pieces = [(0, C, 0, len(wall))]
        else:
            # case by case: (cell range, wall-node range) for the prisms and for the tets of each case
            po, to, wo = (np.asarray(mesh.meta[k], dtype=np.int64) for k in ("prism_off", "tet_off", "wall_off"))
            pieces = []
            for i in range(len(po) - 1):
                if wo[i + 1] == wo[i] and (po[i + 1] > po[i] or to[i + 1] > to[i]):
                    raise ValueError(f"case {i} has no WALL boundary nodes for the wall distance")
                pieces.append((po[i], po[i + 1], wo[i], wo[i + 1]))
                pieces.append((P + to[i], P + to[i + 1], wo[i], wo[i + 1]))
        for c0, c1, w0, w1 in pieces:
            if c1 <= c0:
                continue
            _lib.check(self.lib.afr_wall_distance(ctypes.c_void_p(xyz[c0:c1].data_ptr()), int(c1 - c0) * T.ns,
                                                  ctypes.c_void_p(wall_d[w0:w1].data_ptr()), int(w1 - w0),
                                                  ctypes.c_void_p(self.d_wall[c0:c1].data_ptr()),
                                                  0 if dt == torch.float32 else 1, self._stream()),
                       "afr_wall_distance")
        torch.cuda.current_stream(self.device).synchronize()

    def sst_fields(self, k0, omega0):
        """Uniform k, omega (C, ns) and zero nu_t / mu_t of the SST model."""
        shape = (self.n_cells, self.ns)
        k = torch.full(shape, float(k0), dtype=self.dtype, device=self.device)
        w = torch.full(shape, float(omega0), dtype=self.dtype, device=self.device)
        return k, w, torch.zeros(shape, dtype=self.dtype, device=self.device), \
            torch.zeros(shape, dtype=self.dtype, device=self.device)

    def sst_step(self, U, k, w, dt, nu_t, mu_t, mu=1.8e-5, rates=None):
        """SST step: k, omega updated in place, or (rates given) dk/dt, domega/dt written to rates (C, ns, 2)."""
        r = ctypes.c_void_p(rates.data_ptr()) if rates is not None else None
        _lib.check(self.lib.afr_sst_step(self.handle, ctypes.c_void_p(U.data_ptr()), ctypes.c_void_p(k.data_ptr()),
                                         ctypes.c_void_p(w.data_ptr()), ctypes.c_void_p(self.d_wall.data_ptr()),
                                         ctypes.c_void_p(dt.data_ptr()), ctypes.c_void_p(nu_t.data_ptr()),
                                         ctypes.c_void_p(mu_t.data_ptr()), r, float(mu), self._stream()),
                   "afr_sst_step")

    def sst_rk3_step(self, U, k, w, nu_t, mu_t, mu=1.8e-5, Pr=0.72, Pr_t=0.9, n_vars_total=7, norm=None):
        """One SST pseudo-step (k, omega, nu_t, mu_t and U updated in place); returns the (1,) FP64 norm."""
        self._check_state(U)
        norm = self.norm_dev if norm is None else norm
        self._check_norm(norm)
        _lib.check(self.lib.afr_sst_rk3_step(self.handle, ctypes.c_void_p(U.data_ptr()), ctypes.c_void_p(k.data_ptr()),
                                             ctypes.c_void_p(w.data_ptr()), ctypes.c_void_p(self.d_wall.data_ptr()),
                                             ctypes.c_void_p(nu_t.data_ptr()), ctypes.c_void_p(mu_t.data_ptr()),
                                             float(mu), float(Pr), float(Pr_t), int(n_vars_total),
                                             ctypes.c_void_p(norm.data_ptr()), self._stream()), "afr_sst_rk3_step")
        return norm

    def _gpu_metrics(self, sl):
        """FP64 metric precompute on the device (afr_metrics)."""
        mesh, T, dt, C = self.mesh, self.T, self.dtype, self.n_cells
        f64 = torch.float64
        dv = self.device

        def t64(x):
            return torch.as_tensor(np.ascontiguousarray(x, dtype=np.float64)).to(dv)

        nodes = t64(mesh.nodes)
        pconn = torch.as_tensor(np.ascontiguousarray(mesh.prism_conn.reshape(-1, 6), dtype=np.int32)).to(dv)
        tconn = torch.as_tensor(np.ascontiguousarray(mesh.tet_conn.reshape(-1, 4), dtype=np.int32)).to(dv)
        if mesh.n_prism and int(mesh.prism_conn.max()) >= len(mesh.nodes):
            raise ValueError("prism connectivity out of range")
        if mesh.n_tet and int(mesh.tet_conn.max()) >= len(mesh.nodes):
            raise ValueError("tet connectivity out of range")
        sps, fine, w = t64(T.sps), t64(T.fine), t64(T.w_sps)
        op64 = [t64(T.Op[ct]) for ct in (0, 1)]
        E64 = [t64(e) for e in self.E64]
        tn64 = t64(sl.tn)
        self.det = torch.empty((C, T.ns), dtype=dt, device=dv)
        self.adj = torch.empty((C, T.ns, 3, 3), dtype=dt, device=dv)
        self.adj_fine = torch.empty((C, T.nf, 3, 3), dtype=dt, device=dv)
        self.gcl = torch.empty((C, T.ns, 3), dtype=dt, device=dv)
        self.metric_flux_scale = torch.empty((C, T.ns), dtype=dt, device=dv)
        self.slot_adjrow = torch.empty((C, NSLOT, T.nfp, 3), dtype=dt, device=dv)
        vol = torch.empty(C, dtype=f64, device=dv)
        dmean = torch.empty(C, dtype=f64, device=dv)
        self.slot_fb = torch.empty((C, NSLOT), dtype=torch.int32, device=dv)
        bad = torch.zeros(1, dtype=torch.int32, device=dv)
        g = _lib.AfrGeomDesc()
        g.abi_version, g.order, g.dtype = _lib.ABI_VERSION, self.order, 0 if dt == torch.float32 else 1
        g.n_cells, g.n_prism = C, mesh.n_prism
        p = lambda t: ctypes.c_void_p(t.data_ptr())  # noqa: E731
        g.nodes, g.prism_conn, g.tet_conn = p(nodes), p(pconn), p(tconn)
        g.sps, g.fine, g.w_sps = p(sps), p(fine), p(w)
        for ct in (0, 1):
            g.op64[ct], g.E64[ct] = p(op64[ct]), p(E64[ct])
        g.slot_tn64 = p(tn64)
        g.det, g.adj, g.adj_fine, g.gcl = p(self.det), p(self.adj), p(self.adj_fine), p(self.gcl)
        g.metric_flux_scale, g.slot_adjrow = p(self.metric_flux_scale), p(self.slot_adjrow)
        g.volume, g.det_mean, g.slot_fallback, g.bad_det = p(vol), p(dmean), p(self.slot_fb), p(bad)
        if self.turbulence is not None:
            self.sps_xyz = torch.empty((C, T.ns, 3), dtype=f64, device=dv)
            g.sps_xyz = p(self.sps_xyz)
        _lib.check(self.lib.afr_metrics(ctypes.byref(g), self._stream()), "afr_metrics")
        torch.cuda.current_stream(dv).synchronize()
        if int(bad.item()) != 0:
            raise ValueError("non-positive Jacobian determinant (mesh distortion)")
        self.volume = vol.cpu().numpy()
        self.det_mean = dmean.cpu().numpy()

    @staticmethod
    def _validate(sl, C, T, n_prism):
        """Host-side bounds checks of every index the kernels dereference."""
        nb = sl.nb
        if nb.min() < -1 or nb.max() >= C:
            raise ValueError("slot source cell out of range")
        n_rows = sl.rows.shape[0]
        used = (sl.bc == -1)
        if np.any(used & (nb[..., 0] < 0)):
            raise ValueError("interior slot without a source cell")
        if np.any((sl.blk + 1) * T.nfp > n_rows):
            raise ValueError("slot row block out of range")
        if np.any((sl.bc < -2) | (sl.bc > 7)):
            raise ValueError("unknown boundary code")
        rows = sl.rows_of_bc()
        if np.any((sl.bc >= 0) & ((rows < 0) | (rows >= 8))):
            raise ValueError("boundary-state row out of range (8 rows per case)")
        if np.any(sl.bc[:n_prism] == -2):
            raise ValueError("prism with an unused face slot")

    def __del__(self):
        h = getattr(self, "handle", None)
        if h is not None and h.value:
            try:
                self.lib.afr_mesh_destroy(h)
            except Exception:
                pass

    def _stream(self):
        return ctypes.c_void_p(torch.cuda.current_stream(self.device).cuda_stream)

    def _check_state(self, U):
        if U.dtype != self.dtype or U.device.type != "cuda" or not U.is_contiguous():
            raise ValueError("state must be a contiguous device tensor of the mesh precision")
        if tuple(U.shape[:2]) != (self.n_cells, self.ns) or U.shape[2] < 5:
            raise ValueError(f"state shape {tuple(U.shape)} != ({self.n_cells}, {self.ns}, 5)")

    def primitives(self, U, dQ=None):
        """dQ = primitive(U) - ref_state into dQ (default: the mesh work buffer)."""
        self._check_state(U)
        if self.state_ref is not None:
            return self.prepare_state(U, dQ)[0]
        dQ = self.work_q if dQ is None else dQ
        ref = (ctypes.c_double * 5)(*[float(x) for x in self.ref_state])
        _lib.check(self.lib.afr_primitives(ctypes.c_void_p(U.data_ptr()), ctypes.c_void_p(dQ.data_ptr()),
                                           self.n_cells * self.ns, 0 if self.dtype == torch.float32 else 1,
                                           ref, self._stream()), "afr_primitives")
        return dQ

    def prepare_state(self, U, dQ=None, umean=None):
        """dQ = primitive(U) - ref_state and the per-cell mean |U| (outlier reference)."""
        self._check_state(U)
        dQ = self.work_q if dQ is None else dQ
        umean = self.work_umean if umean is None else umean
        _lib.check(self.lib.afr_prepare_state(self.handle, ctypes.c_void_p(U.data_ptr()),
                                              ctypes.c_void_p(dQ.data_ptr()), ctypes.c_void_p(umean.data_ptr()),
                                              self._stream()), "afr_prepare_state")
        return dQ, umean

    def inviscid_from_primitives(self, dQ, umean, R, suppress=True):
        _lib.check(self.lib.afr_inviscid_from_primitives(self.handle, ctypes.c_void_p(dQ.data_ptr()),
                                                         ctypes.c_void_p(umean.data_ptr()),
                                                         ctypes.c_void_p(R.data_ptr()),
                                                         int(bool(suppress)), self._stream()),
                   "afr_inviscid_from_primitives")
        return R

    def inviscid_residual(self, U, R=None, suppress=True):
        """dU/dt inviscid (C, ns, 5) of the conserved state U (C, ns, 5)."""
        self._check_state(U)
        if U.shape[2] != 5:
            raise ValueError("inviscid_residual takes the 5 Euler variables")
        R = torch.empty_like(U) if R is None else R
        _lib.check(self.lib.afr_inviscid_residual(self.handle, ctypes.c_void_p(U.data_ptr()),
                                                  ctypes.c_void_p(R.data_ptr()), int(bool(suppress)),
                                                  self._stream()), "afr_inviscid_residual")
        return R

    def _check_mut(self, mu_t):
        if mu_t is None:
            return None
        if mu_t.dtype != self.dtype or not mu_t.is_contiguous() or tuple(mu_t.shape) != (self.n_cells, self.ns):
            raise ValueError(f"mu_t must be a contiguous ({self.n_cells}, {self.ns}) tensor of the mesh precision")
        return ctypes.c_void_p(mu_t.data_ptr())

    def viscous_residual(self, U, mu_t=None, R=None, mu=1.8e-5, Pr=0.72, Pr_t=0.9, suppress=True):
        """Viscous BR1 part of dU/dt (C, ns, 5) of the conserved state U."""
        self._check_state(U)
        R = torch.empty_like(U) if R is None else R
        _lib.check(self.lib.afr_viscous_residual(self.handle, ctypes.c_void_p(U.data_ptr()), self._check_mut(mu_t),
                                                 ctypes.c_void_p(R.data_ptr()), float(mu), float(Pr), float(Pr_t),
                                                 int(bool(suppress)), self._stream()), "afr_viscous_residual")
        return R

    def viscous_volume(self, U, mu_t=None, mu=1.8e-5, Pr=0.72, Pr_t=0.9):
        """Prepared state, gradients and viscous volume term of U into the work buffers."""
        self._check_state(U)
        _lib.check(self.lib.afr_viscous_volume(self.handle, ctypes.c_void_p(U.data_ptr()), self._check_mut(mu_t),
                                               float(mu), float(Pr), float(Pr_t), self._stream()),
                   "afr_viscous_volume")

    def viscous_interface(self, U, mu_t, R, mu=1.8e-5, Pr=0.72, Pr_t=0.9, suppress=True, accumulate=True):
        """BR1 interface correction + volume term into R (added when accumulate)."""
        _lib.check(self.lib.afr_viscous_interface(self.handle, ctypes.c_void_p(U.data_ptr()), self._check_mut(mu_t),
                                                  ctypes.c_void_p(R.data_ptr()), float(mu), float(Pr), float(Pr_t),
                                                  int(bool(suppress)), int(bool(accumulate)), self._stream()),
                   "afr_viscous_interface")
        return R

    def local_time_step_nut(self, U, nu_t, dt=None, mu=1.8e-5):
        """Local time step with mu_t = rho(U) nu_t."""
        self._check_state(U)
        dt = torch.empty((self.n_cells, self.ns), dtype=self.dtype, device=self.device) if dt is None else dt
        _lib.check(self.lib.afr_local_time_step_nut(self.handle, ctypes.c_void_p(U.data_ptr()),
                                                    ctypes.c_void_p(nu_t.data_ptr()), ctypes.c_void_p(dt.data_ptr()),
                                                    float(mu), self._stream()), "afr_local_time_step_nut")
        return dt

    def residual(self, U, mu_t=None, R=None, mu=1.8e-5, Pr=0.72, Pr_t=0.9, suppress=True):
        """Navier-Stokes dU/dt = inviscid + viscous (C, ns, 5)."""
        self._check_state(U)
        R = torch.empty_like(U) if R is None else R
        _lib.check(self.lib.afr_residual(self.handle, ctypes.c_void_p(U.data_ptr()), self._check_mut(mu_t),
                                         ctypes.c_void_p(R.data_ptr()), float(mu), float(Pr), float(Pr_t),
                                         int(bool(suppress)), self._stream()), "afr_residual")
        return R

    def local_time_step(self, U, mu_t=None, mu=1.8e-5, dt=None):
        """Local time step (C, ns) of the conserved state U."""
        self._check_state(U)
        dt = torch.empty((self.n_cells, self.ns), dtype=self.dtype, device=self.device) if dt is None else dt
        _lib.check(self.lib.afr_local_time_step(self.handle, ctypes.c_void_p(U.data_ptr()), self._check_mut(mu_t),
                                                ctypes.c_void_p(dt.data_ptr()), float(mu), self._stream()),
                   "afr_local_time_step")
        return dt

    def rk_stage(self, U0, Ua, R, dt, U_out, c0, c1, c2):
        """U_out = c0 U0 + c1 Ua + c2 dt R, positivity, modal filter."""
        ua = ctypes.c_void_p(Ua.data_ptr()) if Ua is not None else None
        _lib.check(self.lib.afr_rk_stage(self.handle, ctypes.c_void_p(U0.data_ptr()), ua, ctypes.c_void_p(R.data_ptr()),
                                         ctypes.c_void_p(dt.data_ptr()), ctypes.c_void_p(U_out.data_ptr()),
                                         float(c0), float(c1), float(c2), self._stream()), "afr_rk_stage")
        return U_out

    def residual_norm_cases(self, R, n_vars_total=5, out=None):
        """Per-case FP64 RMS norms of R as an (n_cases,) device tensor (batched mesh)."""
        out = self.norm_dev if out is None else out
        self._check_norm(out)
        _lib.check(self.lib.afr_residual_norm_cases(self.handle, ctypes.c_void_p(R.data_ptr()), int(n_vars_total),
                                                    ctypes.c_void_p(out.data_ptr()), self._stream()),
                   "afr_residual_norm_cases")
        return out

    def _check_norm(self, out):
        if out.dtype != torch.float64 or out.numel() < self.n_cases or not out.is_cuda:
            raise ValueError(f"norm output must be a float64 device tensor of >= {self.n_cases} values")

    def residual_norm(self, R, n_vars_total=5, out=None):
        """FP64 RMS norm of R as a (1,) device tensor."""
        out = self.norm_dev if out is None else out
        _lib.check(self.lib.afr_residual_norm_f64(self.handle, ctypes.c_void_p(R.data_ptr()), int(n_vars_total),
                                                  ctypes.c_void_p(out.data_ptr()), self._stream()),
                   "afr_residual_norm_f64")
        return out

    def ssp_rk3_step_dt(self, U, dt, mu_t=None, mu=1.8e-5, Pr=0.72, Pr_t=0.9, n_vars_total=5, norm=None):
        """One SSP-RK3 step of U in place with the same dt at every solution point; returns the
        (1,) FP64 device norm of R(U)."""
        self._check_state(U)
        norm = self.norm_dev if norm is None else norm
        self._check_norm(norm)
        _lib.check(self.lib.afr_ssp_rk3_step_dt(self.handle, ctypes.c_void_p(U.data_ptr()), self._check_mut(mu_t),
                                                float(dt), float(mu), float(Pr), float(Pr_t), int(n_vars_total),
                                                ctypes.c_void_p(norm.data_ptr()), self._stream()),
                   "afr_ssp_rk3_step_dt")
        return norm

    def ssp_rk3_step(self, U, mu_t=None, mu=1.8e-5, Pr=0.72, Pr_t=0.9, n_vars_total=5, norm=None):
        """One steady pseudo-step of U in place; returns the (1,) FP64 device tensor of the RMS norm of R(U)."""
        self._check_state(U)
        norm = self.norm_dev if norm is None else norm
        self._check_norm(norm)
        _lib.check(self.lib.afr_ssp_rk3_step(self.handle, ctypes.c_void_p(U.data_ptr()), self._check_mut(mu_t),
                                             float(mu), float(Pr), float(Pr_t), int(n_vars_total),
                                             ctypes.c_void_p(norm.data_ptr()), self._stream()), "afr_ssp_rk3_step")
        return norm
