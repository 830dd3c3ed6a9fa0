"""Order continuation P0 -> P1 -> ... -> P on the GPU.

src/autoflowcfd/core/utils/order_continuation.py:207-425 (``run_order_continuation``, the
reference's default for P >= 2 in FRSolver.solve, src/autoflowcfd/core/fr_solver/solver.py:292):

  * P0 from the uniform free stream: the finite-volume inviscid residual plus the one-point viscous
    residual (the boundary interior penalty), SSP-RK3 with the P0 local time step -- ``P0Mesh``
    (libafr.so: afr_p0_ssp_rk3_step);
  * then each order p = 1 .. P on a GPUFRSolver of that order, the state interpolated from the
    previous order by the tensor-product linear interpolation of the GL points
    (order_continuation.py:14-57; afr_interpolate_order);
  * each phase runs at most max_iter // (P + 1) pseudo-steps, converges below tol 10^(P - p), and
    hands over to the next order after at least 20 steps once its residual fell 100x.

With SST the P0 phase also runs the turbulence step with the one-point operators
(afr_p0_sst_step), from the reference's reset fields k = 1e-6, omega = 1, nu_t = 0 and the per-cell
mean of the target order's wall distance (order_continuation.py:241-272); k, omega, nu_t and the
wall distance are interpolated with the state at every transition (order_continuation.py:118-140),
so every phase, the target one included, sees that cell-constant wall distance.  All orders share
the locality cell order the solvers store, so the interpolation is one kernel on resident arrays.
"""

import ctypes

import numpy as np
import torch

from . import _lib
from .device import DeviceMesh, record_incidence
from .geometry import det_adj, jacobian
from .mesh import CUBE_FACE_AXIS, CUBE_FACE_SIDE, PRISM_CUBE_FACES, is_wall

GAMMA = 1.4


def interp_matrix(old_order, new_order):
    """(ns_new, ns_old) linear interpolation / extrapolation on the tensor GL points, ordered
    ia * n^2 + ib * n + ic as the solution points are: per axis, each new point takes the two old
    points of its interval (the first or last interval outside the old range), as scipy's
    RegularGridInterpolator(method='linear', fill_value=None) does in order_continuation.py:14-57."""
    xo = np.polynomial.legendre.leggauss(old_order + 1)[0]
    xn = np.polynomial.legendre.leggauss(new_order + 1)[0]
    no = len(xo)
    W1 = np.zeros((len(xn), no))
    for i, x in enumerate(xn):
        if no == 1:
            W1[i, 0] = 1.0
            continue
        j = int(np.clip(np.searchsorted(xo, x) - 1, 0, no - 2))
        t = (x - xo[j]) / (xo[j + 1] - xo[j])
        W1[i, j], W1[i, j + 1] = 1.0 - t, t
    return np.kron(np.kron(W1, W1), W1)


def _quad_lower(mesh, cells, codes, tri_nodes):
    """Whether each record's triangle is the lower half of its owner's prism quad (the half that
    holds the quad's first corner and the diagonal of the triangulation rule,
    face_flux_points_data.py:145-217)."""
    cn = mesh.prism_conn[cells]
    vs = np.argsort(cn[:, :3], axis=1, kind="stable")
    ws = vs + 3
    quad = np.array([PRISM_CUBE_FACES[c] for c in (0, 1, 2)])[codes]
    e0, e1 = quad[:, 0], quad[:, 1]

    def same(a, b, c, d):
        return ((a == c) & (b == d)) | ((a == d) & (b == c))
    d01 = same(e0, e1, vs[:, 0], vs[:, 1])
    d12 = same(e0, e1, vs[:, 1], vs[:, 2])
    da = np.where(d01, vs[:, 0], np.where(d12, vs[:, 1], vs[:, 0]))
    db = np.where(d01, ws[:, 1], np.where(d12, ws[:, 2], ws[:, 2]))
    i0, i1, i2, i3 = quad.T
    std = same(da, db, i0, i2)
    loc = np.where(std[:, None], np.stack([i0, i1, i2], 1), np.stack([i0, i1, i3], 1))
    lower = np.sort(np.take_along_axis(cn, loc, 1), axis=1)
    return np.all(lower == np.sort(tri_nodes, axis=1), axis=1)


def p0_record_sources(mesh):
    """(F, 2) source cells and (F, 2) weights of every record's owner side at the one P0 flux point
    (face_flux_points_merge.py:83-124 with the split of face_flux_points_data.py:145-217): a plain
    interior record takes its neighbour; the primary record of a prism quad group takes the
    group's first neighbour with weight 1 when its triangle is the quad's lower half (the centre
    point lies on the diagonal, which belongs to the lower half) and the second record's neighbour
    with the complementary weight; a quad held by one record (a collapsed-edge wedge's) takes its
    neighbour with weight 1, as any single-source face; boundary and non-primary records have none
    (-1)."""
    fr = mesh.faces
    F, P = fr.n_faces, mesh.n_prism
    src = np.full((F, 2), -1, dtype=np.int64)
    wgt = np.zeros((F, 2))
    inner = fr.neighbor >= 0
    quad = inner & (fr.owner < P) & np.isin(fr.owner_code, (0, 1, 2))
    plain = np.flatnonzero(inner & ~quad)
    src[plain, 0] = fr.neighbor[plain]
    wgt[plain, 0] = 1.0
    qi = np.flatnonzero(quad)
    if len(qi):
        key = fr.owner[qi] * 8 + fr.owner_code[qi]
        order = np.lexsort((qi, key))
        qs, ks = qi[order], key[order]
        start = np.flatnonzero(np.concatenate([[True], ks[1:] != ks[:-1]]))
        size = np.diff(np.append(start, len(qs)))
        if np.any(size > 2):
            raise ValueError("more than two records on one prism quad")
        g0 = qs[start]
        low = _quad_lower(mesh, fr.owner[g0], fr.owner_code[g0], fr.node_ids[g0]).astype(np.float64)
        src[g0, 0] = fr.neighbor[g0]
        wgt[g0, 0] = np.where(size == 1, 1.0, low)
        two = size == 2
        g1 = qs[start[two] + 1]
        src[g0[two], 1] = fr.neighbor[g1]
        wgt[g0[two], 1] = 1.0 - low[two]
    return src, wgt


class P0Mesh:
    """Device arrays of the P0 phase on one mesh (in the order of ``mesh``); with ``d_wall`` (C)
    also those of the P0 SST step.  state_ref: None, or the (5,) conserved offset of the stored
    state (every state this mesh takes then holds U - state_ref, as DeviceMesh's)."""

    def __init__(self, mesh, cell_volume, bc_params, dtype=torch.float32, device="cuda", d_wall=None,
                 state_ref=None):
        self.lib = _lib.load()
        self.dtype, self.device = dtype, torch.device(device)
        if self.device.index is None:
            self.device = torch.device(self.device.type, torch.cuda.current_device())
        C = mesh.n_cells
        self.n_cells = C
        ref = np.zeros((1, 3))
        J = np.zeros((C, 1, 3, 3))
        if mesh.n_prism:
            J[:mesh.n_prism] = jacobian(True, ref, mesh.nodes[mesh.prism_conn])
        if mesh.n_tet:
            J[mesh.n_prism:] = jacobian(False, ref, mesh.nodes[mesh.tet_conn])
        det, adj = det_adj(J)
        det, adj = det[:, 0], adj[:, 0]
        mfs = np.sum(np.linalg.norm(adj, axis=-1), axis=-1)
        fr = mesh.faces
        F = fr.n_faces
        bnd = fr.neighbor < 0
        rows = np.maximum(fr.rows(), 0)
        rec_bc = np.where(bnd, fr.bc | (rows << 4), -1)
        primary = DeviceMesh.owner_primary(mesh)
        axis = CUBE_FACE_AXIS[fr.owner_code]
        side = CUBE_FACE_SIDE[fr.owner_code] > 0
        flags = (primary.astype(np.int64) | (axis << 1) | (side.astype(np.int64) << 3)
                 | ((bnd & is_wall(fr.bc)).astype(np.int64) << 4))
        geom = np.concatenate([fr.normal, fr.area[:, None]], axis=1)
        inc = record_incidence(C, fr)
        if F * 16 >= 2 ** 31:
            raise ValueError("too many face records for the int32 incidence encoding")
        ghost = np.full(C, -1, dtype=np.int64)
        b = np.flatnonzero(bnd)
        ghost[fr.owner[b]] = b            # ascending record order: the last boundary record stays
        dv = self.device

        def st(x):
            return torch.as_tensor(np.ascontiguousarray(x), dtype=dtype).to(dv)

        def i32(x):
            return torch.as_tensor(np.ascontiguousarray(x, dtype=np.int32)).to(dv)

        def f64(x):
            return torch.as_tensor(np.ascontiguousarray(x, dtype=np.float64)).to(dv)
        self.det, self.adj, self.mfs, self.vol1 = st(det), st(adj), st(mfs), st(8.0 * det)
        self.cell_volume = f64(cell_volume)
        self.rec_owner, self.rec_nb, self.rec_bc, self.rec_flags = i32(fr.owner), i32(fr.neighbor), i32(rec_bc), i32(flags)
        self.rec_geom = f64(geom)
        self.cell_inc, self.cell_ghost_rec = i32(inc), i32(ghost)
        bcp = np.asarray(bc_params, dtype=np.float64).reshape(-1, 5)
        if bnd.any() and (rows[bnd].max() >= len(bcp) or np.any(fr.bc[bnd] < 0)):
            raise ValueError("a boundary record names a bc table row the table does not hold")
        self.bc_params = f64(bcp)
        self.work_r = torch.empty((C, 5), dtype=dtype, device=dv)
        self.work_s1 = torch.empty((C, 5), dtype=dtype, device=dv)
        self.work_s2 = torch.empty((C, 5), dtype=dtype, device=dv)
        self.work_dt = torch.empty((C,), dtype=dtype, device=dv)
        self.work_part = torch.zeros((4096,), dtype=torch.float64, device=dv)
        self.norm = torch.zeros((1,), dtype=torch.float64, device=dv)
        d = _lib.AfrP0Desc()
        d.abi_version, d.dtype, d.n_cells, d.n_records = _lib.ABI_VERSION, 0 if dtype == torch.float32 else 1, C, F
        p = lambda t: ctypes.c_void_p(t.data_ptr())  # noqa: E731
        d.det, d.adj, d.mfs, d.vol1, d.cell_volume = p(self.det), p(self.adj), p(self.mfs), p(self.vol1), p(self.cell_volume)
        d.rec_owner, d.rec_nb, d.rec_bc, d.rec_flags = p(self.rec_owner), p(self.rec_nb), p(self.rec_bc), p(self.rec_flags)
        d.rec_geom, d.cell_inc, d.cell_ghost_rec = p(self.rec_geom), p(self.cell_inc), p(self.cell_ghost_rec)
        d.bc_params = p(self.bc_params)
        d.work_r, d.work_s1, d.work_s2 = p(self.work_r), p(self.work_s1), p(self.work_s2)
        d.work_dt, d.work_part = p(self.work_dt), p(self.work_part)
        self.sst = d_wall is not None
        if self.sst:
            src, wgt = p0_record_sources(mesh)
            self.rec_src, self.rec_w = i32(src), f64(wgt)
            self.d_wall = torch.as_tensor(d_wall, dtype=dtype).reshape(C).to(dv).contiguous()
            self.work_sst = torch.zeros((C, 8), dtype=torch.float64, device=dv)
            d.rec_src, d.rec_w, d.d_wall, d.work_sst = p(self.rec_src), p(self.rec_w), p(self.d_wall), p(self.work_sst)
        self.state_ref = None if state_ref is None else np.asarray(state_ref, dtype=np.float64).reshape(5)
        if self.state_ref is not None:
            self.state_ref_dev = f64(self.state_ref)
            d.state_ref = p(self.state_ref_dev)
        self._desc = d

    def to_storage(self, U_abs):
        """Stored form (this mesh's precision) of a conserved (C, 1, 5) state."""
        X = torch.as_tensor(U_abs, device=self.device).double()
        if self.state_ref is not None:
            X = X - self.state_ref_dev
        return X.to(self.dtype).contiguous()

    def to_absolute(self, U):
        """Conserved state (FP64) of a state stored by this mesh."""
        X = U.double()
        return X if self.state_ref is None else X + self.state_ref_dev

    def _stream(self):
        return ctypes.c_void_p(torch.cuda.current_stream(self.device).cuda_stream)

    def _check(self, U):
        if U.shape != (self.n_cells, 1, 5) or U.dtype != self.dtype or not U.is_contiguous() or U.device != self.device:
            raise ValueError(f"P0 state must be a contiguous ({self.n_cells}, 1, 5) {self.dtype} tensor on {self.device}")

    def residual(self, U, mu=1.8e-5, mu_t=None, R=None):
        self._check(U)
        R = torch.empty_like(U) if R is None else R
        mt = None if mu_t is None else ctypes.c_void_p(mu_t.data_ptr())
        _lib.check(self.lib.afr_p0_residual(ctypes.byref(self._desc), ctypes.c_void_p(U.data_ptr()), mt, float(mu),
                                            ctypes.c_void_p(R.data_ptr()), self._stream()), "afr_p0_residual")
        return R

    def step(self, U, mu=1.8e-5, mu_t=None, n_vars_total=5):
        """One P0 SSP-RK3 pseudo-step of U in place; returns the (1,) FP64 device norm."""
        self._check(U)
        mt = None if mu_t is None else ctypes.c_void_p(mu_t.data_ptr())
        _lib.check(self.lib.afr_p0_ssp_rk3_step(ctypes.byref(self._desc), ctypes.c_void_p(U.data_ptr()), mt, float(mu),
                                                int(n_vars_total), ctypes.c_void_p(self.norm.data_ptr()),
                                                self._stream()), "afr_p0_ssp_rk3_step")
        return self.norm


    def sst_step(self, U, k, omega, nu_t, mu_t, mu=1.8e-5, n_vars_total=7):
        """One P0 pseudo-step with SST: k, omega, nu_t, mu_t (C,) and U updated in place; returns the
        (1,) FP64 device norm."""
        if not self.sst:
            raise ValueError("P0Mesh built without a wall distance: no SST step")
        self._check(U)
        for x in (k, omega, nu_t, mu_t):
            if x.numel() != self.n_cells or x.dtype != self.dtype or not x.is_contiguous() or x.device != self.device:
                raise ValueError(f"P0 SST fields must be contiguous ({self.n_cells},) {self.dtype} tensors")
        vp = lambda t: ctypes.c_void_p(t.data_ptr())  # noqa: E731
        _lib.check(self.lib.afr_p0_sst_step(ctypes.byref(self._desc), vp(U), vp(k), vp(omega), vp(nu_t), vp(mu_t),
                                            float(mu), int(n_vars_total), vp(self.norm), self._stream()),
                   "afr_p0_sst_step")
        return self.norm


def interpolate(X, old_order, new_order, out=None):
    """X (C, ns_old, nv) -> (C, ns_new, nv) on the device (afr_interpolate_order)."""
    lib = _lib.load()
    C, ns_old, nv = X.shape
    if ns_old != (old_order + 1) ** 3:
        raise ValueError("X does not hold the solution points of old_order")
    ns_new = (new_order + 1) ** 3
    W = torch.as_tensor(interp_matrix(old_order, new_order)).to(X.device)
    out = torch.empty((C, ns_new, nv), dtype=X.dtype, device=X.device) if out is None else out
    X = X.contiguous()
    _lib.check(lib.afr_interpolate_order(ctypes.c_void_p(X.data_ptr()), ctypes.c_void_p(out.data_ptr()),
                                         ctypes.c_void_p(W.data_ptr()), C, ns_old, ns_new, nv,
                                         0 if X.dtype == torch.float32 else 1,
                                         ctypes.c_void_p(torch.cuda.current_stream(X.device).cuda_stream)),
               "afr_interpolate_order")
    return out


def solve_with_order_continuation(make_solver, order, max_iter=1000, tol=1e-6, n_min=20, drop=100.0,
                                  history=None):
    """run_order_continuation (order_continuation.py:207-425) on the GPU.

    make_solver(p) -> GPUFRSolver of order p (the same mesh, free stream, turbulence model and
    storage order).  Returns (solver of the target order holding the final state, SolverResult-like
    dict)."""
    from .solver import _cons
    target = make_solver(order)
    sst = target.turb_model_name == "SST"
    fs = target.freestream
    dm = target.dm
    C, dv, dt = dm.n_cells, target.device, dm.dtype
    Q_free = np.array([fs["rho_inf"], fs["vel_inf"], 0.0, 0.0, fs["p_inf"]])
    d0 = None
    if sst:
        # the reference's P0 wall distance: the mean over each cell's solution points
        d0 = dm.d_wall.double().mean(dim=1).to(dt).contiguous()
    # the P0 state is stored as the target's is (offset from the same reference, or the state itself)
    p0 = P0Mesh(dm.mesh, dm.volume, dm.bc_params.cpu().numpy(), dtype=dt, device=dv, d_wall=d0,
                state_ref=None if dm.state_ref is None else dm.state_ref[0])
    U = p0.to_storage(torch.as_tensor(np.broadcast_to(_cons(Q_free), (C, 1, 5)).copy()))
    if sst:
        # turbulence fields reset to the model's defaults (order_continuation.py:245-259)
        k = torch.full((C, 1), 1e-6, dtype=dt, device=dv)
        w = torch.full((C, 1), 1.0, dtype=dt, device=dv)
        nut = torch.zeros((C, 1), dtype=dt, device=dv)
        mut = torch.zeros((C, 1), dtype=dt, device=dv)
        dwall = d0.reshape(C, 1).clone()
    n_vars = 7 if sst else 5
    orders = list(range(order + 1))
    phase_max = max_iter // len(orders)
    total, final, converged = 0, float("inf"), False
    hist = {} if history is None else history
    solver = None
    for p in orders:
        if p > 0:
            U = interpolate(U, p - 1, p)
            solver = target if p == order else make_solver(p)
            sr = solver.dm.state_ref
            if (sr is None) != (p0.state_ref is None) or (sr is not None and not np.array_equal(sr[0], p0.state_ref)):
                raise ValueError("make_solver must store the state of every order the same way (offset_storage)")
            solver.U.copy_(U)
            U = solver.U
            if sst:
                ns = (p + 1) ** 3
                k, w, nut, dwall = (interpolate(x.reshape(C, -1, 1), p - 1, p).reshape(C, ns)
                                    for x in (k, w, nut, dwall))
                solver.k.copy_(k)
                solver.omega.copy_(w)
                solver.nu_t.copy_(nut)
                rho = torch.clamp(solver.dm.to_absolute(U)[..., 0], min=1e-10).to(dt)
                solver.mu_t.copy_(nut * rho)
                solver.dm.d_wall.copy_(dwall)
                k, w, nut = solver.k, solver.omega, solver.nu_t
        phase_tol = tol * 10.0 ** (order - p)
        h = []
        for i in range(phase_max):
            if p == 0:
                n = p0.sst_step(U, k.view(C), w.view(C), nut.view(C), mut.view(C), mu=target.mu_molecular,
                                n_vars_total=n_vars) if sst else p0.step(U, mu=target.mu_molecular)
                res = float(n.item())
            else:
                res = solver.step()
            h.append(res)
            total += 1
            final = res
            if res < phase_tol:
                converged = True
                break
            if p < order and i >= n_min and h[0] > 0 and h[0] / max(res, 1e-30) >= drop:
                break
        hist[p] = h
        if p == order and converged:
            break
        converged = False
    return target, {"converged": converged, "iterations": total, "final_residual": final, "residual_history": hist}
