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

Laminar path (turbulence model NONE); the P0 phase of the SST transport is not built, so an SST
solve raises instead of silently starting at P1.  All orders share the locality cell order the
solvers store, so the interpolation is one kernel on resident arrays.
"""

import ctypes

import numpy as np
import torch

from . import _lib
from .device import DeviceMesh, record_incidence
from .geometry import det_adj, jacobian
from .mesh import CUBE_FACE_AXIS, CUBE_FACE_SIDE

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


class P0Mesh:
    """Device arrays of the P0 phase on one mesh (in the order of ``mesh``)."""

    def __init__(self, mesh, cell_volume, bc_params, dtype=torch.float32, device="cuda"):
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
        flags = primary.astype(np.int64) | (axis << 1) | (side.astype(np.int64) << 3)
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
        self._desc = d

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

    make_solver(p) -> GPUFRSolver of order p (the same mesh, free stream and storage order).
    Returns (solver of the target order holding the final state, SolverResult-like dict)."""
    from .solver import _cons
    target = make_solver(order)
    if target.turb_model_name != "NONE":
        raise NotImplementedError("order continuation: the P0 phase of the SST transport is not built")
    fs = target.freestream
    C = target.dm.n_cells
    Q_free = np.array([fs["rho_inf"], fs["vel_inf"], 0.0, 0.0, fs["p_inf"]])
    p0 = P0Mesh(target.dm.mesh, target.dm.volume, target.dm.bc_params.cpu().numpy(), dtype=target.dm.dtype,
                device=target.device)
    U = torch.as_tensor(np.broadcast_to(_cons(Q_free), (C, 1, 5)).copy(), dtype=target.dm.dtype).to(target.device)
    orders = list(range(order + 1))
    phase_max = max_iter // len(orders)
    total, final, converged = 0, float("inf"), False
    hist = {} if history is None else history
    solver = None
    for p in orders:
        if p > 0:
            U = interpolate(U, p - 1, p)
            solver = target if p == order else make_solver(p)
            solver.U.copy_(U)
            U = solver.U
        phase_tol = tol * 10.0 ** (order - p)
        h = []
        for i in range(phase_max):
            res = float(p0.step(U, mu=target.mu_molecular).item()) if p == 0 else solver.step()
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
