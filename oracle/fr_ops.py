"""FP64 FR operator tables (oracle restatement).

Restates, in plain NumPy/SciPy:
  * Gauss-Legendre points         -- src/autoflowcfd/fr/quadrature_points.py:16-28
  * Jacobi recurrence / gradient  -- src/autoflowcfd/fr/collapsed_basis.py:51-107
  * collapsed triangle mode       -- src/autoflowcfd/fr/collapsed_basis.py:110-148
  * tet / prism modal bases       -- src/autoflowcfd/fr/collapsed_basis.py:151-230
  * D = Va V^-1 (LU solve)        -- src/autoflowcfd/fr/collapsed_basis.py:233-270
  * over-integration operators    -- src/autoflowcfd/fr/collapsed_basis.py:290-375
  * boundary extrapolation E      -- src/autoflowcfd/fr/collapsed_basis.py:386-470
  * exponential modal filter      -- src/autoflowcfd/fr/modal_filter.py:47-128
  * Radau correction derivatives  -- src/autoflowcfd/fr/matrix_operators.py:135-221
  * distribute mapping            -- src/autoflowcfd/core/fr_operators/face_kernels.py:68-121
                                     (with src/autoflowcfd/core/fr_residual/inviscid.py:82-89)
  * operator bundle               -- src/autoflowcfd/fr/operators.py:105-206

Ordering conventions (same as the reference): solution points and modes are
flattened ``ia*n1d^2 + ib*n1d + ic``; face flux points on the face
``axis=side`` are flattened ``(other_axes[0] outer, other_axes[1] inner)``.
"""

from dataclasses import dataclass
from typing import Dict, Tuple

import numpy as np
from scipy.linalg import lu_factor, lu_solve

OVERINTEGRATION_MAX_ORDER = 3  # src/autoflowcfd/fr/collapsed_basis.py:290
FILTER_ALPHA = -np.log(np.finfo(np.float64).eps)  # src/autoflowcfd/fr/modal_filter.py:44
FILTER_ORDER = 4  # src/autoflowcfd/fr/modal_filter.py:45


def gauss_legendre(n: int):
    """src/autoflowcfd/fr/quadrature_points.py:16-28."""
    return np.polynomial.legendre.leggauss(n)


def jacobi(x, alpha, beta, n):
    """Unnormalised Jacobi P_n^(alpha,beta)(x), three-term recurrence.

    src/autoflowcfd/fr/collapsed_basis.py:51-95.
    """
    x = np.asarray(x, dtype=np.float64)
    p0 = np.ones_like(x)
    if n == 0:
        return p0
    p1 = 0.5 * ((alpha - beta) + (alpha + beta + 2.0) * x)
    if n == 1:
        return p1
    pm1, p = p0, p1
    for k in range(1, n):
        a1 = 2.0 * (k + 1) * (k + alpha + beta + 1) * (2 * k + alpha + beta)
        a2 = (2 * k + alpha + beta + 1) * (alpha ** 2 - beta ** 2)
        a3 = (2 * k + alpha + beta) * (2 * k + alpha + beta + 1) * (2 * k + alpha + beta + 2)
        a4 = 2.0 * (k + alpha) * (k + beta) * (2 * k + alpha + beta + 2)
        pn = ((a2 + a3 * x) * p - a4 * pm1) / a1
        pm1, p = p, pn
    return p


def djacobi(x, alpha, beta, n):
    """d/dx P_n^(alpha,beta) = (n+a+b+1)/2 P_{n-1}^(a+1,b+1).

    src/autoflowcfd/fr/collapsed_basis.py:98-107.
    """
    if n == 0:
        return np.zeros_like(np.asarray(x, dtype=np.float64))
    return 0.5 * (n + alpha + beta + 1.0) * jacobi(x, alpha + 1.0, beta + 1.0, n - 1)


def _tri_mode(a, b, i, j):
    """g_ij(a,b) = P_i(a) ((1-b)/2)^i P_j^(2i+1,0)(b) and its a/b derivatives.

    src/autoflowcfd/fr/collapsed_basis.py:110-148.
    """
    f = jacobi(a, 0.0, 0.0, i)
    df = djacobi(a, 0.0, 0.0, i)
    h = (1.0 - b) / 2.0
    pj = jacobi(b, float(2 * i + 1), 0.0, j)
    dpj = djacobi(b, float(2 * i + 1), 0.0, j)
    if i == 0:
        w = np.ones_like(h)
        dw = np.zeros_like(h)
    else:
        w = h ** i
        dw = -0.5 * i * h ** (i - 1)
    g = w * pj
    dg = dw * pj + w * dpj
    return f * g, df * g, f * dg


def modal_basis(cell_type: str, a, b, c, order: int):
    """(V, Va, Vb, Vc) of the collapsed tensor basis at points (a,b,c).

    tet:   psi_ijk = g_ij(a,b) ((1-c)/2)^(i+j) P_k^(2i+2j+2,0)(c)
           -- src/autoflowcfd/fr/collapsed_basis.py:151-194
    prism: phi_ijk = g_ij(a,b) P_k(c)
           -- src/autoflowcfd/fr/collapsed_basis.py:197-230
    """
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    c = np.asarray(c, dtype=np.float64)
    n1d = order + 1
    nm = n1d ** 3
    npt = a.shape[0]
    V = np.zeros((npt, nm))
    Va = np.zeros((npt, nm))
    Vb = np.zeros((npt, nm))
    Vc = np.zeros((npt, nm))
    hc = (1.0 - c) / 2.0
    for i in range(n1d):
        for j in range(n1d):
            g, ga, gb = _tri_mode(a, b, i, j)
            if cell_type == "tet":
                p = i + j
                if p == 0:
                    w = np.ones_like(hc)
                    dw = np.zeros_like(hc)
                else:
                    w = hc ** p
                    dw = -0.5 * p * hc ** (p - 1)
            for k in range(n1d):
                flat = i * n1d * n1d + j * n1d + k
                if cell_type == "tet":
                    pk = jacobi(c, float(2 * i + 2 * j + 2), 0.0, k)
                    dpk = djacobi(c, float(2 * i + 2 * j + 2), 0.0, k)
                    hk = w * pk
                    dhk = dw * pk + w * dpk
                elif cell_type == "prism":
                    hk = jacobi(c, 0.0, 0.0, k)
                    dhk = djacobi(c, 0.0, 0.0, k)
                else:
                    raise ValueError(cell_type)
                V[:, flat] = g * hk
                Va[:, flat] = ga * hk
                Vb[:, flat] = gb * hk
                Vc[:, flat] = g * dhk
    return V, Va, Vb, Vc


def tensor_points(n1d: int) -> np.ndarray:
    """Tensor GL points, flattened ia*n1d^2+ib*n1d+ic (operators.py:149-150)."""
    x, _ = gauss_legendre(n1d)
    aa, bb, cc = np.meshgrid(x, x, x, indexing="ij")
    return np.column_stack([aa.ravel(), bb.ravel(), cc.ravel()])


def _right_solve(V_at_pts, V_nodes):
    """X = V_at_pts @ V_nodes^{-1} via LU of V_nodes^T (collapsed_basis.py:264-269)."""
    lu = lu_factor(V_nodes.T)
    return lu_solve(lu, V_at_pts.T).T


def diff_matrices(cell_type: str, order: int, ref_pts: np.ndarray) -> np.ndarray:
    """D[:, :, m] = d/dxi_m, shape (n, n, 3).  collapsed_basis.py:233-270."""
    V, Va, Vb, Vc = modal_basis(cell_type, ref_pts[:, 0], ref_pts[:, 1], ref_pts[:, 2], order)
    lu = lu_factor(V.T)
    Da = lu_solve(lu, Va.T).T
    Db = lu_solve(lu, Vb.T).T
    Dc = lu_solve(lu, Vc.T).T
    return np.stack([Da, Db, Dc], axis=-1)


def face_ref_grid(n1d: int, axis: int, side: float) -> np.ndarray:
    """FP grid of face ``axis=side``: src/autoflowcfd/fr/face_flux_points.py:154-161."""
    x, _ = gauss_legendre(n1d)
    other = [a for a in range(3) if a != axis]
    g1, g2 = np.meshgrid(x, x, indexing="ij")
    pts = np.zeros((n1d * n1d, 3))
    pts[:, axis] = side
    pts[:, other[0]] = g1.ravel()
    pts[:, other[1]] = g2.ravel()
    return pts


def boundary_extrap(cell_type: str, order: int, axis: int, side: float) -> np.ndarray:
    """E (n_fp, n_sps): SP values -> FP values on face axis=side.

    src/autoflowcfd/fr/collapsed_basis.py:386-470.
    """
    n1d = order + 1
    sps = tensor_points(n1d)
    fp = face_ref_grid(n1d, axis, side)
    Vs = modal_basis(cell_type, sps[:, 0], sps[:, 1], sps[:, 2], order)[0]
    Vf = modal_basis(cell_type, fp[:, 0], fp[:, 1], fp[:, 2], order)[0]
    return _right_solve(Vf, Vs)


def overintegration(cell_type: str, order: int):
    """(ref_fine, c2f, D_fine, f2c) -- src/autoflowcfd/fr/collapsed_basis.py:293-375."""
    over = min(2 * order, OVERINTEGRATION_MAX_ORDER)
    nf1 = over + 1
    ref_c = tensor_points(order + 1)
    ref_f = tensor_points(nf1)
    Vc_c = modal_basis(cell_type, ref_c[:, 0], ref_c[:, 1], ref_c[:, 2], order)[0]
    Vf_f = modal_basis(cell_type, ref_f[:, 0], ref_f[:, 1], ref_f[:, 2], over)[0]
    Vc_f = modal_basis(cell_type, ref_f[:, 0], ref_f[:, 1], ref_f[:, 2], order)[0]
    Vf_c = modal_basis(cell_type, ref_c[:, 0], ref_c[:, 1], ref_c[:, 2], over)[0]
    c2f = _right_solve(Vc_f, Vc_c)
    D_fine = diff_matrices(cell_type, over, ref_f)
    f2c = _right_solve(Vf_c, Vf_f)
    return ref_f, c2f, D_fine, f2c


def modal_filter(cell_type: str, order: int) -> np.ndarray:
    """V diag(sigma) V^{-1}, sigma = exp(-alpha*(max(i,j,k)/P)^8).

    src/autoflowcfd/fr/modal_filter.py:47-128 (explicit inverse, as there).
    """
    n1d = order + 1
    if order == 0:
        return np.eye(1)
    sps = tensor_points(n1d)
    V = modal_basis(cell_type, sps[:, 0], sps[:, 1], sps[:, 2], order)[0]
    sig = np.zeros(n1d ** 3)
    for i in range(n1d):
        for j in range(n1d):
            for k in range(n1d):
                eta = max(i, j, k) / order
                sig[i * n1d * n1d + j * n1d + k] = np.exp(-FILTER_ALPHA * eta ** (2 * FILTER_ORDER))
    return V @ np.diag(sig) @ np.linalg.inv(V)


def _radau_coeffs(n: int, right: bool) -> np.ndarray:
    """Monomial coefficients of the degree-n Radau/VCJH correction polynomial.

    src/autoflowcfd/fr/matrix_operators.py:135-176.
    """
    def moment(j):
        return 0.0 if j % 2 == 1 else 2.0 / (j + 1)

    A, rhs = [], []
    if right:
        A.append([1.0] * (n + 1)); rhs.append(1.0)
        A.append([(-1.0) ** k for k in range(n + 1)]); rhs.append(0.0)
    else:
        A.append([(-1.0) ** k for k in range(n + 1)]); rhs.append(1.0)
        A.append([1.0] * (n + 1)); rhs.append(0.0)
    for m in range(n - 1):
        A.append([moment(m + k) for k in range(n + 1)]); rhs.append(0.0)
    return np.linalg.solve(np.array(A), np.array(rhs))


def correction_derivatives(n: int) -> Tuple[np.ndarray, np.ndarray]:
    """(g_L'(x_i), g_R'(x_i)) at the n GL points.  matrix_operators.py:179-221."""
    x, _ = gauss_legendre(n)
    cl = _radau_coeffs(n, right=False)
    cr = _radau_coeffs(n, right=True)
    dl = np.array([cl[k] * k for k in range(1, n + 1)])
    dr = np.array([cr[k] * k for k in range(1, n + 1)])

    def pv(cf, xx):
        out = np.zeros_like(xx)
        for k, cc in enumerate(cf):
            out = out + cc * xx ** k
        return out

    return pv(dl, x), pv(dr, x)


def distribute_from_face(fp_data: np.ndarray, n1d: int, axis: int, g_prime: np.ndarray) -> np.ndarray:
    """(n1d^2, ...) face data -> (n1d^3, ...) SP contributions, outer product with g'.

    src/autoflowcfd/core/fr_residual/inviscid.py:82-89.
    """
    trailing = fp_data.shape[1:]
    grid = fp_data.reshape((n1d, n1d) + trailing)
    expanded = np.tensordot(g_prime, grid, axes=0)
    return np.moveaxis(expanded, 0, axis).reshape((n1d ** 3,) + trailing)


def distribute_mapping(n1d: int):
    """Index form of distribute_from_face: (fp_of_sp (3,n_sps), coord_of_sp (3,n_sps)).

    src/autoflowcfd/core/fr_operators/face_kernels.py:68-121 (probe derivation).
    """
    nfp, nsp = n1d * n1d, n1d ** 3
    fp_of = np.full((3, nsp), -1, dtype=np.int64)
    co_of = np.full((3, nsp), -1, dtype=np.int64)
    for axis in range(3):
        for i in range(nfp):
            for p in range(n1d):
                fpd = np.zeros((nfp, 1)); fpd[i, 0] = 1.0
                g = np.zeros(n1d); g[p] = 1.0
                r = distribute_from_face(fpd, n1d, axis, g)[:, 0]
                hit = np.flatnonzero(np.abs(r - 1.0) < 1e-12)
                assert len(hit) == 1
                fp_of[axis, hit[0]] = i
                co_of[axis, hit[0]] = p
    return fp_of, co_of


@dataclass
class Operators:
    """Bundle of per-order tables; celltype index 0 = prism, 1 = tet."""
    order: int
    n1d: int
    n_sps: int
    n_fp: int
    n_fine: int
    D: np.ndarray          # (2, n_sps, n_sps, 3)
    E: np.ndarray          # (2, 3, 2, n_fp, n_sps)  [celltype, axis, side_idx]
    c2f: np.ndarray        # (2, n_fine, n_sps)
    D_fine: np.ndarray     # (2, n_fine, n_fine, 3)
    f2c: np.ndarray        # (2, n_sps, n_fine)
    filt: np.ndarray       # (2, n_sps, n_sps)
    g_left: np.ndarray     # (n1d,)
    g_right: np.ndarray    # (n1d,)
    fp_of_sp: np.ndarray   # (3, n_sps)
    coord_of_sp: np.ndarray  # (3, n_sps)
    ref_sps: np.ndarray    # (n_sps, 3)
    ref_fine: np.ndarray   # (n_fine, 3)
    w_sps: np.ndarray      # (n_sps,) tensor GL weights


_CACHE: Dict[int, Operators] = {}


def generate_operators(order: int) -> Operators:
    """src/autoflowcfd/fr/operators.py:105-206.  Order 0 gives the one-point operators the order
    continuation runs its P0 phase's turbulence step with (order_continuation.py:241,276:
    generate_fr_operators(0)): zero derivatives, unit extrapolation and filter, g' = -1/2, +1/2;
    the P0 mean-flow residual is the finite-volume one (oracle.continuation)."""
    if order in _CACHE:
        return _CACHE[order]
    if order < 0:
        raise ValueError("negative order")
    n1d = order + 1
    ref = tensor_points(n1d)
    types = ("prism", "tet")
    D = np.stack([diff_matrices(t, order, ref) for t in types])
    E = np.zeros((2, 3, 2, n1d * n1d, n1d ** 3))
    for ti, t in enumerate(types):
        for axis in range(3):
            for si, side in enumerate((-1.0, 1.0)):
                E[ti, axis, si] = boundary_extrap(t, order, axis, side)
    oi = [overintegration(t, order) for t in types]
    gl, gr = correction_derivatives(n1d)
    fp_of, co_of = distribute_mapping(n1d)
    _, w = gauss_legendre(n1d)
    wx, wy, wz = np.meshgrid(w, w, w, indexing="ij")
    ops = Operators(
        order=order, n1d=n1d, n_sps=n1d ** 3, n_fp=n1d * n1d, n_fine=oi[0][0].shape[0],
        D=D, E=E,
        c2f=np.stack([o[1] for o in oi]), D_fine=np.stack([o[2] for o in oi]),
        f2c=np.stack([o[3] for o in oi]),
        filt=np.stack([modal_filter(t, order) for t in types]),
        g_left=gl, g_right=gr, fp_of_sp=fp_of, coord_of_sp=co_of,
        ref_sps=ref, ref_fine=oi[0][0], w_sps=(wx * wy * wz).ravel(),
    )
    _CACHE[order] = ops
    return ops
