"""FP64 FR operator tables for the device path (host-side set-up).

Same operator definitions as the reference's
src/autoflowcfd/fr/operators.py:105-206 (collapsed Karniadakis-Sherwin tensor
bases of src/autoflowcfd/fr/collapsed_basis.py:51-470, Radau correction
derivatives of src/autoflowcfd/fr/matrix_operators.py:135-221, exponential
modal filter of src/autoflowcfd/fr/modal_filter.py:47-128).  The device path
additionally folds the over-integration restriction into the fine-point
divergence, Op = f2c @ [D_fine_a | D_fine_b | D_fine_c] (exact by
associativity), so the volume term is two shared-operator GEMMs:
Q_fine = c2f @ Q and div = Op @ F~_fine.
"""

from dataclasses import dataclass

import threading

import numpy as np
from scipy.linalg import lu_factor, lu_solve

OVER_MAX = 3            # src/autoflowcfd/fr/collapsed_basis.py:290
_ALPHA = -np.log(np.finfo(np.float64).eps)   # modal_filter.py:44


def gl(n):
    return np.polynomial.legendre.leggauss(n)


def _jac(x, a, b, n):
    p0 = np.ones_like(x)
    if n == 0:
        return p0
    p1 = 0.5 * ((a - b) + (a + b + 2.0) * x)
    for k in range(1, n):
        c1 = 2.0 * (k + 1) * (k + a + b + 1) * (2 * k + a + b)
        c2 = (2 * k + a + b + 1) * (a * a - b * b)
        c3 = (2 * k + a + b) * (2 * k + a + b + 1) * (2 * k + a + b + 2)
        c4 = 2.0 * (k + a) * (k + b) * (2 * k + a + b + 2)
        p0, p1 = p1, ((c2 + c3 * x) * p1 - c4 * p0) / c1
    return p1


def _djac(x, a, b, n):
    return np.zeros_like(x) if n == 0 else 0.5 * (n + a + b + 1.0) * _jac(x, a + 1.0, b + 1.0, n - 1)


def basis(ctype, pts, order, deriv=True):
    """Collapsed modal basis (V, dV/da, dV/db, dV/dc) at pts (n,3); modes i*n1^2+j*n1+k."""
    a, b, c = pts[:, 0], pts[:, 1], pts[:, 2]
    n1 = order + 1
    out = np.zeros((4, len(a), n1 ** 3))
    hb = 0.5 * (1.0 - b)
    hc = 0.5 * (1.0 - c)
    for i in range(n1):
        fa, dfa = _jac(a, 0.0, 0.0, i), _djac(a, 0.0, 0.0, i)
        wb = hb ** i if i else np.ones_like(b)
        dwb = -0.5 * i * hb ** (i - 1) if i else np.zeros_like(b)
        for j in range(n1):
            pj, dpj = _jac(b, 2.0 * i + 1, 0.0, j), _djac(b, 2.0 * i + 1, 0.0, j)
            g, ga, gb = fa * wb * pj, dfa * wb * pj, fa * (dwb * pj + wb * dpj)
            for k in range(n1):
                if ctype == 1:  # tet
                    p = i + j
                    wc = hc ** p if p else np.ones_like(c)
                    dwc = -0.5 * p * hc ** (p - 1) if p else np.zeros_like(c)
                    pk, dpk = _jac(c, 2.0 * p + 2, 0.0, k), _djac(c, 2.0 * p + 2, 0.0, k)
                    h, dh = wc * pk, dwc * pk + wc * dpk
                else:           # prism
                    h, dh = _jac(c, 0.0, 0.0, k), _djac(c, 0.0, 0.0, k)
                m = (i * n1 + j) * n1 + k
                out[0, :, m] = g * h
                out[1, :, m] = ga * h
                out[2, :, m] = gb * h
                out[3, :, m] = g * dh
    return out


def tensor_pts(n1):
    x, _ = gl(n1)
    A, B, C = np.meshgrid(x, x, x, indexing="ij")
    return np.stack([A.ravel(), B.ravel(), C.ravel()], axis=1)


def face_pts(n1, axis, side):
    x, _ = gl(n1)
    other = [d for d in range(3) if d != axis]
    G1, G2 = np.meshgrid(x, x, indexing="ij")
    p = np.zeros((n1 * n1, 3))
    p[:, axis] = side
    p[:, other[0]] = G1.ravel()
    p[:, other[1]] = G2.ravel()
    return p


class NodalInterp:
    """X @ V_nodes^{-1} via a cached LU of V_nodes^T (collapsed_basis.py:264-269)."""

    def __init__(self, ctype, order):
        self.ctype, self.order = ctype, order
        self.nodes = tensor_pts(order + 1)
        self.lu = lu_factor(basis(ctype, self.nodes, order)[0].T)

    def rows(self, V_at):
        return lu_solve(self.lu, V_at.T).T

    def at(self, pts):
        return self.rows(basis(self.ctype, pts, self.order)[0])


def radau_gprime(n):
    """(g_L', g_R') at the n GL points, matrix_operators.py:135-221."""
    mom = [0.0 if j % 2 else 2.0 / (j + 1) for j in range(2 * n + 2)]
    out = []
    for right in (False, True):
        A = [[(-1.0) ** k for k in range(n + 1)], [1.0] * (n + 1)]
        rhs = [0.0, 1.0] if right else [1.0, 0.0]
        for m in range(n - 1):
            A.append([mom[m + k] for k in range(n + 1)])
            rhs.append(0.0)
        coef = np.linalg.solve(np.array(A), np.array(rhs))
        x, _ = gl(n)
        out.append(sum(k * coef[k] * x ** (k - 1) for k in range(1, n + 1)))
    return out[0], out[1]


@dataclass
class Tables:
    order: int
    n1: int
    ns: int
    nf: int      # fine points
    nfp: int
    D: np.ndarray        # (2, ns, ns, 3)
    E: np.ndarray        # (2, 3, 2, nfp, ns)
    c2f: np.ndarray      # (2, nf, ns)
    Dfine: np.ndarray    # (2, nf, nf, 3)
    f2c: np.ndarray      # (2, ns, nf)
    Op: np.ndarray       # (2, ns, 3*nf): f2c @ D_fine, column index m*nf + j
    filt: np.ndarray     # (2, ns, ns)
    gL: np.ndarray
    gR: np.ndarray
    fp_of_sp: np.ndarray     # (3, ns)
    coord_of_sp: np.ndarray  # (3, ns)
    w_sps: np.ndarray
    sps: np.ndarray
    fine: np.ndarray
    interp: tuple            # NodalInterp per cell type


_CACHE = {}
# Host work that calls LAPACK (the table generation, the location of flux points) runs under this
# lock: solves set up on several host threads would otherwise enter the BLAS library at once.
HOST_LOCK = threading.RLock()


def tables(order: int) -> Tables:
    if order in _CACHE:
        return _CACHE[order]
    with HOST_LOCK:
        if order not in _CACHE:
            _CACHE[order] = _make_tables(order)
    return _CACHE[order]


def _make_tables(order: int) -> Tables:
    n1 = order + 1
    sps = tensor_pts(n1)
    over = min(2 * order, OVER_MAX)
    fine = tensor_pts(over + 1)
    D = np.zeros((2, n1 ** 3, n1 ** 3, 3))
    E = np.zeros((2, 3, 2, n1 * n1, n1 ** 3))
    c2f = np.zeros((2, len(fine), n1 ** 3))
    Dfine = np.zeros((2, len(fine), len(fine), 3))
    f2c = np.zeros((2, n1 ** 3, len(fine)))
    filt = np.zeros((2, n1 ** 3, n1 ** 3))
    interps = []
    for ct in (0, 1):
        ni = NodalInterp(ct, order)
        interps.append(ni)
        Bs = basis(ct, sps, order)
        for m in range(3):
            D[ct, :, :, m] = ni.rows(Bs[m + 1])
        for ax in range(3):
            for si, sd in enumerate((-1.0, 1.0)):
                E[ct, ax, si] = ni.at(face_pts(n1, ax, sd))
        c2f[ct] = ni.at(fine)
        nf_i = NodalInterp(ct, over)
        Bf = basis(ct, fine, over)
        for m in range(3):
            Dfine[ct, :, :, m] = nf_i.rows(Bf[m + 1])
        f2c[ct] = nf_i.at(sps)
        sig = np.array([np.exp(-_ALPHA * (max(i, j, k) / order) ** 8)
                        for i in range(n1) for j in range(n1) for k in range(n1)])
        V = Bs[0]
        filt[ct] = V @ np.diag(sig) @ np.linalg.inv(V)
    Op = np.stack([np.concatenate([f2c[ct] @ Dfine[ct, :, :, m] for m in range(3)], axis=1) for ct in (0, 1)])
    gL, gR = radau_gprime(n1)
    # distribute mapping: SP s on axis gets g'[coord along axis] * face value at the other two coords
    fp_of = np.zeros((3, n1 ** 3), dtype=np.int64)
    co_of = np.zeros((3, n1 ** 3), dtype=np.int64)
    for s in range(n1 ** 3):
        ijk = (s // (n1 * n1), (s // n1) % n1, s % n1)
        for ax in range(3):
            other = [d for d in range(3) if d != ax]
            fp_of[ax, s] = ijk[other[0]] * n1 + ijk[other[1]]
            co_of[ax, s] = ijk[ax]
    _, w = gl(n1)
    W = (w[:, None, None] * w[None, :, None] * w[None, None, :]).ravel()
    t = Tables(order=order, n1=n1, ns=n1 ** 3, nf=len(fine), nfp=n1 * n1, D=D, E=E, c2f=c2f, Dfine=Dfine,
               f2c=f2c, Op=Op, filt=filt, gL=gL, gR=gR, fp_of_sp=fp_of, coord_of_sp=co_of, w_sps=W,
               sps=sps, fine=fine, interp=tuple(interps))
    return t


def _tri_basis(pts_ab, order, deriv=False):
    """Collapsed triangle modes g_ij(a,b) (i,j <= order) at (a,b) points: (V, Va, Vb) (n, n1^2)."""
    a, b = pts_ab[:, 0], pts_ab[:, 1]
    n1 = order + 1
    V, Va, Vb = (np.zeros((len(a), n1 * n1)) for _ in range(3))
    hb = 0.5 * (1.0 - b)
    for i in range(n1):
        fa, dfa = _jac(a, 0.0, 0.0, i), _djac(a, 0.0, 0.0, i)
        wb = hb ** i if i else np.ones_like(b)
        dwb = -0.5 * i * hb ** (i - 1) if i else np.zeros_like(b)
        for j in range(n1):
            pj, dpj = _jac(b, 2.0 * i + 1, 0.0, j), _djac(b, 2.0 * i + 1, 0.0, j)
            V[:, i * n1 + j] = fa * wb * pj
            Va[:, i * n1 + j] = dfa * wb * pj
            Vb[:, i * n1 + j] = fa * (dwb * pj + wb * dpj)
    return V, Va, Vb


def _line_basis(x, order):
    n1 = order + 1
    return (np.stack([_jac(x, 0.0, 0.0, k) for k in range(n1)], 1),
            np.stack([_djac(x, 0.0, 0.0, k) for k in range(n1)], 1))


def _rs(V_at, V_nodes):
    return np.linalg.solve(V_nodes.T, V_at.T).T


@dataclass
class PrismFactors:
    """Prism operators in (a,b)-triangle x c-line factored form (FP64)."""
    c2f_ab: np.ndarray   # (nabf, nab)
    c2f_c: np.ndarray    # (nf1, n1)
    A_a: np.ndarray      # (nab, nabf) = f2c_ab @ D_ab,a (fine)
    A_b: np.ndarray      # (nab, nabf)
    B_c: np.ndarray      # (n1, nf1)   = f2c_c @ D_c (fine)
    f2c_c: np.ndarray    # (n1, nf1)
    f2c_ab: np.ndarray   # (nab, nabf)
    E_ab: np.ndarray     # (3, n1, nab): quad slots a=-1, a=+1, b=-1; row = free (a|b) index
    e_c: np.ndarray      # (2, n1): c=-1, c=+1
    D_ab: np.ndarray     # (2, nab, nab) coarse d/da, d/db on the triangle
    D_c: np.ndarray      # (n1, n1) coarse d/dc

    def packed(self):
        return np.concatenate([x.ravel() for x in (self.c2f_ab, self.c2f_c, self.A_a, self.A_b, self.B_c,
                                                   self.f2c_c, self.f2c_ab, self.E_ab, self.e_c, self.D_ab,
                                                   self.D_c)])


def prism_factors(order: int) -> PrismFactors:
    n1 = order + 1
    over = min(2 * order, OVER_MAX)
    nf1 = over + 1
    x, _ = gl(n1)
    xf, _ = gl(nf1)
    A, B = np.meshgrid(x, x, indexing="ij")
    ab = np.stack([A.ravel(), B.ravel()], 1)
    Af, Bf = np.meshgrid(xf, xf, indexing="ij")
    abf = np.stack([Af.ravel(), Bf.ravel()], 1)
    Vab = _tri_basis(ab, order)[0]
    Vab_f_P = _tri_basis(abf, order)[0]
    Vo_f, Vo_fa, Vo_fb = _tri_basis(abf, over)
    Vo_c = _tri_basis(ab, over)[0]
    Vc = _line_basis(x, order)[0]
    Vc_fP = _line_basis(xf, order)[0]
    Vco_f, dVco_f = _line_basis(xf, over)
    Vco_c = _line_basis(x, over)[0]
    c2f_ab = _rs(Vab_f_P, Vab)
    c2f_c = _rs(Vc_fP, Vc)
    f2c_ab = _rs(Vo_c, Vo_f)
    f2c_c = _rs(Vco_c, Vco_f)
    Dfa, Dfb = _rs(Vo_fa, Vo_f), _rs(Vo_fb, Vo_f)
    Dfc = _rs(dVco_f, Vco_f)
    E_ab = np.zeros((3, n1, n1 * n1))
    for si, (fix, val) in enumerate(((0, -1.0), (0, 1.0), (1, -1.0))):
        p = np.zeros((n1, 2))
        p[:, fix] = val
        p[:, 1 - fix] = x
        E_ab[si] = _rs(_tri_basis(p, order)[0], Vab)
    e_c = np.stack([_rs(_line_basis(np.array([s]), order)[0], Vc)[0] for s in (-1.0, 1.0)])
    Vab_c, Vab_a, Vab_b = _tri_basis(ab, order)
    Vcc, dVcc = _line_basis(x, order)
    return PrismFactors(c2f_ab=c2f_ab, c2f_c=c2f_c, A_a=f2c_ab @ Dfa, A_b=f2c_ab @ Dfb, B_c=f2c_c @ Dfc,
                        f2c_c=f2c_c, f2c_ab=f2c_ab, E_ab=E_ab, e_c=e_c,
                        D_ab=np.stack([_rs(Vab_a, Vab_c), _rs(Vab_b, Vab_c)]), D_c=_rs(dVcc, Vcc))
