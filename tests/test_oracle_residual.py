"""Oracle residuals pinned by the reference's analytic invariants.

tests/unit/test_fr_residual_inviscid.py:110-200: free-stream preservation
(P2 max|R|/p_inf < 3e-5, P3 < 5e-3), linear shear u = 30 + y (P2 mass residual
< 1e-4); tests/unit/test_fr_viscous_flux.py:55-80: uniform flow has zero
viscous residual (max|R|/1e5 < 1e-6), finite with mu_t;
tests/unit/test_fr_gradients.py:52-80: gradients of linear functions are exact.
"""

import numpy as np
import pytest

from cases import mixed_mesh, prim_to_cons
from autoflowcfd_amd.mesh import mesh_from_cells
from oracle import fr_ops
from oracle.faces import build_flat_faces
from oracle.geometry import Geometry
from oracle import residual as R

Q_INF = np.array([1.225, 30.0, 5.0, -3.0, 101325.0])
BCP = {c: Q_INF for c in range(8)}


def _setup(order, mesh=None):
    mesh = mixed_mesh() if mesh is None else mesh
    ops = fr_ops.generate_operators(order)
    return mesh, ops, Geometry(mesh, ops), build_flat_faces(mesh, ops)


@pytest.mark.parametrize("order,tol", [(2, 3e-5), (3, 5e-3)])
def test_free_stream_preservation(order, tol):
    mesh, ops, g, fl = _setup(order)
    U = np.broadcast_to(prim_to_cons(Q_INF), (mesh.n_cells, ops.n_sps, 5)).copy()
    res = R.inviscid_residual(U, g, ops, fl, BCP)
    assert np.max(np.abs(res)) / Q_INF[4] < tol


def test_p1_finite():
    mesh, ops, g, fl = _setup(1)
    U = np.broadcast_to(prim_to_cons(Q_INF), (mesh.n_cells, ops.n_sps, 5)).copy()
    assert np.all(np.isfinite(R.inviscid_residual(U, g, ops, fl, BCP)))


def test_linear_shear_flow():
    mesh, ops, g, fl = _setup(2)
    y = g.sps_xyz[..., 1]
    Q = np.zeros(y.shape + (5,))
    Q[..., 0], Q[..., 1], Q[..., 4] = 1.225, 30.0 + y, 101325.0
    res = R.inviscid_residual(prim_to_cons(Q), g, ops, fl, BCP)
    assert np.max(np.abs(res[..., 0])) < 1e-4


@pytest.mark.parametrize("order", (2, 3))
def test_viscous_uniform_flow_zero(order):
    mesh, ops, g, fl = _setup(order)
    U = np.broadcast_to(prim_to_cons(Q_INF), (mesh.n_cells, ops.n_sps, 5)).copy()
    res = R.viscous_residual(U, g, ops, fl, BCP, 1.8e-5)
    assert np.max(np.abs(res)) / 1e5 < 1e-6
    res = R.viscous_residual(U, g, ops, fl, BCP, 1.8e-5, mu_t=np.full((mesh.n_cells, ops.n_sps), 1e-4))
    assert np.all(np.isfinite(res))


@pytest.mark.parametrize("order,tol", [(1, 1e-9), (2, 1e-9), (3, 1e-6)])
def test_linear_gradient_exact(order, tol):
    nodes = np.array([[0, 0, 0], [1, 0, 0], [0.2, 1, 0], [0.1, 0.3, 1],
                      [5, 0, 0], [6, 0, 0], [5, 1, 0], [5.3, 0.2, 1], [6.1, 0.1, 1], [5.2, 1.3, 1]], dtype=float)
    mesh = mesh_from_cells(nodes, np.array([[4, 5, 6, 7, 8, 9]]), np.array([[0, 1, 2, 3]]))
    ops = fr_ops.generate_operators(order)
    g = Geometry(mesh, ops)
    a = np.array([2.0, -3.0, 5.0])
    phi = g.sps_xyz @ a + 7.0
    grad = R.physical_gradient(phi[..., None], g, ops)[:, :, 0, :]
    assert np.max(np.abs(grad - a)) < tol
