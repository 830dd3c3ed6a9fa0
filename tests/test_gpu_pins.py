"""The device residuals (FP64) against the independent per-cell identities of tests/test_oracle_pins.py.

These check the HIP kernels directly against properties that do not go through the oracle: for a
cell-wise constant state, every cell's integrated inviscid residual equals minus the net AUSM+up
outflow through its faces computed from the node coordinates; for cell-wise linear fields, every
cell's integrated viscous residual equals the BR1 flux plus the boundary penalty integrated with a
physical face quadrature.  Same meshes and tolerances as the oracle pins.
"""

import numpy as np
import pytest

from test_oracle_pins import _bc_params, _setup, inviscid_identity_error, viscous_identity_error

pytestmark = pytest.mark.gpu


def _device(mesh, order):
    import torch
    from autoflowcfd_amd.device import DeviceMesh
    table = np.zeros((8, 5))
    for code, q in _bc_params().items():
        table[code] = q
    return DeviceMesh(mesh, order, dtype=torch.float64, bc_params=table)


@pytest.mark.parametrize("kind,order,tol", [
    ("prism", 1, 1e-12), ("prism", 2, 1e-12), ("prism-rotated", 1, 1e-12), ("prism-rotated", 2, 1e-12),
    ("tet", 2, 1e-7), ("tet", 3, 1e-6), ("mixed", 2, 1e-6)])
def test_device_inviscid_cell_conservation(kind, order, tol):
    import torch
    print(f"[gpu] inviscid identity {kind} P{order}", flush=True)
    mesh, classify, ops, g, fl = _setup(kind, order)
    dm = _device(mesh, order)

    def res(U, _bcp):
        Ud = torch.as_tensor(U, dtype=torch.float64, device="cuda").contiguous()
        return dm.inviscid_residual(Ud, suppress=False).cpu().numpy()
    err = inviscid_identity_error(mesh, classify, ops, g, fl, residual=res)
    assert err < tol, f"{err:.3e}"


@pytest.mark.parametrize("kind,order,tol", [
    ("prism", 1, 1e-10), ("prism", 2, 1e-10), ("prism-rotated", 1, 1e-10), ("prism-rotated", 2, 1e-10),
    ("tet", 3, 1e-4)])
def test_device_viscous_cell_conservation(kind, order, tol):
    import torch
    print(f"[gpu] viscous identity {kind} P{order}", flush=True)
    mesh, classify, ops, g, fl = _setup(kind, order)
    dm = _device(mesh, order)

    def res(U, _bcp, mu):
        Ud = torch.as_tensor(U, dtype=torch.float64, device="cuda").contiguous()
        return dm.viscous_residual(Ud, mu=mu, suppress=False).cpu().numpy()
    err = viscous_identity_error(mesh, classify, ops, g, fl, residual=res)
    assert err < tol, f"{err:.3e}"
