"""GPU FP64 metric precompute vs the host FP64 closed forms (and the oracle geometry)."""

import numpy as np
import pytest

from cases import mixed_mesh, tet_box, wall_box

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("name,make,order", [("mixed", mixed_mesh, 2), ("wallbox", lambda: wall_box(seed=1), 2),
                                             ("tetbox", lambda: tet_box(3, seed=2), 1)])
def test_device_metrics_match_host(name, make, order):
    import torch
    from autoflowcfd_amd.device import DeviceMesh
    from oracle import fr_ops
    from oracle.geometry import Geometry
    print(f"[gpu] metrics {name} P{order}", flush=True)
    mesh = make()
    g = DeviceMesh(mesh, order, dtype=torch.float64, metrics="gpu")
    h = DeviceMesh(mesh, order, dtype=torch.float64, metrics="host")
    for k in ("det", "adj", "adj_fine", "metric_flux_scale", "slot_adjrow"):
        a, b = getattr(g, k).cpu().numpy(), getattr(h, k).cpu().numpy()
        assert np.allclose(a, b, rtol=1e-12, atol=1e-12 * np.abs(b).max()), k
    # the GCL defect is a near-cancelling sum: compare on the scale of its terms
    term_scale = np.abs(g.T.Op).sum(axis=-1).max() * np.abs(h.adj_fine.cpu().numpy()).max()
    assert np.abs(g.gcl.cpu().numpy() - h.gcl.cpu().numpy()).max() < 1e-13 * term_scale
    assert np.array_equal(g.slot_fb.cpu().numpy(), h.slot_fb.cpu().numpy())
    assert np.allclose(g.volume, h.volume, rtol=1e-13)
    o = Geometry(mesh, fr_ops.generate_operators(order))
    assert np.allclose(g.det.cpu().numpy(), o.det, rtol=1e-12)
    assert np.allclose(g.volume, o.volume, rtol=1e-12)
