"""GPU: afr_residual_norm_f64_noalloc (partial sums in the library's own device array) against
afr_residual_norm_f64 (partial sums in the mesh's work_part buffer), through the C ABI.

Two calls in a row, the second on a smaller mesh, so that the second reads an array that still
holds the first call's partial sums: both must give exactly the bits of the workspace entry point,
in FP32 and FP64 storage."""

import ctypes

import numpy as np
import pytest

from cases import wall_box

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("prec", ["f32", "f64"])
def test_noalloc_norm_equals_workspace_norm(prec):
    import torch
    from autoflowcfd_amd.device import DeviceMesh
    print(f"[gpu] no-alloc norm {prec}", flush=True)
    dt = torch.float32 if prec == "f32" else torch.float64
    meshes = [wall_box(n=(12, 10, 8), seed=1), wall_box(n=(3, 2, 2), seed=2)]  # 1920, then 24 prisms
    rng = np.random.default_rng(7)
    for mesh in meshes:
        dm = DeviceMesh(mesh, 2, dtype=dt)
        R = torch.as_tensor(rng.standard_normal((mesh.n_cells, dm.ns, 5)) * 1e3, dtype=dt, device="cuda").contiguous()
        ref = torch.zeros(1, dtype=torch.float64, device="cuda")
        out = torch.full((1,), np.nan, dtype=torch.float64, device="cuda")
        lib = dm.lib
        st = dm._stream()
        for n_total in (5, 7):
            rc = lib.afr_residual_norm_f64(dm.handle, ctypes.c_void_p(R.data_ptr()), n_total,
                                           ctypes.c_void_p(ref.data_ptr()), st)
            assert rc == 0
            rc = lib.afr_residual_norm_f64_noalloc(dm.handle, ctypes.c_void_p(R.data_ptr()), n_total,
                                                   ctypes.c_void_p(out.data_ptr()), st)
            assert rc == 0
            torch.cuda.synchronize()
            a, b = float(ref.item()), float(out.item())
            host = float(np.sqrt(np.sum(R.double().cpu().numpy() ** 2) / (R.numel() / 5 * n_total)))
            assert a == b, (a, b)
            assert abs(b - host) <= 1e-12 * host


def test_noalloc_norm_rejects_bad_arguments():
    import torch
    from autoflowcfd_amd import _lib
    from autoflowcfd_amd.device import DeviceMesh
    dm = DeviceMesh(wall_box(n=(3, 2, 2), seed=2), 1, dtype=torch.float32)
    R = torch.zeros((dm.n_cells, dm.ns, 5), dtype=torch.float32, device="cuda")
    out = torch.zeros(1, dtype=torch.float64, device="cuda")
    assert dm.lib.afr_residual_norm_f64_noalloc(dm.handle, ctypes.c_void_p(R.data_ptr()), 4,
                                                ctypes.c_void_p(out.data_ptr()), dm._stream()) != 0
    buf = ctypes.create_string_buffer(256)
    _lib.load().afr_last_error(buf, 256)
    assert b"n_vars_total" in buf.value
    assert dm.lib.afr_residual_norm_f64_noalloc(None, ctypes.c_void_p(R.data_ptr()), 5,
                                                ctypes.c_void_p(out.data_ptr()), dm._stream()) != 0
