"""GPU: the kernels' tile variants (afr_mesh_set_variant) compute the same values bit for bit,
so the autotuner may pick any of them per mesh without changing a result (and a batched case
stays bit-identical to itself alone, whatever each mesh picked).

* prism inviscid kernel, slot sources staged in LDS or read from global memory: the residual at
  P1 / P2, FP32 / FP64, on a graded wall box and on rotated prisms (alignment fallback);
* SST cell kernel, two prisms or one cell per wavefront: a whole SST pseudo-step (k, omega, nu_t,
  mu_t, U and the FP64 norm) at P1 / P2, FP32 / FP64, and on the odd 23-prism mesh (a lone last
  prism in the paired variant);
* the autotuner over (variant, grid) leaves results unchanged and reports its choice."""

import numpy as np
import pytest

from cases import freestream, odd_prism_mesh, rotated_prism_box, smooth_state, wall_box

pytestmark = pytest.mark.gpu


def _dev(mesh, order, dtype, turbulence=None):
    import torch
    from autoflowcfd_amd.device import DeviceMesh, bc_params_table
    from autoflowcfd_amd.fr_tables import tables
    from autoflowcfd_amd.geometry import map_ref
    Q0 = freestream(0.15, 4.0)
    dm = DeviceMesh(mesh, order, dtype=dtype, bc_params=bc_params_table(Q0), turbulence=turbulence)
    xyz = np.zeros((mesh.n_cells, tables(order).ns, 3))
    if mesh.n_prism:
        xyz[:mesh.n_prism] = map_ref(True, tables(order).sps, mesh.nodes[mesh.prism_conn])
    if mesh.n_tet:
        xyz[mesh.n_prism:] = map_ref(False, tables(order).sps, mesh.nodes[mesh.tet_conn])
    U = torch.as_tensor(smooth_state(xyz, Q0, amp=2e-2, seed=3), dtype=dtype, device="cuda").contiguous()
    return dm, U, xyz


@pytest.mark.parametrize("order", [1, 2])
@pytest.mark.parametrize("prec", ["f32", "f64"])
@pytest.mark.parametrize("kind", ["wallbox", "rotated"])
def test_prism_inviscid_variants_are_bit_identical(kind, prec, order):
    import torch
    from autoflowcfd_amd.autotune import set_variant
    print(f"[gpu] prism inviscid variants {kind} {prec} P{order}", flush=True)
    dtype = torch.float32 if prec == "f32" else torch.float64
    mesh = wall_box(seed=4) if kind == "wallbox" else rotated_prism_box(seed=2)
    dm, U, _ = _dev(mesh, order, dtype)
    mut = torch.full(U.shape[:2], 1e-4, dtype=dtype, device="cuda")
    out = []
    for v in (0, 1):
        set_variant(dm, "prism_inviscid", v)
        out.append(dm.residual(U, mut, torch.empty_like(U)).clone())
    assert torch.equal(out[0], out[1])
    assert float(out[0].abs().max()) > 0


@pytest.mark.parametrize("order", [1, 2])
@pytest.mark.parametrize("prec", ["f32", "f64"])
@pytest.mark.parametrize("kind", ["wallbox", "odd"])
def test_sst_cell_variants_are_bit_identical(kind, prec, order):
    import torch
    from autoflowcfd_amd.autotune import set_variant
    print(f"[gpu] SST cell variants {kind} {prec} P{order}", flush=True)
    dtype = torch.float32 if prec == "f32" else torch.float64
    mesh = wall_box(seed=5) if kind == "wallbox" else odd_prism_mesh()
    dm, U0, xyz = _dev(mesh, order, dtype, turbulence="SST")
    x = xyz / np.maximum(np.ptp(xyz.reshape(-1, 3), axis=0), 1e-30)
    k0 = torch.as_tensor(0.39 * (1.0 + 0.2 * np.sin(2 * np.pi * x[..., 0])), dtype=dtype, device="cuda")
    w0 = torch.as_tensor(2660.0 * (1.0 + 0.2 * np.cos(2 * np.pi * x[..., 1])), dtype=dtype, device="cuda")
    res = []
    for v in (0, 1):
        set_variant(dm, "sst_cell", v)
        U, k, w = U0.clone(), k0.clone().contiguous(), w0.clone().contiguous()
        nut, mut = torch.zeros_like(k), torch.zeros_like(k)
        norms = [float(dm.sst_rk3_step(U, k, w, nut, mut).item()) for _ in range(2)]
        res.append((norms, U, k, w, nut, mut))
    assert res[0][0] == res[1][0]
    for a, b in zip(res[0][1:], res[1][1:]):
        assert torch.equal(a, b)


def test_autotune_over_variants_keeps_results():
    import torch
    from autoflowcfd_amd.autotune import autotune, get_variant
    print("[gpu] autotune over (variant, grid)", flush=True)
    dm, U, xyz = _dev(wall_box(seed=6), 2, torch.float32, turbulence="SST")
    k, w = dm.sst_fields(0.39, 2660.0)[:2]
    nut, mut = torch.zeros_like(k), torch.zeros_like(k)
    ref = dm.residual(U, mut, torch.empty_like(U)).clone()
    # the SST phase writes nu_t and mu_t: tune on copies (as CaseBatch.tune does)
    fields = tuple(x.clone() for x in (k, w, nut, mut))
    res = autotune(dm, U, fields[3], reps=1, candidates=(0, 1024), sst_fields=fields, reuse=False)
    assert {"prism_inviscid", "sst_cell"} <= set(res)
    assert res["prism_inviscid"]["variant"] == get_variant(dm, "prism_inviscid")
    assert res["sst_cell"]["variant"] == get_variant(dm, "sst_cell")
    assert len(res["prism_inviscid"]["ms"]) == 4
    assert torch.equal(dm.residual(U, mut, torch.empty_like(U)), ref)
