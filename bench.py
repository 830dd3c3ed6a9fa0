"""Benchmark: FR residual evaluation on the graded wall-bounded prism box (one MI355X).

Workload (default): 82x82x82 hex box on [0,4]x[0,1]x[0,1] (length unit 0.2879 m),
y-graded toward the no-slip wall with ratio 1.1, split into 2 prisms per hex
(1,102,736 prisms), P2 (27 solution points per cell), FP32 storage, Mach 0.15
freestream with a seeded smooth perturbation.  One step = one residual
evaluation of the state resident in HBM.

Prints one JSON line.  --gpus N with N > 1 exits with a message: this build runs
on one GPU.
"""

import argparse
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)

L_UNIT = 0.2879          # m: mu = rho U L / Re = 1.8e-5 at M 0.15, Re 1e6 per unit length
RHO, P_INF = 1.225, 101325.0
A_INF = float(np.sqrt(1.4 * P_INF / RHO))
HBM_PEAK = 8.0e12        # B/s, MI355X HBM3E


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def wall_box_mesh(n, seed=0):
    from autoflowcfd_amd.mesh import (BC_INLET, BC_OUTLET, BC_SYMMETRY, BC_WALL, hex_box)
    bc = {"y0": BC_WALL, "x0": BC_INLET, "x1": BC_OUTLET, "y1": BC_SYMMETRY, "z0": BC_SYMMETRY, "z1": BC_SYMMETRY}
    return hex_box(n, n, n, 4.0 * L_UNIT, 1.0 * L_UNIT, 1.0 * L_UNIT, split="prism", grading_y=1.1, seed=seed,
                   bc_by_plane=bc)


def freestream(mach=0.15):
    return np.array([RHO, mach * A_INF, 0.0, 0.0, P_INF])


def cons(Q):
    rho, u, v, w, p = (Q[..., i] for i in range(5))
    E = p / (0.4 * rho) + 0.5 * (u * u + v * v + w * w)
    return np.stack([rho, rho * u, rho * v, rho * w, rho * E], -1)


def device_state(mesh, ns, Q0, dtype, seed=0):
    """Freestream + seeded smooth perturbation of the cell centroids + small per-SP variation."""
    import torch
    rng = np.random.default_rng(seed)
    cen = np.concatenate([mesh.nodes[mesh.prism_conn].mean(axis=1), mesh.nodes[mesh.tet_conn].mean(axis=1)])
    x = cen / np.maximum(np.ptp(cen, axis=0), 1e-30)
    Q = np.broadcast_to(Q0, (mesh.n_cells, 5)).copy()
    scale = np.array([Q0[0], Q0[1], Q0[1], Q0[1], Q0[4]])
    for v in range(5):
        k = rng.uniform(0.5, 2.0, 3)
        Q[:, v] += 1e-3 * scale[v] * np.sin(2 * np.pi * (x @ k) + rng.uniform(0, 6.28))
    U = torch.as_tensor(cons(Q), dtype=torch.float64, device="cuda")
    g = torch.Generator(device="cuda").manual_seed(seed)
    noise = 1e-5 * torch.randn((mesh.n_cells, ns, 5), generator=g, device="cuda", dtype=torch.float64)
    return (U[:, None, :] * (1.0 + noise)).to(dtype).contiguous()


def algorithmic_bytes_inviscid(n_sp, n_fine_pts, n_faces, nfp, w=4):
    """Compulsory traffic of one inviscid residual launch: read dQ, write R (5 vars),
    det(J) per solution point, adj(J) at the fine points, per physical face the
    flux-point normals and face ids/codes (the SURVEY traffic model, inviscid terms)."""
    return n_sp * (5 * w * 2 + w) + n_fine_pts * 9 * w + n_faces * (nfp * 3 * w + 16)


def cpu_baseline(order, n_sample=10, budget_s=20.0):
    """FP64 oracle (NumPy restatement of the reference CPU path), single thread, on a
    bounded sample: an n_sample^3 hex version of the same graded wall box."""
    from threadpoolctl import threadpool_limits
    sys.path.insert(0, os.path.join(HERE, "tests"))
    from oracle import fr_ops
    from oracle.faces import build_flat_faces
    from oracle.geometry import Geometry
    from oracle.residual import inviscid_residual
    mesh = wall_box_mesh(n_sample, seed=1)
    ops = fr_ops.generate_operators(order)
    g = Geometry(mesh, ops)
    fl = build_flat_faces(mesh, ops)
    Q0 = freestream()
    from autoflowcfd_amd.device import bc_params_table
    bcp = bc_params_table(Q0)
    U = cons(np.broadcast_to(Q0, g.sps_xyz.shape[:-1] + (5,)).copy())
    U *= 1.0 + 1e-3 * np.sin(g.sps_xyz[..., :1] / L_UNIT)
    with threadpool_limits(limits=1):
        inviscid_residual(U, g, ops, fl, {c: bcp[c] for c in range(8)})
        times = []
        t_end = time.perf_counter() + budget_s
        while time.perf_counter() < t_end or not times:
            t0 = time.perf_counter()
            inviscid_residual(U, g, ops, fl, {c: bcp[c] for c in range(8)})
            times.append(time.perf_counter() - t0)
    t = float(np.median(times))
    n_sp = mesh.n_cells * ops.n_sps
    return {"value": n_sp * 5 / t / 1e9, "unit": "GDOF/s", "cores": 1, "kind": "port",
            "sample": f"FP64 NumPy oracle inviscid residual, {n_sample}^3-hex graded wall box "
                      f"({mesh.n_cells} prisms, P{order}), median of {len(times)} evaluations, "
                      f"{t:.3f} s each"}


def traffic_from_profiles(name):
    p = os.path.join(HERE, "profiles", f"{name}.json")
    if os.path.exists(p):
        with open(p) as f:
            return json.load(f).get("bytes_per_launch")
    return None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--n", type=int, default=82, help="hex cells per direction")
    ap.add_argument("--order", type=int, default=2)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-sample-n", type=int, default=10)
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}); this build has no multi-GPU path.",
              file=sys.stderr)
        sys.exit(2)

    import torch
    from autoflowcfd_amd.device import DeviceMesh, bc_params_table

    t0 = time.time()
    mesh = wall_box_mesh(args.n)
    log(f"[bench] mesh {mesh.n_cells} prisms, {mesh.faces.n_faces} face records ({time.time() - t0:.1f}s)")
    Q0 = freestream()
    t0 = time.time()
    dm = DeviceMesh(mesh, args.order, dtype=torch.float32, bc_params=bc_params_table(Q0), metrics="gpu")
    torch.cuda.synchronize()
    log(f"[bench] device set-up {time.time() - t0:.1f}s, {dm.slots.n_blocks} row blocks")
    U = device_state(mesh, dm.ns, Q0, torch.float32)
    R = torch.empty_like(U)
    dQ, umean = dm.work_q, dm.work_umean

    for _ in range(args.warmup):
        dm.prepare_state(U, dQ, umean)
        dm.inviscid_from_primitives(dQ, umean, R)
    torch.cuda.synchronize()
    log("[bench] warm-up done")

    ev = [[torch.cuda.Event(enable_timing=True) for _ in range(3)] for _ in range(args.steps)]
    torch.cuda.synchronize()
    t_start = time.perf_counter()
    for k in range(args.steps):
        ev[k][0].record()
        dm.prepare_state(U, dQ, umean)
        ev[k][1].record()
        dm.inviscid_from_primitives(dQ, umean, R)
        ev[k][2].record()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t_start
    t_prim = float(np.mean([ev[k][0].elapsed_time(ev[k][1]) for k in range(args.steps)])) * 1e-3
    t_inv = float(np.mean([ev[k][1].elapsed_time(ev[k][2]) for k in range(args.steps)])) * 1e-3
    ms_per_step = wall / args.steps * 1e3
    log(f"[bench] {args.steps} steps: {ms_per_step:.3f} ms/step (prepare {t_prim * 1e3:.3f} ms, "
        f"inviscid {t_inv * 1e3:.3f} ms)")
    if not bool(torch.isfinite(R).all()):
        raise RuntimeError("non-finite residual")

    n_sp = mesh.n_cells * dm.ns
    n_phys_faces = int(((~mesh.faces.is_boundary).sum() + mesh.faces.is_boundary.sum()))
    # physical faces: prism quads are two triangle records
    is_quad = np.isin(mesh.faces.owner_code, (0, 1, 2)) & (mesh.faces.owner < mesh.n_prism)
    n_phys_faces = int(n_phys_faces - is_quad.sum() // 2)
    b_alg = algorithmic_bytes_inviscid(n_sp, mesh.n_cells * dm.nf, n_phys_faces, dm.nfp)
    achieved = b_alg / t_inv
    value = n_sp * 5 / ms_per_step * 1e3 / 1e9
    line = {
        "metric": "P2 residual GDOF/s + %HBM roofline",
        "value": value,
        "unit": "GDOF/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"inviscid FR residual, {args.n}^3-hex graded wall box split into "
                               f"{mesh.n_cells} prisms, P{args.order}, {n_sp} solution points x 5 vars",
                   "cells": mesh.n_cells, "order": args.order, "n_vars": 5},
        "roofline": {"bound": "hbm", "achieved": achieved / 1e9, "peak": HBM_PEAK / 1e9, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK, "traffic": traffic_from_profiles("inviscid_c2"),
                     "kernel": f"k_prism_inviscid<{args.order},float>", "kernel_ms": t_inv * 1e3,
                     "algorithmic_bytes": b_alg},
    }
    if not args.no_cpu_baseline:
        log("[bench] CPU baseline (oracle) ...")
        line["cpu_baseline"] = cpu_baseline(args.order, n_sample=args.cpu_sample_n)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
