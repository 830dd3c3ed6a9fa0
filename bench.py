"""Benchmark: P2 SST RANS residual on the graded wall-bounded prism box (one MI355X).

Workload (default): 82x82x82 hex box on [0,4]x[0,1]x[0,1] (length unit 0.2879 m),
y-graded toward the no-slip wall with ratio 1.1, split into 2 prisms per hex
(1,102,736 prisms), P2 (27 solution points per cell), FP32 storage, Mach 0.15,
Re 1e6 per unit length, k and omega from 1% intensity and viscosity ratio 10,
seeded smooth perturbation.  One step = one full residual evaluation of the state
resident in HBM: local time step, SST sources + transport (with the k/omega
update), inviscid + viscous mean-flow residual.  Full SSP-RK3 pseudo-steps are
timed afterwards and reported beside it.

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
MU = 1.8e-5
K_INF = 0.3908           # 1.5 (0.01 U)^2, 1% turbulence intensity
OMEGA_INF = 2660.0       # rho k / (10 mu), viscosity ratio 10
RHO, P_INF = 1.225, 101325.0
A_INF = float(np.sqrt(1.4 * P_INF / RHO))
HBM_PEAK = 8.0e12        # B/s, MI355X HBM3E


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def wall_box_mesh(n, seed=0, topology="host"):
    from autoflowcfd_amd.mesh import (BC_INLET, BC_OUTLET, BC_SYMMETRY, BC_WALL, hex_box)
    bc = {"y0": BC_WALL, "x0": BC_INLET, "x1": BC_OUTLET, "y1": BC_SYMMETRY, "z0": BC_SYMMETRY, "z1": BC_SYMMETRY}
    return hex_box(n, n, n, 4.0 * L_UNIT, 1.0 * L_UNIT, 1.0 * L_UNIT, split="prism", grading_y=1.1, seed=seed,
                   bc_by_plane=bc, topology=topology)


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
    flux-point normals and face ids/codes."""
    return n_sp * (5 * w * 2 + w) + n_fine_pts * 9 * w + n_faces * (nfp * 3 * w + 16)


def algorithmic_bytes_residual(n_sp, n_fine_pts, n_faces, nfp, n_v=7, w=4, sst=True):
    """B_res of one full residual evaluation: read U, write R (n_v vars), det + adj(J) at the
    solution points, adj(J) at the fine points, wall distance (SST), face normals and ids."""
    return (n_sp * n_v * w * 2 + n_sp * 10 * w + n_fine_pts * 9 * w + (n_sp * w if sst else 0)
            + n_faces * (nfp * 3 * w + 16))


def algorithmic_bytes_visc_volume(n_sp, n_cells, w=4):
    """Compulsory traffic of the viscous volume launch: read U (5), mu_t, det, adj (9) per
    solution point; write dQ (5), the 18 used values of the interface pack, the volume
    residual (5) per point, and the outlier field mean (5) and h per cell."""
    return n_sp * (5 + 1 + 1 + 9 + 5 + 18 + 5) * w + n_cells * 6 * w


def algorithmic_bytes_dt_local(n_sp, n_cells, w=4):
    """Compulsory traffic of the local-time-step launch: read U (5), nu_t, det and the metric
    flux scale per solution point, the volume per cell; write dt per point."""
    return n_sp * (5 + 1 + 1 + 1 + 1) * w + n_cells * w


def algorithmic_bytes_sst(n_sp, n_cells, n_records, w=4):
    """Compulsory traffic of the SST phase (sources, transport, update rates: k_sst_cell,
    k_sst_face, k_sst_update): read U (5), k, omega, wall distance, dt, det and adj (9) per
    solution point, write dk/dt, domega/dt, nu_t and mu_t per point; per face record its unit
    normal (3) and owner, owner slot and flags; per cell its eight record incidences."""
    return n_sp * (5 + 1 + 1 + 1 + 1 + 1 + 9 + 4) * w + n_records * (3 * w + 3 * 4) + n_cells * 8 * 4


def algorithmic_bytes_visc_iface(n_sp, n_cells, nfp, w=4):
    """Compulsory traffic of the viscous interface launch: read the 18 pack values, the volume
    residual (5) and det per point, read-modify-write R (5); per cell the five face slots'
    extrapolated adjugate rows (nfp x 3), normals (3), six slot ints, h and the field mean (5)."""
    return n_sp * (18 + 5 + 1 + 10) * w + n_cells * (5 * (nfp * 3 + 3) * w + 5 * 6 * 4 + 6 * w)


def _cpu_sample(order, n_sample, budget_s, seed):
    """One single-threaded worker of the CPU baseline: (solution points, median seconds, evaluations)
    of the FP64 oracle's full P2 SST RANS residual on an n_sample^3-hex graded wall box."""
    from threadpoolctl import threadpool_limits
    from oracle import fr_ops, sst
    from oracle.faces import build_flat_faces
    from oracle.geometry import Geometry
    from oracle.residual import BC_WALL, inviscid_residual, viscous_residual
    from oracle.time_step import local_time_step, update_primitives
    from autoflowcfd_amd.device import bc_params_table
    mesh = wall_box_mesh(n_sample, seed=seed)
    ops = fr_ops.generate_operators(order)
    g = Geometry(mesh, ops)
    fl = build_flat_faces(mesh, ops)
    Q0 = freestream()
    bcp = bc_params_table(Q0)
    bcd = {c: bcp[c] for c in range(8)}
    U = cons(np.broadcast_to(Q0, g.sps_xyz.shape[:-1] + (5,)).copy())
    U *= 1.0 + 1e-3 * np.sin(g.sps_xyz[..., :1] / L_UNIT)
    d = sst.wall_distance(g.sps_xyz, mesh.nodes[mesh.wall_nodes])
    k = np.full(U.shape[:2], K_INF)
    w = np.full(U.shape[:2], OMEGA_INF)
    wall = fl.bc == BC_WALL

    def residual():
        Q = update_primitives(U)
        dt = local_time_step(Q, g, order, MU)
        _, _, nut, _, _ = sst.turbulence_step(U, k, w, g, ops, fl, d, dt, MU, wall)
        mut = Q[..., 0] * nut
        return inviscid_residual(U, g, ops, fl, bcd) + viscous_residual(U, g, ops, fl, bcd, MU, mu_t=mut)

    with threadpool_limits(limits=1):
        residual()
        times = []
        t_end = time.perf_counter() + budget_s
        while time.perf_counter() < t_end or not times:
            t0 = time.perf_counter()
            residual()
            times.append(time.perf_counter() - t0)
    return mesh.n_cells * ops.n_sps, float(np.median(times)), len(times)


def cpu_baseline(order, n_sample=6, budget_s=20.0, workers=None):
    """FP64 oracle (NumPy restatement of the reference CPU path) on the host cores: one
    single-threaded process per core, each timing full SST RANS residual evaluations of its own
    seeded n_sample^3-hex graded wall box for budget_s; the throughput is their sum (the cases
    are independent, as the reference's one-case-per-process runs are)."""
    import multiprocessing as mp
    n_host = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    workers = workers or max(1, min(16, n_host))
    with mp.get_context("spawn").Pool(workers) as pool:
        res = pool.starmap(_cpu_sample, [(order, n_sample, budget_s, 1 + i) for i in range(workers)])
    gdof = sum(n_sp * 7 / t for n_sp, t, _ in res) / 1e9
    one = res[0][0] * 7 / res[0][1] / 1e9
    return {"value": gdof, "unit": "GDOF/s", "cores": workers, "kind": "port", "single_core_value": one,
            "host_cpus": n_host,
            "sample": f"FP64 NumPy oracle, P{order} SST RANS residual (turbulence sources + transport + inviscid + "
                      f"viscous) on {workers} single-threaded processes at once, each its own seeded {n_sample}^3-hex "
                      f"graded wall box ({res[0][0] // ((order + 1) ** 3)} prisms), median of "
                      f"{min(r[2] for r in res)}-{max(r[2] for r in res)} evaluations of {budget_s:.0f} s per process "
                      f"({res[0][1]:.3f} s each on one core)"}


def residual_kernel_names(P, knz, ns, v_inv=0, v_sst=0):
    """{phase: [kernel names as the profiler prints them]} of one prism-mesh residual evaluation at
    order P with knz-wide interpolation rows, ns stored points per cell and the tile variants in
    force (afr_mesh_set_variant: prism variant v stages the slot sources when v = 1; SST cell
    variant v pairs prisms when v = 0)."""
    return {
        "dt_local": [f"afr_step::k_dt_cells<{ns}, float>" if ns in (1, 8, 27, 64) else "afr_step::k_dt_local<float>"],
        "sst": [f"afr_sst::k_sst_cell<{P}, 0, float, {1 - v_sst}>", f"afr_sst::k_sst_face<{P}, {knz}, float>",
                f"afr_sst::k_sst_update<{P}, 0, float>"],
        "visc_volume": [f"afr_visc::k_visc_volume<{P}, 0, float>"],
        "inviscid": [f"afr_prism::k_prism_inviscid<{P}, {knz}, float, {v_inv}>"],
        "visc_interface": [f"afr_visc::k_visc_iface<{P}, 0, {knz}, float>"],
    }


def traffic_from_profiles(kernel, workload="wallbox82_p2_sst_fp32"):
    """HBM bytes per launch of a kernel from the committed rocprofv3 PMC summary (FETCH_SIZE and
    WRITE_SIZE scaled by a plain copy's), or None."""
    p = os.path.join(HERE, "profiles", f"traffic_{workload}.json")
    if os.path.exists(p):
        with open(p) as f:
            k = json.load(f).get("kernels", {}).get(kernel)
        return None if k is None else k.get("bytes_per_launch")
    return None


def served_leg(args):
    """Served stream vs the plain one-at-a-time GPU path on the same seeded requests: solves/h,
    p50 / p99 latency, and the largest Cd difference between the two (same-result check).  The
    server steps its resident solves of each order as one batched mesh assembled on the device
    (continuous batching); both paths use the server's common reference state, so every solve's
    residual history must be the same bit for bit."""
    import torch
    from autoflowcfd_amd.serve import SolveServer, latency_stats, solve_alone, synthetic_requests
    lo, hi = (float(x) for x in args.serve_hex.split(","))
    turb = args.serve_turb.upper()
    reqs = synthetic_requests(args.serve_requests, seed=args.serve_seed, hex_range=(lo, hi), rate=args.serve_rate,
                              max_iter=args.serve_max_iter, rel_tol=1e-6, topology="gpu")
    n_sp = sum(r.case.mesh.n_cells * (r.order + 1) ** 3 for r in reqs)
    log(f"[bench] served leg: {len(reqs)} {turb} requests, {n_sp} solution points in total")
    batched = args.serve_mode == "batched"
    mk = lambda n, hw: SolveServer(dtype=torch.float32, turb_model=turb, max_active=n, host_workers=hw,  # noqa: E731
                                   batched=batched)
    # warm the code paths once (library load, first batch / graph) outside both timings
    warm = mk(2, 0)
    warm.run(reqs[:1])
    ref = warm.ref_state if batched else None
    solve_alone(reqs[:1], dtype=torch.float32, turb_model=turb, ref_state=ref)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    plain = solve_alone(reqs, dtype=torch.float32, turb_model=turb, ref_state=ref)
    log(f"[bench] plain path done in {time.perf_counter() - t0:.1f}s")
    srv = mk(args.serve_active, args.serve_host_workers)
    t0 = time.perf_counter()
    served = srv.run(reqs)
    log(f"[bench] served path ({args.serve_mode}) done in {time.perf_counter() - t0:.1f}s "
        f"(peak {srv.peak_active} resident solves, {srv.rebatches} batch rebuilds)")
    srv.close()
    sp, ss = latency_stats(plain), latency_stats(served)
    busy = [r.finished - r.admitted - r.setup_s for r in plain]
    log(f"[bench] plain: mean set-up {np.mean([r.setup_s for r in plain]):.3f}s, mean stepping {np.mean(busy):.3f}s "
        f"per solve; served: mean set-up {np.mean([r.setup_s for r in served]):.3f}s")
    dcd = max_abs_diff([a.coefficients.get("Cd", 0.0) for a in served], [b.coefficients.get("Cd", 0.0) for b in plain])
    same = all(a.residual_history == b.residual_history for a, b in zip(served, plain))
    return {
        "workload": f"{len(reqs)} steady {'laminar NS' if turb == 'NONE' else 'SST RANS'} solves, graded wall boxes of "
                    f"{lo:g}-{hi:g} hex (log-uniform) split into prisms, P1/P2, Mach 0.05-0.3, yaw 0-10 deg, Poisson "
                    f"arrivals {args.serve_rate}/s, stop at res/res0 < 1e-6 or {args.serve_max_iter} pseudo-steps; "
                    f"server mode: {args.serve_mode}",
        "solves_per_hour": ss["solves_per_hour"], "p50_latency_s": ss["p50_s"], "p99_latency_s": ss["p99_s"],
        "plain_solves_per_hour": sp["solves_per_hour"], "plain_p50_latency_s": sp["p50_s"],
        "plain_p99_latency_s": sp["p99_s"], "speedup_vs_plain": ss["solves_per_hour"] / sp["solves_per_hour"],
        "max_abs_dCd_vs_plain": dcd, "histories_identical": same, "peak_resident_solves": srv.peak_active,
        "batch_rebuilds": srv.rebatches, "iterations": [r.iterations for r in served],
    }


def max_abs_diff(a, b):
    """Largest |a_i - b_i|, 0 where both are the same non-finite value, inf where only one is."""
    out = 0.0
    for x, y in zip(a, b):
        if np.isfinite(x) and np.isfinite(y):
            out = max(out, abs(x - y))
        elif not (np.isnan(x) and np.isnan(y)) and x != y:
            out = float("inf")
    return out


def batched_leg(args):
    """Batched multi-case form vs the plain one-at-a-time path on the same seeded SST cases (fixed
    step count, set-up included in both): P1 and P2 cases in one MixedBatch (one HIP graph per
    pseudo-step holding both order groups), solves/h and the largest Cd difference to the plain
    path."""
    import torch
    from autoflowcfd_amd.batch import MixedBatch, static_ref_state
    from autoflowcfd_amd.serve import solve_alone, synthetic_requests
    lo, hi = (float(x) for x in args.batch_hex.split(","))
    reqs = synthetic_requests(args.batch_cases, seed=args.serve_seed + 7, hex_range=(lo, hi), rate=1e9,
                              max_iter=args.batch_steps, rel_tol=0.0, topology="gpu", stop_nonfinite=False)
    ref = static_ref_state([r.case for r in reqs])
    n_sp = sum(r.case.mesh.n_cells * (r.order + 1) ** 3 for r in reqs)
    log(f"[bench] batched leg: {len(reqs)} SST cases, {n_sp} solution points, {args.batch_steps} pseudo-steps each")
    solve_alone(reqs[:1], dtype=torch.float32, turb_model="SST", ref_state=ref, use_graph=True)   # warm the paths
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    plain = solve_alone(reqs, dtype=torch.float32, turb_model="SST", ref_state=ref, use_graph=True)
    t_plain = time.perf_counter() - t0
    t0 = time.perf_counter()
    b = MixedBatch([r.case for r in reqs], [r.order for r in reqs], dtype=torch.float32, turb_model="SST",
                   ref_state=ref, host_workers=args.batch_host_workers)
    t_tune = time.perf_counter()
    tuned = b.tune() if args.batch_tune else None
    t_tune = time.perf_counter() - t_tune
    t_setup = time.perf_counter() - t0
    b.capture_graph()                                       # one eager warm-up step, counted
    t1 = time.perf_counter()
    for _ in range(args.batch_steps - 1):
        b.step(sync=False)
    torch.cuda.synchronize()
    t_steps = time.perf_counter() - t1
    coef = b.coefficients()
    t_batch = time.perf_counter() - t0
    log(f"[bench] plain {t_plain:.2f}s, batched {t_batch:.2f}s (set-up {t_setup:.2f}s incl. autotuning "
        f"{t_tune:.2f}s, "
        f"{t_steps / max(args.batch_steps - 1, 1) * 1e3:.2f} ms per mixed pseudo-step)")
    if any(r.iterations != args.batch_steps for r in plain):
        raise RuntimeError("a plain solve did not run the fixed step count")
    dcd = max_abs_diff([coef[r.rid]["Cd"] for r in plain], [r.coefficients["Cd"] for r in plain])
    nonfinite = int(sum(not np.isfinite(r.final_residual) for r in plain))
    del b
    torch.cuda.empty_cache()
    return {"workload": f"{len(reqs)} SST cases, graded wall boxes of {lo:g}-{hi:g} hex (log-uniform, aspect 2-6) "
                        f"split into prisms, P1/P2 drawn 50/50 in one mixed batch (one HIP graph per pseudo-step), "
                        f"Mach 0.05-0.3, yaw 0-10 deg, {args.batch_steps} pseudo-steps each, set-up included",
            "solves_per_hour": len(reqs) / t_batch * 3600.0, "plain_solves_per_hour": len(reqs) / t_plain * 3600.0,
            "speedup_vs_plain": t_plain / t_batch, "max_abs_dCd_vs_plain": dcd,
            "cases_nonfinite_at_end": nonfinite,
            "ms_per_mixed_pseudo_step": t_steps / max(args.batch_steps - 1, 1) * 1e3,
            "autotune_s": t_tune,
            "autotuned": None if tuned is None else {f"P{p}": {k: {"grid": v["grid"], "variant": v["variant"]}
                                                                for k, v in ch.items()} for p, ch in tuned.items()},
            "solution_points": n_sp}


def tgv_noise(xyz, v0, seed=7, modes=4):
    """Seeded smooth velocity noise of rms 1e-3 v0: per component a sum of periodic Fourier modes
    (wavenumbers 1..2, random phases and amplitudes).  Point-wise white noise is not usable at P3:
    the collapsed-coordinate derivative rows reach 6e7 and turn it into gradients of 1e6."""
    rng = np.random.default_rng(seed)
    out = np.zeros(xyz.shape)
    for c in range(3):
        for _ in range(modes):
            k = rng.integers(1, 3, 3)
            ph = rng.uniform(0, 2 * np.pi, 3)
            amp = rng.standard_normal()
            out[..., c] += amp * np.sin(k[0] * xyz[..., 0] + ph[0]) * np.sin(k[1] * xyz[..., 1] + ph[1]) * \
                np.sin(k[2] * xyz[..., 2] + ph[2])
    return 1e-3 * v0 * out / np.sqrt(modes / 8.0)


def tgv_p3_leg(args):
    """P3 Navier-Stokes Taylor-Green vortex: triply periodic box of side 2 pi, N^3 hex x 6 tets,
    FP64 (the P3 path is FP64 only), Mach 0.1 (V0 = 0.1 a_inf), Re 1600, the initial field of
    tests/validation/test_tgv.py:96-112 of the reference plus smooth seeded noise of 1e-3 V0, SSP-RK3
    with one global dt frozen from the initial state (CFL 0.1 on V^(1/3) with |u| + a, the viscous
    bound and the minimum local time step: the global dt of test_tgv.py:70-92).  Reports the time
    per step (3 residual evaluations), GDOF/s per residual evaluation, and the kinetic energy and
    enstrophy (mean over solution points of 1/2 rho |u|^2 and 1/2 rho |curl u|^2, from the
    device's own velocity gradients) along the run."""
    import torch
    from autoflowcfd_amd.fr_tables import tables
    from autoflowcfd_amd.geometry import map_ref
    from autoflowcfd_amd.mesh import hex_box
    from autoflowcfd_amd.solver import GPUFRSolver
    n, steps = args.tgv_p3, args.tgv_steps
    gam, L = 1.4, 2.0 * np.pi
    a_inf = float(np.sqrt(gam * P_INF / RHO))
    v0 = 0.1 * a_inf
    mu = RHO * v0 * 1.0 / 1600.0
    t0 = time.time()
    mesh = hex_box(n, n, n, L, L, L, split="tet", periodic=True, permute=False)
    s = GPUFRSolver(mesh, order=3, rho_inf=RHO, vel_inf=v0, p_inf=P_INF, mu_molecular=mu, dtype=torch.float64,
                    global_time_step=True)
    xyz = map_ref(False, tables(3).sps, mesh.nodes[mesh.tet_conn])
    x, y, z = xyz[..., 0], xyz[..., 1], xyz[..., 2]
    noise = tgv_noise(xyz, v0)
    u = v0 * np.sin(x) * np.cos(y) * np.cos(z) + noise[..., 0]
    v = -v0 * np.cos(x) * np.sin(y) * np.cos(z) + noise[..., 1]
    w = noise[..., 2]
    p = P_INF + (RHO * v0 ** 2 / 16.0) * (np.cos(2 * x) + np.cos(2 * y)) * (np.cos(2 * z) + 2.0)
    U0 = np.stack([np.full_like(u, RHO), RHO * u, RHO * v, RHO * w,
                   p / (gam - 1.0) + 0.5 * RHO * (u * u + v * v + w * w)], axis=-1)
    del xyz, x, y, z, noise, u, v, w, p
    s.set_state_from_cpu(U0)
    h = np.abs(s.cell_volumes())[:, None] ** (1.0 / 3.0)
    vel = U0[..., 1:4] / U0[..., :1]
    pr = (gam - 1.0) * (U0[..., 4] - 0.5 * RHO * np.sum(vel ** 2, axis=-1))
    vm = np.linalg.norm(vel, axis=-1)
    dt_glob = np.minimum(0.1 * h / (vm + np.sqrt(gam * pr / RHO)), 0.25 * 0.1 * RHO * h ** 2 / mu).min()
    dt_loc = float(s.dm.local_time_step(s.U, None, mu=mu).min().item())
    dt = float(min(dt_glob, dt_loc))
    del U0, vel, pr, vm, h
    n_sp = mesh.n_cells * s.dm.ns
    log(f"[tgv-p3] {mesh.n_cells} tets, {n_sp} solution points, set-up {time.time() - t0:.1f}s, dt {dt:.3e}")

    def diagnostics():
        U = s.U
        ke = float((0.5 * (U[..., 1:4] ** 2).sum(-1) / U[..., 0]).mean().item())
        s.dm.viscous_volume(U, None, mu=mu)
        g = s.dm.work_vpack[..., 5:14].reshape(U.shape[0], U.shape[1], 3, 3)   # [i][j] = du_i/dx_j
        om2 = (g[..., 2, 1] - g[..., 1, 2]) ** 2 + (g[..., 0, 2] - g[..., 2, 0]) ** 2 + (g[..., 1, 0] - g[..., 0, 1]) ** 2
        ens = float((0.5 * U[..., 0] * om2).mean().item())
        return ke, ens, bool(torch.isfinite(U).all().item())

    hist = [(0,) + diagnostics()]
    every = max(steps // 10, 1)
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    t_steps = 0.0
    for i in range(steps):
        ev0.record()
        s.step(dt)
        ev1.record()
        ev1.synchronize()
        t_steps += ev0.elapsed_time(ev1)
        if (i + 1) % every == 0 or i + 1 == steps:
            hist.append((i + 1,) + diagnostics())
            log(f"[tgv-p3] step {i + 1}: KE {hist[-1][1]:.6e} enstrophy {hist[-1][2]:.6e} finite {hist[-1][3]}")
            if not hist[-1][3]:
                break
    ms = t_steps / max(hist[-1][0], 1)
    return {"config": {"workload": f"P3 NS Taylor-Green vortex, Mach 0.1, Re 1600, {n}^3 hex x 6 tets "
                                   f"({mesh.n_cells} cells, {n_sp} solution points), FP64, global dt"},
            "steps": hist[-1][0], "dt": dt, "ms_per_step": ms,
            "gdof_per_s_per_residual": n_sp * 5 * 3 / (ms * 1e-3) / 1e9,
            "history": [{"step": a, "ke": b, "enstrophy": c, "finite": d} for a, b, c, d in hist]}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--n", type=int, default=82, help="hex cells per direction")
    ap.add_argument("--order", type=int, default=2)
    ap.add_argument("--rk-steps", type=int, default=2, help="full SSP-RK3 pseudo-steps timed after the residuals")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-sample-n", type=int, default=6)
    ap.add_argument("--no-serve", action="store_true", help="skip the served-stream leg")
    ap.add_argument("--no-autotune", action="store_true", help="keep the built-in persistent grid sizes")
    ap.add_argument("--storage", choices=("locality", "input"), default="locality",
                    help="device cell storage order: Morton order of the centroids, or the mesh's own order")
    ap.add_argument("--grids", default="", help="explicit launch shapes instead of autotuning: kernel=grid[:variant],...")
    ap.add_argument("--serve-only", action="store_true", help="run only the served-stream leg (prints its JSON)")
    ap.add_argument("--serve-requests", type=int, default=12)
    ap.add_argument("--serve-hex", default="2e3,2e4", help="hex-count range of the served requests (lo,hi)")
    ap.add_argument("--serve-turb", default="none", choices=("none", "sst"))
    ap.add_argument("--serve-mode", default="batched", choices=("batched", "streams"),
                    help="batched: one device-assembled batch per order; streams: one graph per solve")
    ap.add_argument("--serve-max-iter", type=int, default=200)
    ap.add_argument("--serve-rate", type=float, default=20.0, help="Poisson arrival rate (requests/s)")
    ap.add_argument("--serve-active", type=int, default=8, help="resident solves at most")
    ap.add_argument("--serve-seed", type=int, default=1)
    ap.add_argument("--serve-host-workers", type=int, default=0,
                    help="processes building the slot tables with NumPy (0: on the GPU)")
    ap.add_argument("--batch-cases", type=int, default=64)
    ap.add_argument("--batch-steps", type=int, default=100)
    ap.add_argument("--batch-hex", default="2e3,2e4", help="hex-count range of the batched cases (lo,hi)")
    ap.add_argument("--batch-tune", type=int, default=1, help="autotune the batch's launch shapes (1) or not (0)")
    ap.add_argument("--batch-host-workers", type=int, default=0,
                    help="processes building the slot tables with NumPy (0: on the GPU)")
    ap.add_argument("--no-batch", action="store_true", help="skip the batched multi-case leg")
    ap.add_argument("--batch-only", action="store_true", help="run only the batched multi-case leg (prints its JSON)")
    ap.add_argument("--tgv-p3", type=int, default=0,
                    help="run only the P3 Taylor-Green vortex leg on N^3 hex x 6 tets (prints its JSON)")
    ap.add_argument("--tgv-steps", type=int, default=200)
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}); this build has no multi-GPU path.",
              file=sys.stderr)
        sys.exit(2)

    import torch
    from autoflowcfd_amd.device import DeviceMesh, bc_params_table

    if args.tgv_p3:
        print(json.dumps(tgv_p3_leg(args)), flush=True)
        return
    if args.batch_only:
        print(json.dumps(batched_leg(args)), flush=True)
        return
    if args.serve_only:
        out = served_leg(args)
        if not args.no_batch:
            out["batched"] = batched_leg(args)
        print(json.dumps(out), flush=True)
        return
    t0 = time.time()
    mesh = wall_box_mesh(args.n, topology="gpu")
    log(f"[bench] mesh {mesh.n_cells} prisms, {mesh.faces.n_faces} face records ({time.time() - t0:.1f}s)")
    if args.storage == "locality":
        # the solver stores cells in the Morton order of their centroids (the input order is random)
        from autoflowcfd_amd.device import morton_order
        from autoflowcfd_amd.mesh import renumber_cells
        mesh, _ = renumber_cells(mesh, morton_order(mesh.nodes[mesh.prism_conn].mean(axis=1)))
    Q0 = freestream()
    t0 = time.time()
    # FP32 storage as the solver keeps it: the offset from the free stream
    dm = DeviceMesh(mesh, args.order, dtype=torch.float32, bc_params=bc_params_table(Q0), metrics="gpu",
                    turbulence="SST", state_ref=cons(Q0))
    torch.cuda.synchronize()
    log(f"[bench] device set-up {time.time() - t0:.1f}s, {dm.slots.n_blocks} row blocks")
    U = dm.to_storage(device_state(mesh, dm.ns, Q0, torch.float64))
    k, w, nut, mut = dm.sst_fields(K_INF, OMEGA_INF)
    R = torch.empty_like(U)
    rates = torch.empty((mesh.n_cells, dm.ns, 2), dtype=torch.float32, device="cuda")
    dt = dm.work_dt

    def residual_step():
        # one full P2 SST RANS residual evaluation of the resident state: local time step,
        # SST sources + transport (dk/dt, domega/dt of the point-implicit update), mean-flow residual
        dm.local_time_step_nut(U, nut, dt)
        dm.sst_step(U, k, w, dt, nut, mut, mu=MU, rates=rates)
        dm.residual(U, mut, R, mu=MU)

    for _ in range(args.warmup):
        residual_step()
    torch.cuda.synchronize()
    log("[bench] warm-up done")
    tuned = None
    if args.grids:
        from autoflowcfd_amd.autotune import set_grid
        tuned = {}
        from autoflowcfd_amd.autotune import set_variant
        for item in args.grids.split(","):
            name, val = item.split("=")
            cap, _, var = val.partition(":")
            set_variant(dm, name.strip(), int(var or 0))
            set_grid(dm, name.strip(), int(cap))
            tuned[name.strip()] = {"grid": int(cap), "variant": int(var or 0)}
        log("[bench] grids set: " + args.grids)
    elif not args.no_autotune:
        from autoflowcfd_amd.autotune import autotune
        tuned = autotune(dm, U, mut, mu=MU, sst_fields=(k, w, nut, mut))
        log("[bench] autotuned grids: " + ", ".join(f"{n} {v['grid']}" + (f" (variant {v['variant']})" if v.get('variant') else "")
                                                    for n, v in tuned.items()))
        for _ in range(args.warmup):
            residual_step()
        torch.cuda.synchronize()

    n_ev = 6
    ev = [[torch.cuda.Event(enable_timing=True) for _ in range(n_ev)] for _ in range(args.steps)]
    torch.cuda.synchronize()
    t_start = time.perf_counter()
    for i in range(args.steps):
        e = ev[i]
        e[0].record()
        dm.local_time_step_nut(U, nut, dt)
        e[1].record()
        dm.sst_step(U, k, w, dt, nut, mut, mu=MU, rates=rates)
        e[2].record()
        dm.viscous_volume(U, mut, mu=MU)
        e[3].record()
        dm.inviscid_from_primitives(dm.work_q, dm.work_umean, R)
        e[4].record()
        dm.viscous_interface(U, mut, R, mu=MU)
        e[5].record()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t_start
    ph = np.array([[ev[i][j].elapsed_time(ev[i][j + 1]) for j in range(n_ev - 1)] for i in range(args.steps)]).mean(0)
    ms_per_step = wall / args.steps * 1e3
    names = ["dt_local", "sst", "visc_volume", "inviscid", "visc_interface"]
    log("[bench] residual: %.3f ms/step (%s)" % (ms_per_step, ", ".join(f"{n} {t:.3f}" for n, t in zip(names, ph))))
    if not bool(torch.isfinite(R).all()):
        raise RuntimeError("non-finite residual")

    # full SSP-RK3 pseudo-steps (3 mean-flow residuals + SST + stages + FP64 norm)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.rk_steps):
        dm.sst_rk3_step(U, k, w, nut, mut, mu=MU)
    torch.cuda.synchronize()
    ms_rk = (time.perf_counter() - t0) / max(args.rk_steps, 1) * 1e3
    norm = float(dm.norm_dev.item())
    log(f"[bench] SSP-RK3 SST pseudo-step {ms_rk:.3f} ms, residual RMS {norm:.4e}")

    n_sp = mesh.n_cells * dm.ns
    fr = mesh.faces
    is_quad = np.isin(fr.owner_code, (0, 1, 2)) & (fr.owner < mesh.n_prism)
    n_phys_faces = int(fr.n_faces - is_quad.sum() // 2)
    b_inv = algorithmic_bytes_inviscid(n_sp, mesh.n_cells * dm.nf, n_phys_faces, dm.nfp)
    # per-phase rooflines (prism-only mesh: one launch per phase, three for SST), timed with HIP
    # events on the launch stream; PMC traffic from the committed rocprofv3 summary
    P, knz = args.order, dm.rows_knz
    from autoflowcfd_amd.autotune import get_variant
    v_inv, v_sst = get_variant(dm, "prism_inviscid"), get_variant(dm, "sst_cell")
    kn = residual_kernel_names(P, knz, dm.ns, v_inv, v_sst)
    phase_bytes = {"dt_local": algorithmic_bytes_dt_local(n_sp, mesh.n_cells),
                   "sst": algorithmic_bytes_sst(n_sp, mesh.n_cells, fr.n_faces),
                   "visc_volume": algorithmic_bytes_visc_volume(n_sp, mesh.n_cells), "inviscid": b_inv,
                   "visc_interface": algorithmic_bytes_visc_iface(n_sp, mesh.n_cells, dm.nfp)}
    kr = {}
    for phase, parts in kn.items():
        kname, b = " + ".join(parts), phase_bytes[phase]
        parts = parts if len(parts) > 1 else None
        t = float(ph[names.index(phase)]) * 1e-3
        tr = [traffic_from_profiles(k) for k in (parts or [kname])]
        kr[kname] = {"phase": phase, "bound": "hbm", "achieved": b / t / 1e9, "peak": HBM_PEAK / 1e9,
                     "unit": "GB/s", "frac": b / t / HBM_PEAK,
                     "traffic": None if any(x is None for x in tr) else float(sum(tr)), "kernel_ms": t * 1e3,
                     "algorithmic_bytes": b}
    dominant = max(kr, key=lambda k: kr[k]["kernel_ms"])
    b_res = algorithmic_bytes_residual(n_sp, mesh.n_cells * dm.nf, n_phys_faces, dm.nfp)
    t_res = ms_per_step * 1e-3
    res_traffic = [v["traffic"] for v in kr.values()]
    value = n_sp * 7 / t_res / 1e9
    line = {
        "metric": "P2 residual GDOF/s + %HBM roofline; served solves/h & p99 latency vs plain",
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
        "config": {"workload": f"P{args.order} SST RANS residual (AUSM+up inviscid + BR1 viscous + SST sources and "
                               f"transport), {args.n}^3-hex graded wall box split into {mesh.n_cells} prisms, "
                               f"{n_sp} solution points x 7 variables",
                   "cells": mesh.n_cells, "order": args.order, "n_vars": 7,
                   "state_storage": "FP32 offsets from the free stream (conserved, FP64 reference)"},
        # headline: one full residual evaluation (all seven kernels) against the HBM roofline
        "roofline": {"bound": "hbm", "achieved": b_res / t_res / 1e9, "peak": HBM_PEAK / 1e9, "unit": "GB/s",
                     "frac": b_res / t_res / HBM_PEAK,
                     "traffic": None if any(x is None for x in res_traffic) else float(sum(res_traffic)),
                     "algorithmic_bytes": b_res, "ms": t_res * 1e3,
                     "kernel": "full residual evaluation (" + ", ".join(kr) + ")",
                     "dominant_kernel": dominant},
        "kernel_rooflines": kr,
        "phase_ms": {n: float(t) for n, t in zip(names, ph)},
        "ms_per_rk3_pseudo_step": ms_rk,
        "autotuned_grids": None if tuned is None else {n: v["grid"] for n, v in tuned.items()},
        "autotuned_variants": None if tuned is None else {n: v.get("variant", 0) for n, v in tuned.items()},
    }
    if not args.no_serve:
        del U, R, rates, k, w, nut, mut, dm
        torch.cuda.empty_cache()
        line["served"] = served_leg(args)
    if not args.no_batch:
        line["batched"] = batched_leg(args)
    if not args.no_cpu_baseline:
        log("[bench] CPU baseline (oracle) ...")
        line["cpu_baseline"] = cpu_baseline(args.order, n_sample=args.cpu_sample_n)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
