"""Command line: ``python -m autoflowcfd_amd solve steady INPUT --backend gpu ...``.

Flags mirror src/autoflowcfd/cli/solve_steady_command.py:24-40 for the single-GPU path
(INPUT: a .npz mesh or wallbox:N, see autoflowcfd_amd.meshio).
"""

import argparse
import json
import os
import sys


def main(argv=None):
    ap = argparse.ArgumentParser(prog="autoflowcfd_amd")
    sub = ap.add_subparsers(dest="cmd", required=True)
    solve = sub.add_parser("solve").add_subparsers(dest="mode", required=True)
    st = solve.add_parser("steady", help="steady FR solve on one MI355X")
    st.add_argument("input_file")
    st.add_argument("--backend", choices=["gpu"], default="gpu")
    st.add_argument("--order", type=int, default=2, choices=[1, 2])
    st.add_argument("--turbulence-model", choices=["none", "sst"], default="sst")
    st.add_argument("--max-iter", type=int, default=1000)
    st.add_argument("--tol", type=float, default=1e-6)
    st.add_argument("--output", "-o", dest="output_dir", default="./results")
    st.add_argument("--checkpoint-interval", type=int, default=100)
    st.add_argument("--reference-area", type=float, default=None)
    st.add_argument("--gpu-device", type=int, default=0)
    st.add_argument("--rho-inf", type=float, default=1.225)
    st.add_argument("--vel-inf", type=float, default=33.33)
    st.add_argument("--p-inf", type=float, default=101325.0)
    st.add_argument("--k-inf", type=float, default=1e-6)
    st.add_argument("--omega-inf", type=float, default=1.0)
    args = ap.parse_args(argv)

    from .meshio import load_mesh
    from .solver import GPUFRSolver
    mesh = load_mesh(args.input_file)
    print(f"Mesh: {mesh.n_cells} cells ({mesh.n_prism} prisms, {mesh.n_tet} tets) | Backend: gpu | "
          f"Order: P{args.order} | Method: rk3", flush=True)
    solver = GPUFRSolver(mesh, order=args.order, device_id=args.gpu_device, rho_inf=args.rho_inf,
                         vel_inf=args.vel_inf, p_inf=args.p_inf,
                         turb_model="SST" if args.turbulence_model == "sst" else "NONE",
                         k_inf=args.k_inf, omega_inf=args.omega_inf)
    os.makedirs(args.output_dir, exist_ok=True)
    done = 0
    result = None
    while done < args.max_iter:
        chunk = min(args.checkpoint_interval or args.max_iter, args.max_iter - done)
        result = solver.solve(max_iter=chunk, tol=args.tol)
        done += chunk
        solver.save_checkpoint(os.path.join(args.output_dir, "checkpoint.npz"))
        if result["converged"] or not result["final_residual"] < float("inf"):
            break
    summary = {"converged": result["converged"], "iterations": result["iterations"],
               "final_residual": result["final_residual"]}
    if args.reference_area:
        summary.update(solver.compute_aerodynamic_coefficients(args.reference_area))
    print(json.dumps(summary), flush=True)
    return 0 if result["converged"] or result["final_residual"] < float("inf") else 1


if __name__ == "__main__":
    sys.exit(main())
