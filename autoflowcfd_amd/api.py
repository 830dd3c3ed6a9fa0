"""run_steady: the AutoFlowCFDAPI.run_steady surface (src/autoflowcfd/api.py:123-198) on the GPU backend.

The mesh argument is the reference's ``VolumeMeshData`` (read by its attributes, with its boundary
groups: autoflowcfd_amd.boundary), an autoflowcfd_amd.mesh.Mesh, or a path readable by
autoflowcfd_amd.meshio.load_mesh.  With SST the wall distance is taken to the wall nodes of the
WALL groups (src/autoflowcfd/cli/solve_wall_distance.py:30-60).  Returns a SolverResult; the solver
is kept on the returned object as ``result.solver`` for force / state queries.

For P >= 2 the solve runs the order continuation P0 -> P1 -> .. -> P (laminar and SST), the
reference's default in FRSolver.solve (src/autoflowcfd/core/fr_solver/solver.py:236,292);
order_continuation=False solves at the target order from the free stream.
"""

from .solver import GPUFRSolver, SolverResult


def run_steady(volume_mesh, backend: str = "gpu", order: int = 2, turbulence_model: str = "sst",
               max_iter: int = 1000, dt: float = 1e-4, tol: float = 1e-6, rho_inf: float = 1.225,
               vel_inf: float = 33.33, p_inf: float = 101325.0, threads: int = -1, output_dir: str = "./results",
               **kwargs) -> SolverResult:
    if backend != "gpu":
        raise ValueError("this package is the GPU backend; backend must be 'gpu'")
    if isinstance(volume_mesh, str):
        from .meshio import load_mesh
        volume_mesh = load_mesh(volume_mesh)
    from .boundary import as_mesh
    volume_mesh = as_mesh(volume_mesh, topology="gpu")
    tm = {"none": "NONE", "sst": "SST"}.get(str(turbulence_model).lower())
    if tm is None:
        raise NotImplementedError(f"turbulence model {turbulence_model!r}: this backend implements none and sst")
    cont = kwargs.pop("order_continuation", None)
    if cont is None:
        cont = order >= 2

    def make(p):
        return GPUFRSolver(volume_mesh, order=p, rho_inf=rho_inf, vel_inf=vel_inf, p_inf=p_inf, turb_model=tm,
                           **kwargs)
    if cont and order >= 1:
        from .continuation import solve_with_order_continuation
        solver, out = solve_with_order_continuation(make, order, max_iter=max_iter, tol=tol)
    else:
        solver = make(order)
        out = solver.solve(max_iter=max_iter, dt=dt, tol=tol)
    res = SolverResult(converged=out["converged"], iterations=out["iterations"], final_residual=out["final_residual"])
    res.solver = solver
    return res
