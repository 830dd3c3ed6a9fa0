"""GPUFRSolver: the steady FR solver on one MI355X.

Drop-in for the reference's ``GPUFRSolver`` (src/autoflowcfd/core/gpu/gpu_solver.py:36-453,
src/autoflowcfd/core/gpu/gpu_solver_io.py:13-84): same constructor arguments, ``step(dt)``
returning the residual RMS, ``solve(...)`` returning the reference's result dict,
``get_state_cpu`` / ``set_state_from_cpu``, ``save_checkpoint`` / ``load_checkpoint`` and
``cleanup``.  The arithmetic follows the reference CPU solver (src/autoflowcfd/core/fr_solver/
step.py:15-221): local time step (cfl.py), SST k-omega step once per pseudo-step, SSP-RK3 mean
flow with positivity and filter after every stage, FP64 RMS norm over all n_vars components.

All state stays in HBM between steps; every step is one sequence of HIP launches on the
current torch stream (libafr.so), and the only host synchronisation per step is the FP64
norm read-back that the ``step`` contract returns.
"""

import os
import time
from dataclasses import dataclass

import numpy as np
import torch

from .boundary import as_mesh, boundary_table
from .device import DeviceMesh, bc_params_table, locality_renumbered
from .mesh import BC_FARFIELD, BC_INLET, BC_OUTLET

GAMMA = 1.4


@dataclass
class SolverResult:
    """src/autoflowcfd/core/fr_solver/state.py:13-17."""
    converged: bool
    iterations: int
    final_residual: float


def _cons(Q):
    rho, u, v, w, p = (Q[..., i] for i in range(5))
    E = p / ((GAMMA - 1.0) * rho) + 0.5 * (u * u + v * v + w * w)
    return np.stack([rho, rho * u, rho * v, rho * w, rho * E], axis=-1)


class GPUFRSolver:
    """Steady P1/P2 FR solver (Euler / Navier-Stokes / SST RANS) on one GPU."""

    def __init__(self, mesh, ops=None, order: int = 2, n_vars: int = 5, device_id: int = 0,
                 time_scheme: str = "ssp_rk3", cfl: float = 1.0, rho_inf: float = 1.225, vel_inf: float = 33.33,
                 p_inf: float = 101325.0, mu_molecular: float = 1.8e-5, boundary_ghost_provider=None,
                 turb_model: str = "NONE", dtype=torch.float32, k_inf: float = 1e-6, omega_inf: float = 1.0,
                 bc_overrides=None, Pr: float = 0.72, Pr_t: float = 0.9, ref_state=None, slots=None,
                 global_time_step: bool = False, storage: str = "locality", offset_storage=None):
        """mesh: autoflowcfd_amd.mesh.Mesh with face records; ops is accepted for signature
        compatibility (the tables are generated for ``order``).  boundary_ghost_provider /
        bc_overrides: tabulated reference states {BC code: (5,) primitive state} (FARFIELD, INLET)
        or {BC_OUTLET: p}; the default is the freestream along +x (boundary.py:52-145).
        k_inf / omega_inf: initial turbulence fields (reference defaults 1e-6 and 1.0, sst.py:47-48).
        slots: precomputed face-slot tables of the mesh at this order (face_slots.slots_for).
        global_time_step: step(dt) advances every solution point by the given dt (the global-dt
        runs of tests/validation/test_tgv.py:70-92) instead of the local time step.
        storage: "locality" stores the cells on the device in the Morton order of their centroids
        (neighbour gathers then share cache lines and pages); "input" keeps the mesh's order.
        Either way every array this class takes or returns (states, residuals, checkpoints) is
        in the mesh's own cell order; ``U`` and ``dm`` are the device-order internals.  With
        storage="locality", ``slots`` must be the tables of the renumbered mesh.
        ref_state: constant primitive state of the perturbation form (default: the FARFIELD state);
        a case stepped alone with the ref_state of a batch reproduces its batched trajectory bit
        for bit (batch.CaseBatch).
        offset_storage: store the state on the device as its offset from a conserved reference
        (DeviceMesh state_ref): True for the FARFIELD state (bc_overrides may set it, e.g. a yawed
        free stream; else rho_inf, vel_inf along +x, p_inf), a (5,) conserved state for that
        state, None for True with FP32 storage (it removes the free-stream rounding floor) and
        False with FP64.  ``U`` then holds the offset; get_state_cpu, set_state_from_cpu,
        checkpoints and the residual entry points take and return the conserved state.
        """
        if time_scheme.lower().replace("-", "_") != "ssp_rk3":
            raise NotImplementedError(f"time_scheme {time_scheme!r}: this backend implements ssp_rk3")
        if not torch.cuda.is_available():
            raise RuntimeError("GPUFRSolver needs a visible GPU (HIP device)")
        self.device = torch.device(f"cuda:{device_id}")
        torch.cuda.set_device(self.device)
        self.mesh, self.ops, self.order, self.device_id = mesh, ops, order, device_id
        self.turb_model_name = str(turb_model).upper()
        if self.turb_model_name not in ("NONE", "SST"):
            raise NotImplementedError(f"turbulence model {turb_model!r}: this backend implements NONE and SST")
        self.n_vars = 7 if self.turb_model_name == "SST" else n_vars
        self.mu_molecular, self.Pr, self.Pr_t = mu_molecular, Pr, Pr_t
        self.cfl = cfl
        self.freestream = {"rho_inf": rho_inf, "vel_inf": vel_inf, "p_inf": p_inf}
        Q_free = np.array([rho_inf, vel_inf, 0.0, 0.0, p_inf])
        mesh = as_mesh(mesh, topology="gpu")
        self.mesh = mesh
        provider = boundary_ghost_provider if hasattr(boundary_ghost_provider, "group_code") else None
        overrides = {} if provider is not None else dict(boundary_ghost_provider or {})
        overrides.update(bc_overrides or {})
        # boundary groups (reference meshes) or a reference provider: one table row per distinct state
        bcp = boundary_table(mesh, rho_inf, vel_inf, p_inf, bc_overrides=overrides, provider=provider)
        if bcp is not None and ref_state is None:
            ref_state = Q_free   # the table's rows follow the groups, not the type codes
        if bcp is None:
            Q_inlet = overrides.get(BC_INLET)
            p_out = overrides.get(BC_OUTLET)
            if BC_FARFIELD in overrides:
                Q_free = np.asarray(overrides[BC_FARFIELD], dtype=np.float64)
            p_out = None if p_out is None else float(np.asarray(p_out).ravel()[-1])
            bcp = bc_params_table(Q_free, Q_inlet, p_out)
        if storage == "locality":
            dev_mesh, new_of_old = locality_renumbered(mesh)
        elif storage == "input":
            dev_mesh, new_of_old = mesh, np.arange(mesh.n_cells)
        else:
            raise ValueError(f"storage {storage!r}: 'locality' or 'input'")
        self._new_of_old = torch.as_tensor(new_of_old, dtype=torch.int64, device=self.device)
        self._old_of_new = torch.as_tensor(np.argsort(new_of_old), dtype=torch.int64, device=self.device)
        U_inf = _cons(np.array([rho_inf, vel_inf, 0.0, 0.0, p_inf]))
        if offset_storage is None:
            offset_storage = dtype == torch.float32
        if isinstance(offset_storage, (bool, np.bool_)):
            Q_far = np.asarray(overrides.get(BC_FARFIELD, Q_free), dtype=np.float64).reshape(5)
            state_ref = _cons(Q_far) if offset_storage else None
        else:
            state_ref = np.asarray(offset_storage, dtype=np.float64).reshape(5)
        self.dm = DeviceMesh(dev_mesh, order, dtype=dtype, device=self.device, bc_params=bcp, ref_state=ref_state,
                             slots=slots, turbulence="SST" if self.turb_model_name == "SST" else None,
                             state_ref=state_ref)
        C, ns = self.dm.n_cells, self.dm.ns
        U0 = torch.as_tensor(np.broadcast_to(U_inf, (C, ns, 5)).copy(), device=self.device)
        self.U = self.dm.to_storage(U0)
        if self.turb_model_name == "SST":
            self.k, self.omega, self.nu_t, self.mu_t = self.dm.sst_fields(k_inf, omega_inf)
        else:
            self.k = self.omega = self.nu_t = self.mu_t = None
        self.residual_history = []
        self.iteration = 0
        self._graph = None
        self._graph_norm = None
        self.global_time_step = bool(global_time_step)
        if self.global_time_step and self.turb_model_name == "SST":
            raise NotImplementedError("global_time_step is implemented for the laminar path")

    # ------------------------------------------------------------------ stepping
    def _launch_step(self):
        """Enqueue one pseudo-step on the current stream; returns the (1,) FP64 device norm."""
        if self.turb_model_name == "SST":
            return self.dm.sst_rk3_step(self.U, self.k, self.omega, self.nu_t, self.mu_t, mu=self.mu_molecular,
                                        Pr=self.Pr, Pr_t=self.Pr_t, n_vars_total=self.n_vars)
        return self.dm.ssp_rk3_step(self.U, None, mu=self.mu_molecular, Pr=self.Pr, Pr_t=self.Pr_t,
                                    n_vars_total=self.n_vars)

    def capture_graph(self, capture_error_mode: str = "global"):
        """Record one pseudo-step into a HIP graph; later steps replay it (one launch per step).

        The pseudo-step has no host synchronisation, no allocation and no memset (the FP64 norm
        partials are fully overwritten), so it captures as is; buffers are the solver's own.
        capture_error_mode "thread_local" lets other host threads keep launching meanwhile."""
        torch.cuda.synchronize(self.device)
        side = torch.cuda.Stream(self.device)
        side.wait_stream(torch.cuda.current_stream(self.device))
        with torch.cuda.stream(side):
            self._launch_step()          # warm-up launch outside the graph (first-use set-up)
        torch.cuda.current_stream(self.device).wait_stream(side)
        torch.cuda.synchronize(self.device)
        # the warm-up advanced the state by one step: count it
        self.residual_history.append(float(self.dm.norm_dev.item()))
        self.iteration += 1
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g, capture_error_mode=capture_error_mode):
            self._graph_norm = self._launch_step()
        self._graph = g
        return g

    def step(self, dt: float = 0.0) -> float:
        """One pseudo-step; returns the residual RMS.  dt is used only with global_time_step
        (every point advances by dt); otherwise the local time step is used."""
        if self.global_time_step:
            norm = self.dm.ssp_rk3_step_dt(self.U, dt, None, mu=self.mu_molecular, Pr=self.Pr, Pr_t=self.Pr_t,
                                           n_vars_total=self.n_vars)
            res = float(norm.item())
            self.residual_history.append(res)
            self.iteration += 1
            return res
        if self._graph is not None:
            self._graph.replay()
            norm = self._graph_norm
        else:
            norm = self._launch_step()
        res = float(norm.item())
        self.residual_history.append(res)
        self.iteration += 1
        return res

    def solve(self, max_iter: int = 1000, dt: float = 1e-4, tol: float = 1e-6, output_interval: int = 10):
        """Steady loop (gpu_solver.py:403-453): stop at res < tol, on a non-finite residual or max_iter."""
        converged = False
        final = float("inf")
        for i in range(max_iter):
            t0 = time.time()
            final = self.step(dt)
            if output_interval and (i == 0 or (i + 1) % output_interval == 0):
                print(f"GPU Iter {i + 1}: Residual = {final:.6e} | Time/step: {time.time() - t0:.3f}s", flush=True)
            if final < tol:
                converged = True
                break
            if not np.isfinite(final):
                break
        return {"converged": converged, "iterations": self.iteration, "final_residual": final,
                "residual_history": self.residual_history}

    # ------------------------------------------------------------------ residuals (reference method names)
    def cell_volumes(self):
        """FP64 cell volumes (numpy), mesh cell order."""
        return self.dm.volume[self._new_of_old.cpu().numpy()]

    def _to_dev(self, x):
        """Mesh cell order -> device cell order (rows)."""
        return x.index_select(0, self._old_of_new).contiguous()

    def _to_user(self, x):
        """Device cell order -> mesh cell order (rows)."""
        return x.index_select(0, self._new_of_old).contiguous()

    def _trial(self, U_trial):
        """The stored state, or the stored form of a conserved trial state given in mesh order."""
        if U_trial is None:
            return self.U
        return self.dm.to_storage(self._to_dev(torch.as_tensor(U_trial, device=self.device)))

    def compute_inviscid_residual_gpu(self, U_trial=None):
        """Inviscid dU/dt (device tensor, mesh cell order) of the state or of U_trial (mesh order)."""
        return self._to_user(self.dm.inviscid_residual(self._trial(U_trial)))

    def compute_viscous_residual_gpu(self, U_trial=None, mu_t_field=None):
        """Viscous dU/dt (device tensor, mesh cell order); U_trial and mu_t_field in mesh order."""
        U = self._trial(U_trial)
        mut = None if mu_t_field is None else self._to_dev(mu_t_field)
        return self._to_user(self.dm.viscous_residual(U, mut, mu=self.mu_molecular, Pr=self.Pr, Pr_t=self.Pr_t))

    # ------------------------------------------------------------------ forces
    def compute_aerodynamic_coefficients(self, reference_area: float, include_viscous: bool = True):
        """Cd, Cl, Cs from the FP64 WALL force integral (fr_coefficients.py:27-152): fixed axes,
        Cd = F_x / (q A), Cl = F_z / (q A), Cs = F_y / (q A), q = rho_inf vel_inf^2 / 2."""
        mu_t = self.mu_t if self.turb_model_name == "SST" else None
        F = self.dm.wall_forces(self.U, mu_t, mu=self.mu_molecular).cpu().numpy()
        f = F[:3] + (F[3:] if include_viscous else 0.0)
        q = 0.5 * self.freestream["rho_inf"] * self.freestream["vel_inf"] ** 2
        d = max(q * reference_area, 1e-300)
        return {"Cd": float(f[0] / d), "Cl": float(f[2] / d), "Cs": float(f[1] / d),
                "F_pressure": F[:3].tolist(), "F_viscous": F[3:].tolist()}

    # ------------------------------------------------------------------ state I/O (gpu_solver_io.py:13-84)
    def _primitives(self, U):
        rho = np.maximum(U[..., 0], 1e-10)
        u, v, w = U[..., 1] / rho, U[..., 2] / rho, U[..., 3] / rho
        p = np.maximum((U[..., 4] - rho * 0.5 * (u * u + v * v + w * w)) * 0.4, 1.0)
        return np.stack([rho, u, v, w, p], axis=-1)

    def get_state_cpu(self):
        """{'U', 'Q'} as FP64 NumPy arrays (n_cells, n_sps, n_vars); with SST the last two
        components are rho k, rho omega (U) and k, omega (Q)."""
        U = self._to_user(self.dm.to_absolute(self.U)).cpu().numpy()
        Q = self._primitives(U)
        if self.n_vars > 5:
            k = self._to_user(self.k).double().cpu().numpy() if self.k is not None else np.zeros(U.shape[:2])
            w = self._to_user(self.omega).double().cpu().numpy() if self.omega is not None else np.zeros(U.shape[:2])
            U = np.concatenate([U, (Q[..., 0] * k)[..., None], (Q[..., 0] * w)[..., None]], axis=-1)
            Q = np.concatenate([Q, k[..., None], w[..., None]], axis=-1)
        return {"U": U, "Q": Q}

    def set_state_from_cpu(self, U_np: np.ndarray):
        U_np = np.asarray(U_np, dtype=np.float64)
        if U_np.shape[:2] != tuple(self.U.shape[:2]):
            raise ValueError(f"state shape {U_np.shape} does not match ({self.U.shape[0]}, {self.U.shape[1]}, ...)")
        dev = lambda a: self._to_dev(torch.as_tensor(np.ascontiguousarray(a), device=self.device))  # noqa: E731
        self.U.copy_(self.dm.to_storage(dev(U_np[..., :5])))
        if U_np.shape[-1] >= 7 and self.k is not None:
            rho = np.maximum(U_np[..., 0], 1e-10)
            self.k.copy_(dev(U_np[..., 5] / rho).to(self.k.dtype))
            self.omega.copy_(dev(U_np[..., 6] / rho).to(self.omega.dtype))

    def save_checkpoint(self, path: str):
        """NumPy .npz with the reference's HDF5 layout (U, Q, k, omega, residual_history + attrs)."""
        st = self.get_state_cpu()
        data = {"U": st["U"], "Q": st["Q"], "iteration": np.int64(self.iteration),
                "n_cells": np.int64(self.dm.n_cells), "n_sps": np.int64(self.dm.ns), "order": np.int64(self.order),
                "time_scheme": np.array("ssp_rk3"), "cfl": np.float64(self.cfl),
                "residual_history": np.asarray(self.residual_history, dtype=np.float64)}
        if self.k is not None:
            data["k"] = self._to_user(self.k).double().cpu().numpy()
            data["omega"] = self._to_user(self.omega).double().cpu().numpy()
            data["nu_t"] = self._to_user(self.nu_t).double().cpu().numpy()
        d = os.path.dirname(os.path.abspath(path))
        os.makedirs(d, exist_ok=True)
        np.savez(path, **data)

    def load_checkpoint(self, path: str):
        with np.load(path, allow_pickle=False) as f:
            if int(f["n_cells"]) != self.dm.n_cells or int(f["n_sps"]) != self.dm.ns:
                raise ValueError("checkpoint does not match this mesh / order")
            dev = lambda a: self._to_dev(torch.as_tensor(np.ascontiguousarray(a), device=self.device))  # noqa: E731
            self.U.copy_(self.dm.to_storage(dev(np.asarray(f["U"][..., :5], dtype=np.float64))))
            self.iteration = int(f["iteration"])
            self.residual_history = f["residual_history"].tolist()
            if self.k is not None and "k" in f:
                self.k.copy_(dev(f["k"]).to(self.k.dtype))
                self.omega.copy_(dev(f["omega"]).to(self.omega.dtype))
                if "nu_t" in f:
                    self.nu_t.copy_(dev(f["nu_t"]).to(self.nu_t.dtype))

    def cleanup(self):
        """Release the device buffers of this solver."""
        for name in ("U", "k", "omega", "nu_t", "mu_t"):
            setattr(self, name, None)
        self.dm = None
        torch.cuda.empty_cache()
