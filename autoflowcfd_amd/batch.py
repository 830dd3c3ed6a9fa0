"""Batched multi-case form: independent steady cases of one order stepped as one mesh.

The reference solves one case per process (src/autoflowcfd/core/gpu/gpu_solver.py:36-453, one
``GPUFRSolver`` per mesh).  Small cases leave most of an MI355X idle: a P2 case of 16k hex
cells is ~130 workgroups per launch against 256 CUs.  ``CaseBatch`` packs many cases into the
disjoint union of their meshes (``mesh.concat_meshes``): cells are numbered prisms first then
tets, case by case (CSR-style offsets ``prism_off`` / ``tet_off`` / wall-record offsets), so
every kernel of a pseudo-step is ONE launch over all cases, and one HIP graph replays a whole
batched pseudo-step.

What stays per case:
  - boundary states: a (n_cases, 8, 5) table indexed by the per-cell case id at every ghost
    state (FARFIELD / INLET / OUTLET of each case's own free stream, e.g. its Mach and yaw);
  - initial state and SST k, omega of each case's free stream;
  - wall distance (computed case by case: cases may overlap in space);
  - residual norm (FP64, reduced case by case in the order of a single-case reduction) and
    wall forces (per-case FP64 sums over the case's own wall records).
What is common: order, storage precision, viscosity, Pr, Pr_t and the constant reference state
of the perturbation form.  With the same reference state, a case stepped alone
(``GPUFRSolver(..., ref_state=batch.ref_state)``) follows its batched trajectory bit for bit:
every kernel computes each cell from that cell's own data, so packing changes only where a
cell lives, never what is computed for it.
"""

from dataclasses import dataclass

import numpy as np
import torch

from .device import DeviceMesh, bc_params_table
from .face_slots import concat_slots
from .fr_tables import tables
from .mesh import concat_meshes, locality_renumbered

GAMMA = 1.4


@dataclass
class Case:
    """One steady case: a mesh with face records and its free stream (primitive rho, u, v, w, p)."""
    mesh: object
    Q_free: np.ndarray
    k_inf: float = 1e-6
    omega_inf: float = 1.0
    reference_area: float = 1.0
    name: str = ""


def cons(Q):
    """Primitive (..., 5) -> conserved (..., 5)."""
    Q = np.asarray(Q, dtype=np.float64)
    rho, u, v, w, p = (Q[..., i] for i in range(5))
    E = p / ((GAMMA - 1.0) * rho) + 0.5 * (u * u + v * v + w * w)
    return np.stack([rho, rho * u, rho * v, rho * w, rho * E], axis=-1)


def static_ref_state(cases):
    """Common reference state of a batch: mean free-stream density and pressure at rest.

    The perturbation form needs one constant state per mesh; leaving the velocity out keeps it
    equally close to every case whatever its speed and direction."""
    Q = np.array([np.asarray(c.Q_free, dtype=np.float64) for c in cases])
    return np.array([Q[:, 0].mean(), 0.0, 0.0, 0.0, Q[:, 4].mean()])


def free_stream_coefficients(F, Q_free, reference_area, include_viscous=True):
    """Cd, Cl, Cs from a (6,) FP64 wall force (fixed axes; src/autoflowcfd/postprocess/
    fr_coefficients.py:27-152): Cd = F_x / (q A), Cl = F_z / (q A), Cs = F_y / (q A),
    q = rho |u_inf|^2 / 2."""
    Q = np.asarray(Q_free, dtype=np.float64)
    q = 0.5 * Q[0] * float(np.sum(Q[1:4] ** 2))
    f = F[:3] + (F[3:] if include_viscous else 0.0)
    d = max(q * reference_area, 1e-300)
    return {"Cd": float(f[0] / d), "Cl": float(f[2] / d), "Cs": float(f[1] / d),
            "F_pressure": F[:3].tolist(), "F_viscous": F[3:].tolist()}


def case_slots(meshes, orders, host_workers=0):
    """Slot tables of every case in the locality order GPUFRSolver stores it: on the GPU
    (slots_gpu.build_slots_gpu) by default, or with NumPy (face_slots.slots_for) in ``host_workers``
    spawned processes; the tables are the same bit for bit either way."""
    from .face_slots import slots_for
    if not host_workers:
        from .fr_tables import tables
        from .mesh import locality_renumbered
        from .slots_gpu import build_slots_gpu
        return [build_slots_gpu(locality_renumbered(m)[0], tables(p)) for m, p in zip(meshes, orders)]
    if len(meshes) > 1:
        import multiprocessing
        from concurrent.futures import ProcessPoolExecutor
        with ProcessPoolExecutor(max_workers=host_workers, mp_context=multiprocessing.get_context("spawn")) as pool:
            return list(pool.map(slots_for, meshes, orders))
    return [slots_for(m, p) for m, p in zip(meshes, orders)]


class CaseBatch:
    """Many independent steady cases of one order on one GPU, one launch per kernel per stage."""

    def __init__(self, cases, order=2, dtype=torch.float32, turb_model="SST", mu_molecular=1.8e-5, Pr=0.72,
                 Pr_t=0.9, device_id=0, ref_state=None, slots_list=None, host_workers=0, offset_storage=None):
        """slots_list: each case's slot tables at this order, else built here: on the GPU, or in
        ``host_workers`` spawned processes when > 0 (NumPy only, never the GPU).
        offset_storage: store each case's state as its offset from the case's free stream
        (DeviceMesh state_ref); None: with FP32 storage.  ``U`` then holds the offsets;
        case_state / set_case_state take and return the conserved state."""
        if not torch.cuda.is_available():
            raise RuntimeError("CaseBatch needs a visible GPU (HIP device)")
        if not cases:
            raise ValueError("empty batch")
        self.cases = list(cases)
        self.n_cases = len(self.cases)
        self.order, self.dtype = order, dtype
        self.device = torch.device(f"cuda:{device_id}")
        torch.cuda.set_device(self.device)
        self.turb = str(turb_model).upper()
        if self.turb not in ("NONE", "SST"):
            raise NotImplementedError(f"turbulence model {turb_model!r}: this backend implements NONE and SST")
        self.n_vars = 7 if self.turb == "SST" else 5
        self.mu, self.Pr, self.Pr_t = mu_molecular, Pr, Pr_t
        self.ref_state = static_ref_state(self.cases) if ref_state is None else np.asarray(ref_state, np.float64)
        # every case stored in the locality order GPUFRSolver uses for it alone
        parts, self._new_of_old = [], []
        for c in self.cases:
            m, n_o = locality_renumbered(c.mesh)
            parts.append(m)
            self._new_of_old.append(n_o)
        mesh, cell_case, po, to, _ = concat_meshes(parts)
        self.mesh, self.cell_case, self.prism_off, self.tet_off = mesh, cell_case, po, to
        bcp = np.stack([bc_params_table(np.asarray(c.Q_free, dtype=np.float64)) for c in self.cases])
        # slot tables case by case: each cell gets exactly the rows its case alone has
        T = tables(order)
        if slots_list is None:
            slots_list = case_slots([c.mesh for c in self.cases], [order] * self.n_cases, host_workers)
        slots = concat_slots(list(slots_list), [p.n_prism for p in parts], [p.n_tet for p in parts], T.nfp)
        Ucase = cons(np.array([c.Q_free for c in self.cases], dtype=np.float64))       # (n_cases, 5)
        if offset_storage is None:
            offset_storage = dtype == torch.float32
        self.dm = DeviceMesh(mesh, order, dtype=dtype, device=self.device, bc_params=bcp, ref_state=self.ref_state,
                             slots=slots, turbulence="SST" if self.turb == "SST" else None, cell_case=cell_case,
                             state_ref=Ucase if offset_storage else None)
        ns = self.dm.ns
        U0 = np.ascontiguousarray(np.broadcast_to(Ucase[cell_case][:, None, :], (mesh.n_cells, ns, 5)))
        self.U = self.dm.to_storage(torch.as_tensor(U0))
        if self.turb == "SST":
            kc = np.array([c.k_inf for c in self.cases])[cell_case]
            wc = np.array([c.omega_inf for c in self.cases])[cell_case]
            shape = (mesh.n_cells, ns)
            self.k = torch.as_tensor(np.broadcast_to(kc[:, None], shape).copy(), dtype=dtype).to(self.device)
            self.omega = torch.as_tensor(np.broadcast_to(wc[:, None], shape).copy(), dtype=dtype).to(self.device)
            self.nu_t = torch.zeros(shape, dtype=dtype, device=self.device)
            self.mu_t = torch.zeros(shape, dtype=dtype, device=self.device)
        else:
            self.k = self.omega = self.nu_t = self.mu_t = None
        self.norms = torch.zeros((self.n_cases,), dtype=torch.float64, device=self.device)
        self.forces = torch.zeros((self.n_cases, 6), dtype=torch.float64, device=self.device)
        self.history = []
        self.iteration = 0
        self._graph = None

    # ------------------------------------------------------------------ stepping
    def _launch_step(self):
        """Enqueue one batched pseudo-step on the current stream (per-case norms into self.norms)."""
        if self.turb == "SST":
            return self.dm.sst_rk3_step(self.U, self.k, self.omega, self.nu_t, self.mu_t, mu=self.mu, Pr=self.Pr,
                                        Pr_t=self.Pr_t, n_vars_total=self.n_vars, norm=self.norms)
        return self.dm.ssp_rk3_step(self.U, None, mu=self.mu, Pr=self.Pr, Pr_t=self.Pr_t, n_vars_total=self.n_vars,
                                    norm=self.norms)

    def tune(self, reps=1, candidates=None):
        """Autotune this batch's kernels (tile variant and persistent grid per kernel, for its order,
        precision and size) on copies of the state: the batch's own fields are not touched, and
        results do not depend on the choice.  Call before capture_graph (a graph keeps the launch
        shapes it was captured with)."""
        from .autotune import autotune
        U = self.U.clone()
        sst = None
        if self.turb == "SST":
            sst = tuple(x.clone() for x in (self.k, self.omega, self.nu_t, self.mu_t))
        return autotune(self.dm, U, None if sst is None else sst[3], mu=self.mu, sst_fields=sst, reps=reps,
                        candidates=candidates)

    def capture_graph(self):
        """Record one batched pseudo-step into a HIP graph (after one eager warm-up step, counted)."""
        torch.cuda.synchronize(self.device)
        side = torch.cuda.Stream(self.device)
        side.wait_stream(torch.cuda.current_stream(self.device))
        with torch.cuda.stream(side):
            self._launch_step()
        torch.cuda.current_stream(self.device).wait_stream(side)
        torch.cuda.synchronize(self.device)
        self.history.append(self.norms.cpu().numpy().copy())
        self.iteration += 1
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g):
            self._launch_step()
        self._graph = g
        return g

    def step(self, sync=True):
        """One pseudo-step of every case; returns the per-case residual norms (host, FP64) when
        sync, else None (the norms stay in self.norms on the device)."""
        if self._graph is not None:
            self._graph.replay()
        else:
            self._launch_step()
        self.iteration += 1
        if not sync:
            return None
        n = self.norms.cpu().numpy().copy()
        self.history.append(n)
        return n

    def run(self, n_steps, sync_every=1):
        """n_steps pseudo-steps; the norms are read back every ``sync_every`` steps."""
        for i in range(n_steps):
            self.step(sync=((i + 1) % sync_every == 0) or i == n_steps - 1)
        return np.array(self.history)

    # ------------------------------------------------------------------ per-case results
    def case_cells(self, i):
        """Cell indices of case i in the batched mesh, in the case mesh's own cell order."""
        P = self.mesh.n_prism
        dev = np.concatenate([np.arange(self.prism_off[i], self.prism_off[i + 1]),
                              P + np.arange(self.tet_off[i], self.tet_off[i + 1])])
        return dev[self._new_of_old[i]]

    def case_state(self, i):
        """Conserved state (C_i, ns, 5) of case i and, with SST, k and omega (host, FP64), in the
        case mesh's own cell order."""
        idx = torch.as_tensor(self.case_cells(i), device=self.device)
        out = {"U": (self.U.index_select(0, idx).double() + self._case_offset(i)).cpu().numpy()}
        if self.turb == "SST":
            out["k"] = self.k.index_select(0, idx).double().cpu().numpy()
            out["omega"] = self.omega.index_select(0, idx).double().cpu().numpy()
        return out

    def set_case_state(self, i, U):
        """Set the conserved state (C_i, ns, 5) of case i (the case mesh's own cell order)."""
        idx = torch.as_tensor(self.case_cells(i), device=self.device)
        X = torch.as_tensor(np.ascontiguousarray(U, dtype=np.float64)).to(self.device) - self._case_offset(i)
        self.U.index_copy_(0, idx, X.to(self.dtype))

    def _case_offset(self, i):
        """(5,) FP64 device offset of case i's stored state (zeros without offsets)."""
        if self.dm.state_ref is None:
            return torch.zeros((5,), dtype=torch.float64, device=self.device)
        return self.dm.state_ref_dev[i]

    def wall_forces(self):
        """(n_cases, 6) FP64 host array: pressure force, viscous force of every case."""
        mu_t = self.mu_t if self.turb == "SST" else None
        self.dm.wall_forces_cases(self.U, mu_t, mu=self.mu, out=self.forces)
        return self.forces.cpu().numpy()

    def coefficients(self, include_viscous=True):
        F = self.wall_forces()
        return [free_stream_coefficients(F[i], c.Q_free, c.reference_area, include_viscous)
                for i, c in enumerate(self.cases)]


class MixedBatch:
    """Cases of different orders (P1 and P2) stepped together.

    Cases are grouped by order into one ``CaseBatch`` each (its kernels are specialised for one
    order: per-order register and LDS budgets, so P1 work is not run at P2 occupancy), all with one
    common reference state, and ONE HIP graph per pseudo-step holds every group's step on its own
    branch: the branches fork from the capturing stream onto default-priority side streams and
    join back, so the order groups run concurrently on the device and a whole mixed pseudo-step is
    one graph launch.  ``order_of[i]`` is case i's order; results come back in case order.  A case
    stepped alone with ``ref_state`` follows its trajectory here bit for bit."""

    def __init__(self, cases, orders, dtype=torch.float32, turb_model="SST", mu_molecular=1.8e-5, Pr=0.72,
                 Pr_t=0.9, device_id=0, ref_state=None, host_workers=0):
        if len(cases) != len(orders) or not cases:
            raise ValueError("one order per case, at least one case")
        self.cases, self.order_of = list(cases), [int(p) for p in orders]
        self.n_cases = len(self.cases)
        self.device = torch.device(f"cuda:{device_id}")
        self.ref_state = static_ref_state(self.cases) if ref_state is None else np.asarray(ref_state, np.float64)
        self.members = {}
        self.groups = {}
        slots = case_slots([c.mesh for c in self.cases], self.order_of, host_workers)
        for p in sorted(set(self.order_of)):
            idx = [i for i, q in enumerate(self.order_of) if q == p]
            self.members[p] = idx
            self.groups[p] = CaseBatch([self.cases[i] for i in idx], order=p, dtype=dtype, turb_model=turb_model,
                                       mu_molecular=mu_molecular, Pr=Pr, Pr_t=Pr_t, device_id=device_id,
                                       ref_state=self.ref_state, slots_list=[slots[i] for i in idx])
        self.history = []
        self.iteration = 0
        self._graph = None
        self._side = {}

    def _launch_step(self):
        """Enqueue one pseudo-step of every group: the first on the current stream, the others on
        side streams forked from it and joined back."""
        cur = torch.cuda.current_stream(self.device)
        orders = sorted(self.groups)
        for p in orders[1:]:
            s = self._side.setdefault(p, torch.cuda.Stream(self.device))
            s.wait_stream(cur)
            with torch.cuda.stream(s):
                self.groups[p]._launch_step()
        self.groups[orders[0]]._launch_step()
        for p in orders[1:]:
            cur.wait_stream(self._side[p])

    def tune(self, reps=1, candidates=None):
        """Autotune every order group on its own mesh (CaseBatch.tune); {order: choices}."""
        return {p: b.tune(reps, candidates) for p, b in self.groups.items()}

    def capture_graph(self):
        """One eager warm-up step (counted), then record one mixed pseudo-step into a HIP graph."""
        torch.cuda.synchronize(self.device)
        side = torch.cuda.Stream(self.device)
        side.wait_stream(torch.cuda.current_stream(self.device))
        with torch.cuda.stream(side):
            self._launch_step()
        torch.cuda.current_stream(self.device).wait_stream(side)
        torch.cuda.synchronize(self.device)
        self.history.append(self.norms())
        self.iteration += 1
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g):
            self._launch_step()
        self._graph = g
        return g

    def norms(self):
        """Per-case FP64 residual norms of the last step (host), in case order."""
        out = np.zeros(self.n_cases)
        for p, b in self.groups.items():
            out[self.members[p]] = b.norms.cpu().numpy()
        return out

    def step(self, sync=True):
        if self._graph is not None:
            self._graph.replay()
        else:
            self._launch_step()
        self.iteration += 1
        if not sync:
            return None
        n = self.norms()
        self.history.append(n)
        return n

    def run(self, n_steps, sync_every=1):
        for i in range(n_steps):
            self.step(sync=((i + 1) % sync_every == 0) or i == n_steps - 1)
        return np.array(self.history)

    def _where(self, i):
        p = self.order_of[i]
        return self.groups[p], self.members[p].index(i)

    def case_state(self, i):
        b, j = self._where(i)
        return b.case_state(j)

    def set_case_state(self, i, U):
        b, j = self._where(i)
        b.set_case_state(j, U)

    def coefficients(self, include_viscous=True):
        out = [None] * self.n_cases
        for p, b in self.groups.items():
            for j, c in enumerate(b.coefficients(include_viscous)):
                out[self.members[p][j]] = c
        return out


class ResidentBatch:
    """Resident solves of one order (GPUFRSolver objects, each holding its own device mesh and
    state) stepped as ONE batched mesh assembled on the device (DeviceMesh.concat): nothing is
    rebuilt on the host, so membership can change as solves arrive and retire.  The members' states
    are gathered into the batch's buffers at construction and written back to each member by
    ``write_back`` (before the batch is replaced, or when a member retires).  Every member must have
    been built with the same ``ref_state``; each then follows its trajectory alone bit for bit."""

    def __init__(self, solvers, ref_state):
        from .device import DeviceMesh
        if not solvers:
            raise ValueError("empty batch")
        self.members = list(solvers)
        s0 = self.members[0]
        self.turb = s0.turb_model_name
        self.order, self.device = s0.order, s0.device
        self.mu, self.Pr, self.Pr_t, self.n_vars = s0.mu_molecular, s0.Pr, s0.Pr_t, s0.n_vars
        for s in self.members:
            if (s.turb_model_name, s.order, s.mu_molecular, s.Pr, s.Pr_t) != (self.turb, self.order, self.mu,
                                                                                self.Pr, self.Pr_t):
                raise ValueError("members differ in turbulence model, order, viscosity or Prandtl numbers")
        bcp = np.stack([s.dm.bc_params.cpu().numpy().reshape(8, 5) for s in self.members])
        self.dm = DeviceMesh.concat([s.dm for s in self.members], bcp, ref_state)
        self.idx = [self.dm.part_cells(i) for i in range(len(self.members))]
        self.U = self._gather(lambda s: s.U)
        if self.turb == "SST":
            self.k, self.omega = self._gather(lambda s: s.k), self._gather(lambda s: s.omega)
            self.nu_t, self.mu_t = self._gather(lambda s: s.nu_t), self._gather(lambda s: s.mu_t)
        self.norms = torch.zeros((len(self.members),), dtype=torch.float64, device=self.device)
        self._graph = None

    def _gather(self, get):
        x0 = get(self.members[0])
        out = torch.empty((self.dm.n_cells,) + tuple(x0.shape[1:]), dtype=x0.dtype, device=self.device)
        for i, s in enumerate(self.members):
            out.index_copy_(0, self.idx[i], get(s))
        return out

    def write_back(self, which=None):
        """Copy the batch's state of the members ``which`` (default all) into their own buffers."""
        for i in (range(len(self.members)) if which is None else which):
            s, ix = self.members[i], self.idx[i]
            s.U.copy_(self.U.index_select(0, ix))
            if self.turb == "SST":
                for name in ("k", "omega", "nu_t", "mu_t"):
                    getattr(s, name).copy_(getattr(self, name).index_select(0, ix))

    def _launch_step(self):
        if self.turb == "SST":
            return self.dm.sst_rk3_step(self.U, self.k, self.omega, self.nu_t, self.mu_t, mu=self.mu, Pr=self.Pr,
                                        Pr_t=self.Pr_t, n_vars_total=self.n_vars, norm=self.norms)
        return self.dm.ssp_rk3_step(self.U, None, mu=self.mu, Pr=self.Pr, Pr_t=self.Pr_t, n_vars_total=self.n_vars,
                                    norm=self.norms)

    def capture_graph(self, capture_error_mode="thread_local"):
        """Record one batched pseudo-step into a HIP graph (nothing is run: every step counts)."""
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g, capture_error_mode=capture_error_mode):
            self._launch_step()
        self._graph = g
        return g

    def launch(self):
        """Enqueue one pseudo-step of every member on the current stream (per-member norms in
        self.norms)."""
        if self._graph is not None:
            self._graph.replay()
        else:
            self._launch_step()
        for s in self.members:
            s.iteration += 1
