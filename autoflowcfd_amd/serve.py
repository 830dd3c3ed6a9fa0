"""Served steady solves on one GPU: continuous batching of independent requests.

The reference runs one steady solve per process (src/autoflowcfd/cli/solve_steady_command.py:
94-128 -> GPUFRSolver.solve, src/autoflowcfd/core/gpu/gpu_solver.py:403-453).  ``SolveServer``
accepts a stream of requests (mesh, free stream, order) as they arrive and keeps every admitted
solve resident in HBM until it retires:

  * admission: a request is admitted when an active slot is free and its estimated device
    bytes fit the HBM budget (bytes per solution point, measured on the solves already built);
  * set-up overlaps the running solves: the host face-slot tables of every waiting request
    are built ahead in worker processes (NumPy, no GPU), and the device part (tables upload,
    metrics, first pseudo-step, HIP-graph capture of the pseudo-step) runs on a worker
    thread with its own stream;
  * every active solve has its own stream (default priority); one pseudo-step is one graph
    replay followed by an asynchronous copy of its FP64 norm to pinned host memory and an
    event; the scheduler polls the events and relaunches each solve as soon as its previous
    step is done, so the steps of all active solves overlap on the device;
  * a solve retires at ``res / res_0 < rel_tol`` (res_0 = first pseudo-step residual), at a
    non-finite residual or at ``max_iter`` -- decided after every step, exactly as the plain
    loop ``solve_alone`` decides it, so each result is the same alone and in the mix.

``solve_alone`` is the plain one-case-at-a-time GPU path: build, step with a host read-back
of the norm after every step, retire, next request.
"""

import multiprocessing
import time
from concurrent.futures import ProcessPoolExecutor, ThreadPoolExecutor
from dataclasses import dataclass, field
from typing import List

import numpy as np
import torch

from .batch import Case, cons, free_stream_coefficients
from .face_slots import slots_for
from .mesh import BC_FARFIELD, BC_INLET


@dataclass
class SolveRequest:
    """One steady solve: case (mesh + free stream), order, arrival time in seconds after the
    server starts, and its stopping rule."""
    rid: int
    case: Case
    order: int = 2
    arrival: float = 0.0
    max_iter: int = 2000
    rel_tol: float = 1e-6
    stop_nonfinite: bool = True     # False: a fixed step count runs on even through a non-finite residual


@dataclass
class SolveResult:
    rid: int
    order: int
    n_cells: int
    iterations: int
    converged: bool
    final_residual: float
    residual_history: List[float]
    coefficients: dict
    arrival: float
    admitted: float
    finished: float
    setup_s: float = 0.0          # set-up time once started: tables, metrics, first step, graph capture

    @property
    def latency(self):
        return self.finished - self.arrival


def _stop(hist, it, req):
    """(stop, converged) after a step: relative drop, non-finite residual or iteration cap."""
    r = hist[-1]
    if not np.isfinite(r) and req.stop_nonfinite:
        return True, False
    if hist[0] > 0 and r / hist[0] < req.rel_tol:
        return True, True
    return it >= req.max_iter, False


def build_solver(req, dtype=torch.float32, turb_model="NONE", mu=1.8e-5, ref_state=None, slots=None):
    """GPUFRSolver of a request: free stream of the case at every FARFIELD / INLET face and as
    the uniform initial state (``ref_state``: the perturbation form's constant state, default
    the free stream)."""
    from .solver import GPUFRSolver
    c = req.case
    Q = np.asarray(c.Q_free, dtype=np.float64)
    s = GPUFRSolver(c.mesh, order=req.order, rho_inf=Q[0], vel_inf=float(np.linalg.norm(Q[1:4])), p_inf=Q[4],
                    turb_model=turb_model, dtype=dtype, k_inf=c.k_inf, omega_inf=c.omega_inf, mu_molecular=mu,
                    bc_overrides={BC_FARFIELD: Q, BC_INLET: Q}, ref_state=ref_state, slots=slots)
    U0 = np.broadcast_to(cons(Q), (s.dm.n_cells, s.dm.ns, 5))
    s.U.copy_(s.dm.to_storage(torch.as_tensor(np.ascontiguousarray(U0))))
    return s


def _coefficients(solver, case):
    mu_t = solver.mu_t if solver.turb_model_name == "SST" else None
    if not solver.dm.n_wall_rec:
        return {}
    F = solver.dm.wall_forces(solver.U, mu_t, mu=solver.mu_molecular).cpu().numpy()
    return free_stream_coefficients(F, case.Q_free, case.reference_area)


def solve_alone(requests, dtype=torch.float32, turb_model="NONE", mu=1.8e-5, clock=time.perf_counter,
                use_graph=False, ref_state=None):
    """The plain path: requests one at a time in arrival order; each step reads the norm back
    (GPUFRSolver.step, eager launches unless use_graph).  Returns results with arrival / start /
    finish times on one clock."""
    out = []
    t0 = clock()
    for req in sorted(requests, key=lambda r: (r.arrival, r.rid)):
        while clock() - t0 < req.arrival:
            time.sleep(min(0.001, max(req.arrival - (clock() - t0), 0.0)))
        start = clock() - t0
        s = build_solver(req, dtype, turb_model, mu, ref_state=ref_state)
        hist = []
        if use_graph:
            s.capture_graph()
            hist.append(s.residual_history[-1])
        else:
            hist.append(s.step())
        setup = clock() - t0 - start
        while not _stop(hist, len(hist), req)[0]:
            hist.append(s.step())
        conv = _stop(hist, len(hist), req)[1]
        coef = _coefficients(s, req.case)
        torch.cuda.synchronize()
        out.append(SolveResult(req.rid, req.order, s.dm.n_cells, len(hist), conv, hist[-1], hist, coef,
                               req.arrival, start, clock() - t0, setup))
        s.cleanup()
    return out


def _noop(_):
    return None


@dataclass
class _Active:
    req: SolveRequest
    solver: object
    stream: torch.cuda.Stream
    norm_host: torch.Tensor
    event: torch.cuda.Event
    admitted: float
    nbytes: int
    hist: List[float] = field(default_factory=list)
    in_flight: bool = False
    setup_s: float = 0.0


class SolveServer:
    """Continuous batching of steady solves on one GPU (see the module docstring)."""

    def __init__(self, dtype=torch.float32, turb_model="NONE", mu=1.8e-5, max_active=8, hbm_budget=None,
                 use_graph=True, clock=time.perf_counter, host_workers=0, batched=False, ref_state=None,
                 setup_threads=3):
        """batched: step the resident solves of each order as ONE batched mesh assembled on the device
        (batch.ResidentBatch), rebuilt whenever solves join or retire, instead of one graph per solve
        on its own stream; every solve is then built with the common ``ref_state`` (default: the
        static state rho 1.225, p 101325 at rest), and a plain run with the same ref_state gives
        the same results bit for bit.  setup_threads: solves set up concurrently in batched mode (no
        graph capture there, so set-ups may overlap each other)."""
        if not torch.cuda.is_available():
            raise RuntimeError("SolveServer needs a visible GPU (HIP device)")
        self.dtype, self.turb, self.mu = dtype, turb_model, mu
        self.batched = bool(batched)
        self.setup_threads = max(1, int(setup_threads))
        self.ref_state = (np.array([1.225, 0.0, 0.0, 0.0, 101325.0]) if ref_state is None
                          else np.asarray(ref_state, dtype=np.float64))
        self.rebatches = 0
        self.max_active = int(max_active)
        free, _ = torch.cuda.mem_get_info()
        self.hbm_budget = int(0.8 * free) if hbm_budget is None else int(hbm_budget)
        self.use_graph = use_graph
        self.clock = clock
        self.bytes_per_point = {}      # order -> measured device bytes per solution point
        self.peak_active = 0
        self.peak_bytes = 0
        self.loop_rounds = 0           # scheduler rounds, and those in which nothing was ready
        self.idle_rounds = 0
        self.steps_launched = 0
        # host set-up processes: fresh interpreters (spawn) that import NumPy only, never the GPU
        self.host_pool = None
        if host_workers > 0:
            self.host_pool = ProcessPoolExecutor(max_workers=host_workers,
                                                 mp_context=multiprocessing.get_context("spawn"))
            list(self.host_pool.map(_noop, range(host_workers)))     # start the workers now

    def close(self):
        if self.host_pool is not None:
            self.host_pool.shutdown(wait=True)
            self.host_pool = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def estimate_bytes(self, req):
        ns = (req.order + 1) ** 3
        n_pts = req.case.mesh.n_cells * ns
        # before any solve of this order is measured: a generous 2 KiB per solution point
        return int(self.bytes_per_point.get(req.order, 2048.0) * n_pts * 1.1)

    def _prepare(self, req, host_future):
        """Worker thread: build the solve, run its first pseudo-step and capture the step graph."""
        t_start = self.clock()
        slots = host_future.result() if host_future is not None else None
        stream = torch.cuda.Stream()
        before = torch.cuda.memory_allocated()
        with torch.cuda.stream(stream):
            s = build_solver(req, self.dtype, self.turb, self.mu, slots=slots)
            if self.use_graph:
                s.capture_graph(capture_error_mode="thread_local")    # warm-up step counted
                first = s.residual_history[-1]
            else:
                first = s.step()
        torch.cuda.synchronize()
        used = max(torch.cuda.memory_allocated() - before, 1)
        return s, stream, first, used, self.clock() - t_start

    def _prepare_batched(self, req, host_future):
        """Worker thread (batched mode): build the solve with the common ref_state and run its first
        pseudo-step on the thread's stream; no graph (the batch is stepped as a whole)."""
        t_start = self.clock()
        slots = host_future.result() if host_future is not None else None
        stream = torch.cuda.Stream()
        before = torch.cuda.memory_allocated()
        with torch.cuda.stream(stream):
            s = build_solver(req, self.dtype, self.turb, self.mu, ref_state=self.ref_state, slots=slots)
            norm = s._launch_step()
            host = torch.zeros(1, dtype=torch.float64, pin_memory=True)
            host.copy_(norm[:1], non_blocking=True)
        stream.synchronize()
        first = float(host[0])
        s.residual_history.append(first)
        s.iteration += 1
        used = max(torch.cuda.memory_allocated() - before, 1)
        return s, stream, first, used, self.clock() - t_start

    def run(self, requests) -> List[SolveResult]:
        if self.batched:
            return self._run_batched(requests)
        reqs = sorted(requests, key=lambda r: (r.arrival, r.rid))
        t0 = self.clock()
        now = lambda: self.clock() - t0  # noqa: E731
        waiting, active, results = [], [], []
        host = {}                # rid -> future of the host slot tables (or None)
        self._graveyard = []     # retired solvers awaiting release
        i_next = 0
        used_bytes = 0
        prep = None              # (future, request, admitted time, reserved bytes)
        pool = ThreadPoolExecutor(max_workers=1)
        try:
            while len(results) < len(reqs):
                t = now()
                while i_next < len(reqs) and reqs[i_next].arrival <= t:
                    r_new = reqs[i_next]
                    waiting.append(r_new)
                    # host tables built ahead in the worker processes, off the GPU thread
                    host[r_new.rid] = (self.host_pool.submit(slots_for, r_new.case.mesh, r_new.order)
                                       if self.host_pool is not None else None)
                    i_next += 1
                if prep is None and self._graveyard:
                    self._graveyard.clear()      # no capture in flight: release retired solvers
                # admission: one set-up at a time, slots and memory permitting
                if prep is None and waiting and len(active) < self.max_active:
                    need = self.estimate_bytes(waiting[0])
                    if used_bytes + need <= self.hbm_budget or not active:
                        req = waiting.pop(0)
                        used_bytes += need
                        prep = (pool.submit(self._prepare, req, host.pop(req.rid)), req, t, need)
                if prep is not None and prep[0].done():
                    fut, req, t_adm, need = prep
                    prep = None
                    s, stream, first, used, setup = fut.result()
                    ns = (req.order + 1) ** 3
                    self.bytes_per_point[req.order] = max(self.bytes_per_point.get(req.order, 0.0),
                                                          used / (s.dm.n_cells * ns))
                    used_bytes += used - need
                    a = _Active(req, s, stream, torch.zeros(1, dtype=torch.float64, pin_memory=True),
                                torch.cuda.Event(), t_adm, used, hist=[first], setup_s=setup)
                    stop, conv = _stop(a.hist, 1, req)
                    if stop:
                        results.append(self._retire(a, conv, now))
                        used_bytes -= a.nbytes
                    else:
                        active.append(a)
                    self.peak_active = max(self.peak_active, len(active))
                    self.peak_bytes = max(self.peak_bytes, used_bytes)
                progressed = False
                for a in list(active):
                    if a.in_flight:
                        if not a.event.query():
                            continue
                        a.in_flight = False
                        a.hist.append(float(a.norm_host[0]))
                        stop, conv = _stop(a.hist, len(a.hist), a.req)
                        if stop:
                            active.remove(a)
                            results.append(self._retire(a, conv, now))
                            used_bytes -= a.nbytes
                            progressed = True
                            continue
                    self._launch(a)
                    progressed = True
                self.loop_rounds += 1
                if not progressed:
                    self.idle_rounds += 1
                    time.sleep(0.0002)
        finally:
            pool.shutdown(wait=True)
            self._graveyard.clear()
        return sorted(results, key=lambda r: r.rid)

    def _run_batched(self, requests) -> List[SolveResult]:
        """The scheduler of the batched mode: per order, the active solves are one ResidentBatch
        stepped on that order's stream; after every step the per-solve norms come back by one
        asynchronous copy, retiring solves are written back and retired by the plain path's rule,
        and the batch is rebuilt (device concatenation) when solves have joined or left."""
        from .batch import ResidentBatch
        reqs = sorted(requests, key=lambda r: (r.arrival, r.rid))
        t0 = self.clock()
        now = lambda: self.clock() - t0  # noqa: E731
        waiting, results = [], []
        host = {}
        i_next = 0
        used_bytes = 0
        preps = []
        groups = {}
        pool = ThreadPoolExecutor(max_workers=self.setup_threads)

        def group(p):
            if p not in groups:
                groups[p] = {"active": [], "batch": None, "stream": torch.cuda.Stream(), "dirty": False,
                             "event": torch.cuda.Event(), "in_flight": False,
                             "host": torch.zeros(max(self.max_active, 1), dtype=torch.float64, pin_memory=True)}
            return groups[p]

        def n_active():
            return sum(len(g["active"]) for g in groups.values())

        try:
            while len(results) < len(reqs):
                t = now()
                while i_next < len(reqs) and reqs[i_next].arrival <= t:
                    r_new = reqs[i_next]
                    waiting.append(r_new)
                    host[r_new.rid] = (self.host_pool.submit(slots_for, r_new.case.mesh, r_new.order)
                                       if self.host_pool is not None else None)
                    i_next += 1
                while (waiting and len(preps) < self.setup_threads
                       and n_active() + len(preps) < self.max_active):
                    need = self.estimate_bytes(waiting[0])
                    if not (used_bytes + need <= self.hbm_budget or not (n_active() or preps)):
                        break
                    req = waiting.pop(0)
                    used_bytes += need
                    preps.append((pool.submit(self._prepare_batched, req, host.pop(req.rid)), req, t, need))
                for prep in [q for q in preps if q[0].done()]:
                    preps.remove(prep)
                    fut, req, t_adm, need = prep
                    s, _, first, used, setup = fut.result()
                    ns = (req.order + 1) ** 3
                    self.bytes_per_point[req.order] = max(self.bytes_per_point.get(req.order, 0.0),
                                                          used / (s.dm.n_cells * ns))
                    used_bytes += used - need
                    a = _Active(req, s, None, None, None, t_adm, used, hist=[first], setup_s=setup)
                    stop, conv = _stop(a.hist, 1, req)
                    if stop:
                        results.append(self._retire_batched(a, conv, now, torch.cuda.current_stream()))
                        used_bytes -= a.nbytes
                    else:
                        g = group(req.order)
                        g["active"].append(a)
                        g["dirty"] = True
                    self.peak_active = max(self.peak_active, n_active())
                    self.peak_bytes = max(self.peak_bytes, used_bytes)
                progressed = False
                for p, g in groups.items():
                    if g["in_flight"]:
                        if not g["event"].query():
                            continue
                        g["in_flight"] = False
                        b = g["batch"]
                        for i, a in enumerate(b.members_active):
                            a.hist.append(float(g["host"][i]))
                            stop, conv = _stop(a.hist, len(a.hist), a.req)
                            if stop:
                                with torch.cuda.stream(g["stream"]):
                                    b.write_back([i])
                                results.append(self._retire_batched(a, conv, now, g["stream"]))
                                used_bytes -= a.nbytes
                                g["active"].remove(a)
                                g["dirty"] = True
                        progressed = True
                    if g["dirty"]:
                        with torch.cuda.stream(g["stream"]):
                            if g["batch"] is not None:
                                old = g["batch"]
                                stay = [i for i, a in enumerate(old.members_active) if a in g["active"]]
                                old.write_back(stay)
                            g["batch"] = None
                            if g["active"]:
                                b = ResidentBatch([a.solver for a in g["active"]], self.ref_state)
                                b.members_active = list(g["active"])
                                g["batch"] = b
                                self.rebatches += 1
                        g["dirty"] = False
                    if g["batch"] is not None and not g["in_flight"]:
                        b = g["batch"]
                        n = len(b.members_active)
                        with torch.cuda.stream(g["stream"]):
                            b.launch()
                            g["host"][:n].copy_(b.norms, non_blocking=True)
                            g["event"].record(g["stream"])
                        g["in_flight"] = True
                        self.steps_launched += n
                        progressed = True
                self.loop_rounds += 1
                if not progressed:
                    self.idle_rounds += 1
                    time.sleep(0.0002)
        finally:
            pool.shutdown(wait=True)
        return sorted(results, key=lambda r: r.rid)

    def _retire_batched(self, a, converged, now, stream):
        s = a.solver
        with torch.cuda.stream(stream):
            coef = _coefficients(s, a.req.case)
        stream.synchronize()
        s.residual_history = list(a.hist)
        return SolveResult(a.req.rid, a.req.order, s.dm.n_cells, len(a.hist), converged, a.hist[-1], list(a.hist),
                           coef, a.req.arrival, a.admitted, now(), a.setup_s)

    def _launch(self, a):
        """Enqueue one pseudo-step of an active solve on its stream, with the norm read-back."""
        s = a.solver
        with torch.cuda.stream(a.stream):
            if s._graph is not None:
                s._graph.replay()
                norm = s._graph_norm
            else:
                norm = s._launch_step()
            a.norm_host.copy_(norm[:1], non_blocking=True)
            a.event.record(a.stream)
        s.iteration += 1
        a.in_flight = True
        self.steps_launched += 1

    def _retire(self, a, converged, now):
        s = a.solver
        with torch.cuda.stream(a.stream):
            coef = _coefficients(s, a.req.case)
        a.stream.synchronize()
        s.residual_history = list(a.hist)
        r = SolveResult(a.req.rid, a.req.order, s.dm.n_cells, len(a.hist), converged, a.hist[-1], list(a.hist), coef,
                        a.req.arrival, a.admitted, now(), a.setup_s)
        # the solver (device buffers, HIP graph) is released later by the scheduler, at a point
        # where no graph capture runs on the set-up thread: destroying a graph during another
        # thread's capture is not permitted
        self._graveyard.append(s)
        return r


def latency_stats(results):
    """p50 / p99 latency (s) and solves per hour over the span first arrival -> last finish."""
    lat = np.array([r.latency for r in results])
    span = max(r.finished for r in results) - min(r.arrival for r in results)
    return {"p50_s": float(np.percentile(lat, 50)), "p99_s": float(np.percentile(lat, 99)),
            "solves_per_hour": float(len(results) / max(span, 1e-9) * 3600.0), "span_s": float(span)}


def synthetic_requests(n_requests, seed=0, hex_range=(2e3, 2e4), orders=(1, 2), mach_range=(0.05, 0.3),
                       yaw_range=(0.0, 10.0), aspect_range=(2.0, 6.0), rate=4.0, max_iter=2000, rel_tol=1e-6,
                       length=0.2879, rho=1.225, p_inf=101325.0, topology="host", stop_nonfinite=True):
    """A seeded stream of steady wall-box solves: hex count log-uniform over hex_range, aspect
    ratio (box length / height) uniform, y-graded wall boxes (ratio 1.1, no-slip y = 0, inlet,
    outlet, symmetry) split into prisms, order drawn uniformly from ``orders``, Mach and yaw
    uniform (free stream (U cos yaw, 0, U sin yaw)), Poisson arrivals at ``rate`` per second.
    k, omega of every case from 1% turbulence intensity and viscosity ratio 10.  topology: "host" or
    "gpu" face topology of the meshes (the same records)."""
    from .mesh import BC_INLET, BC_OUTLET, BC_SYMMETRY, BC_WALL, hex_box
    rng = np.random.default_rng(seed)
    bc = {"y0": BC_WALL, "x0": BC_INLET, "x1": BC_OUTLET, "y1": BC_SYMMETRY, "z0": BC_SYMMETRY, "z1": BC_SYMMETRY}
    a_inf = float(np.sqrt(1.4 * p_inf / rho))
    t = 0.0
    out = []
    for j in range(n_requests):
        H = float(np.exp(rng.uniform(np.log(hex_range[0]), np.log(hex_range[1]))))
        ar = float(rng.uniform(*aspect_range))
        m = max(int(round((H / ar) ** (1.0 / 3.0))), 2)
        nx = max(int(round(ar * m)), 2)
        order = int(rng.choice(orders))
        mach = float(rng.uniform(*mach_range))
        yaw = np.deg2rad(float(rng.uniform(*yaw_range)))
        U = mach * a_inf
        Q = np.array([rho, U * np.cos(yaw), 0.0, U * np.sin(yaw), p_inf])
        k = 1.5 * (0.01 * U) ** 2
        mesh = hex_box(nx, m, m, ar * length, length, length, split="prism", grading_y=1.1, seed=seed * 1000 + j,
                       bc_by_plane=bc, topology=topology)
        case = Case(mesh=mesh, Q_free=Q, k_inf=k, omega_inf=rho * k / (10 * 1.8e-5), reference_area=length * length,
                    name=f"wallbox{nx}x{m}x{m}")
        out.append(SolveRequest(rid=j, case=case, order=order, arrival=t, max_iter=max_iter, rel_tol=rel_tol,
                                stop_nonfinite=stop_nonfinite))
        t += float(rng.exponential(1.0 / rate))
    return out
