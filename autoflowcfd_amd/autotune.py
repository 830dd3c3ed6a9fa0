"""Autotuner of the launch shapes of the residual kernels on one mesh.

Every residual kernel walks its cells with a persistent grid; how many workgroups that grid
has trades latency hiding against tail effects and L2 sharing, and two kernels also come in two
tile shapes (``VARIANTS``: the prism inviscid kernel with or without its five slot sources staged
in LDS, the SST cell kernel with two prisms or one cell per wavefront, which trade LDS and lanes
against occupancy).  The best choice depends on the kernel, the order, the precision and the
mesh size.  ``autotune`` times each kernel's phase with HIP events on the launch stream for every
(tile variant, grid size) candidate (results do not depend on either: only the time does) and
leaves the fastest set on the mesh (``afr_mesh_set_variant``, ``afr_mesh_set_grid``).  Choices
are cached per (device, order, dtype, kernel, size bucket) in memory and, when ``cache_path`` is
given, in a JSON file.
"""

import json
import math
import os

import torch

from . import _lib

KERNELS = {"prism_inviscid": 0, "tet_inviscid": 1, "visc_volume": 2, "visc_iface": 3, "sst_cell": 4,
           "sst_face": 5, "sst_update": 6}
# workgroups per launch, as multiples of the compute-unit count: a kernel whose grid equals the
# workgroups resident at once (CUs x resident workgroups per CU, e.g. 6 x 256 = 1536 on a 256-CU
# device at 6 waves per SIMD) walks each XCD's cell range in one contiguous sweep, so neighbour
# data is reused from L2 as the sweep passes; larger grids hide latency better.  0 = built-in default.
CU_MULTIPLES = (1, 2, 3, 4, 5, 6, 7, 8, 12, 16, 24, 32, 64, 96)


def candidates_for(device=None):
    """Grid candidates of this device: 0 and CU_MULTIPLES x its compute-unit count."""
    n_cu = torch.cuda.get_device_properties(device if device is not None else torch.cuda.current_device()) \
        .multi_processor_count
    return (0,) + tuple(int(n_cu) * k for k in CU_MULTIPLES)

VARIANTS = {"prism_inviscid": (0, 1), "sst_cell": (0, 1)}   # tile shapes (afr_mesh_set_variant)

_memo = {}


def _key(dm, name):
    bucket = int(round(math.log2(max(dm.n_cells, 1))))
    dev = torch.cuda.get_device_name(dm.device)
    return f"{dev}|P{dm.order}|{str(dm.dtype).split('.')[-1]}|{name}|2^{bucket}"


def set_grid(dm, name, cap):
    _lib.check(dm.lib.afr_mesh_set_grid(dm.handle, KERNELS[name], int(cap)), "afr_mesh_set_grid")


def get_grid(dm, name):
    return int(dm.lib.afr_mesh_get_grid(dm.handle, KERNELS[name]))


def set_variant(dm, name, variant):
    _lib.check(dm.lib.afr_mesh_set_variant(dm.handle, KERNELS[name], int(variant)), "afr_mesh_set_variant")


def get_variant(dm, name):
    return int(dm.lib.afr_mesh_get_variant(dm.handle, KERNELS[name]))


def _apply(dm, name, choice):
    """choice: {"grid": g, "variant": v} (or a bare grid size)."""
    if not isinstance(choice, dict):
        choice = {"grid": int(choice), "variant": 0}
    set_variant(dm, name, choice.get("variant", 0))
    set_grid(dm, name, choice["grid"])
    return choice


def _time(fn, reps):
    st = torch.cuda.current_stream()
    fn()                                         # warm-up at this setting
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(2)]
    ev[0].record(st)
    for _ in range(reps):
        fn()
    ev[1].record(st)
    ev[1].synchronize()
    return ev[0].elapsed_time(ev[1]) / reps


def autotune(dm, U, mu_t=None, mu=1.8e-5, sst_fields=None, reps=3, candidates=None, cache_path=None,
             reuse=True):
    """Pick the grid size of every residual kernel present on this mesh.

    U: the state to time on (device, resident); mu_t: eddy viscosity or None; sst_fields:
    (k, omega, nu_t, mu_t) to tune the SST kernels (None skips them); reuse: take a choice
    already made in this process or stored in ``cache_path`` for the same key instead of timing
    (False times every candidate again and replaces the stored choice).  Returns
    {kernel: {"grid": chosen, "variant": v, "ms": {candidate: time} or None when reused}}; the
    choice stays set on the mesh."""
    if candidates is None:
        candidates = candidates_for(dm.device)
    cache = {}
    if cache_path and os.path.exists(cache_path):
        with open(cache_path) as f:
            cache = json.load(f)
    R = torch.empty_like(U)
    phases = {}
    if dm.mesh.n_prism:
        phases["prism_inviscid"] = lambda: dm.inviscid_from_primitives(dm.work_q, dm.work_umean, R)
    if dm.mesh.n_tet:
        phases["tet_inviscid"] = lambda: dm.inviscid_from_primitives(dm.work_q, dm.work_umean, R)
    phases["visc_volume"] = lambda: dm.viscous_volume(U, mu_t, mu=mu)
    phases["visc_iface"] = lambda: dm.viscous_interface(U, mu_t, R, mu=mu)
    if sst_fields is not None and dm.turbulence == "SST":
        k, w, nut, mut = sst_fields
        rates = torch.empty((dm.n_cells, dm.ns, 2), dtype=dm.dtype, device=dm.device)
        dt = dm.work_dt
        dm.local_time_step_nut(U, nut, dt, mu=mu)
        sst = lambda: dm.sst_step(U, k, w, dt, nut, mut, mu=mu, rates=rates)  # noqa: E731
        for name in ("sst_cell", "sst_face", "sst_update"):
            phases[name] = sst                   # the SST call times its three kernels; one varies
    dm.viscous_volume(U, mu_t, mu=mu)            # dQ / field means for the inviscid phase
    out = {}
    for name, fn in phases.items():
        key = _key(dm, name)
        if reuse and (key in _memo or key in cache):
            choice = _apply(dm, name, _memo.get(key, cache.get(key)))
            out[name] = dict(choice, ms=None, cached=True)
            continue
        times = {}
        for var in VARIANTS.get(name, (0,)):
            set_variant(dm, name, var)
            for cap in candidates:
                set_grid(dm, name, cap)
                times[(var, cap)] = _time(fn, reps)
        var, cap = min(times, key=times.get)
        choice = _apply(dm, name, {"grid": cap, "variant": var})
        _memo[key] = choice
        cache[key] = choice
        out[name] = dict(choice, ms={f"v{v}/g{g}": t for (v, g), t in times.items()})
    dm.viscous_volume(U, mu_t, mu=mu)
    torch.cuda.synchronize(dm.device)
    if cache_path:
        os.makedirs(os.path.dirname(os.path.abspath(cache_path)), exist_ok=True)
        with open(cache_path, "w") as f:
            json.dump(cache, f, indent=1)
    return out
