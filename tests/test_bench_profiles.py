"""CPU: the committed PMC traffic table covers every kernel the bench line names for the residual
workload (82^3-hex wall box split into prisms, P2, 9-wide rows, 27 stored points per cell, default
tile variants), so that the line's per-phase and residual traffic are never silently null after a
kernel is renamed or re-templated."""

import json
import os

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_traffic_table_covers_bench_kernels():
    import bench
    with open(os.path.join(ROOT, "profiles", "traffic_wallbox82_p2_sst_fp32.json")) as f:
        table = json.load(f)
    names = bench.residual_kernel_names(2, 9, 27)
    assert set(names) == {"dt_local", "sst", "visc_volume", "inviscid", "visc_interface"}
    for phase, parts in names.items():
        for k in parts:
            assert k in table["kernels"], (phase, k)
            assert bench.traffic_from_profiles(k) == pytest.approx(table["kernels"][k]["bytes_per_launch"])
            assert table["kernels"][k]["bytes_per_launch"] > 0 and table["kernels"][k]["avg_ns"] > 0
