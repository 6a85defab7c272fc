"""bench.py's single-GPU contract and its lookup of measured traffic (CPU).

The bench times one domain on one GPU; asked for more it must refuse before
touching a device.  Its ``roofline.traffic`` is the rocprofv3 byte count of
the dominant kernel from the committed summaries under profiles/, and the
kernel names it looks up are the byte model's (the template arguments of the
flux kernels folded into their u- or v-face grid)."""
from __future__ import annotations

import json
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
WORKLOAD = "two_fronts_1440x1080x75"


def test_more_than_one_gpu_is_refused():
    p = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=300, cwd=ROOT)
    assert p.returncode == 2, p.stderr[-2000:]
    line = json.loads(p.stdout.strip().splitlines()[-1])
    assert "exactly one GPU" in line["error"]


def _summary(form):
    with open(os.path.join(ROOT, "profiles", f"kernels_{WORKLOAD}_{form}.json")) as f:
        return json.load(f)["kernels"]


@pytest.mark.parametrize("form", ["fp64", "fp32_store"])
@pytest.mark.parametrize("kernel,y", [("k_neutral_sweep_uv", None), ("k_flux_x", "false"),
                                      ("k_flux_y", "true"), ("k_tracer_update", None)])
def test_measured_traffic_is_the_committed_fetch_plus_write(form, kernel, y):
    sys.path.insert(0, ROOT)
    import bench

    got = bench.measured_traffic(kernel, WORKLOAD, form)
    base = "k_flux" if y else kernel
    rows = [e for k, e in _summary(form).items()
            if k.split("<")[0] == base and (y is None or k.rstrip(">").split(",")[-1].strip() == y)]
    assert len(rows) == 1, (kernel, form)
    e = rows[0]
    assert got == e["traffic_bytes"]
    assert got == pytest.approx(e["fetch_bytes"] + e["write_bytes"], rel=1e-12)
    # at least the kernel's own algorithmic bytes reach HBM
    assert got >= e["alg_bytes"]


def test_measured_traffic_of_an_unprofiled_case_is_none():
    sys.path.insert(0, ROOT)
    import bench

    assert bench.measured_traffic("k_neutral_sweep_uv", "no_such_workload") is None
    assert bench.measured_traffic("no_such_kernel", WORKLOAD) is None
