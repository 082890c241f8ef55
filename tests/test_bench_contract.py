"""CPU: the parts of bench.py that need no GPU.

* `roofline.traffic` is read from the committed rocprofv3 summary
  (profiles/fused_step_4096.md): every figure there equals its table row's
  bytes per launch, is within 1.0-1.2x the kernel's algorithmic bytes at
  4096^2 (72 B/cell hydro, 16 B/cell thermal), and a grid or tile variant
  the summary does not cover gives None instead of a borrowed figure.
* More than one GPU is refused (exit status 2) before any GPU work.
* The batched workload is the 64-terrain draw its profiles and tests quote:
  sides in {256, 384, ..., 2048}, 101,728,256 cells, coefficients in their
  ranges, the same on every call."""
from __future__ import annotations

import os
import re
import subprocess
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
import bench  # noqa: E402

PROFILE = os.path.join(ROOT, "profiles", "fused_step_4096.md")


def _table_rows():
    """(kernel, variant) -> traffic bytes per launch, from the summary's table."""
    rows = {}
    pat = re.compile(r"^\| `k_(fused_\w+)<([^>]*)>` \|(.*)\|\s*$")
    for line in open(PROFILE):
        m = pat.match(line.strip())
        if not m:
            continue
        cells = [c.strip() for c in m.group(3).split("|")]
        if len(cells) < 5:
            continue
        args = [a.strip() for a in m.group(2).split(",")]
        rows[(m.group(1), tuple(int(a) for a in args[:-2]))] = int(cells[4].replace(",", ""))
    return rows


def test_traffic_figures_match_the_profile_table():
    rows = _table_rows()
    assert len(rows) >= 2
    n = 4096
    seen = set()
    for line in open(PROFILE):
        m = re.match(r"^- traffic\[(\w+) (\d+) ([\dx]+)\]: (\d+)$", line.strip())
        if not m:
            continue
        kernel, grid, variant = m.group(1), int(m.group(2)), tuple(int(v) for v in m.group(3).split("x"))
        got = bench.profiled_traffic(kernel, grid, variant)
        assert got == int(m.group(4))
        assert grid == n
        ratio = got / (bench.ALGO_BYTES[kernel] * n * n)
        assert 1.0 <= ratio <= 1.2, (kernel, variant, ratio)
        # the table names hydro variants by (TX, TY, HA, NT), thermal ones by (TX, TY, NT)
        key = (kernel, variant if kernel == "fused_hydro" else (variant[0], variant[1], variant[3]))
        if key in rows:
            assert rows[key] == got, key
            seen.add(key)
    assert ("fused_hydro", (64, 16, 2, 512)) in seen and ("fused_thermal", (64, 64, 1024)) in seen


def test_traffic_is_none_where_the_profile_has_no_figure():
    assert bench.profiled_traffic("fused_hydro", 2048, (64, 16, 2, 512)) is None
    assert bench.profiled_traffic("fused_hydro", 4096, (32, 16, 2, 512)) is None
    assert bench.profiled_traffic("step_sweep", 4096, (64, 16, 2, 512)) is None


def test_more_than_one_gpu_is_refused():
    env = dict(os.environ)
    for argv, extra in ((["--gpus", "2"], {}), (["--gpus", "8"], {}), ([], {"WORLD_SIZE": "2"})):
        r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), *argv], cwd=ROOT,
                           env={**env, **extra}, capture_output=True, text=True, timeout=120)
        assert r.returncode == 2, (argv, r.stderr)
        assert "exactly one GPU" in r.stderr
        assert r.stdout == ""


def test_batched_workload_is_the_64_terrain_draw():
    sides, coeffs = bench.batched_workload()
    assert len(sides) == len(coeffs) == 64
    assert set(sides) <= set(range(256, 2049, 128))
    assert sum(s * s for s in sides) == 101_728_256
    assert min(sides) == 256 and max(sides) == 2048
    for c in coeffs:
        assert 0.02 <= c["Kc"] <= 0.1 and 0.01 <= c["Ks"] <= 0.05 and 0.01 <= c["Kd"] <= 0.05
        assert 0.005 <= c["Ke"] <= 0.02 and 30.0 <= c["talus_deg"] <= 45.0
        assert c["slippage"] == 0.1 and c["G"] == 9.81 and c["dt"] == 0.005
    again = bench.batched_workload()
    assert again[0] == sides and again[1] == coeffs
    assert np.unique(sides).size > 8           # a spread of sizes, not one repeated
