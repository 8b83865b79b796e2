"""Closed-loop example programs on the GPU (examples/Makefile builds them).

- examples/unicycle_closed_loop: this project's harness over the MPPIDriver
  API, writing the trajectory CSV of the reference's examples.  With one
  driver per host thread (one context and stream each), the rows of every
  simulation must not depend on how many threads share the GPU: 1 and 4
  threads give byte-identical CSV rows.  Most runs reach the goal.
- examples/_ref/*: the reference's own unicycle examples
  (examples/nonholonomic_vehicle*/main.cpp), compiled unchanged against this
  project's headers and library with the bundled unicycle plugin as their
  plugin path.  They run to completion, write their CSV and reach the goal.
"""
import csv
import importlib.util
import subprocess
from pathlib import Path

import pytest

pytestmark = pytest.mark.gpu

ROOT = Path(__file__).resolve().parent.parent
HARNESS = ROOT / "examples" / "unicycle_closed_loop"
REF = ROOT / "examples" / "_ref"
PLUGIN = ROOT / "mppi_metal_amd" / "plugins" / "unicycle.hip"
HEADER = ["sim_id", "step", "x", "y", "theta", "v", "a_cmd", "omega_cmd", "best_cost"]


def _rows(path):
    with open(path) as f:
        r = list(csv.reader(f))
    assert r[0] == HEADER
    return sorted(r[1:], key=lambda row: (int(row[0]), int(row[1])))


def _harness(tmp_path, mode, sims, threads, steps, tag):
    if not HARNESS.exists():
        pytest.skip("examples/unicycle_closed_loop is not built (make -C examples)")
    out = tmp_path / f"{mode}_{tag}.csv"
    print(f"[examples] {mode} sims={sims} threads={threads} steps={steps}", flush=True)
    p = subprocess.run([str(HARNESS), "--plugin", str(PLUGIN), "--mode", mode, "--sims", str(sims),
                        "--threads", str(threads), "--steps", str(steps), "--out", str(out)],
                       capture_output=True, text=True, timeout=240)
    assert p.returncode == 0, p.stderr
    reached = int(p.stdout.split("/")[0])
    return _rows(out), reached


@pytest.mark.parametrize("mode,sims,steps", [("simulate", 16, 85), ("step", 8, 300)])
def test_drivers_on_several_threads_match_one(require_gpu, tmp_path, mode, sims, steps):
    one, r1 = _harness(tmp_path, mode, sims, 1, steps, "t1")
    four, r4 = _harness(tmp_path, mode, sims, 4, steps, "t4")
    assert one == four
    assert sorted({int(r[0]) for r in one}) == list(range(sims))
    assert r1 == r4 and r1 >= sims - 1, f"{r1}/{sims} reached the goal"


def test_batch_agents_reach_the_goal(require_gpu, tmp_path):
    rows, reached = _harness(tmp_path, "batch", 64, 1, 100, "b")
    assert reached >= 58, f"{reached}/64 reached the goal"
    first = [r for r in rows if r[1] == "0"]
    assert len(first) == 64 and all(float(r[2]) == 0.0 and float(r[3]) == 0.0 for r in first)


@pytest.mark.timeout(600)
@pytest.mark.parametrize("name,min_reached", [("nonholonomic_vehicle_batch", 0.9),
                                              ("nonholonomic_vehicle_concurrent", 0.9),
                                              ("nonholonomic_vehicle", 0.9)])
def test_reference_example_runs_unchanged(require_gpu, tmp_path, name, min_reached):
    exe = REF / name
    if not exe.exists():
        pytest.skip(f"examples/_ref/{name} is not built (make -C examples reference, needs the reference sources)")
    print(f"[examples] reference {name}", flush=True)
    p = subprocess.run([str(exe)], cwd=tmp_path, capture_output=True, text=True, timeout=540)
    assert p.returncode == 0, p.stderr[-2000:]
    out = tmp_path / ("trajectory_batch.csv" if name.endswith("batch") else "trajectory.csv")
    rows = _rows(out)
    sims = sorted({int(r[0]) for r in rows})
    goal_hits = 0
    last = {}
    for r in rows:
        last[int(r[0])] = r
    for s, r in last.items():
        dx, dy = float(r[2]) - 5.0, float(r[3]) - 5.0
        goal_hits += (dx * dx + dy * dy) ** 0.5 < 0.5
    n = 512 if name.endswith("batch") else 1000
    assert len(sims) == n
    assert goal_hits >= min_reached * n, f"{goal_hits}/{n} ended near the goal"
    # the plotter reads the reference examples' CSV too
    spec = importlib.util.spec_from_file_location("plot_trajectory", ROOT / "examples" / "plot_trajectory.py")
    plot = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(plot)
    runs = plot.load(out)
    assert plot.summary(runs)["runs"] == n
    plot.plot(runs, tmp_path / "t.png")
