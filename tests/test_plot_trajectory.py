"""The trajectory CSV reader and plotter (examples/plot_trajectory.py) on a
synthetic CSV in the examples' format: runs and steps parsed in order, the
goal summary, and a PNG written (no GPU)."""
import importlib.util
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parent.parent


def _module():
    spec = importlib.util.spec_from_file_location("plot_trajectory", ROOT / "examples" / "plot_trajectory.py")
    m = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(m)
    return m


def _write(path, runs):
    with open(path, "w") as f:
        f.write("sim_id,step,x,y,theta,v,a_cmd,omega_cmd,best_cost\n")
        for sim, pts in runs.items():
            for step in reversed(range(len(pts))):        # out of order on purpose
                x, y = pts[step]
                f.write(f"{sim},{step},{x},{y},0.1,1.5,0.2,-0.1,{10.0 / (step + 1)}\n")


def test_load_summary_and_plot(tmp_path):
    m = _module()
    t = np.linspace(0, 1, 40)
    runs = {0: [(5 * s, 5 * s) for s in t],                 # ends at the goal
            1: [(5 * s, 5 * s - 0.05) for s in t],          # ends 0.05 away
            2: [(2 * s, 1 * s) for s in t[:25]]}            # stops short
    path = tmp_path / "trajectory.csv"
    _write(path, runs)
    loaded = m.load(path)
    assert sorted(loaded) == [0, 1, 2]
    assert np.array_equal(loaded[2][:, 0], np.arange(25))
    assert loaded[0][-1, 1] == pytest.approx(5.0)
    s = m.summary(loaded)
    assert s == {"runs": 3, "reached": 2, "mean_steps": pytest.approx((40 + 40 + 25) / 3)}
    out = tmp_path / "t.png"
    m.plot(loaded, out)
    assert out.stat().st_size > 10000 and out.read_bytes()[:8] == b"\x89PNG\r\n\x1a\n"


def test_rejects_other_headers(tmp_path):
    m = _module()
    p = tmp_path / "bad.csv"
    p.write_text("a,b,c\n1,2,3\n")
    with pytest.raises(ValueError):
        m.load(p)
