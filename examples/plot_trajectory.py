"""Plot and summarise a trajectory CSV (sim_id,step,x,y,theta,v,a_cmd,omega_cmd,best_cost),
as written by examples/unicycle_closed_loop and by the reference's unicycle
examples (examples/nonholonomic_vehicle*/main.cpp; the reference's plotter is
scripts/visualize_trajectory.py:14-45): paths in the plane, commanded
controls, speed and heading, and the best sample cost per step.

    python examples/plot_trajectory.py trajectory.csv [out.png] [--goal X Y R]
"""
from __future__ import annotations

import argparse
import csv
from collections import defaultdict

import numpy as np

COLUMNS = ["sim_id", "step", "x", "y", "theta", "v", "a_cmd", "omega_cmd", "best_cost"]


def load(path):
    """{sim_id: array [steps, 8]} with columns step, x, y, theta, v, a_cmd, omega_cmd, best_cost,
    rows in step order."""
    runs = defaultdict(list)
    with open(path) as f:
        rd = csv.reader(f)
        header = next(rd)
        if header != COLUMNS:
            raise ValueError(f"{path}: header {header} is not {COLUMNS}")
        for row in rd:
            runs[int(row[0])].append([float(v) for v in row[1:]])
    return {s: np.array(sorted(r, key=lambda v: v[0])) for s, r in sorted(runs.items())}


def summary(runs, goal=(5.0, 5.0), radius=0.2):
    """Number of runs, how many end within `radius` (plus one step of travel)
    of the goal, and the mean number of logged steps."""
    final = np.array([r[-1, 1:3] for r in runs.values()])
    dist = np.hypot(final[:, 0] - goal[0], final[:, 1] - goal[1])
    return {"runs": len(runs), "reached": int(np.sum(dist < radius + 0.1)),
            "mean_steps": float(np.mean([len(r) for r in runs.values()]))}


def plot(runs, out, goal=(5.0, 5.0)):
    import matplotlib
    matplotlib.use("Agg")
    import matplotlib.pyplot as plt
    fig, ax = plt.subplots(2, 2, figsize=(13, 9))
    alpha = max(0.03, min(0.6, 20.0 / max(len(runs), 1)))
    for r in runs.values():
        step = r[:, 0]
        ax[0, 0].plot(r[:, 1], r[:, 2], color="tab:blue", alpha=alpha, lw=0.8)
        ax[0, 1].plot(step, r[:, 5], color="tab:red", alpha=alpha, lw=0.8)
        ax[0, 1].plot(step, r[:, 6], color="tab:orange", alpha=alpha, lw=0.8)
        ax[1, 0].plot(step, r[:, 4], color="tab:green", alpha=alpha, lw=0.8)
        ax[1, 0].plot(step, r[:, 3], color="tab:purple", alpha=alpha, lw=0.8)
        ax[1, 1].plot(step, r[:, 7], color="tab:brown", alpha=alpha, lw=0.8)
    ax[0, 0].plot([goal[0]], [goal[1]], "k*", ms=12)
    ax[0, 0].set(title="paths", xlabel="x", ylabel="y", aspect="equal")
    ax[0, 1].set(title="commands: a (red), omega (orange)", xlabel="step")
    ax[1, 0].set(title="speed (green), heading (purple)", xlabel="step")
    ax[1, 1].set(title="best sample cost", xlabel="step", yscale="log")
    fig.tight_layout()
    fig.savefig(out, dpi=110)
    plt.close(fig)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("csv")
    ap.add_argument("out", nargs="?", default="trajectory.png")
    ap.add_argument("--goal", nargs=3, type=float, default=(5.0, 5.0, 0.2), metavar=("X", "Y", "R"))
    args = ap.parse_args()
    runs = load(args.csv)
    s = summary(runs, goal=args.goal[:2], radius=args.goal[2])
    plot(runs, args.out, goal=args.goal[:2])
    print(f"{s['runs']} runs, {s['reached']} reached the goal, {s['mean_steps']:.1f} steps on average; "
          f"plot written to {args.out}")


if __name__ == "__main__":
    main()
