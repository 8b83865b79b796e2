"""Benchmark: plain MPPI on one MI355X, unicycle with obstacles.

Workload (the one the metric is quoted on): kinematic unicycle, fp32,
K = 65536 samples, H = 64, 16 random circular obstacles, sigma = [0.5, 0.3],
lambda = 1, bounds +-[2, 1], Philox seed 42, closed loop from x0 = 0 for
--steps MPPI steps (default 500), state and warm start resident on the device
(MPPIDriver::simulate).  One "step" is one MPPI iteration: rollout of all K
samples over H, softmax weights, weighted-noise reduce, plant step and shift.

Printed as one JSON line:
  value            sample-steps/s = K * H * steps / wall time of the timed
                   closed loop
  step_p50/p99_ms  host wall time of a blocking MPPIDriver-style step()
                   (uploads x0 and the nominal, downloads u_next), 200 steps
  roofline         the rollout kernel (the dominant one) against the VALU
                   lane-op peak, from HIP-event kernel times; reduce_roofline
                   the reduce kernel against HBM bandwidth
  cpu_baseline     the CPU oracle (scalar C++, same algorithm and noise
                   stream) on the host's cores over a bounded sample of steps
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

VALU_PEAK_LANE_OPS = 256 * 64 * 2.4e9     # 256 CUs x 64 lanes/clk x 2.4 GHz = 39.3e12
HBM_PEAK_BPS = 8.0e12


def valu_ops_per_sample_step(n_obs: int) -> int:
    """Algorithmic VALU lane-op slots per sample-step of the unicycle rollout
    (transcendentals weighted x4): Philox4x32-10 60, uniform conversion 4,
    Box-Muller 4 + 4*4, scale and clamp 6, unicycle 11 + 2*4, goal and
    control cost 7, per obstacle 7 + 4."""
    return 92 + 7 * n_obs + 4 * (6 + n_obs)


def reduce_bytes(K: int, H: int, C: int, n_chunks: int) -> int:
    """Algorithmic bytes of one reduce launch: costs read and weights written
    (8K), nominal read and u_next written (8HC), workgroup partials written and
    read (2 * 16 * n_chunks * ceil(HC/4))."""
    nb = (H * C + 3) // 4
    return 8 * K + 8 * H * C + 2 * 16 * n_chunks * nb


def reduce_chunks(K: int, H: int, C: int, N: int = 1, spt: int = 0) -> int:
    nb = (H * C + 3) // 4
    groups = (nb + 7) // 8
    if not spt:
        spt = max(1, min(64, -(-(K * groups * N) // (256 * 2048))))
    return -(-K // (256 * spt))


def cpu_baseline(n_obs, H, K, steps, threads):
    sys.path.insert(0, str(ROOT / "tests"))
    from problems import unicycle_oracle
    ora, sc = unicycle_oracle(K, H, n_obs)
    x0 = np.zeros(4, np.float32)
    t0 = time.perf_counter()
    ora.simulate(steps, 0, x0, threads=threads)
    dt = time.perf_counter() - t0
    return K * H * steps / dt, dt


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=500)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--latency-steps", type=int, default=200)
    ap.add_argument("--cpu-steps", type=int, default=20)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--K", type=int, default=65536)
    ap.add_argument("--H", type=int, default=64)
    ap.add_argument("--obstacles", type=int, default=16)
    ap.add_argument("--spt", type=int, default=0, help="reduce samples per thread (0: built-in)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU: this build has no multi-GPU path "
              "(agents or samples would shard with no or one small exchange per step).",
              file=sys.stderr)
        sys.exit(2)

    from mppi_metal_amd import HipMPPI, scenes

    K, H, C, n_obs = args.K, args.H, 2, args.obstacles
    sc = scenes.unicycle_scene(n_obs, 42)
    mb, cb = sc.model_blob(), sc.cost_blob()
    hip = HipMPPI(4, C, H, K, sc.lam, sc.sigma, sc.u_min, sc.u_max)
    hip.link("unicycle.hip", "mppi_dynamics", "unicycle_stage_cost_obstacles",
             "unicycle_terminal_cost", len(mb), len(cb), 42)
    hip.bind_params(mb, cb)
    if args.spt:
        hip.set_reduce_samples_per_thread(args.spt)
    x0 = np.zeros(4, np.float32)
    u0 = np.zeros((H, C), np.float32)
    print(f"[bench] workload unicycle K={K} H={H} obstacles={n_obs}; warmup {args.warmup}",
          file=sys.stderr, flush=True)

    # Warm-up (code object already loaded by link()).
    hip.simulate(max(1, args.warmup), 0, x0, u0)

    # Timed closed loop: exactly `steps` MPPI steps, device resident.
    t0 = time.perf_counter()
    st, ct, kt = hip.simulate(args.steps, 0, x0, u0)
    elapsed = time.perf_counter() - t0
    value = K * H * args.steps / elapsed
    print(f"[bench] timed {args.steps} steps in {elapsed:.3f}s", file=sys.stderr, flush=True)

    # Kernel times from HIP events on the library's stream.
    hip.profile(True)
    prof_steps = min(args.steps, 100)
    hip.simulate(prof_steps, 0, x0, u0)
    roll_ms, red_ms, launches = hip.profile_read()
    hip.profile(False)
    roll_s = roll_ms / launches / 1e3
    red_s = red_ms / launches / 1e3
    w_ss = valu_ops_per_sample_step(n_obs)
    achieved_ops = K * H * w_ss / roll_s
    n_chunks = reduce_chunks(K, H, C, spt=args.spt)
    rb = reduce_bytes(K, H, C, n_chunks)

    # Blocking step latency (host round trip per step).
    lat = []
    u_nom = u0.copy()
    for i in range(10 + args.latency_steps):
        t1 = time.perf_counter()
        u, d = hip.step(i, x0, u_nom)
        t2 = time.perf_counter()
        if i >= 10:
            lat.append((t2 - t1) * 1e3)
        u_nom = np.concatenate([u[1:], u[-1:]], axis=0)
    lat = np.array(lat)

    cpu = None
    if not args.no_cpu:
        threads = min(16, os.cpu_count() or 1)
        cps, cdt = cpu_baseline(n_obs, H, K, args.cpu_steps, threads)
        cpu = {"value": cps, "unit": "sample-steps/s", "cores": threads, "kind": "port",
               "sample": f"CPU oracle (scalar C++, FP32, same noise stream), first {args.cpu_steps} "
                         f"closed-loop steps of the same workload, {threads} threads, {cdt:.1f}s"}

    out = {
        "metric": "MPPI sample-steps/s (+ step p50/p99 ms)",
        "value": value,
        "unit": "sample-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": elapsed / args.steps * 1e3,
        "step_p50_ms": float(np.percentile(lat, 50)),
        "step_p99_ms": float(np.percentile(lat, 99)),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic scene from Philox seed 42",
        "config": {"workload": "plain MPPI closed loop, kinematic unicycle", "K": K, "H": H,
                   "obstacles": n_obs, "closed_loop_steps": args.steps, "lambda": 1.0,
                   "sigma": [0.5, 0.3]},
        "roofline": {"bound": "valu", "kernel": "mppi_rollout",
                     "achieved": achieved_ops / 1e12, "peak": VALU_PEAK_LANE_OPS / 1e12,
                     "unit": "T lane-op/s", "frac": achieved_ops / VALU_PEAK_LANE_OPS,
                     "ops_per_sample_step": w_ss, "avg_launch_us": roll_s * 1e6, "traffic": None},
        "reduce_roofline": {"bound": "hbm", "kernel": "mppi_reduce",
                            "achieved": rb / red_s / 1e9, "peak": HBM_PEAK_BPS / 1e9,
                            "unit": "GB/s", "frac": rb / red_s / HBM_PEAK_BPS,
                            "bytes_per_launch": rb, "avg_launch_us": red_s * 1e6},
        "cpu_baseline": cpu,
        "final_closed_loop_cost": float(kt[-1]),
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
