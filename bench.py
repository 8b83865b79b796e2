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
                   lane-op peak, from kernel-exact HIP event times, and against
                   the plain-instruction issue ceiling measured on the device at
                   start-up (mppi_hip_calibrate); reduce_roofline
                   the reduce kernel against HBM bandwidth
  cpu_baseline     the CPU oracle (scalar C++, same algorithm and noise
                   stream) on the host's cores over a bounded sample of steps

Two more workloads are reported in the same line (skip with --no-batch,
--no-surrogate):
  batch            256-agent dynamic-bicycle pursuit with uneven K (log-uniform
                   1024..16384) and H (16..96), 32 threat zones, one
                   graph-replayed step of all agents with the autotuned launch
                   shape; against the plain GPU path driven one agent at a time
                   (one context per agent, blocking step() each, as a user of
                   the reference's single-agent driver would) and against a
                   uniform batch padded to the largest K and H (the
                   reference's batch form needs one K and H for all agents).
  serving          quadrotor (state 12, control 4, full 4x4 SPD control
                   covariance), 1024 agent requests arriving as a Poisson
                   process at 2000 req/s, K log-uniform 4096..32768, H
                   uniform 30..80; the library's serving loop admits arrived
                   requests in order against an HBM budget and solves every
                   admitted agent in one batched closed-loop step (state and
                   warm start on the device); p50/p99 latency from arrival to
                   control, per numeric mode (MIXED is the workload's mode),
                   and a burst of all 1024 requests at once for throughput
                   under a 32 MB budget (admission in waves).
  surrogate        the learned bicycle dynamics (MLP 8-128-128-6, trained on
                   the GPU on 4M plugin transitions) inside the rollout at
                   K = 262144, H = 64, 500 closed-loop steps, against plugin
                   dynamics from the same start and seed: sample-steps/s of
                   both and the gap of the mean realised closed-loop cost.
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

# Published MI355X peaks.  FP32 vector: 157.3 TFLOP/s, an FMA counted as two
# flops, i.e. 78.6e12 lane-instructions/s (256 CUs x 4 SIMDs x 32 lanes x
# 2.4 GHz), the most any VALU instruction stream can issue.  Dense bf16 and
# fp16 matrix cores: 2.5 PFLOP/s.
VALU_PEAK_LANE_OPS = 157.3e12 / 2
MFMA_PEAK_16 = 2.5e15
HBM_PEAK_BPS = 8.0e12
COUNTERS = ROOT / "profiles" / "mppi_counters_unicycle_pursuit_surrogate.md"
# Issue ceilings measured on this device at start-up (mppi_hip_calibrate):
# plain FP32 VALU lane-ops/s and FP16 32x32x16 matrix flops/s.
CEILING = {"valu": None, "mfma": None, "mixed_over_sum": None}


def calibrate():
    """Measure the device's issue ceilings once (mppi_hip_calibrate)."""
    import ctypes
    from mppi_metal_amd import _native
    lib = _native.load()
    v, m, x = ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
    rc = lib.mppi_hip_calibrate(0, ctypes.byref(v), ctypes.byref(m), ctypes.byref(x))
    if rc != 0:
        raise RuntimeError(f"mppi_hip_calibrate failed ({rc})")
    CEILING["valu"], CEILING["mfma"], CEILING["mixed_over_sum"] = v.value, m.value, x.value
    print(f"[bench] measured ceilings: VALU {v.value / 1e12:.2f} T lane-op/s (plain FP32), "
          f"MFMA {m.value / 1e12:.1f} TFLOP/s (fp16 32x32x16), one MFMA per 16 FMAs in one stream "
          f"{x.value:.2f} of their separate times", file=sys.stderr, flush=True)


def against_ceiling(rate: float, kind: str) -> dict:
    """`rate` (lane-op/s or flop/s) against the measured issue ceiling of
    `kind` ("valu" or "mfma")."""
    c = CEILING[kind]
    unit = "T lane-op/s" if kind == "valu" else "TFLOP/s"
    how = ("plain (unpacked) FP32 FMAs, independent streams, 8 wavefronts per SIMD" if kind == "valu"
           else "v_mfma_f32_32x32x16_f16, independent accumulators, 8 wavefronts per SIMD")
    return {"measured": None if c is None else c / 1e12, "unit": unit,
            "frac": None if c is None else rate / c, "how": "mppi_hip_calibrate: " + how}


def noise_ops(C: int) -> dict:
    """Algorithmic VALU lane-ops per sample-step of the control noise (C
    normals), transcendentals weighted x4: Philox4x32-10 per four normals
    (10 rounds of 2 low and 2 high products and 4 xors; key bumps are
    uniform), 3 per uniform word (shift, or, convert and scale), per
    Box-Muller pair log, sqrt, sin, cos (x4 each) and 5 multiplies, and 3 per
    control (scale by sigma, add the nominal, clamp)."""
    return {"philox": 80 * C / 4, "uniform": 3 * C, "box_muller": 21 * C / 2, "control": 3 * C}


def unicycle_w_ss(C: int = 2, fused: bool = True) -> dict:
    """Algorithmic VALU lane-ops per sample-step of the unicycle rollout with
    its 16-slot obstacle gate: noise; dynamics 26 (sincos x4 each, two
    position updates of 2 products and an add, two FMAs, the heading wrap 8,
    the speed clamp 2); goal and control cost 9; the gate 4 per slot (two
    FMAs, max of the magnitudes, running min) over all 16 slots (the exact
    pass runs only where a sample is in a square: none at this workload); the
    cost accumulation 1; and in the one-pass step, 2 per control to keep and
    average the noise."""
    w = noise_ops(C)
    w.update(dynamics=26, goal_cost=9, obstacle_gate=4 * 16, accumulate=1, keep_noise=2 * C if fused else 0)
    w["total"] = sum(w.values())
    return w


def bicycle_w_ss(substeps: int = 2, zones: int = 32, C: int = 2, one_launch: bool = False) -> dict:
    """Algorithmic VALU lane-ops per sample-step of the dynamic-bicycle
    pursuit rollout: noise; per Euler substep 40 (sincos x4 each, the speed
    floor 1, 1/u x4, the two tyre forces 7, the body-frame rates 12, six
    state FMAs, the clamp 2); the tracking and control cost 9 (the target's
    position is uniform); the zone test through the workgroup's zone grid 14
    (two band indices of 5 each, the set intersection with the zones every
    wavefront tests 2, the wavefront's any-candidate test 2; the exact pass,
    which runs only in wavefronts where some sample has a candidate zone, is
    not counted); accumulation 1.  The one-launch step also regenerates the
    noise for its weighted sum: the noise again and 2 per control."""
    w = noise_ops(C)
    w.update(dynamics=40 * substeps, cost=9, zone_test=14, accumulate=1)
    if one_launch:
        w["noise_sums"] = sum(noise_ops(C).values()) + 2 * C
    w["total"] = sum(w.values())
    return w


# Matrix-core flops per sample-step of the 8-128-128-6 surrogate: the
# algorithm's 2 (8*128 + 128*128 + 128*6) and the issued ones: layers 1 and 2
# exactly (32x32x8 and 32x32x16 tiles), the output layer as 16x16x32
# products that hold 8 output rows and, per sample, 16 of their 32 K values
# (2 * 16 * 128).
SURROGATE_FLOPS = 2 * (8 * 128 + 128 * 128 + 128 * 6)
SURROGATE_FLOPS_ISSUED = 2 * (8 * 128 + 128 * 128 + 16 * 128)


def profile_line(key: str) -> dict:
    """The `key kernel value kernel value ...` line of the committed counter
    summary, as {kernel: float}; empty without it."""
    if not COUNTERS.exists():
        return {}
    for line in COUNTERS.read_text().splitlines():
        if line.startswith(key + " "):
            f = line.split()[1:]
            return {k: float(v) for k, v in zip(f[0::2], f[1::2])}
    return {}


def issued(kernel: str, sample_steps: float, seconds: float):
    """Counter-derived VALU issue of `kernel` (named kernel@workload): the
    VALU instructions it executes per sample-step (SQ_INSTS_VALU x 64 lanes
    over the sample-steps of a launch, from the committed rocprofv3 summary;
    MFMA instructions are counted separately there), times the sample-steps of
    a launch over the launch's duration measured here, against the FP32
    vector peak; or None without a summary."""
    per = profile_line("valu_instructions_per_sample_step").get(kernel)
    if per is None:
        return None
    rate = per * sample_steps / seconds
    return {"valu_per_sample_step": per, "achieved": rate / 1e12, "peak": VALU_PEAK_LANE_OPS / 1e12,
            "unit": "T lane-op/s", "frac": rate / VALU_PEAK_LANE_OPS, "source": "profiles/" + COUNTERS.name}


def reduce_bytes(K: int, H: int, C: int, n_chunks: int) -> int:
    """Algorithmic bytes of one reduce launch: costs read and weights written
    (8K), nominal read and u_next written (8HC), workgroup partials written and
    read (2 * 16 * n_chunks * ceil(HC/4))."""
    nb = (H * C + 3) // 4
    return 8 * K + 8 * H * C + 2 * 16 * n_chunks * nb


def combine_bytes(K: int, H: int, C: int) -> int:
    """Algorithmic bytes of one combine launch of the one-pass step: the
    tiles' weighted noise sums (4 HC per 256-sample tile) and statistics (16
    per tile) read, the nominal read, the update and the next nominal
    written (4 HC each)."""
    tiles = -(-K // 256)
    return tiles * (4 * H * C + 16) + 3 * 4 * H * C


def reduce_chunks(K: int, H: int, C: int, N: int = 1, spt: int = 0, cus: int = 256) -> int:
    """Sample chunks of the reduce launch (the library's default_spt rule)."""
    nb = (H * C + 3) // 4
    groups = (nb + 7) // 8
    if not spt:
        spt = max(1, min(64, -(-(K * groups * N) // (256 * 2 * cus))))
    return -(-K // (256 * spt))


def measured_traffic(kernel: str):
    """HBM bytes per launch of `kernel` (named kernel@workload), from the committed
    rocprofv3 FETCH_SIZE / WRITE_SIZE summary (scaled by the copy-kernel
    calibration recorded there), or None."""
    v = profile_line("traffic_bytes_per_launch").get(kernel)
    return None if v is None else int(v)


def host_cpus():
    """The host's CPUs as this process sees them: nproc (the affinity mask),
    os.cpu_count(), the cgroup CPU quota, the model name; `usable` is what a
    worker pool may use (affinity, capped by the quota)."""
    aff = len(os.sched_getaffinity(0))
    quota = None
    try:
        q, per = Path("/sys/fs/cgroup/cpu.max").read_text().split()
        if q != "max":
            quota = int(q) / int(per)
    except (OSError, ValueError):
        pass
    model = None
    try:
        for line in Path("/proc/cpuinfo").read_text().splitlines():
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    usable = aff if quota is None else max(1, min(aff, int(quota)))
    return {"nproc": aff, "os_cpu_count": os.cpu_count(), "cgroup_cpu_quota": quota, "model": model,
            "usable": usable}


def cpu_baseline(n_obs, H, K, steps, threads):
    """sample-steps/s of the CPU oracle over the first `steps` closed-loop
    steps of the workload, on `threads` threads."""
    sys.path.insert(0, str(ROOT / "tests"))
    from problems import unicycle_oracle
    ora, sc = unicycle_oracle(K, H, n_obs)
    x0 = np.zeros(4, np.float32)
    t0 = time.perf_counter()
    ora.simulate(steps, 0, x0, threads=threads)
    dt = time.perf_counter() - t0
    return K * H * steps / dt, dt


def small_workload(hip_cls, threads, steps=200, warm=10, no_cpu=False):
    """The CPU-runnable workload (unicycle K = 1024, H = 32, 8 obstacles, seed
    42, 200 closed-loop steps from x0 = 0), on the CPU oracle and on the GPU:
    sample-steps/s of the closed loop, p50/p99 of a blocking step (warm start
    shifted on the host between steps; 10 warm-up steps excluded) and the
    mean realised closed-loop cost."""
    from mppi_metal_amd import scenes
    K, H, n_obs = 1024, 32, 8
    sc = scenes.unicycle_scene(n_obs, 42)
    mb, cb = sc.model_blob(), sc.cost_blob()
    x0 = np.zeros(4, np.float32)
    u0 = np.zeros((H, 2), np.float32)
    out = {"workload": "plain MPPI closed loop, kinematic unicycle, K=1024, H=32, 8 obstacles, seed 42",
           "steps": steps}

    def latencies(step_fn):
        lat, u_nom = [], u0.copy()
        for i in range(warm + steps):
            t1 = time.perf_counter()
            u = step_fn(i, u_nom)
            t2 = time.perf_counter()
            if i >= warm:
                lat.append((t2 - t1) * 1e3)
            u_nom = np.concatenate([u[1:], u[-1:]], axis=0)
        return np.array(lat)

    hip = hip_cls(4, 2, H, K, sc.lam, sc.sigma, sc.u_min, sc.u_max)
    hip.link("unicycle.hip", "mppi_dynamics", "unicycle_stage_cost_obstacles", "unicycle_terminal_cost",
             len(mb), len(cb), 42)
    hip.bind_params(mb, cb)
    sys.path.insert(0, str(ROOT / "tests"))
    from problems import unicycle_oracle
    ora, _ = unicycle_oracle(K, H, n_obs)
    hip.simulate(warm, 0, x0, u0)
    t0 = time.perf_counter()
    st, ct, kt, tt = hip.simulate_batch(1, steps, 0, x0, u0)     # one agent: simulate() plus its end state
    el = time.perf_counter() - t0
    cost = ora.closed_loop_cost(st[:, 0], ct[:, 0], tt[0], 0)
    lat = latencies(lambda i, un: hip.step(i, x0, un)[0])
    out["gpu"] = {"value": K * H * steps / el, "unit": "sample-steps/s", "ms_per_step": el / steps * 1e3,
                  "step_p50_ms": float(np.percentile(lat, 50)), "step_p99_ms": float(np.percentile(lat, 99)),
                  "mean_realised_cost": float(np.mean(cost))}
    hip.close()
    if not no_cpu:
        t0 = time.perf_counter()
        so, co, ko, to = ora.simulate(steps, 0, x0, threads=threads)
        el = time.perf_counter() - t0
        cost_o = ora.closed_loop_cost(so[:, 0], co[:, 0], to[0], 0)
        lat = latencies(lambda i, un: ora.step(i, x0, un, threads=threads)[0])
        out["cpu"] = {"value": K * H * steps / el, "unit": "sample-steps/s", "ms_per_step": el / steps * 1e3,
                      "step_p50_ms": float(np.percentile(lat, 50)), "step_p99_ms": float(np.percentile(lat, 99)),
                      "mean_realised_cost": float(np.mean(cost_o)), "threads": threads, "kind": "port"}
    return out


def bench_batch(args):
    """Uneven 256-agent pursuit: batched step vs per-agent plain path."""
    from mppi_metal_amd import HipBatch, HipMPPI, scenes
    sc = scenes.bicycle_pursuit_scene(256, 42)
    mb = sc.model.blob()
    N = len(sc.K)
    work = int(np.sum(sc.K.astype(np.int64) * sc.H.astype(np.int64)))   # sample-steps per step
    hip = HipMPPI(6, 2, 8, 256, sc.lam, sc.sigma, sc.u_min, sc.u_max)
    hip.link("bicycle.hip", "bicycle_dynamics", "bicycle_pursuit_cost", "bicycle_terminal_cost",
             len(mb), scenes.BICYCLE_COST_BYTES, 42)
    b = HipBatch(hip, sc.K, sc.H)
    b.set_params(mb, sc.cost_blob_packed(), 0, scenes.BICYCLE_COST_BYTES)
    b.reset(0, sc.x0)
    if args.batch_tuning:
        f = [int(v) for v in args.batch_tuning.split(",")]
        b.set_tuning(block=f[0], persistent=f[1], wgs_per_cu=f[2], reduce_spt=f[3], order=f[4], launches=f[5])
        tune_ms = None
    else:
        tune_ms = b.autotune(3)
    tuning = b.get_tuning()
    b.reset(0, sc.x0)
    b.run(5, closed_loop=True)
    T = args.batch_steps
    t0 = time.perf_counter()
    b.run(T, closed_loop=True)
    el = time.perf_counter() - t0
    batch_rate = work * T / el
    print(f"[bench] batch: {T} steps in {el:.3f}s, tuning {tuning}", file=sys.stderr, flush=True)

    # The persistent one-launch form (a grid of 4 workgroups per CU takes
    # tiles from a ticket counter; each agent's last tile forms its update),
    # timed the same way beside the chosen shape.
    forms = {}
    for name, (pers, wgs, launches) in (("persistent_one_launch", (1, 4, 1)), ("table_one_launch", (0, 0, 1)),
                                        ("persistent_two_launches", (1, 4, 2))):
        b.set_tuning(block=256, persistent=pers, wgs_per_cu=wgs, reduce_spt=tuning["reduce_spt"],
                     order=tuning["order"], launches=launches)
        b.reset(0, sc.x0)
        b.run(3, closed_loop=True)
        t0 = time.perf_counter()
        b.run(T, closed_loop=True)
        el_f = time.perf_counter() - t0
        forms[name] = {"value": work * T / el_f, "ms_per_step": el_f / T * 1e3, "tuning": b.get_tuning()}
    b.set_tuning(**tuning)
    b.reset(0, sc.x0)
    b.run(3, closed_loop=True)

    # Kernel times from events attached to each dispatch (while profiling the
    # batch launches its steps directly rather than replaying its graph).
    hip.profile(True)
    b.run(min(T, 20), closed_loop=True)
    r_ms, d_ms, n_l = hip.profile_read()
    hip.profile(False)
    one = tuning["launches"] == 1
    kern = "mppi_{}_{}_b{}".format("step" if one else "rollout", "persist" if tuning["persistent"] else "table",
                                   tuning["block"])
    w = bicycle_w_ss(2, 32, 2, one)
    k_s = r_ms / n_l / 1e3
    rate = work * w["total"] / k_s
    roofline = {"bound": "valu", "kernel": kern, "achieved": rate / 1e12, "peak": VALU_PEAK_LANE_OPS / 1e12,
                "unit": "T lane-op/s", "frac": rate / VALU_PEAK_LANE_OPS,
                "issue_ceiling": against_ceiling(rate, "valu"), "ops_per_sample_step": w["total"],
                "ops_breakdown": w, "sample_steps_per_launch": work, "avg_launch_us": k_s * 1e6,
                "second_launch_us": None if one else d_ms / n_l * 1e3,
                "issued": issued(kern + "@pursuit", work, k_s),
                "traffic": measured_traffic(kern + "@pursuit"),
                "traffic_source": "rocprofv3 FETCH_SIZE + WRITE_SIZE, profiles/" + COUNTERS.name}

    # Plain GPU path, one agent at a time (blocking step per agent).
    ctxs = []
    for a in range(N):
        h = HipMPPI(6, 2, int(sc.H[a]), int(sc.K[a]), sc.lam, sc.sigma, sc.u_min, sc.u_max)
        h.link("bicycle.hip", "bicycle_dynamics", "bicycle_pursuit_cost", "bicycle_terminal_cost",
               len(mb), scenes.BICYCLE_COST_BYTES, 42)
        h.bind_params(mb, sc.cost_blobs[a])
        ctxs.append(h)
    noms = [np.zeros((int(sc.H[a]), 2), np.float32) for a in range(N)]
    for a in range(N):
        ctxs[a].step(0, sc.x0[a], noms[a])
    Ts = args.serial_steps
    t0 = time.perf_counter()
    for s in range(Ts):
        for a in range(N):
            u, _ = ctxs[a].step(s, sc.x0[a], noms[a])
    el_s = time.perf_counter() - t0
    serial_rate = work * Ts / el_s
    for h in ctxs:
        h.close()

    # Reference-style uniform batch padded to the largest K and H.
    Kmax, Hmax = int(sc.K.max()), int(sc.H.max())
    pad = HipMPPI(6, 2, Hmax, Kmax, sc.lam, sc.sigma, sc.u_min, sc.u_max)
    pad.link("bicycle.hip", "bicycle_dynamics", "bicycle_pursuit_cost", "bicycle_terminal_cost",
             len(mb), scenes.BICYCLE_COST_BYTES, 42)
    pad.bind_params(mb, sc.cost_blobs[0])
    pad.simulate_batch(N, 2, 0, sc.x0, outputs=())
    Tp = args.serial_steps
    t0 = time.perf_counter()
    pad.simulate_batch(N, Tp, 0, sc.x0, outputs=())
    el_p = time.perf_counter() - t0
    padded_rate = work * Tp / el_p
    pad.close()
    hip.close()
    return {
        "workload": "256-agent dynamic-bicycle pursuit, uneven K (2^U[10,14]) and H (U[16,96]), "
                    "32 threat zones, closed loop",
        "agents": N, "sample_steps_per_step": work, "steps": T,
        "value": batch_rate, "unit": "sample-steps/s", "ms_per_step": el / T * 1e3,
        "tuning": tuning, "autotune_best_ms": tune_ms, "roofline": roofline,
        "plain_per_agent": {"value": serial_rate, "ms_per_step": el_s / Ts * 1e3, "steps": Ts,
                            "how": "one context per agent, blocking step() per agent"},
        "padded_uniform_batch": {"value_useful": padded_rate, "ms_per_step": el_p / Tp * 1e3,
                                 "K": Kmax, "H": Hmax, "steps": Tp,
                                 "how": "simulate_batch with every agent at the largest K and H"},
        "speedup_vs_plain_per_agent": batch_rate / serial_rate,
        "speedup_vs_padded_batch": batch_rate / padded_rate,
        "other_launch_forms": forms,
    }


def bench_serving(args):
    """Quadrotor serving loop: Poisson arrivals and a burst."""
    from mppi_metal_amd import HipMPPI, HipServer, scenes
    rq = scenes.quadrotor_requests(args.requests, seed=42, rate_hz=2000.0)
    cov = scenes.quadrotor_covariance(42)
    mb = scenes.QuadParams().blob()
    blobs = [rq.cost_blob(i) for i in range(len(rq.K))]
    hover = []
    for h in rq.H:
        u = np.zeros((int(h), 4), np.float32)
        u[:, 0] = 9.81
        hover.append(u)
    work = float(np.sum(rq.K.astype(np.float64) * rq.H))
    out = {"workload": "quadrotor serving, full 4x4 covariance, Poisson 2000 req/s, K 2^U[12,15], H U[30,80]",
           "requests": len(rq.K), "sample_steps": work, "modes": {}}
    for name, mode in (("mixed", 2), ("fp16", 1), ("fp32", 0)):
        hip = HipMPPI(12, 4, 8, 256, scenes.QUAD_LAMBDA, cov.ravel(), scenes.QUAD_U_MIN, scenes.QUAD_U_MAX,
                      covariance_full=True, numeric_mode=mode)
        hip.link("quadrotor.hip", "quad_dynamics", "quad_stage_cost", "quad_terminal_cost", len(mb),
                 scenes.QUAD_COST_BYTES, 42)
        hip.bind_params(mb, None)
        srv = HipServer(hip, 1024, 32 << 20)
        srv.replay(rq.arrival_s[:32] - rq.arrival_s[0], rq.K[:32], rq.H[:32], rq.x0[:32], blobs[:32],
                   u_nominal=hover[:32])                                   # warm-up
        lat, st = srv.replay(rq.arrival_s, rq.K, rq.H, rq.x0, blobs, u_nominal=hover)
        zero = np.zeros(len(rq.K))
        blat, bst = srv.replay(zero, rq.K, rq.H, rq.x0, blobs, u_nominal=hover)
        srv.close()
        hip.close()
        out["modes"][name] = {
            "poisson": {"p50_ms": float(np.median(lat) * 1e3), "p99_ms": float(np.quantile(lat, 0.99) * 1e3),
                        "throughput": st["sample_steps"] / st["seconds"], "seconds": st["seconds"],
                        "batches": st["batches"], "max_batch": st["max_batch"]},
            "burst": {"throughput": bst["sample_steps"] / bst["seconds"], "seconds": bst["seconds"],
                      "p50_ms": float(np.median(blat) * 1e3), "p99_ms": float(np.quantile(blat, 0.99) * 1e3),
                      "batches": bst["batches"], "max_batch": bst["max_batch"],
                      "admission_waits": bst["admission_waits"],
                      "max_workspace_mb": bst["max_workspace_bytes"] / 2 ** 20, "budget_mb": 32},
            "unit": "sample-steps/s",
        }
        print(f"[bench] serving {name}: p50 {out['modes'][name]['poisson']['p50_ms']:.3f} ms "
              f"p99 {out['modes'][name]['poisson']['p99_ms']:.3f} ms, burst "
              f"{out['modes'][name]['burst']['throughput'] / 1e9:.2f} G/s", file=sys.stderr, flush=True)
    return out


def bench_surrogate(args, substeps=2):
    """Learned bicycle dynamics in the rollout vs plugin dynamics (the
    bicycle integrated with `substeps` Euler steps per MPPI step).  Speed on
    pursuit agent 0; the closed-loop cost gap on agents 0..n-1 of the scene
    (own start, target and threat zones; plant = plugin dynamics, same seed),
    each agent's gap (surrogate - plain) / plain and their mean."""
    from mppi_metal_amd import HipMPPI, scenes
    n_agents = max(1, args.surrogate_agents)
    sc = scenes.bicycle_pursuit_scene(n_agents, 42)
    sc.model.substeps = substeps
    mb = sc.model.blob()
    K, H, T = 262144, 64, args.surrogate_steps

    def context(agent):
        cb = sc.cost_blobs[agent]
        h = HipMPPI(6, 2, H, K, sc.lam, sc.sigma, sc.u_min, sc.u_max)
        h.link("bicycle.hip", "bicycle_dynamics", "bicycle_pursuit_cost", "bicycle_terminal_cost",
               len(mb), len(cb), 42)
        h.bind_params(mb, cb)
        return h

    hip = context(0)
    t0 = time.perf_counter()
    recipe = dict(scenes.BICYCLE_SURROGATE_TRAINING, steps=args.train_steps)
    if args.train_transitions:
        recipe["num_transitions"] = args.train_transitions
    hist = hip.surrogate_train(**scenes.BICYCLE_SURROGATE_ENVELOPE, **recipe)
    train_s = time.perf_counter() - t0
    _, rel = hip.surrogate_evaluate(1 << 16, seed=99)
    print(f"[bench] surrogate ({substeps} substeps) trained in {train_s:.2f}s, loss {hist[0]:.3e} -> "
          f"{hist[-20:].mean():.3e}", file=sys.stderr, flush=True)
    P, norm = hip.surrogate_export()
    un = np.zeros((H, 2), np.float32)
    res, gaps = {}, []
    for agent in range(n_agents):
        h = hip if agent == 0 else context(agent)
        if agent:
            h.surrogate_import(P, norm)
        x0 = sc.x0[agent]
        cost = {}
        for name, use in (("plain", False), ("surrogate", True)):
            h.surrogate_use(use)
            if agent == 0:
                h.simulate_batch(1, 3, 0, x0, un, outputs=())
                h.profile(True)
            t0 = time.perf_counter()
            st, ct, _, tt = h.simulate_batch(1, T, 0, x0, un)
            el = time.perf_counter() - t0
            nxt = np.concatenate([st[1:, 0], tt])
            cost[name] = float(scenes.pursuit_realised_cost(sc, agent, nxt, ct[:, 0], np.arange(T)).mean())
            if agent == 0:
                roll_ms, red_ms, launches = h.profile_read()
                h.profile(False)
                res[name] = {"value": K * H * T / el, "ms_per_step": el / T * 1e3,
                             "rollout_us": roll_ms / launches * 1e3, "reduce_us": red_ms / launches * 1e3,
                             "mean_realised_cost": cost[name]}
                print(f"[bench] {name}: {T} steps in {el:.3f}s, mean cost {cost[name]:.3f}", file=sys.stderr,
                      flush=True)
        h.surrogate_use(False)
        if agent:
            h.close()
        gaps.append((cost["surrogate"] - cost["plain"]) / cost["plain"])
    hip.close()
    print(f"[bench] closed-loop gaps over {n_agents} agents: mean {np.mean(gaps):+.4f}, "
          + " ".join(f"{g:+.4f}" for g in gaps), file=sys.stderr, flush=True)
    sur = res["surrogate"]
    sur_s = sur["rollout_us"] * 1e-6
    flops = K * H * SURROGATE_FLOPS / sur_s
    key = "mppi_rollout_surrogate@surrogate"
    roofline = {"bound": "mfma", "kernel": "mppi_rollout_surrogate", "achieved": flops / 1e12,
                "peak": MFMA_PEAK_16 / 1e12, "unit": "TFLOP/s", "frac": flops / MFMA_PEAK_16,
                "issue_ceiling": against_ceiling(K * H * SURROGATE_FLOPS_ISSUED / sur_s, "mfma"),
                "flops_per_sample_step": SURROGATE_FLOPS, "sample_steps_per_launch": K * H,
                "avg_launch_us": sur["rollout_us"],
                "issued_flops_per_sample_step": SURROGATE_FLOPS_ISSUED,
                "issued_frac": K * H * SURROGATE_FLOPS_ISSUED / sur_s / MFMA_PEAK_16,
                "valu_issued": issued(key, K * H, sur_s),
                "lds_bank_conflict_frac": profile_line("lds_bank_conflict_cycles_over_lds_cycles").get(key),
                "traffic": measured_traffic(key),
                "traffic_source": "rocprofv3 FETCH_SIZE + WRITE_SIZE, profiles/" + COUNTERS.name}
    return {
        "workload": "learned dynamic-bicycle surrogate (MLP 8-128-128-6, tanh hidden layers as a binary16 "
                    "polynomial, fp16 MFMA, fp32 accumulate), "
                    f"pursuit, closed loop, plugin dynamics with {substeps} Euler substeps per step",
        "K": K, "H": H, "closed_loop_steps": T, "substeps": substeps,
        "train": {"transitions": recipe["num_transitions"], "batch": recipe["batch"], "steps": recipe["steps"],
                  "learning_rate": recipe["learning_rate"], "seconds": train_s,
                  "loss_first": float(hist[0]), "loss_last": float(hist[-20:].mean()),
                  "one_step_relative_rmse": [float(v) for v in rel]},
        "plain": res["plain"], "surrogate": sur,
        "speed_ratio_surrogate_vs_plain": sur["value"] / res["plain"]["value"],
        "closed_loop_cost_gap": float(np.mean(gaps)),
        "closed_loop_cost_gap_per_agent": [float(g) for g in gaps],
        "closed_loop_cost_gap_how": f"mean over pursuit agents 0..{n_agents - 1} of (surrogate - plain) / plain, "
                                    f"mean realised cost of {T} closed-loop steps each, plant = plugin dynamics",
        "roofline": roofline if substeps == 2 else None,
    }


def copy_calibration():
    """Three clones of a 256 MiB tensor (the runtime's copy kernel): 256 MiB
    read and 256 MiB written each, to scale the HBM byte counters."""
    import torch
    x = torch.ones(64 * 1024 * 1024, dtype=torch.float32, device="cuda")
    torch.cuda.synchronize()
    for _ in range(3):
        y = x.clone()
        torch.cuda.synchronize()
    print("copy calibration done", float(y[0]), file=sys.stderr)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=500)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--latency-steps", type=int, default=200)
    ap.add_argument("--cpu-steps", type=int, default=200)
    ap.add_argument("--cpu-steps-1t", type=int, default=3)
    ap.add_argument("--no-small", action="store_true",
                    help="skip the K = 1024, H = 32 closed loop (GPU and CPU oracle, 200 steps)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--K", type=int, default=65536)
    ap.add_argument("--H", type=int, default=64)
    ap.add_argument("--obstacles", type=int, default=16)
    ap.add_argument("--spt", type=int, default=0, help="reduce samples per thread (0: built-in)")
    ap.add_argument("--no-batch", action="store_true")
    ap.add_argument("--batch-steps", type=int, default=100)
    ap.add_argument("--serial-steps", type=int, default=5)
    ap.add_argument("--batch-tuning", default="",
                    help="block,persistent,wgs_per_cu,reduce_spt,order,launches for the batch instead of "
                         "the autotuner's choice (profiling runs fix it so that traced and untraced runs agree)")
    ap.add_argument("--no-surrogate", action="store_true")
    ap.add_argument("--no-costly", action="store_true",
                    help="skip the surrogate against 64-substep dynamics")
    ap.add_argument("--no-serving", action="store_true")
    ap.add_argument("--requests", type=int, default=1024)
    ap.add_argument("--surrogate-steps", type=int, default=500)
    ap.add_argument("--train-steps", type=int, default=40000)
    ap.add_argument("--train-transitions", type=int, default=0, help="0: the workload's recipe")
    ap.add_argument("--surrogate-agents", type=int, default=8,
                    help="pursuit agents over which the closed-loop cost gap is averaged")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run three device-to-device copies of 256 MiB (the reference kernel "
                         "for scaling the FETCH_SIZE / WRITE_SIZE counters) and exit")
    args = ap.parse_args()
    if args.copy_calibration:
        copy_calibration()
        return
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU: this build has no multi-GPU path "
              "(agents or samples would shard with no or one small exchange per step).",
              file=sys.stderr)
        sys.exit(2)

    from mppi_metal_amd import HipMPPI, scenes

    calibrate()
    K, H, C, n_obs = args.K, args.H, 2, args.obstacles
    default_workload = (K, H, n_obs) == (65536, 64, 16)
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

    # Kernel times from start/stop events attached to each profiled dispatch
    # (the library launches with hipExtModuleLaunchKernel while profiling).
    from mppi_metal_amd import _native
    fused = H * C <= 128 and K < 2 ** 18 and -(-K // 256) <= 256
    hip.profile(True)
    prof_steps = min(args.steps, 100)
    hip.simulate(prof_steps, 0, x0, u0)
    roll_ms, red_ms, launches = hip.profile_read()
    hip.profile(False)
    roll_s = roll_ms / launches / 1e3
    # one agent: the one-pass kernels that take the agent's descriptor from
    # their arguments
    roll_kernel = "mppi_rollout_fused_one" if fused else "mppi_rollout"
    second_kernel = "mppi_fused_combine_one" if fused else "mppi_reduce"
    second_s = red_ms / launches / 1e3     # the combine launch (fused) or the reduce
    w_ss = unicycle_w_ss(C, fused)
    achieved_ops = K * H * w_ss["total"] / roll_s

    # The two-kernel step (rollout, then the noise-regenerating reduce).
    hip.set_step_path(_native.PATH_TWO_KERNEL)
    hip.simulate(max(1, args.warmup), 0, x0, u0)
    t0 = time.perf_counter()
    hip.simulate(args.steps, 0, x0, u0)
    el2 = time.perf_counter() - t0
    hip.profile(True)
    hip.simulate(prof_steps, 0, x0, u0)
    roll2_ms, red2_ms, launches2 = hip.profile_read()
    hip.profile(False)
    hip.set_step_path(_native.PATH_AUTO)
    roll2_s = roll2_ms / launches2 / 1e3
    red_s = red2_ms / launches2 / 1e3
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
    hosts = host_cpus()
    if not args.no_cpu:
        threads = hosts["usable"]
        cps, cdt = cpu_baseline(n_obs, H, K, args.cpu_steps, threads)
        one, one_dt = cpu_baseline(n_obs, H, K, args.cpu_steps_1t, 1)
        cpu = {"value": cps, "unit": "sample-steps/s", "cores": threads, "kind": "port",
               "sample": f"CPU oracle (scalar C++ at -O3 -march=native, FP32, same noise stream and "
                         f"plugin source), first {args.cpu_steps} closed-loop steps of the same workload "
                         f"on {threads} threads ({cdt:.1f} s); 1 thread: first {args.cpu_steps_1t} steps "
                         f"({one_dt:.1f} s)",
               "one_thread": {"value": one, "steps": args.cpu_steps_1t, "seconds": one_dt},
               "host": hosts, "gpu_over_cpu": value / cps, "gpu_over_one_thread": value / one}
    small = None if args.no_small else small_workload(HipMPPI, hosts["usable"], no_cpu=args.no_cpu)

    batch = None if args.no_batch else bench_batch(args)
    surrogate = None if args.no_surrogate else bench_surrogate(args)
    # The same workload on dynamics that cost more than the network (64
    # Euler substeps per step): where the surrogate pays.
    surrogate_costly = None if args.no_surrogate or args.no_costly else bench_surrogate(args, substeps=64)
    serving = None if args.no_serving else bench_serving(args)

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
        "step_path": "one pass (rollout keeping the noise, then combine)" if fused else "two kernels",
        "roofline": {"bound": "valu", "kernel": roll_kernel,
                     "achieved": achieved_ops / 1e12, "peak": VALU_PEAK_LANE_OPS / 1e12,
                     "unit": "T lane-op/s", "frac": achieved_ops / VALU_PEAK_LANE_OPS,
                     "issue_ceiling": against_ceiling(achieved_ops, "valu"),
                     "ops_per_sample_step": w_ss["total"], "ops_breakdown": w_ss,
                     "sample_steps_per_launch": K * H, "avg_launch_us": roll_s * 1e6,
                     "issued": issued(roll_kernel + "@unicycle", K * H, roll_s) if default_workload else None,
                     "traffic": measured_traffic(roll_kernel + "@unicycle") if default_workload else None,
                     "traffic_source": "rocprofv3 FETCH_SIZE + WRITE_SIZE, profiles/" + COUNTERS.name},
        "second_launch": {"kernel": second_kernel, "avg_launch_us": second_s * 1e6,
                          "traffic": measured_traffic(second_kernel + "@unicycle") if default_workload else None,
                          "algorithmic_bytes": combine_bytes(K, H, C) if fused else rb},
        "two_kernel_path": {"value": K * H * args.steps / el2, "ms_per_step": el2 / args.steps * 1e3,
                            "rollout_us": roll2_s * 1e6, "reduce_us": red_s * 1e6,
                            "rollout_frac": K * H * unicycle_w_ss(C, False)["total"] / roll2_s / VALU_PEAK_LANE_OPS,
                            "rollout_issued": issued("mppi_rollout@unicycle", K * H, roll2_s)
                            if default_workload else None},
        "reduce_roofline": {"bound": "hbm", "kernel": "mppi_reduce (two-kernel path)",
                            "achieved": rb / red_s / 1e9, "peak": HBM_PEAK_BPS / 1e9,
                            "unit": "GB/s", "frac": rb / red_s / HBM_PEAK_BPS,
                            "bytes_per_launch": rb, "avg_launch_us": red_s * 1e6,
                            "traffic": measured_traffic("mppi_reduce@unicycle") if default_workload else None},
        "cpu_baseline": cpu,
        "small_workload": small,
        "final_closed_loop_cost": float(kt[-1]),
        "batch": batch,
        "surrogate": surrogate,
        "surrogate_costly_dynamics": surrogate_costly,
        "serving": serving,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
