"""Shared problem builders: the same workload for the HIP backend and the
oracle (test infrastructure)."""
from __future__ import annotations

import struct

import numpy as np

import oracle_lib as ol
from mppi_metal_amd import HipMPPI, scenes

SEED = 42


def unicycle(K, H, n_obs, seed=SEED, stage="unicycle_stage_cost_obstacles",
             terminal="unicycle_terminal_cost", lam=1.0, cov=None):
    sc = scenes.unicycle_scene(n_obs, seed)
    mb, cb = sc.model_blob(), sc.cost_blob()
    sigma = np.array(sc.sigma, np.float32) if cov is None else np.asarray(cov, np.float32).ravel()
    hip = HipMPPI(4, 2, H, K, lam, sigma, sc.u_min, sc.u_max, covariance_full=cov is not None)
    hip.link("unicycle.hip", "mppi_dynamics", stage, terminal, len(mb), len(cb), seed)
    hip.bind_params(mb, cb)
    ora = ol.Problem("unicycle", 4, 2, H, K, lam, sigma, sc.u_min, sc.u_max, seed,
                     "mppi_dynamics", stage, terminal, mb, cb, cov_full=cov is not None)
    return hip, ora, sc


def unicycle_oracle(K, H, n_obs, seed=SEED, stage="unicycle_stage_cost_obstacles",
                    terminal="unicycle_terminal_cost", lam=1.0):
    """The oracle side only (no GPU needed)."""
    sc = scenes.unicycle_scene(n_obs, seed)
    mb, cb = sc.model_blob(), sc.cost_blob()
    ora = ol.Problem("unicycle", 4, 2, H, K, lam, np.array(sc.sigma, np.float32), sc.u_min,
                     sc.u_max, seed, "mppi_dynamics", stage, terminal, mb, cb)
    return ora, sc


def double_integrator(K=64, H=10, seed=0, stage="mppi_stage_cost", terminal="", sigma=(1.0, 1.0),
                      bound=1.0, lam=1.0):
    lo, hi = [-bound, -bound], [bound, bound]
    hip = HipMPPI(4, 2, H, K, lam, sigma, lo, hi)
    hip.link("double_integrator.hip", "mppi_dynamics", stage, terminal, 0, 0, seed)
    ora = ol.Problem("double_integrator", 4, 2, H, K, lam, sigma, lo, hi, seed, "mppi_dynamics",
                     stage, terminal)
    return hip, ora


def bicycle(K, H, seed=SEED, agent=0, substeps=2, lam=None):
    """Single-agent dynamic-bicycle pursuit: agent `agent` of the pursuit
    scene (start state, moving target, 32 threat zones)."""
    sc = scenes.bicycle_pursuit_scene(max(agent + 1, 1), seed)
    sc.model.substeps = substeps
    mb, cb = sc.model.blob(), sc.cost_blobs[agent]
    lam = sc.lam if lam is None else lam
    sigma = np.array(sc.sigma, np.float32)
    hip = HipMPPI(6, 2, H, K, lam, sigma, sc.u_min, sc.u_max)
    hip.link("bicycle.hip", "bicycle_dynamics", "bicycle_pursuit_cost", "bicycle_terminal_cost",
             len(mb), len(cb), seed)
    hip.bind_params(mb, cb)
    ora = ol.Problem("bicycle", 6, 2, H, K, lam, sigma, sc.u_min, sc.u_max, seed, "bicycle_dynamics",
                     "bicycle_pursuit_cost", "bicycle_terminal_cost", mb, cb)
    return hip, ora, sc


BICYCLE_ENVELOPE = scenes.BICYCLE_SURROGATE_ENVELOPE
BICYCLE_TRAINING = scenes.BICYCLE_SURROGATE_TRAINING


def quadrotor(K, H, numeric_mode=0, seed=SEED, goal=(2.0, -1.0, 1.5)):
    """Quadrotor with the random SPD 4x4 control covariance (full-covariance
    noise), hovering start; the oracle rounds its noise to binary16 when the
    GPU does (FP16 mode)."""
    cov = scenes.quadrotor_covariance(seed)
    mb, cb = scenes.QuadParams().blob(), scenes.quad_cost_blob(goal)
    hip = HipMPPI(12, 4, H, K, scenes.QUAD_LAMBDA, cov.ravel(), scenes.QUAD_U_MIN, scenes.QUAD_U_MAX,
                  covariance_full=True, numeric_mode=numeric_mode)
    hip.link("quadrotor.hip", "quad_dynamics", "quad_stage_cost", "quad_terminal_cost", len(mb), len(cb), seed)
    hip.bind_params(mb, cb)
    ora = ol.Problem("quadrotor", 12, 4, H, K, scenes.QUAD_LAMBDA, cov.ravel(), scenes.QUAD_U_MIN,
                     scenes.QUAD_U_MAX, seed, "quad_dynamics", "quad_stage_cost", "quad_terminal_cost", mb, cb,
                     cov_full=True, noise_f16=numeric_mode == 1)
    return hip, ora


PM3_SIGMA = (0.8, 0.6, 0.7)
PM3_LO, PM3_HI = (-2.0, -2.0, -1.5), (2.0, 2.0, 1.5)


def pm3_blobs(dt=0.05, process_std=0.2, goal=(1.0, -0.5, 0.8), w_pos=1.0, w_vel=0.1, w_control=0.05,
              w_terminal=5.0):
    """Model and cost blobs of the three-axis point mass (plugins/point_mass3.hip)."""
    mb = struct.pack("<4f", dt, process_std, 0.0, 0.0)
    cb = struct.pack("<8f", *goal, w_pos, w_vel, w_control, w_terminal, 0.0)
    return mb, cb


def point_mass3(K, H, seed=SEED, lam=2.0, process_std=0.2, numeric_mode=0):
    """Three-control point mass whose dynamics draws process noise from its
    random handle: noise blocks straddle timesteps (C = 3), and the GPU's and
    the oracle's dynamics streams must agree."""
    mb, cb = pm3_blobs(process_std=process_std)
    hip = HipMPPI(6, 3, H, K, lam, PM3_SIGMA, PM3_LO, PM3_HI, numeric_mode=numeric_mode)
    hip.link("point_mass3.hip", "pm3_dynamics", "pm3_stage_cost", "pm3_terminal_cost", len(mb), len(cb), seed)
    hip.bind_params(mb, cb)
    ora = point_mass3_oracle(K, H, seed, lam, process_std)
    return hip, ora


def point_mass3_oracle(K, H, seed=SEED, lam=2.0, process_std=0.2):
    mb, cb = pm3_blobs(process_std=process_std)
    return ol.Problem("point_mass3", 6, 3, H, K, lam, np.array(PM3_SIGMA, np.float32),
                      np.array(PM3_LO, np.float32), np.array(PM3_HI, np.float32), seed, "pm3_dynamics",
                      "pm3_stage_cost", "pm3_terminal_cost", mb, cb)


def rel_err(a, b, floor=0.0):
    """max |a - b| / max(|b|, floor), elementwise."""
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.max(np.abs(a - b) / np.maximum(np.abs(b), floor)))


def seq_rel_err(a, b):
    """Error of a control sequence relative to its largest magnitude."""
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-30))


def update_rel_err(u, u_ref, u_nom, mag):
    """Elementwise error of an update u = clamp(u_nom + sum_k w_k eps_k)
    relative to the magnitude of what was summed, |u_nom| + sum_k w_k
    |eps_k| (`mag`, Problem.reduce_abs): the conditioning of the sum, so
    the bound holds for components that cancel to near zero as for any
    other."""
    u = np.asarray(u, np.float64)
    r = np.asarray(u_ref, np.float64)
    scale = np.abs(np.asarray(u_nom, np.float64)).reshape(r.shape) + np.asarray(mag, np.float64).reshape(r.shape)
    return float(np.max(np.abs(u - r) / np.maximum(scale, 1e-30)))


# ------------------------------------------------------ GPU-vs-oracle step --
# FP32 bars: sample costs and the best cost bit-exact; weights 1e-5 relative
# elementwise; u_next 1e-5 of the sequence's largest magnitude (tree versus
# sequential sums) and, elementwise, 1e-5 relative to the magnitude of its
# sum (update_rel_err); mean cost and ESS 1e-5 relative.  Errors against the
# FP64 oracle are printed and the costs' bounded.
def report(name, **errs):
    print(f"[{name}] " + " ".join(f"{k}={v:.3e}" for k, v in errs.items()), flush=True)


def check_step(hip, ora, step, x0, u_nom, name, check_fp64=True):
    u_g, d_g = hip.step(step, x0, u_nom)
    c_g, w_g = hip.read_costs(), hip.read_weights()
    u_o, d_o, c_o, w_o = ora.step(step, x0, u_nom, threads=8)
    ulps = np.abs(c_g.view(np.int32).astype(np.int64) - c_o.view(np.int32).astype(np.int64))
    assert np.array_equal(c_g, c_o), f"costs differ: {np.count_nonzero(ulps)} samples, max {ulps.max()} ulp"
    assert d_g[0] == d_o[0]
    assert rel_err(w_g, w_o, 1e-37) <= 1e-5
    assert abs(w_g.sum() - 1.0) <= 1e-5
    assert seq_rel_err(u_g, u_o) <= 1e-5
    assert rel_err(d_g[1:], d_o[1:]) <= 1e-5
    u_elem = update_rel_err(u_g, u_o, u_nom, ora.reduce_abs(step, w_o.astype(np.float64)))
    assert u_elem <= 1e-5
    errs = dict(u_vs_fp32=seq_rel_err(u_g, u_o), u_elem_vs_fp32=u_elem,
                w_vs_fp32=rel_err(w_g, w_o, 1e-37), ess_vs_fp32=rel_err(d_g[2:], d_o[2:]))
    if check_fp64:
        u_d, d_d, c_d, w_d = ora.step(step, x0, np.asarray(u_nom, np.float64), threads=8, fp64=True)
        errs.update(cost_vs_fp64=rel_err(c_g, c_d), u_vs_fp64=seq_rel_err(u_g, u_d),
                    oracle32_u_vs_fp64=seq_rel_err(u_o, u_d))
        assert errs["cost_vs_fp64"] < 1e-3
    report(name, **errs)
    return u_g, d_g
