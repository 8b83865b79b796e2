"""Shared problem builders: the same workload for the HIP backend and the
oracle (test infrastructure)."""
from __future__ import annotations

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
