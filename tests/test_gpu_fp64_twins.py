"""FP32 plugins against their FP64 twins, through the C ABI.

The oracle restates the dynamic bicycle and the quadrotor in double precision
(oracle/mppi_oracle.cpp: the same equations, FP64 noise and FP64 sums).  The
GPU's FP32 costs are bit-identical to the FP32 oracle's, so their error
against the FP64 twin is the FP32 arithmetic's own: it must stay below 1e-3
relative on every sample.  The GPU's u_next is held to the FP32 oracle's own
distance from the FP64 answer: u_gpu - u_fp64 <= 2 (u_oracle32 - u_fp64) +
1e-5, in units of the sequence's largest magnitude (the GPU's tree sums and
the oracle's sequential sums round differently, neither is closer by
construction).
"""
import numpy as np
import pytest

from problems import bicycle, quadrotor, rel_err, seq_rel_err

pytestmark = pytest.mark.gpu


def _report(name, **v):
    print(f"[{name}] " + " ".join(f"{k}={x:.3e}" for k, x in v.items()), flush=True)


def _check(hip, ora, step, x0, u_nom, name):
    u_g, d_g = hip.step(step, x0, u_nom)
    c_g = hip.read_costs()
    u_o, d_o, c_o, _ = ora.step(step, x0, u_nom, threads=8)
    u_d, d_d, c_d, _ = ora.step(step, x0, u_nom.astype(np.float64), threads=8, fp64=True)
    assert np.array_equal(c_g, c_o)
    e_cost = rel_err(c_g, c_d)
    e_g, e_o = seq_rel_err(u_g, u_d), seq_rel_err(u_o, u_d)
    _report(name, cost_vs_fp64=e_cost, u_vs_fp64=e_g, oracle32_u_vs_fp64=e_o,
            best_cost_vs_fp64=abs(float(d_g[0]) - d_d[0]) / abs(d_d[0]))
    assert e_cost <= 1e-3
    assert e_g <= 2.0 * e_o + 1e-5
    assert abs(float(d_g[0]) - d_d[0]) <= 1e-3 * abs(d_d[0])


@pytest.mark.parametrize("K,H,substeps", [(4096, 64, 2), (2048, 32, 8)])
def test_bicycle_fp32_vs_fp64_twin(require_gpu, K, H, substeps):
    hip, ora, sc = bicycle(K, H, substeps=substeps)
    u_nom = (np.random.default_rng(3).standard_normal((H, 2)) * [0.1, 1.0]).astype(np.float32)
    _check(hip, ora, 2, sc.x0[0], u_nom, f"bicycle K={K} H={H} substeps={substeps}")


def test_quadrotor_fp32_vs_fp64_twin(require_gpu):
    K, H = 4096, 40
    hip, ora = quadrotor(K, H, numeric_mode=0)
    u_nom = np.zeros((H, 4), np.float32)
    u_nom[:, 0] = 9.81
    x0 = np.array([0.5, -0.3, 0.2, 0.1, 0.0, -0.2, 0.05, -0.03, 0.1, 0.0, 0.0, 0.0], np.float32)
    _check(hip, ora, 3, x0, u_nom, "quadrotor FP32, full covariance")
