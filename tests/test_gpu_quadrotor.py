"""GPU parity of the quadrotor workload (state 12, control 4, full 4x4 SPD
control covariance) in the three numeric modes, through the C ABI.

FP32   sample costs bit-exact against the oracle, weights 1e-5 relative,
       u_next 1e-5 of its largest magnitude (as every FP32 path).
MIXED  fast hardware transcendentals, fused multiply-adds, approximate
       division (the plugin's trigonometry and the noise path).  Against the
       FP32 oracle: u_next within 1e-3 of its largest magnitude, weights
       within 1e-3 in total variation (sum |w_gpu - w_oracle|), best cost and
       mean cost within 1e-4 relative.
FP16   MIXED with the control noise rounded to binary16; the oracle rounds its
       noise the same way, so the same bars hold.
Every mode's u_next error against the FP64 oracle is printed beside the FP32
oracle's own.
"""
import numpy as np
import pytest

from problems import quadrotor, rel_err, seq_rel_err

pytestmark = pytest.mark.gpu

X0 = np.array([0.5, -0.3, 0.2, 0.1, 0.0, -0.2, 0.05, -0.03, 0.1, 0.0, 0.0, 0.0], np.float32)


def _u_nom(H):
    u = np.zeros((H, 4), np.float32)
    u[:, 0] = 9.81
    return u


def _report(name, **v):
    print(f"[{name}] " + " ".join(f"{k}={x:.3e}" for k, x in v.items()), flush=True)


@pytest.mark.parametrize("mode,name", [(0, "fp32"), (2, "mixed"), (1, "fp16")])
def test_quadrotor_full_covariance(require_gpu, mode, name):
    K, H = 4096, 40
    hip, ora = quadrotor(K, H, numeric_mode=mode)
    u_nom = _u_nom(H)
    u_g, d_g = hip.step(3, X0, u_nom)
    c_g, w_g = hip.read_costs(), hip.read_weights()
    u_o, d_o, c_o, w_o = ora.step(3, X0, u_nom, threads=8)
    u_d, _, _, _ = ora.step(3, X0, u_nom.astype(np.float64), threads=8, fp64=True)
    tv = float(np.abs(w_g.astype(np.float64) - w_o).sum())
    _report(f"quadrotor {name}", u_vs_fp32=seq_rel_err(u_g, u_o), weights_tv=tv,
            cost_rel_max=float(np.max(np.abs(c_g - c_o) / np.abs(c_o))),
            u_vs_fp64=seq_rel_err(u_g, u_d), oracle32_u_vs_fp64=seq_rel_err(u_o, u_d))
    assert np.all(np.isfinite(c_g)) and abs(float(w_g.sum()) - 1.0) <= 1e-5
    if mode == 0:
        assert np.array_equal(c_g, c_o)
        assert rel_err(w_g, w_o, 1e-37) <= 1e-5
        assert seq_rel_err(u_g, u_o) <= 1e-5
    else:
        assert seq_rel_err(u_g, u_o) <= 1e-3
        assert tv <= 1e-3
        assert rel_err(d_g[:2], d_o[:2]) <= 1e-4


def test_mixed_mode_differs_from_fp32_only_by_rounding(require_gpu):
    """MIXED really runs other code (its costs are not bit-identical to
    FP32's) yet follows the same noise stream (per-sample costs agree to
    1e-4)."""
    K, H = 4096, 40
    h32, _ = quadrotor(K, H, numeric_mode=0)
    hmx, _ = quadrotor(K, H, numeric_mode=2)
    u_nom = _u_nom(H)
    h32.step(0, X0, u_nom)
    hmx.step(0, X0, u_nom)
    c32, cmx = h32.read_costs(), hmx.read_costs()
    rel = np.abs(cmx.astype(np.float64) - c32) / np.abs(c32)
    assert not np.array_equal(c32, cmx)
    assert rel.max() <= 1e-4


def test_closed_loop_reaches_goal_in_every_mode(require_gpu):
    for mode in (0, 2, 1):
        hip, _ = quadrotor(4096, 40, numeric_mode=mode)
        st, ct, kt, tt = hip.simulate_batch(1, 150, 0, X0, _u_nom(40))
        err0 = np.linalg.norm(X0[:3] - np.array([2.0, -1.0, 1.5]))
        err = np.linalg.norm(tt[0, :3] - np.array([2.0, -1.0, 1.5]))
        print(f"[closed loop mode {mode}] position error {err0:.3f} -> {err:.3f}", flush=True)
        assert np.all(np.isfinite(st)) and err < 0.25 * err0
