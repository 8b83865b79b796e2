"""GPU parity at the edges of the step's shapes, through the C ABI, at the
single-step bars of the other parity tests (sample costs and best cost
bit-exact, weights 1e-5 relative, u_next 1e-5 of its largest magnitude):

  three controls   C = 3, so four-normal noise blocks straddle timesteps,
                   with H*C not a multiple of 4 (ragged last block) and K
                   not a multiple of the 256-sample workgroup; the dynamics
                   draws process noise from its random handle
  one-pass edge    H*C = 126 and 128 (the one-pass step's limit) and
                   129, 130 (past it, where the two-kernel step runs),
                   requested and automatic
  largest horizon  H*C = 4096, the library's limit
  tiny K           K = 1 (a single sample takes all the weight) and
                   K = 257 (one sample in the last workgroup)
  closed loop      the plant's process noise (plant stream) in simulate
"""
import numpy as np
import pytest

from problems import check_step, point_mass3, rel_err, seq_rel_err, unicycle

pytestmark = pytest.mark.gpu

PATHS = pytest.mark.parametrize("path", [1, 2], ids=["two_kernel", "fused"])

X0_PM3 = np.array([0.2, -0.1, 0.4, 0.3, 0.0, -0.2], np.float32)


def _u_nom(H, C, seed=5, scale=0.3):
    return (np.random.default_rng(seed).standard_normal((H, C)) * scale).astype(np.float32)


@PATHS
@pytest.mark.parametrize("K,H", [(1000, 7), (300, 9), (4096, 16)])
def test_three_controls_ragged(require_gpu, path, K, H):
    hip, ora = point_mass3(K, H)
    hip.set_step_path(path)
    for step in (0, 3):
        check_step(hip, ora, step, X0_PM3, _u_nom(H, 3), f"point mass K={K} H={H} step {step}")


# (control_dim, horizon): H*C on both sides of the one-pass limit of 128
@pytest.mark.parametrize("path", [0, 2], ids=["auto", "fused_requested"])
@pytest.mark.parametrize("C,H", [(2, 64), (2, 65), (3, 42), (3, 43)])
def test_one_pass_limit(require_gpu, path, C, H):
    if C == 2:
        hip, ora, _ = unicycle(1500, H, 8)
        x0 = np.array([0.5, -0.5, 0.3, 0.4], np.float32)
    else:
        hip, ora = point_mass3(1500, H)
        x0 = X0_PM3
    hip.set_step_path(path)
    check_step(hip, ora, 2, x0, _u_nom(H, C), f"H*C={H * C} path {path}")


def test_largest_horizon(require_gpu):
    H = 2048                                       # H*C = 4096
    hip, ora, sc = unicycle(384, H, 8, lam=1e6)   # costs ~1e6..1e8: ESS ~ 48
    u_nom = _u_nom(H, 2, scale=0.05)
    check_step(hip, ora, 1, np.zeros(4, np.float32), u_nom, "H*C=4096 K=384")


@PATHS
def test_single_sample(require_gpu, path):
    hip, ora, sc = unicycle(1, 16, 8)
    hip.set_step_path(path)
    u, d = check_step(hip, ora, 4, np.zeros(4, np.float32), _u_nom(16, 2), "K=1")
    w = hip.read_weights()
    assert w[0] == 1.0 and d[2] == 1.0 and d[0] == d[1]


@PATHS
def test_one_sample_in_last_workgroup(require_gpu, path):
    hip, ora, sc = unicycle(257, 24, 8)
    hip.set_step_path(path)
    check_step(hip, ora, 6, np.zeros(4, np.float32), _u_nom(24, 2), "K=257")


# The closed-loop bar of the other tests: step 0 at 1e-5; later steps at 2e-3,
# since the GPU's tree sums and the oracle's sequential sums leave last-bit
# differences in the plant state that the softmax amplifies.
CLOSED_LOOP_TOL = 2e-3


@PATHS
def test_simulate_with_process_noise(require_gpu, path):
    H = 11
    hip, ora = point_mass3(2048, H)
    hip.set_step_path(path)
    st, ct, kt = hip.simulate(5, 2, X0_PM3, _u_nom(H, 3))
    so, co, ko, _ = ora.simulate(5, 2, X0_PM3, u_nom_packed=_u_nom(H, 3), threads=8)
    assert kt[0] == ko[0, 0]
    scale = float(np.max(np.abs(co)))
    assert np.max(np.abs(ct[0] - co[0, 0])) <= 1e-5 * scale
    assert np.max(np.abs(st - so[:, 0])) <= CLOSED_LOOP_TOL
    assert np.max(np.abs(ct - co[:, 0])) <= CLOSED_LOOP_TOL * scale
    assert rel_err(kt, ko[:, 0]) <= CLOSED_LOOP_TOL
    # the plant's noise moves the state: without it the loop differs
    quiet_hip, _ = point_mass3(2048, H, process_std=0.0)
    quiet_hip.set_step_path(path)
    sq, _, _ = quiet_hip.simulate(5, 2, X0_PM3, _u_nom(H, 3))
    assert seq_rel_err(sq[1:], st[1:]) > 1e-3


def test_one_pass_wait_give_up_is_reported(require_gpu):
    """The one-pass rollout's bounded wait: when a consumer wavefront runs out
    of polls it flags its agent and the call fails with MPPI_HIP_E_DEVICE
    instead of returning results computed from missing noise.  The host makes
    every wait target unreachable (and the bound short), so every wait gives
    up and every kernel still ends normally; the next calls, with the wait
    restored, are correct again."""
    from mppi_metal_amd import MppiHipError
    hip, ora, sc = unicycle(2048, 32, 8)
    hip.set_step_path(2)
    x0 = np.zeros(4, np.float32)
    hip.set_wait_limit(64, force_timeout=True)
    with pytest.raises(MppiHipError, match="gave up waiting"):
        hip.step(0, x0, _u_nom(32, 2))
    with pytest.raises(MppiHipError, match="gave up waiting"):
        hip.simulate(3, 0, x0, _u_nom(32, 2))
    hip.set_wait_limit(0, force_timeout=False)
    check_step(hip, ora, 0, x0, _u_nom(32, 2), "after a reported give-up")
    st, ct, kt = hip.simulate(3, 0, x0, _u_nom(32, 2))
    assert np.all(np.isfinite(kt))
