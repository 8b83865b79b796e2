"""GPU tests of the uneven multi-agent batch (mppi_hip_batch_*), through the C
ABI, against the CPU oracle and against the single-agent path.

Agent a of a batch uses noise counter word 3 = a, so it must reproduce the
oracle's step for agent index a with its own K_a, H_a and parameter blob:
sample costs bit-exact, u_next within 1e-5 of its largest magnitude.  Every
launch shape (two launches or one, block size, tile table or persistent
ticket queue, samples per thread) must give bit-identical costs.
"""
import numpy as np
import pytest

import oracle_lib as ol
from mppi_metal_amd import HipBatch, HipMPPI, scenes
from problems import seq_rel_err, unicycle

pytestmark = pytest.mark.gpu

TUNINGS = [
    dict(block=256, persistent=0, reduce_spt=0),
    dict(block=64, persistent=0, reduce_spt=1),
    dict(block=128, persistent=0, reduce_spt=2, order=1),
    dict(block=256, persistent=0, reduce_spt=8),
    # persistent grids taking tiles from the ticket counters
    dict(block=128, persistent=1, wgs_per_cu=2, reduce_spt=2),
    dict(block=256, persistent=1, wgs_per_cu=4, reduce_spt=8),
    # one launch per step: tile table and persistent grid
    dict(block=256, persistent=0, launches=1),
    dict(block=64, persistent=0, launches=1, order=1),
    dict(block=128, persistent=1, wgs_per_cu=2, launches=1),
    dict(block=256, persistent=1, wgs_per_cu=4, launches=1),
]

def pursuit(n_agents, k_log2=(8, 11), h_range=(8, 40), seed=7):
    sc = scenes.bicycle_pursuit_scene(n_agents, seed, k_log2=k_log2, h_range=h_range)
    mb = sc.model.blob()
    hip = HipMPPI(6, 2, 8, 256, sc.lam, sc.sigma, sc.u_min, sc.u_max)
    hip.link("bicycle.hip", "bicycle_dynamics", "bicycle_pursuit_cost", "bicycle_terminal_cost",
             len(mb), scenes.BICYCLE_COST_BYTES, seed)
    b = HipBatch(hip, sc.K, sc.H)
    b.set_params(mb, sc.cost_blob_packed(), 0, scenes.BICYCLE_COST_BYTES)
    return hip, b, sc, mb


def oracle_for(sc, mb, a, seed=7):
    return ol.Problem("bicycle", 6, 2, int(sc.H[a]), int(sc.K[a]), sc.lam, np.array(sc.sigma, np.float32),
                      sc.u_min, sc.u_max, seed, "bicycle_dynamics", "bicycle_pursuit_cost",
                      "bicycle_terminal_cost", mb, sc.cost_blobs[a])


@pytest.mark.parametrize("tuning", TUNINGS, ids=lambda t: "-".join(f"{k}{v}" for k, v in t.items()))
def test_uneven_batch_matches_oracle(require_gpu, tuning):
    hip, b, sc, mb = pursuit(12)
    b.set_tuning(**tuning)
    rng = np.random.default_rng(3)
    u_nom = (rng.standard_normal(b.seq_total) * 0.1).astype(np.float32)
    b.reset(5, sc.x0, u_nom)
    b.run(1, closed_loop=False)
    x, u, d = b.read()
    for a in range(b.N):
        ora = oracle_for(sc, mb, a)
        un_a = u_nom[b.seq_off[a]:b.seq_off[a] + b.seq[a]]
        u_o, d_o, c_o, w_o = ora.step(5, sc.x0[a], un_a, agent=a, threads=8)
        c_g = hip.read_costs(a, int(sc.K[a]))
        assert np.array_equal(c_g, c_o), f"agent {a}: costs differ"
        assert d[a, 0] == d_o[0]
        assert seq_rel_err(b.u_of(u, a), u_o) <= 1e-5, f"agent {a}"


def test_uniform_batch_equals_simulate_batch(require_gpu):
    hip, ora, sc = unicycle(1024, 24, 8)
    N, T = 6, 4
    x0 = np.zeros((N, 4), np.float32)
    x0[:, 0] = np.linspace(-2, 2, N)
    st, ct, kt, tt = hip.simulate_batch(N, T, 3, x0)
    for tuning in TUNINGS:
        b = HipBatch(hip, [1024] * N, [24] * N)
        b.set_params(sc.model_blob(), sc.cost_blob())
        b.set_tuning(**tuning)
        b.reset(3, x0)
        sb, cb, kb = b.run(T, closed_loop=True, history=True)
        assert np.array_equal(kb[0], kt[0])
        assert np.max(np.abs(sb - st)) <= 1e-4 and np.max(np.abs(cb - ct)) <= 1e-4
        xb, _, _ = b.read()
        assert np.max(np.abs(xb - tt)) <= 1e-4
        b.close()


def test_invalid_tunings_are_rejected(require_gpu):
    from mppi_metal_amd import MppiHipError
    hip, b, sc, mb = pursuit(4)
    with pytest.raises(MppiHipError, match="launches must be 1 or 2"):
        b.set_tuning(block=128, launches=3)
    with pytest.raises(MppiHipError, match="wgs_per_cu"):
        b.set_tuning(block=128, persistent=1, wgs_per_cu=0)
    assert b.get_tuning()["launches"] == 2
    big = HipBatch(hip, [1 << 18, 1024], [8, 8])
    with pytest.raises(MppiHipError, match="K < 2\\^18"):
        big.set_tuning(block=256, launches=1)
    big.close()


@pytest.mark.parametrize("tuning", [dict(block=256, persistent=1, wgs_per_cu=2, reduce_spt=0),
                                    dict(block=128, persistent=1, wgs_per_cu=4, launches=1)],
                         ids=["two_launch", "one_launch"])
def test_persistent_tickets_across_replays_and_layouts(require_gpu, tuning):
    """The ticket counters are never reset by the kernels: each launch takes
    exactly n_tiles + grid tickets.  Many graph replays in a row, then a
    layout change (new tile count and grid), must keep every step equal to
    the tile-table form's."""
    hip, b, sc, mb = pursuit(12)
    ref = HipBatch(hip, sc.K, sc.H)
    ref.set_params(mb, sc.cost_blob_packed(), 0, scenes.BICYCLE_COST_BYTES)
    for t in (tuning, dict(tuning, wgs_per_cu=tuning["wgs_per_cu"] + 1), dict(tuning, block=64)):
        b.set_tuning(**t)
        ref.set_tuning(block=t["block"], persistent=0, reduce_spt=t.get("reduce_spt", 0),
                       launches=t.get("launches", 2))
        for bb in (b, ref):
            bb.reset(2, sc.x0)
        s1, c1, k1 = b.run(9, closed_loop=True, history=True)
        s2, c2, k2 = ref.run(9, closed_loop=True, history=True)
        assert np.array_equal(k1, k2) and np.array_equal(s1, s2) and np.array_equal(c1, c2), t
    ref.close()


def test_one_launch_weights_match_two_launch(require_gpu):
    """The one-launch form forms per-sample weights only when they are read;
    they equal the two-launch form's (the same formula on the same costs and
    normalisers)."""
    hip, b, sc, mb = pursuit(6)
    b.reset(1, sc.x0)
    b.set_tuning(block=256, launches=2)
    b.run(1, closed_loop=False)
    w2 = [hip.read_weights(a, int(sc.K[a])) for a in range(b.N)]
    b.set_tuning(block=256, launches=1)
    b.run(1, closed_loop=False)
    for a in range(b.N):
        w1 = hip.read_weights(a, int(sc.K[a]))
        assert np.max(np.abs(w1 - w2[a])) <= 1e-6 * max(1.0, float(np.max(w2[a])))
        assert abs(float(w1.sum()) - 1.0) <= 1e-5


def test_graph_replay_is_stable(require_gpu):
    # The same open-loop step replayed from the same state gives identical results.
    hip, b, sc, mb = pursuit(20)
    b.reset(0, sc.x0)
    b.run(1, closed_loop=False)
    _, u1, d1 = b.read()
    b.run(3, closed_loop=False)
    _, u2, d2 = b.read()
    assert np.array_equal(u1, u2) and np.array_equal(d1, d2)


def test_autotune_picks_a_valid_shape(require_gpu):
    hip, b, sc, mb = pursuit(64)
    b.reset(0, sc.x0)
    ms = b.autotune(2)
    t = b.get_tuning()
    assert ms > 0 and t["block"] in (64, 128, 256)
    print(f"[autotune] best {ms:.3f} ms/step with {t}")
    b.reset(0, sc.x0)
    st, ct, kt = b.run(3, closed_loop=True, history=True)
    assert np.all(np.isfinite(st)) and np.all(np.isfinite(kt))
    assert np.all(ct[..., 0] >= -0.5 - 1e-6) and np.all(ct[..., 0] <= 0.5 + 1e-6)


@pytest.mark.parametrize("tuning", [dict(block=256, reduce_spt=8),
                                    dict(block=256, persistent=1, wgs_per_cu=2, launches=1)],
                         ids=["two_launch_table", "one_launch_persistent"])
def test_full_size_pursuit_spot_checks(require_gpu, tuning):
    """The 256-agent pursuit at its benchmark sizes (K = round(2^U[10, 14]),
    H uniform in [16, 96], 32 threat zones, seed 42): the agents with the
    largest K, the largest H (H*C up to 192: 48 noise blocks, 6 reduce groups)
    and the smallest K and H, against the oracle's step for that agent."""
    sc = scenes.bicycle_pursuit_scene(256, 42)
    mb = sc.model.blob()
    hip = HipMPPI(6, 2, 8, 256, sc.lam, sc.sigma, sc.u_min, sc.u_max)
    hip.link("bicycle.hip", "bicycle_dynamics", "bicycle_pursuit_cost", "bicycle_terminal_cost",
             len(mb), scenes.BICYCLE_COST_BYTES, 42)
    b = HipBatch(hip, sc.K, sc.H)
    b.set_params(mb, sc.cost_blob_packed(), 0, scenes.BICYCLE_COST_BYTES)
    b.set_tuning(**tuning)
    rng = np.random.default_rng(11)
    u_nom = (rng.standard_normal(b.seq_total) * 0.2).astype(np.float32)
    b.reset(3, sc.x0, u_nom)
    b.run(1, closed_loop=False)
    x, u, d = b.read()
    picks = sorted({int(np.argmax(sc.K)), int(np.argmax(sc.H)), int(np.argmin(sc.K)), int(np.argmin(sc.H))})
    assert int(sc.H.max()) * 2 >= 180 and int(sc.K.max()) >= 12000
    for a in picks:
        ora = oracle_for(sc, mb, a, seed=42)
        un_a = u_nom[b.seq_off[a]:b.seq_off[a] + b.seq[a]]
        u_o, d_o, c_o, _ = ora.step(3, sc.x0[a], un_a, agent=a, threads=8)
        c_g = hip.read_costs(a, int(sc.K[a]))
        assert np.array_equal(c_g, c_o), f"agent {a} (K={sc.K[a]}, H={sc.H[a]}): costs differ"
        assert d[a, 0] == d_o[0]
        err = seq_rel_err(b.u_of(u, a), u_o)
        print(f"[full-size pursuit {tuning}] agent {a} K={sc.K[a]} H={sc.H[a]} u_vs_fp32={err:.3e}")
        assert err <= 1e-5, f"agent {a}"
    # the closed loop at full size: every agent's history starts at its x0
    # and stays finite
    b.reset(0, sc.x0)
    st, ct, kt = b.run(3, closed_loop=True, history=True)
    assert np.array_equal(st[0], sc.x0)
    assert np.all(np.isfinite(st)) and np.all(np.isfinite(kt))
    b.close()
    hip.close()
