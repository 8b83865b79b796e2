"""GPU tests of the uneven multi-agent batch (mppi_hip_batch_*), through the C
ABI, against the CPU oracle and against the single-agent path.

Agent a of a batch uses noise counter word 3 = a, so it must reproduce the
oracle's step for agent index a with its own K_a, H_a and parameter blob:
sample costs bit-exact, u_next within 1e-5 of its largest magnitude.  Every
launch shape (block size, tile table or persistent queue, samples per
thread) must give bit-identical costs.
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
]

# The persistent work-queue mode is switched off in the library (set_tuning
# rejects it): one of its launches did not complete on gfx950 and the cause
# has not been found.  Its parity cases stay here, skipped, for when it is.
PERSISTENT_OFF = "persistent work-queue mode switched off: a launch did not complete on gfx950, cause not found"
PERSISTENT_TUNINGS = [
    dict(block=128, persistent=1, wgs_per_cu=2, reduce_spt=2),
    dict(block=256, persistent=1, wgs_per_cu=4, reduce_spt=8),
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


@pytest.mark.parametrize("tuning", TUNINGS + [pytest.param(t, marks=pytest.mark.skip(reason=PERSISTENT_OFF))
                                              for t in PERSISTENT_TUNINGS],
                         ids=lambda t: "-".join(f"{k}{v}" for k, v in t.items()))
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


def test_persistent_mode_is_rejected(require_gpu):
    from mppi_metal_amd import MppiHipError
    hip, b, sc, mb = pursuit(4)
    with pytest.raises(MppiHipError, match="switched off"):
        b.set_tuning(block=128, persistent=1, wgs_per_cu=2)
    assert b.get_tuning()["persistent"] == 0


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
