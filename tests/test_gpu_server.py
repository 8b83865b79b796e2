"""GPU tests of serving (mppi_hip_server_*) on the quadrotor with full 4x4
covariance, through the C ABI.

A solve of several admitted agents must reproduce, agent by agent, the
oracle's step for that agent's own K, H, cost blob and agent index (its id):
sample costs bit-exact, first control within 1e-5 of the update's largest
magnitude, best cost exact.  The device-resident closed loop (plant step and
warm-start shift between solves) follows the oracle's closed loop at the same
bars as the other closed-loop tests (step 0 at 1e-5, later 2e-3).
"""
import numpy as np
import pytest

import oracle_lib as ol
from mppi_metal_amd import HipMPPI, HipServer, MppiHipError, scenes
from problems import seq_rel_err

pytestmark = pytest.mark.gpu

CLOSED_LOOP_TOL = 2e-3
SEED = 42


def _ctx(mode=0):
    cov = scenes.quadrotor_covariance(SEED)
    mb = scenes.QuadParams().blob()
    hip = HipMPPI(12, 4, 8, 256, scenes.QUAD_LAMBDA, cov.ravel(), scenes.QUAD_U_MIN, scenes.QUAD_U_MAX,
                  covariance_full=True, numeric_mode=mode)
    hip.link("quadrotor.hip", "quad_dynamics", "quad_stage_cost", "quad_terminal_cost", len(mb),
             scenes.QUAD_COST_BYTES, SEED)
    hip.bind_params(mb, None)
    return hip, cov, mb


def _oracle(K, H, cov, mb, cb):
    return ol.Problem("quadrotor", 12, 4, H, K, scenes.QUAD_LAMBDA, cov.ravel(), scenes.QUAD_U_MIN,
                      scenes.QUAD_U_MAX, SEED, "quad_dynamics", "quad_stage_cost", "quad_terminal_cost", mb, cb,
                      cov_full=True)


def _hover(H):
    u = np.zeros((H, 4), np.float32)
    u[:, 0] = 9.81
    return u


AGENTS = [  # (K, H, x0 position, goal)
    (4096, 30, (0.5, -0.3, 0.2), (2.0, -1.0, 1.5)),
    (1500, 44, (-1.0, 2.0, 0.0), (0.0, 0.0, 1.0)),
    (6000, 36, (3.0, 1.0, -1.0), (-2.0, 2.0, 0.5)),
    (777, 52, (0.0, 0.0, 0.0), (1.0, 1.0, 1.0)),
]


def _x0(p):
    x = np.zeros(12, np.float32)
    x[:3] = p
    return x


def test_batched_solve_matches_oracle_per_agent(require_gpu):
    hip, cov, mb = _ctx()
    srv = HipServer(hip, 16, 64 << 20)
    ids = [srv.admit(K, H, _x0(p), _hover(H), scenes.quad_cost_blob(g)) for K, H, p, g in AGENTS]
    u0, diag = srv.solve(ids)
    for i, ((K, H, p, g), aid) in enumerate(zip(AGENTS, ids)):
        ora = _oracle(K, H, cov, mb, scenes.quad_cost_blob(g))
        u_o, d_o, c_o, w_o = ora.step(0, _x0(p), _hover(H), agent=aid, threads=8)
        assert np.array_equal(hip.read_costs(i, K), c_o), f"agent {aid}: costs differ"
        assert diag[i, 0] == d_o[0]
        scale = float(np.max(np.abs(u_o)))
        assert np.max(np.abs(u0[i] - u_o[0])) <= 1e-5 * scale, f"agent {aid}"
    srv.close()


def test_closed_loop_stays_on_device(require_gpu):
    hip, cov, mb = _ctx()
    K, H, p, g = AGENTS[0]
    cb = scenes.quad_cost_blob(g)
    srv = HipServer(hip, 4, 16 << 20)
    aid = srv.admit(K, H, _x0(p), _hover(H), cb)
    assert aid == 0
    T = 4
    us, xs = [], []
    for _ in range(T):
        u0, _ = srv.solve([aid])
        x, _ = srv.read(aid, H)
        us.append(u0[0])
        xs.append(x)
    ora = _oracle(K, H, cov, mb, cb)
    so, co, ko, to = ora.simulate(T, 0, _x0(p), u_nom_packed=_hover(H), threads=8)
    assert np.max(np.abs(us[0] - co[0, 0])) <= 1e-5 * float(np.max(np.abs(co)))
    assert np.max(np.abs(np.array(us) - co[:, 0])) <= CLOSED_LOOP_TOL * float(np.max(np.abs(co)))
    assert np.max(np.abs(np.array(xs[:-1]) - so[1:, 0])) <= CLOSED_LOOP_TOL
    assert np.max(np.abs(xs[-1] - to[0])) <= CLOSED_LOOP_TOL
    srv.close()


def test_results_do_not_depend_on_the_batch(require_gpu):
    out = []
    for with_others in (False, True):
        hip, _, _ = _ctx()
        srv = HipServer(hip, 8, 64 << 20)
        ids = [srv.admit(K, H, _x0(p), _hover(H), scenes.quad_cost_blob(g)) for K, H, p, g in AGENTS]
        u0, diag = srv.solve(ids if with_others else ids[:1])
        out.append((u0[0].copy(), diag[0].copy()))
        srv.close()
    assert np.array_equal(out[0][0], out[1][0]) and np.array_equal(out[0][1], out[1][1])


def test_admission_respects_the_budget(require_gpu):
    hip, _, _ = _ctx()
    probe = HipServer(hip, 2, 1 << 20)
    w = probe.workspace_bytes(32768, 80)
    probe.close()
    srv = HipServer(hip, 8, int(2.5 * w))
    a = srv.admit(32768, 80, np.zeros(12, np.float32), None, scenes.quad_cost_blob((0, 0, 1)))
    b = srv.admit(32768, 80, np.zeros(12, np.float32), None, scenes.quad_cost_blob((0, 0, 1)))
    with pytest.raises(MppiHipError, match="HBM budget"):
        srv.admit(32768, 80, np.zeros(12, np.float32), None, scenes.quad_cost_blob((0, 0, 1)))
    used, budget, agents = srv.usage()
    assert used == 2 * w and agents == 2 and budget >= used
    srv.release(a)
    with pytest.raises(MppiHipError, match="not admitted"):
        srv.solve([a])                      # released
    c = srv.admit(4096, 30, np.zeros(12, np.float32), None, scenes.quad_cost_blob((0, 0, 1)))
    assert c == a                           # the freed slot is reused
    u0, diag = srv.solve([b, c])
    assert np.all(np.isfinite(u0)) and np.all(np.isfinite(diag))
    with pytest.raises(MppiHipError, match="not admitted"):
        srv.solve([5])                      # never admitted
    with pytest.raises(MppiHipError):
        srv.solve([b, b])                   # duplicate
    srv.close()


@pytest.mark.parametrize("mode", [0, 2, 1], ids=["fp32", "mixed", "fp16"])
def test_replay_serves_every_request(require_gpu, mode):
    hip, _, _ = _ctx(mode)
    rq = scenes.quadrotor_requests(96, seed=3)
    srv = HipServer(hip, 256, 64 << 20)
    lat, st = srv.replay(rq.arrival_s, rq.K, rq.H, rq.x0, [rq.cost_blob(i) for i in range(96)],
                         u_nominal=[_hover(h) for h in rq.H])
    print(f"[replay mode {mode}] p50 {np.median(lat) * 1e3:.3f} ms p99 {np.quantile(lat, 0.99) * 1e3:.3f} ms "
          f"batches {st['batches']} max batch {st['max_batch']} {st['seconds']:.3f} s", flush=True)
    assert np.all(np.isfinite(lat)) and np.all(lat > 0)
    assert st["sample_steps"] == float(np.sum(rq.K.astype(np.float64) * rq.H))
    assert 1 <= st["batches"] <= 96 and st["seconds"] >= rq.arrival_s[-1]
    assert srv.usage()[2] == 0             # every agent released
    srv.close()


def test_full_size_admission_wave_matches_oracle(require_gpu):
    """The serving benchmark's requests at full size (1024 quadrotor requests,
    K log-uniform in [4096, 32768], H uniform in [30, 80]): the first
    admission wave under the benchmark's 32 MB budget, solved as one batch;
    the agents with the largest K, the largest H and the smallest K of the
    wave against the oracle's step for their own K, H, cost blob and id
    (sample costs bit-exact, best cost exact, first control within 1e-5 of
    the update's largest magnitude)."""
    hip, cov, mb = _ctx()
    rq = scenes.quadrotor_requests(1024, seed=42, rate_hz=2000.0)
    srv = HipServer(hip, 1024, 32 << 20)
    ids, req = [], []
    for i in range(len(rq.K)):
        try:
            ids.append(srv.admit(int(rq.K[i]), int(rq.H[i]), rq.x0[i], _hover(int(rq.H[i])), rq.cost_blob(i)))
        except MppiHipError as e:
            assert "HBM budget" in str(e)
            break
        req.append(i)
    assert len(ids) > 100                   # a wave of a few hundred agents
    u0, diag = srv.solve(ids)
    Ks, Hs = rq.K[req], rq.H[req]
    for j in sorted({int(np.argmax(Ks)), int(np.argmax(Hs)), int(np.argmin(Ks))}):
        i, K, H = req[j], int(Ks[j]), int(Hs[j])
        ora = _oracle(K, H, cov, mb, rq.cost_blob(i))
        u_o, d_o, c_o, _ = ora.step(0, rq.x0[i], _hover(H), agent=ids[j], threads=8)
        assert np.array_equal(hip.read_costs(j, K), c_o), f"request {i} (K={K}, H={H}): costs differ"
        assert diag[j, 0] == d_o[0]
        assert np.max(np.abs(u0[j] - u_o[0])) <= 1e-5 * float(np.max(np.abs(u_o))), f"request {i}"
    srv.close()
