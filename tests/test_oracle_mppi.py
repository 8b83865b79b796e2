"""Pins the CPU oracle's MPPI step against facts fixed by the mathematics,
so that a dropped term, a wrong sign or index, or a transposed operand fails
here before the oracle is used to judge the GPU.

  weights  softmax(-c / lambda) from scipy (independent implementation);
           invariants sum w = 1, argmax w = argmin c, 1 <= ESS <= K,
           mean >= min; special cases (equal costs, one dominant sample,
           lambda -> infinity).
  reduce   uniform weights give the sample mean of the noise; one-hot
           weights give that sample's noise; clamping to the bounds.
  rollout  the double integrator is linear, so its costs have a closed form
           given the noise: x_{t+1} = A x_t + B u_t in matrix form; the
           three-axis point mass adds process noise from the dynamics stream
           and is checked from running sums of the same draws.
  noise    normal n = t*C + d of a sample comes from Philox block n // 4,
           lane n % 4; full covariance noise has covariance L L^T = Sigma.
  closed loop  the reference tests' relations: simulate == loop of step +
           plant; batch N=1 == single; terminal(T) == state(T) of T+1.
"""
import numpy as np
import pytest
from scipy.special import softmax

import oracle_lib as ol
from problems import PM3_HI, PM3_LO, PM3_SIGMA, point_mass3_oracle, unicycle_oracle

SIG = np.array([0.7, 0.4], np.float32)
LO, HI = np.array([-1.5, -1.0], np.float32), np.array([1.5, 1.0], np.float32)


def di(K=16, H=6, lam=1.0, seed=3, stage="mppi_stage_cost", terminal="mppi_terminal_cost",
       sigma=SIG, lo=LO, hi=HI, cov_full=False):
    return ol.Problem("double_integrator", 4, 2, H, K, lam, sigma, lo, hi, seed, "mppi_dynamics",
                      stage, terminal, cov_full=cov_full)


def noise64(p, k, step, agent=0):
    """All H*C normals of sample k, FP64, from the documented stream layout."""
    HC = p.H * p.C
    blocks = [ol.normals4(b, k, step, 0, agent, p.s.seed, fp64=True) for b in range((HC + 3) // 4)]
    return np.concatenate(blocks)[:HC].reshape(p.H, p.C)


# ---------------------------------------------------------------- weights --
@pytest.mark.parametrize("fp64", [False, True])
def test_weights_match_softmax_and_invariants(fp64):
    rng = np.random.default_rng(0)
    for K, lam in ((16, 1.0), (257, 3.0), (4096, 0.2)):
        p = di(K=K, lam=lam)
        c = rng.uniform(0.0, 20.0, K).astype(np.float32)
        w, d = p.weights(c, fp64=fp64)
        ref = softmax(-c.astype(np.float64) / np.float64(np.float32(lam)))
        # FP32: the exponent's argument a = (c - min)/lambda carries ~1 ulp of
        # rounding, so exp(-a) is good to about |a| * 2^-23 relative.
        arg = (c.astype(np.float64) - c.min()) / np.float32(lam)
        tol = 1e-12 if fp64 else 2e-6 + 2.4e-7 * arg
        # plus two subnormal steps where the result falls below FLT_MIN
        floor = 0.0 if fp64 else 2 * float(np.finfo(np.float32).smallest_subnormal)
        assert np.all(np.abs(w - ref) <= ref * tol + floor)
        assert abs(w.sum() - 1.0) < (1e-12 if fp64 else 1e-5)
        assert np.argmax(w) == np.argmin(c)
        assert d[0] == c.min()
        assert abs(d[1] - c.astype(np.float64).mean()) <= 1e-5 * c.mean()
        assert 1.0 <= d[2] <= K
        assert abs(d[2] - 1.0 / np.sum(ref ** 2)) <= (1e-9 if fp64 else 1e-4) * d[2]


def test_weights_special_cases():
    p = di(K=1000)
    w, d = p.weights(np.full(1000, 3.25, np.float32))
    assert np.all(w == np.float32(1.0 / 1000)) and abs(d[2] - 1000) < 1e-2
    c = np.full(1000, 50.0, np.float32)
    c[123] = 1.0                                   # lower by 49 lambda
    w, d = p.weights(c)
    assert w[123] > 0.999999 and abs(d[2] - 1.0) < 1e-5
    p_inf = di(K=1000, lam=1e30)                   # lambda -> infinity: uniform
    w, d = p_inf.weights(np.random.default_rng(1).uniform(0, 100, 1000).astype(np.float32))
    assert np.allclose(w, 1e-3, rtol=1e-6)


def test_weights_nonfinite_costs():
    p = di(K=8)
    c = np.array([1, np.nan, 2, np.inf, 0.5, 3, np.nan, 1], np.float32)
    w, d = p.weights(c)
    assert np.all(np.isfinite(w)) and w[1] == 0 and w[3] == 0 and w[6] == 0
    assert abs(w.sum() - 1) < 1e-6 and d[0] == 0.5
    w, d = p.weights(np.full(8, np.nan, np.float32))
    assert np.all(w == np.float32(0.125))


# ----------------------------------------------------------------- reduce --
@pytest.mark.parametrize("fp64", [False, True])
def test_reduce_uniform_weights_is_mean_noise(fp64):
    K, H = 64, 5
    p = di(K=K, H=H, lo=np.array([-50, -50], np.float32), hi=np.array([50, 50], np.float32))
    u_nom = np.random.default_rng(2).uniform(-0.5, 0.5, (H, 2))
    u = p.reduce(9, np.full(K, 1.0 / K), u_nom, fp64=fp64)
    eps = np.mean([noise64(p, k, 9) for k in range(K)], axis=0) * SIG.astype(np.float64)
    assert np.max(np.abs(u - (u_nom + eps))) < (1e-12 if fp64 else 2e-6)


def test_reduce_one_hot_weights_and_clamp():
    K, H = 32, 4
    p = di(K=K, H=H, lo=np.array([-0.1, -0.2], np.float32), hi=np.array([0.1, 0.2], np.float32))
    for j in (0, 17, 31):
        w = np.zeros(K)
        w[j] = 1.0
        u = p.reduce(4, w, np.zeros((H, 2)), fp64=True)
        want = np.clip(noise64(p, j, 4) * SIG.astype(np.float64), p.u_min.astype(np.float64),
                       p.u_max.astype(np.float64))
        assert np.max(np.abs(u - want)) < 1e-12


def test_reduce_abs_is_the_weighted_noise_magnitude():
    """oracle_reduce_abs (the scale of the tests' elementwise update bound):
    sum_k w_k |eps_k| is |eps_j| for one-hot weights, the mean of |eps| for
    uniform weights, and bounds |u - u_nom| (triangle inequality)."""
    K, H = 48, 5
    p = di(K=K, H=H, lo=np.array([-50, -50], np.float32), hi=np.array([50, 50], np.float32))
    sig = SIG.astype(np.float64)
    eps = np.array([noise64(p, k, 6) * sig for k in range(K)])
    for j in (0, 23, 47):
        w = np.zeros(K)
        w[j] = 1.0
        assert np.max(np.abs(p.reduce_abs(6, w) - np.abs(eps[j]))) < 1e-12
    assert np.max(np.abs(p.reduce_abs(6, np.full(K, 1.0 / K)) - np.abs(eps).mean(axis=0))) < 1e-12
    w = np.random.default_rng(4).dirichlet(np.ones(K))
    u_nom = np.zeros((H, 2))
    assert np.all(np.abs(p.reduce(6, w, u_nom, fp64=True) - u_nom) <= p.reduce_abs(6, w) + 1e-12)


# ---------------------------------------------------------------- rollout --
def closed_form_di_costs(p, x0, u_nom, step, with_terminal=True):
    """Costs of the double integrator in matrix form: x_{t+1} = A x_t + B u_t."""
    dt = np.float64(np.float32(0.1))
    A = np.array([[1, 0, dt, 0], [0, 1, 0, dt], [0, 0, 1, 0], [0, 0, 0, 1]])
    B = np.array([[0, 0], [0, 0], [dt, 0], [0, dt]])
    out = []
    for k in range(p.K):
        u = np.clip(u_nom + noise64(p, k, step) * SIG.astype(np.float64), LO, HI)
        x = x0.astype(np.float64)
        c = 0.0
        for t in range(p.H):
            x = A @ x + B @ u[t]
            c += x[0] ** 2 + x[1] ** 2
        if with_terminal:
            c += x[0] ** 2 + x[1] ** 2
        out.append(c)
    return np.array(out)


def test_rollout_double_integrator_closed_form():
    p = di(K=24, H=12)
    x0 = np.array([1.0, -2.0, 0.3, 0.1], np.float32)
    u_nom = np.random.default_rng(4).uniform(-0.5, 0.5, (12, 2))
    want = closed_form_di_costs(p, x0, u_nom, 5)
    got64 = p.rollout(5, x0, u_nom, fp64=True)
    assert np.max(np.abs(got64 - want) / want) < 1e-12
    got32 = p.rollout(5, x0, u_nom.astype(np.float32))
    assert np.max(np.abs(got32 - want) / want) < 1e-5
    # no terminal cost callable: the terminal term disappears
    p2 = di(K=24, H=12, terminal="")
    assert np.max(np.abs(p2.rollout(5, x0, u_nom, fp64=True) -
                         closed_form_di_costs(p2, x0, u_nom, 5, with_terminal=False)) / want) < 1e-12


def closed_form_pm3_costs(p, x0, u_nom, step, agent=0, process_std=0.2):
    """Costs of the three-axis point mass from running sums: with a_t the
    clamped controls and w_t the process normals (draw 0 of the dynamics
    stream at timestep t, lanes 0..2),
        v_t = v_0 + sum_{j<t} (dt a_j + std w_j),   p_{t+1} = p_0 + dt sum_{i<=t} v_i."""
    dt = np.float64(np.float32(0.05))
    std = np.float64(np.float32(process_std))
    goal = np.array([1.0, -0.5, 0.8], np.float32).astype(np.float64)
    w_pos, w_vel, w_u, w_term = 1.0, np.float64(np.float32(0.1)), np.float64(np.float32(0.05)), 5.0
    sig, lo, hi = (np.array(v, np.float32).astype(np.float64) for v in (PM3_SIGMA, PM3_LO, PM3_HI))
    x0 = x0.astype(np.float64)
    out = []
    for k in range(p.K):
        a = np.clip(u_nom + noise64(p, k, step, agent) * sig, lo, hi)
        w = np.array([ol.normals4(t << 8, k, step, 1, agent, p.s.seed, fp64=True)[:3] for t in range(p.H)])
        dv = dt * a + std * w
        v = x0[3:] + np.vstack([np.zeros(3), np.cumsum(dv, axis=0)])    # v_0 .. v_H
        pos = x0[:3] + dt * np.cumsum(v[:-1], axis=0)                   # p_1 .. p_H
        c = (w_pos * np.sum((pos - goal) ** 2) + w_vel * np.sum(v[1:] ** 2) + w_u * np.sum(a ** 2) +
             w_term * np.sum((pos[-1] - goal) ** 2))
        out.append(c)
    return np.array(out)


def test_rollout_point_mass_with_process_noise_closed_form():
    H = 9                                 # H*C = 27: noise blocks straddle timesteps
    p = point_mass3_oracle(K=20, H=H)
    x0 = np.array([0.2, -0.1, 0.4, 0.3, 0.0, -0.2], np.float32)
    u_nom = np.random.default_rng(8).uniform(-0.5, 0.5, (H, 3))
    for agent in (0, 3):
        want = closed_form_pm3_costs(p, x0, u_nom, 4, agent)
        got64 = p.rollout(4, x0, u_nom, agent=agent, fp64=True)
        assert np.max(np.abs(got64 - want) / want) < 1e-12
        got32 = p.rollout(4, x0, u_nom.astype(np.float32), agent=agent)
        assert np.max(np.abs(got32 - want) / want) < 1e-5
    # without process noise the costs change: the dynamics stream is used
    quiet = point_mass3_oracle(K=20, H=H, process_std=0.0)
    assert not np.allclose(quiet.rollout(4, x0, u_nom, fp64=True), want)


def test_rollout_depends_on_step_agent_and_seed():
    p = di(K=16)
    x0 = np.array([1, 1, 0, 0], np.float32)
    u = np.zeros((6, 2), np.float32)
    base = p.rollout(0, x0, u)
    assert not np.array_equal(base, p.rollout(1, x0, u))
    assert not np.array_equal(base, p.rollout(0, x0, u, agent=1))
    assert not np.array_equal(base, di(K=16, seed=4).rollout(0, x0, u))
    assert np.array_equal(base, p.rollout(0, x0, u, threads=4))


def test_full_covariance_noise_has_sigma_covariance():
    cov = np.array([[0.5, 0.2], [0.2, 0.3]], np.float32)
    p = di(K=4096, H=1, sigma=cov, cov_full=True, lo=np.array([-1e3, -1e3], np.float32),
           hi=np.array([1e3, 1e3], np.float32))
    # one-hot reduce at t=0 returns eps_k = L z_k exactly
    L = np.linalg.cholesky(cov.astype(np.float64))
    eps = []
    for k in range(0, 4096, 4):
        w = np.zeros(4096)
        w[k] = 1.0
        e = p.reduce(1, w, np.zeros((1, 2)), fp64=True)[0]
        assert np.max(np.abs(e - L @ noise64(p, k, 1)[0])) < 1e-6
        eps.append(e)
    assert np.max(np.abs(np.cov(np.array(eps).T) - cov)) < 0.05


def test_unicycle_fp32_vs_fp64_costs():
    ora, _ = unicycle_oracle(1024, 32, 8)
    x0 = np.zeros(4, np.float32)
    c32 = ora.rollout(0, x0, np.zeros((32, 2), np.float32), threads=8)
    c64 = ora.rollout(0, x0, np.zeros((32, 2)), threads=8, fp64=True)
    err = np.max(np.abs(c32 - c64) / c64)
    print(f"[unicycle C1] FP32 oracle cost error vs FP64: {err:.3e}")
    assert err < 1e-4


# ------------------------------------------------------------- closed loop --
def test_simulate_equals_step_loop_with_plant():
    p = di(K=64, H=10, sigma=np.array([1, 1], np.float32), lo=np.array([-1, -1], np.float32),
           hi=np.array([1, 1], np.float32), terminal="")
    x = np.array([1, 1, 0, 0], np.float32)
    un = np.zeros((10, 2), np.float32)
    st, ct, kt, tt = p.simulate(5, 0, x)
    dt = np.float32(0.1)
    for s in range(5):
        u, d, _, _ = p.step(s, x, un)
        assert np.array_equal(st[s, 0], x) and np.array_equal(ct[s, 0], u[0]) and kt[s, 0] == d[0]
        x = np.array([x[0] + x[2] * dt, x[1] + x[3] * dt, x[2] + u[0, 0] * dt, x[3] + u[0, 1] * dt],
                     np.float32)
        un = np.concatenate([u[1:], u[-1:]])
    assert np.array_equal(tt[0], x)


def test_batch_agent0_equals_single_and_terminal_relation():
    p = di(K=64, H=10)
    x0 = np.array([[1, 1, 0, 0], [-1, 0.5, 0, 0]], np.float32)
    s1 = p.simulate(4, 0, x0[0])
    sb = p.simulate(4, 0, x0, n_agents=2)
    st1, ct1, kt1, tt1 = s1
    stb, ctb, ktb, ttb = sb
    assert np.array_equal(st1[:, 0], stb[:, 0]) and np.array_equal(ct1[:, 0], ctb[:, 0])
    assert np.array_equal(kt1[:, 0], ktb[:, 0]) and np.array_equal(tt1[0], ttb[0])
    ext = p.simulate(5, 0, x0, n_agents=2)
    assert np.array_equal(sb[3], ext[0][4])


def test_pursuit_realised_cost_matches_oracle_stage_cost():
    """scenes.pursuit_realised_cost (numpy, FP64) restates the bicycle
    pursuit stage cost that the oracle evaluates in FP32 from the plugin."""
    from mppi_metal_amd import scenes
    sc = scenes.bicycle_pursuit_scene(3, 42)
    a = 2
    mb, cb = sc.model.blob(), sc.cost_blobs[a]
    ora = ol.Problem("bicycle", 6, 2, 8, 64, sc.lam, np.array(sc.sigma, np.float32), sc.u_min, sc.u_max, 42,
                     "bicycle_dynamics", "bicycle_pursuit_cost", "bicycle_terminal_cost", mb, cb)
    rng = np.random.default_rng(1)
    n = 200
    z = np.array(sc.zones, np.float32)
    # half the states inside a threat zone, so the hinge term is exercised
    states = np.zeros((n, 6), np.float32)
    states[:, :2] = rng.uniform(-50, 50, (n, 2))
    states[::2, :2] = z[rng.integers(0, len(z), n // 2), :2] + rng.uniform(-1, 1, (n // 2, 2))
    states[:, 2:] = rng.uniform(-2, 2, (n, 4))
    controls = rng.uniform(-1, 1, (n, 2)).astype(np.float32)
    steps = rng.integers(0, 500, n).astype(np.uint32)
    c32 = ora.stage_costs(states, controls, steps)
    c64 = scenes.pursuit_realised_cost(sc, a, states, controls, steps)
    assert np.max(np.abs(c32 - c64) / np.maximum(1.0, np.abs(c64))) <= 1e-4
    assert np.count_nonzero(c64 > scenes.pursuit_realised_cost(sc, a, states, controls, steps, w_zone=0.0)) > 10


def _unicycle_cost32(sc, x, y, u0, u1):
    """unicycle_stage_cost_obstacles restated in numpy FP32, one rounding per
    operation in the plugin's order: goal and control terms, then the hinge
    max(0, r - |p - c|) summed over slots in order (d2 < r*r, r - sqrt(d2);
    a slot with r <= 0 adds nothing)."""
    f = np.float32
    dx, dy = f(x) - f(sc.goal[0]), f(y) - f(sc.goal[1])
    cost = f(sc.pos_weight) * (dx * dx + dy * dy)
    cost = cost + f(sc.control_weight) * (f(u0) * f(u0) + f(u1) * f(u1))
    pen = f(0.0)
    for cx, cy, r in list(sc.obstacles)[:16]:
        ox, oy, r = f(x) - f(cx), f(y) - f(cy), f(r)
        d2 = ox * ox + oy * oy
        if r > 0 and d2 < r * r:
            pen = pen + (r - np.sqrt(d2))
    return cost + f(sc.obstacle_weight) * pen, pen


@pytest.mark.parametrize("obstacles", [
    [(1.0, 2.0, 0.5), (3.0e4, -2.0e4, 0.01), (-7.5, 0.25, 3.0), (1e-3, 1e-3, 1e-3),
     (2.0, -1.0, 0.0), (2.0, -1.0, -0.7), (4.0, 4.0, float("nan")), (0.5, 0.5, 0.25)],
    [(0.0, 0.0, 2.0 ** -61), (1.0, 1.0, 2.0 ** -59), (-3.0, 2.0, 1e-12)],
    [(5.0, 5.0, 2.0 ** 61)],
])
def test_unicycle_obstacle_square_gate_misses_no_sample(obstacles):
    """The obstacle cost's first pass (a square test in coordinates scaled by
    the radius) only decides where the exact hinge is evaluated: on states
    within a few ulps of every circle, on the axes (where square and circle
    touch) and diagonals, including far centres with small radii, radii at
    and beyond the scaled range, zero, negative and NaN radii, the stage
    cost equals the FP32 hinge evaluated for every slot."""
    from mppi_metal_amd import scenes
    sc = scenes.unicycle_scene(4, 42)
    sc.obstacles = obstacles
    mb, cb = sc.model_blob(), sc.cost_blob()
    ora = ol.Problem("unicycle", 4, 2, 4, 16, sc.lam, np.array(sc.sigma, np.float32), sc.u_min, sc.u_max,
                     42, "mppi_dynamics", "unicycle_stage_cost_obstacles", "unicycle_terminal_cost", mb, cb)
    pts = []
    for cx, cy, r in obstacles:
        if not (r > 0.0):
            r = 0.3
        for ang in np.arange(8) * (np.pi / 4):
            for k in range(-6, 7):
                d = np.float64(r) * (1.0 + k * 2.0 ** -23)
                pts.append((cx + d * np.cos(ang), cy + d * np.sin(ang)))
                # the float next to the axis point, toward and away from the centre
                if ang % (np.pi / 2) == 0:
                    px = np.float32(cx + d * np.cos(ang))
                    pts.append((float(np.nextafter(px, np.float32(cx))), cy + d * np.sin(ang)))
    states = np.zeros((len(pts), 4), np.float32)
    states[:, :2] = np.array(pts, np.float64).astype(np.float32)
    controls = np.tile(np.array([0.3, -0.2], np.float32), (len(pts), 1))
    got = ora.stage_costs(states, controls, np.zeros(len(pts), np.uint32))
    want, pen = np.array([_unicycle_cost32(sc, s[0], s[1], 0.3, -0.2) for s in states], np.float32).T
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32))
    # the states straddle the circles: some pay the hinge, some do not
    assert 0 < np.count_nonzero(pen) < len(pts)


def _counted(centres_radii, xy):
    """Bit j of row i: point i counts circle j in the costs' FP32 arithmetic
    (ox = x - cx, oy = y - cy, d2 = ox*ox + oy*oy, each rounded; r > 0 and
    d2 < r*r)."""
    f = np.float32
    out = np.zeros(len(xy), np.uint32)
    for j, (cx, cy, r) in enumerate(centres_radii):
        r = f(r)
        if not (r > 0):
            continue
        with np.errstate(all="ignore"):
            ox = xy[:, 0].astype(f) - f(cx)
            oy = xy[:, 1].astype(f) - f(cy)
            d2 = ox * ox + oy * oy
            hit = d2 < r * r
        out |= hit.astype(np.uint32) << np.uint32(j)
    return out


def _boundary_points(circles, rng, n_random=2000, box=60.0):
    """Points within a few ulps of every circle (on axes and diagonals, and
    the float next to the axis points), near the corners of the squares
    around them, and random points over the scene."""
    pts = []
    for cx, cy, r in circles:
        if not (r > 0.0) or not np.isfinite(r) or r > 1e30:
            r = 0.3
        for ang in np.arange(8) * (np.pi / 4):
            for k in range(-6, 7):
                d = np.float64(r) * (1.0 + k * 2.0 ** -23)
                pts.append((cx + d * np.cos(ang), cy + d * np.sin(ang)))
                if ang % (np.pi / 2) == 0:
                    px = np.float32(cx + d * np.cos(ang))
                    pts.append((float(np.nextafter(px, np.float32(cx))), cy + d * np.sin(ang)))
        for sx in (-1, 1):
            for sy in (-1, 1):
                pts.append((cx + sx * r, cy + sy * r))
    pts += list(map(tuple, rng.uniform(-box, box, (n_random, 2))))
    with np.errstate(all="ignore"):
        return np.array(pts, np.float64).astype(np.float32)


@pytest.mark.parametrize("obstacles", [
    [(1.0, 2.0, 0.5), (3.0e4, -2.0e4, 0.01), (-7.5, 0.25, 3.0), (1e-3, 1e-3, 1e-3),
     (2.0, -1.0, 0.0), (2.0, -1.0, -0.7), (4.0, 4.0, float("nan")), (0.5, 0.5, 0.25)],
    [(0.0, 0.0, 2.0 ** -61), (1.0, 1.0, 2.0 ** -59), (-3.0, 2.0, 1e-12)],
    [(5.0, 5.0, 2.0 ** 61)],
])
def test_unicycle_square_gate_flags_every_counted_obstacle(obstacles):
    """The GPU's first pass of the obstacle cost (the square gate in scaled
    coordinates, the plugin's unicycle_gate) flags, for every point near the
    circles and over the scene, every slot whose hinge the point pays: the
    exact pass it guards then sees them all, so the GPU's costs are the
    definition's.  A gate that flags too few fails here."""
    from mppi_metal_amd import scenes
    sc = scenes.unicycle_scene(4, 42)
    sc.obstacles = obstacles
    xy = _boundary_points(obstacles, np.random.default_rng(5), box=20.0)
    flags = ol.unicycle_gate_flags(sc.cost_blob(), xy)
    counted = _counted(obstacles, xy)
    assert np.all((counted & ~flags) == 0)
    assert np.count_nonzero(counted) > 0


def _zone_grid_case(kind):
    from mppi_metal_amd import scenes
    sc = scenes.bicycle_pursuit_scene(4, 42)
    zones = list(sc.zones)
    if kind == "edges":
        nan, inf = float("nan"), float("inf")
        zones = zones[:20] + [(1.5, 0.2, 0.0), (2.5, 0.3, -1.0), (4.0, 0.0, nan), (inf, 0.0, 1.0),
                              (3.5, nan, 1.0), (1e-3, 1e-3, 1e-12), (-30.0, 10.0, 2.0 ** 61), (7.0, -7.0, 0.05)]
    elif kind == "far":
        # a tiny zone far from the origin: the frame cannot resolve it and
        # every wavefront tests the zones instead
        zones = [(3.0e6, -2.0e6, 0.01), (3.0e6 + 0.05, -2.0e6, 0.02)]
    blob = bytearray(scenes.bicycle_cost_blob((0.0, 0.0), (0.0, 0.0), zones))
    return zones, bytes(blob)


@pytest.mark.parametrize("kind", ["scene", "edges", "far"])
def test_bicycle_zone_grid_lists_every_counted_zone(kind):
    """The GPU's zone pre-filter for the bicycle pursuit cost (per-column
    and per-row zone sets of a grid over the zones' squares, built from the
    plugin's host-callable frame and band functions) lists, for points
    within a few ulps of every zone, near the squares' corners and over the
    scene, every zone whose hinge the point pays; the exact pass runs for
    the listed zones only, so the GPU's costs are the definition's.  On the
    scene's own zones the grid is also selective (most random points list
    no zone), so the check is not met by listing everything."""
    zones, blob = _zone_grid_case(kind)
    xy = _boundary_points(zones, np.random.default_rng(6))
    cand = ol.bicycle_candidates(blob, xy)
    counted = _counted(zones, xy)
    assert np.all((counted & ~cand) == 0)
    assert np.count_nonzero(counted) > 0
    if kind == "scene":
        far = cand[-2000:]
        assert np.count_nonzero(far == 0) > 1000
