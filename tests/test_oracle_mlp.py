"""Pins the surrogate MLP oracle (oracle/mlp_oracle.cpp): its analytic
backpropagation against central finite differences of its own FP64 loss,
and its forward pass against a numpy matrix restatement."""
import ctypes

import numpy as np

import oracle_lib as ol

IN, HID, OUT = 8, 128, 6


def _params(seed=0, spread=(0.4, 0.1, 0.1, 0.1, 0.1, 0.1)):
    """Random parameters; hidden biases centred on 1/2, the centre of the
    activation's linear region."""
    rng = np.random.default_rng(seed)
    shapes = ((HID, IN), (HID,), (HID, HID), (HID,), (OUT, HID), (OUT,))
    centre = (0.0, 0.5, 0.0, 0.5, 0.0, 0.0)
    return [rng.normal(c, s, shape).astype(np.float32) for shape, c, s in zip(shapes, centre, spread)]


# The tanh polynomial's binary16 coefficients (mppi_mlp.h) and the slope it
# is fitted to: Q(y) = s P(s^2 - 1/2), s = 2 y - 1, approximates tanh(A s).
TANH_C = (1.3583984375, -1.0908203125, 1.298828125, -2.537109375, 2.80859375)
TANH_A = 2.77


def _f_exact(z):
    s = 2.0 * np.clip(z, 0.0, 1.0) - 1.0
    v = s * s - 0.5
    c0, c1, c2, c3, c4 = TANH_C
    return s * (c0 + v * (c1 + v * (c2 + v * (c3 + v * c4))))


def _fwd(p, x, bf16=0):
    L = ol.lib()
    L.oracle_mlp_forward.restype = None
    y = np.empty((x.shape[0], OUT), np.float64)
    args = [a.ctypes.data_as(ctypes.POINTER(ctypes.c_float)) for a in p]
    L.oracle_mlp_forward(*args, np.ascontiguousarray(x, np.float32).ctypes.data_as(ctypes.POINTER(ctypes.c_float)),
                         ctypes.c_int(x.shape[0]), ctypes.c_int(bf16), y.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
    return y


def _grad(p, x, t, bf16=0):
    L = ol.lib()
    L.oracle_mlp_grad.restype = ctypes.c_double
    gs = [np.zeros(a.shape, np.float64) for a in p]
    fp = lambda a: a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))
    dp = lambda a: a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
    loss = L.oracle_mlp_grad(*[fp(a) for a in p], fp(np.ascontiguousarray(x, np.float32)),
                             fp(np.ascontiguousarray(t, np.float32)), ctypes.c_int(x.shape[0]),
                             ctypes.c_int(bf16), *[dp(g) for g in gs])
    return loss, gs


def test_forward_matches_numpy():
    p = _params()
    x = np.random.default_rng(1).uniform(-1, 1, (32, IN)).astype(np.float32)
    w1, b1, w2, b2, w3, b3 = [a.astype(np.float64) for a in p]
    h1 = _f_exact(x @ w1.T + b1)
    h2 = _f_exact(h1 @ w2.T + b2)
    y = h2 @ w3.T + b3
    assert np.max(np.abs(_fwd(p, x) - y)) < 1e-12


def test_gradients_match_finite_differences():
    # Pre-activations well inside (0, 1) or beyond it, none within reach of
    # the perturbation of a clamp corner (where f has no second derivative).
    p = _params(2, spread=(0.05, 0.05, 0.04, 0.05, 0.1, 0.1))
    rng = np.random.default_rng(3)
    x = rng.uniform(-1, 1, (16, IN)).astype(np.float32)
    t = rng.normal(0, 1, (16, OUT)).astype(np.float32)
    loss, gs = _grad(p, x, t)
    y = _fwd(p, x)
    assert abs(loss - np.mean((y - t) ** 2)) < 1e-12
    # Perturb float32 parameters by h; compare with the analytic gradient.
    for k, idx in [(0, (5, 3)), (1, (17,)), (2, (40, 77)), (3, (100,)), (4, (2, 64)), (5, (4,))]:
        h = 1e-3
        pp = [a.copy() for a in p]
        base = float(pp[k][idx])
        pp[k][idx] = np.float32(base + h)
        lp, _ = _grad(pp, x, t)
        hp = float(pp[k][idx]) - base
        pp[k][idx] = np.float32(base - h)
        lm, _ = _grad(pp, x, t)
        hm = base - float(pp[k][idx])
        fd = (lp - lm) / (hp + hm)
        assert abs(fd - gs[k][idx]) <= 1e-5 * max(1.0, abs(fd)) + 1e-8, (k, idx, fd, gs[k][idx])


def test_gpu_rounding_changes_little():
    p = _params(4)
    x = np.random.default_rng(5).uniform(-1, 1, (64, IN)).astype(np.float32)
    y64, y16 = _fwd(p, x), _fwd(p, x, bf16=1)
    assert 0 < np.max(np.abs(y64 - y16)) < 0.05


# ---------------------------------------------------------------------------
# The surrogate rollout's pieces (oracle_mlp_forward1, oracle_surrogate_dynamics,
# oracle_rollout_surrogate_range), pinned against numpy restatements.
# ---------------------------------------------------------------------------
def _flat(p):
    return np.concatenate([a.ravel() for a in p]).astype(np.float32)


def _f16(v):
    """Nearest-even binary16 of float64 values (numpy's direct conversion)."""
    return np.asarray(v, np.float64).astype(np.float16).astype(np.float64)


def _np_act(z):
    """The GPU's hidden activation T(z) = Q16(clamp01(r16(z))): numpy's
    binary16 conversion (nearest even, overflow to infinity), every binary16
    operation of Q16 rounded once (its FP64 value is exact)."""
    with np.errstate(over="ignore", invalid="ignore"):
        h = np.asarray(z, np.float32).astype(np.float16).astype(np.float64)
        y = np.where(h > 0, np.minimum(h, 1.0), 0.0)      # NaN, -0 and negatives to 0
    c0, c1, c2, c3, c4 = TANH_C
    s = _f16(2.0 * y - 1.0)
    v = _f16(s * s - 0.5)
    p = _f16(v * c4 + c3)
    for c in (c2, c1, c0):
        p = _f16(v * p + c)
    return _f16(s * p)


def _np_forward1(p, x):
    """binary16 operands, FP64 sums rounded to FP32, FP32 bias, activation T."""
    w1, b1, w2, b2, w3, b3 = p
    f64 = lambda a: np.asarray(a, np.float32).astype(np.float16).astype(np.float64)
    z1 = (f64(x) @ f64(w1).T).astype(np.float32) + b1
    h1 = _np_act(z1)
    z2 = (h1 @ f64(w2).T).astype(np.float32) + b2
    h2 = _np_act(z2)
    return (h2 @ f64(w3).T).astype(np.float32) + b3


def _oracle_act(z):
    z = np.ascontiguousarray(z, np.float32)
    out = np.empty_like(z)
    L = ol.lib()
    L.oracle_mlp_act.restype = None
    L.oracle_mlp_act(_fp(z), _fp(out), ctypes.c_size_t(z.size))
    return out


def test_activation_over_every_binary16_value():
    """T at every binary16 value (infinities and NaN included), just above and
    below each and at random float32 values, against numpy's binary16
    conversions; plus what pins T without numpy: within 4.32e-3 of
    tanh(A (2 y - 1)) at every binary16 y in [0, 1] and of tanh beyond the
    clamp, exactly 0 at y = 1/2, exactly odd about 1/2, saturated outside
    [0, 1], and monotone up to two binary16 steps."""
    v = np.arange(0x10000, dtype=np.uint32).astype(np.uint16).view(np.float16).astype(np.float32)
    fin = v[np.isfinite(v)]
    rng = np.random.default_rng(11)
    for z in (v, np.nextafter(fin, np.float32(np.inf)), np.nextafter(fin, np.float32(-np.inf)),
              rng.uniform(-2, 2, 100000).astype(np.float32),
              np.array([3e38, -3e38, 7e4, -7e4, 65519.0, 65520.0, 1e-30, -1e-30], np.float32)):
        assert np.array_equal(_oracle_act(z), _np_act(z).astype(np.float32), equal_nan=True)
    # every binary16 y in [0, 1] in increasing order
    y = np.arange(0x3C01, dtype=np.uint16).view(np.float16).astype(np.float32)
    t = _oracle_act(y).astype(np.float64)
    yd = y.astype(np.float64)
    assert np.max(np.abs(t - np.tanh(TANH_A * (2.0 * yd - 1.0)))) <= 4.32e-3
    top = float(t[-1])
    assert abs(top - 1.0) <= 4.32e-3 and abs(top - np.tanh(TANH_A)) <= 4.32e-3
    assert _oracle_act(np.array([0.5], np.float32))[0] == 0.0
    # odd about 1/2: for y in [1/2, 1], 1 - y is a binary16 value and T(1 - y) = -T(y)
    hi = y[y >= 0.5]
    assert np.array_equal(_oracle_act((1.0 - hi).astype(np.float32)), -_oracle_act(hi))
    out = np.array([-1e-30, -0.5, -65504.0, -np.inf, np.nan, 1.5, 65504.0, np.inf, 3e38], np.float32)
    assert np.array_equal(_oracle_act(out), np.array([-top] * 5 + [top] * 4, np.float32))
    step = np.spacing(np.abs(t).astype(np.float16)).astype(np.float64)
    # the fit ripples about tanh where tanh is nearly flat: never down by more
    # than two binary16 steps
    assert np.all(np.diff(t) >= -2 * step[1:])
    # a value between binary16 neighbours takes the nearer one, ties the even
    half = np.float32(0.5 + 2.0 ** -12)        # halfway between binary16 0.5 and 0.5 + 2^-11
    above = np.nextafter(half, np.float32(1))
    t = _oracle_act(np.array([half, above, np.float32(0.5 + 2.0 ** -11)], np.float32))
    assert t[0] == 0.0 and t[1] == t[2] > 0.0


def test_activation_derivative_matches_the_function():
    """The backward pass's factor dT/dz (the oracle's FP64 gradients use it)
    against central differences of the exact polynomial across (0, 1), and 0
    outside; tanh's own slope 2 A (1 - tanh^2) within 7 % of its peak 2 A (a
    minimax fit's derivative ripples more than the fit itself)."""
    z = np.linspace(0.01, 0.99, 99)
    h = 1e-6
    fd = (_f_exact(z + h) - _f_exact(z - h)) / (2 * h)
    c0, c1, c2, c3, c4 = TANH_C
    s = 2 * z - 1
    v = s * s - 0.5
    P = c0 + v * (c1 + v * (c2 + v * (c3 + v * c4)))
    dP = c1 + v * (2 * c2 + v * (3 * c3 + v * 4 * c4))
    d = 2 * (P + 2 * s * s * dP)
    assert np.max(np.abs(d - fd)) <= 1e-6
    assert np.max(np.abs(d - 2 * TANH_A * (1 - np.tanh(TANH_A * s) ** 2))) <= 0.07 * 2 * TANH_A
    assert _f_exact(np.array([-0.5, 1.5])).tolist() == _f_exact(np.array([0.0, 1.0])).tolist()


def _fp(a):
    return np.ascontiguousarray(a, np.float32).ctypes.data_as(ctypes.POINTER(ctypes.c_float))


def test_forward1_matches_numpy_f16():
    p = _params(6)
    x = np.random.default_rng(7).uniform(-1, 1, (64, IN)).astype(np.float32)
    L = ol.lib()
    L.oracle_mlp_forward1.restype = None
    P = _flat(p)
    ref = _np_forward1(p, x)
    for s in range(x.shape[0]):
        y = np.empty(OUT, np.float32)
        L.oracle_mlp_forward1(_fp(P), ctypes.c_uint32(IN), ctypes.c_uint32(OUT), _fp(x[s]), _fp(y))
        # numpy sums in another order: a pre-activation can round to the
        # neighbouring float32 and move T by a binary16 step, hence the tolerance.
        assert np.max(np.abs(y - ref[s])) <= 1e-3, s
    # and the FP64 oracle_mlp_forward with rounding agrees to the bias rounding
    assert np.max(np.abs(_fwd(p, x, bf16=1) - ref)) <= 1e-3


def _norm80(center, inv_half, period, mu, sd):
    n = np.zeros(80, np.float32)
    for i, v in enumerate((center, inv_half, period, mu, sd)):
        n[16 * i:16 * i + len(v)] = v
    return n


def test_surrogate_dynamics_normalises_wraps_and_denormalises():
    p = _params(8)
    P = _flat(p)
    D, C = 6, 2
    center = np.array([0, 0, 0, 10, 0, 0, 0, 0], np.float32)
    half = np.array([50, 50, np.pi, 10, 4, 4, 0.5, 4], np.float32)
    inv_half = (1.0 / half).astype(np.float32)
    inv_half[:2] = 0.0                                   # x, y ignored
    period = np.zeros(8, np.float32)
    period[2] = np.float32(2 * np.pi)                    # heading periodic
    mu = np.array([0.1, -0.2, 0.01, 0.3, 0.0, -0.05], np.float32)
    sd = np.array([0.5, 0.5, 0.1, 0.2, 0.3, 0.4], np.float32)
    norm = _norm80(center, inv_half, period, mu, sd)
    L = ol.lib()
    L.oracle_surrogate_dynamics.restype = None
    rng = np.random.default_rng(9)
    for _ in range(20):
        x = np.concatenate([rng.uniform(-80, 80, 2), rng.uniform(-12, 12, 1), rng.uniform(0, 20, 1),
                            rng.uniform(-4, 4, 2)]).astype(np.float32)
        u = rng.uniform(-1, 1, 2).astype(np.float32)
        xo = x.copy()
        L.oracle_surrogate_dynamics(ctypes.c_uint32(D), ctypes.c_uint32(C), _fp(xo), _fp(u), _fp(P), _fp(norm))
        v = np.concatenate([x, u]).astype(np.float32)
        d = (v - center).astype(np.float32)
        wrapped = d[2] - period[2] * np.float32(np.round(d[2] / period[2]))
        assert -np.pi - 1e-5 <= wrapped <= np.pi + 1e-5
        d[2] = wrapped
        xin = (d * inv_half).astype(np.float32)
        y = _np_forward1(p, xin[None, :])[0]
        expect = (x + (y * sd + mu).astype(np.float32)).astype(np.float32)
        assert np.max(np.abs(xo - expect)) <= 1e-3 * np.max(sd) + 1e-6
    # Heading and heading + 2 pi give the same step (periodic input); the
    # ignored position does not change the delta.
    x = np.array([1, 2, 0.3, 8, 0.5, -0.2], np.float32)
    u = np.array([0.1, 1.0], np.float32)
    outs = []
    for shift in ((0, 0, 0), (30, -40, 2 * np.pi)):
        xs = x.copy()
        xs[0] += shift[0]; xs[1] += shift[1]; xs[2] += np.float32(shift[2])
        xo = xs.copy()
        L.oracle_surrogate_dynamics(ctypes.c_uint32(D), ctypes.c_uint32(C), _fp(xo), _fp(u), _fp(P), _fp(norm))
        outs.append(xo - xs)
    assert np.max(np.abs(outs[0] - outs[1])) <= 1e-5


def test_surrogate_rollout_constant_delta_closed_form():
    """W3 = 0: every step adds delta = b3 * sd + mu, so positions move on a
    straight line whatever the controls; the double integrator's quadratic
    stage and terminal costs then have a closed form, evaluated in float32 in
    the same order."""
    D, C, H, K = 4, 2, 12, 8
    Ptot = HID * (D + C) + HID + HID * HID + HID + D * HID + D
    P = np.random.default_rng(10).normal(0, 0.1, Ptot).astype(np.float32)
    w3 = HID * (D + C) + HID + HID * HID + HID
    P[w3:w3 + D * HID] = 0.0
    b3 = np.array([0.5, -0.25, 0.125, 1.0], np.float32)
    P[w3 + D * HID:] = b3
    mu = np.array([0.01, 0.02, 0.0, -0.5], np.float32)
    sd = np.array([0.2, 0.4, 1.0, 0.5], np.float32)
    norm = _norm80(np.zeros(6), np.ones(6), np.zeros(6), mu, sd)
    prob = ol.Problem("double_integrator", D, C, H, K, 1.0, [1.0, 1.0], [-1, -1], [1, 1], 3,
                      "mppi_dynamics", "mppi_stage_cost", "mppi_terminal_cost")
    x0 = np.array([1.0, -2.0, 0.5, 0.0], np.float32)
    un = np.zeros(H * C, np.float32)
    costs = np.full(K, np.nan, np.float32)
    L = ol.lib()
    rc = L.oracle_rollout_surrogate_range(prob.ref, ctypes.c_uint32(0), ctypes.c_uint32(0), _fp(x0), _fp(un),
                                          ctypes.c_uint32(0), ctypes.c_uint32(K), _fp(P), _fp(norm), _fp(costs))
    assert rc == 0
    delta = (b3 * sd + mu).astype(np.float32)
    x = x0.copy()
    total = np.float32(0)
    for _ in range(H):
        x = (x + delta).astype(np.float32)
        total = np.float32(total + np.float32(np.float32(x[0] * x[0]) + np.float32(x[1] * x[1])))
    total = np.float32(total + np.float32(np.float32(x[0] * x[0]) + np.float32(x[1] * x[1])))
    assert np.all(costs == total)
