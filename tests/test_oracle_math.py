"""Accuracy of the shared single-precision elementary functions
(include/mppi_hip/mppi_math.h) against FP64 numpy, and exactness of fmod.
These functions are what make CPU and GPU costs bit-identical; their error
against the true function is bounded here."""
import numpy as np

import oracle_lib as ol


def ulp_err(y32, exact):
    exact = np.asarray(exact, np.float64)
    ulp = np.spacing(np.abs(exact).astype(np.float32)).astype(np.float64)
    ulp = np.maximum(ulp, np.float64(np.finfo(np.float32).smallest_subnormal))
    return np.abs(y32.astype(np.float64) - exact) / ulp


def test_exp_ulp():
    x = np.linspace(-87.0, 88.5, 400001, dtype=np.float32)
    assert ulp_err(ol.mathf(0, x), np.exp(x.astype(np.float64))).max() <= 2.0
    assert ol.mathf(0, np.array([0.0], np.float32))[0] == 1.0
    assert np.isinf(ol.mathf(0, np.array([89.0], np.float32))[0])
    assert ol.mathf(0, np.array([-120.0], np.float32))[0] == 0.0


def test_exp_subnormal_range():
    x = np.linspace(-103.9, -87.4, 20001, dtype=np.float32)
    y = ol.mathf(0, x).astype(np.float64)
    e = np.exp(x.astype(np.float64))
    # subnormal results: absolute error within 2 subnormal steps
    assert np.max(np.abs(y - e)) <= 2 * np.finfo(np.float32).smallest_subnormal * 1.0001


def test_log_ulp():
    x = np.concatenate([np.geomspace(1e-37, 1e37, 400001), np.linspace(0.5, 2.0, 100001)]).astype(np.float32)
    assert ulp_err(ol.mathf(1, x), np.log(x.astype(np.float64))).max() <= 2.0
    # Box-Muller range: u in [2^-24, 1)
    u = ((np.arange(0, 2**23, 37, dtype=np.uint64) * 2 + 1) * 2.0**-24).astype(np.float32)
    assert ulp_err(ol.mathf(1, u), np.log(u.astype(np.float64))).max() <= 2.0


def test_log_every_float_near_one():
    """The logarithm's core is f + f^2 P(f) on the mantissa's f in
    [sqrt(1/2) - 1, sqrt(2) - 1]: every float in [1/2, 2) runs every f the
    polynomial sees (at k = -1, 0 and 1); other binades add k log 2, checked
    on a stride over [2^-30, 2^8).  Bound: one ulp of the double result."""
    lo, hi = np.float32(0.5).view(np.uint32), np.float32(2.0).view(np.uint32)
    x = np.arange(lo, hi, dtype=np.uint32).view(np.float32)
    assert ulp_err(ol.mathf(1, x), np.log(x.astype(np.float64))).max() <= 1.0
    lo, hi = np.float32(2.0**-30).view(np.uint32), np.float32(2.0**8).view(np.uint32)
    x = np.arange(lo, hi, 97, dtype=np.uint32).view(np.float32)
    assert ulp_err(ol.mathf(1, x), np.log(x.astype(np.float64))).max() <= 1.0


def test_sin_cos_abs():
    x = np.linspace(-200.0, 200.0, 400001, dtype=np.float32)
    xd = x.astype(np.float64)
    eps = np.finfo(np.float32).eps
    assert np.max(np.abs(ol.mathf(2, x) - np.sin(xd))) <= 1.5 * eps
    assert np.max(np.abs(ol.mathf(3, x) - np.cos(xd))) <= 1.5 * eps
    small = np.linspace(-0.7, 0.7, 100001, dtype=np.float32)
    assert ulp_err(ol.mathf(2, small[small != 0]), np.sin(small[small != 0].astype(np.float64))).max() <= 2.0


def test_sincospi():
    x = np.linspace(0.0, 2.0, 400001, dtype=np.float32)
    xd = x.astype(np.float64)
    eps = np.finfo(np.float32).eps
    assert np.max(np.abs(ol.mathf(4, x) - np.sin(np.pi * xd))) <= 1.5 * eps
    assert np.max(np.abs(ol.mathf(5, x) - np.cos(np.pi * xd))) <= 1.5 * eps
    # exact special points
    assert ol.mathf(4, np.array([0.5], np.float32))[0] == 1.0
    assert ol.mathf(5, np.array([1.0], np.float32))[0] == -1.0


def test_tanh():
    x = np.linspace(-12.0, 12.0, 400001, dtype=np.float32)
    y = ol.mathf(6, x).astype(np.float64)
    t = np.tanh(x.astype(np.float64))
    assert np.max(np.abs(y - t) / np.maximum(np.abs(t), 1e-30)) <= 4 * np.finfo(np.float32).eps


def test_sqrt_is_correctly_rounded():
    x = np.random.default_rng(0).uniform(0, 1e6, 100000).astype(np.float32)
    assert np.array_equal(ol.mathf(7, x), np.sqrt(x))


def test_fmod_exact():
    rng = np.random.default_rng(3)
    x = rng.uniform(-1e4, 1e4, 200000).astype(np.float32)
    y = rng.uniform(0.1, 50, 200000).astype(np.float32)
    assert np.array_equal(ol.fmodf(x, y), np.fmod(x, y))
    two_pi = np.float32(2.0 * np.float32(3.1415926535))
    th = np.linspace(-10, 10, 100001, dtype=np.float32)
    assert np.array_equal(ol.fmodf(th, np.full_like(th, two_pi)), np.fmod(th, two_pi))


def test_fmod_special_values_bit_for_bit():
    # signed zeros, infinities, NaN, negative and subnormal divisors, and
    # quotients up to 2^200 (the long-division path), against numpy's fmod
    # bit for bit (any NaN matches any NaN)
    f = np.float32
    vals = np.array([0.0, -0.0, np.inf, -np.inf, np.nan, 1e-45, -1e-45, 1e-40, 3.0, -3.0, 6.0, -7.5,
                     2.5, 1e30, -3.4e38, 3.4e38, 0.1, -0.1], f)
    x, y = (a.ravel() for a in np.meshgrid(vals, vals))
    got, want = ol.fmodf(x, y), np.fmod(x, y)
    nan = np.isnan(want)
    assert np.array_equal(np.isnan(got), nan)
    assert np.array_equal(got[~nan].view(np.uint32), want[~nan].view(np.uint32))


def test_max_min_clamp_rules():
    """mppi_maxf / mppi_minf / mppi_clampf on the host: numpy's fmax / fmin
    (a NaN operand yields the other) away from signed zeros, -0 below +0,
    and clamp = min(max(x, lo), hi) with a NaN x giving lo."""
    rng = np.random.default_rng(21)
    x = rng.uniform(-3, 3, 5000).astype(np.float32)
    y = rng.uniform(-3, 3, 5000).astype(np.float32)
    x[::7] = np.nan
    y[::11] = np.nan
    assert np.array_equal(ol.math2f(0, x, y), np.fmax(x, y), equal_nan=True)
    assert np.array_equal(ol.math2f(1, x, y), np.fmin(x, y), equal_nan=True)
    z = np.array([0.0, -0.0], np.float32)
    for a in z:
        for b in z:
            mx = ol.math2f(0, np.array([a]), np.array([b]))[0]
            mn = ol.math2f(1, np.array([a]), np.array([b]))[0]
            assert np.signbit(mx) == (np.signbit(a) and np.signbit(b))
            assert np.signbit(mn) == (np.signbit(a) or np.signbit(b))
    hi = np.full(5000, 1.25, np.float32)
    c = ol.math2f(2, x, hi)
    want = np.where(np.isnan(x), np.float32(-0.5), np.clip(x, -0.5, 1.25))
    assert np.array_equal(c, want)
