"""The shared single-precision elementary functions (include/mppi_hip/mppi_math.h)
evaluated on the GPU inside a rollout, bit for bit against the same source
compiled for the host (the oracle), over random arguments and special values
(signed zeros, infinities, NaN, subnormals, overflow and underflow edges,
long-division fmod quotients).  The range-restricted forms used by Box-Muller
(log of a positive normal number, square root of [2^-96, 2^126] or zero)
are checked over their domains against mppi_logf and the IEEE square root.
The maximum, minimum and clamp (one instruction each on the device, stated
for the host in the header) are checked over random pairs and every pair of
special values, a zero result turned into an infinity of its sign.

A test plugin evaluates function `fn` on input `sample` of its model blob in
the dynamics callable, and the terminal cost returns it, so with H = 1 and a
zero stage cost sample k's cost is 0 + f(x_k) (the addition turns -0 into +0;
the host side applies the same addition).  FP32 mode only: the MIXED and FP16
modes use the hardware's approximate instructions instead.
"""
import numpy as np
import pytest

import oracle_lib as ol
from mppi_metal_amd import HipMPPI

pytestmark = pytest.mark.gpu

MATH_PLUGIN = r'''
#include "mppi_core.h"
// blob: [fn, 0, 0, 0, x_0, y_0, x_1, y_1, ...]; sample k evaluates fn on (x_k, y_k).
MPPI_CALLABLE void eval_fn(float* s, const float* u, const uint8_t* m, mppi_rng r) {
    const float* p = reinterpret_cast<const float*>(m);
    const uint32_t fn = reinterpret_cast<const uint32_t*>(m)[0];
    const float x = p[4 + 2 * r.sample], y = p[5 + 2 * r.sample];
    float v = 0.0f, sn, cs;
    switch (fn) {
        case 0: v = mppi_expf(x); break;
        case 1: v = mppi_logf(x); break;
        case 2: v = mppi_sinf(x); break;
        case 3: v = mppi_cosf(x); break;
        case 4: mppi_sincospif(x, &sn, &cs); v = sn; break;
        case 5: mppi_sincospif(x, &sn, &cs); v = cs; break;
        case 6: v = mppi_tanhf(x); break;
        case 7: v = mppi_sqrtf(x); break;
        case 8: v = mppi_fmodf(x, y); break;
        case 9: v = mppi_log_normalf(x); break;
        case 10: v = mppi_sqrt_boundedf(x); break;
        case 11: v = mppi_maxf(x, y); break;
        case 12: v = mppi_minf(x, y); break;
        default: v = mppi_clampf(x, -0.5f, y); break;
    }
    // a zero result as an infinity of its sign, which the cost's addition keeps
    s[0] = fn >= 11 && v == 0.0f ? 1.0f / v : v;
}
MPPI_CALLABLE float zero_stage(const float* s, const float* u, const uint8_t* m, const uint8_t* c) {
    return 0.0f;
}
MPPI_CALLABLE float value(const float* s, const uint8_t* m, const uint8_t* c) { return s[0]; }
'''

F = np.float32
SPECIAL = np.array([0.0, -0.0, np.inf, -np.inf, np.nan, 1e-45, -1e-45, 1e-40, -1e-40, 1.17549435e-38,
                    1.0, -1.0, 0.5, -0.5, 0.25, 2.0, 88.72, 88.73, -87.33, -103.9, 9.0, -9.0, 0.55,
                    -0.55, 3.14159265, 1.5707963, 6.2831853, 256.0, 1e4, -1e4, 1e30, -3.4e38, 3.4e38],
                   F)


def _inputs(fn, rng):
    if fn == 0:
        x = rng.uniform(-110, 95, 3000)
    elif fn in (1, 7):
        x = np.concatenate([10.0 ** rng.uniform(-45, 38.5, 2500), -rng.uniform(0, 10, 100)])
    elif fn in (2, 3):
        x = np.concatenate([rng.uniform(-20, 20, 2000), rng.uniform(-1e5, 1e5, 1000)])
    elif fn in (4, 5):
        x = np.concatenate([rng.uniform(-4, 4, 2000), rng.uniform(-1e6, 1e6, 500),
                            np.round(rng.uniform(-64, 64, 500)) / 4])
    elif fn == 6:
        x = np.concatenate([rng.uniform(-1, 1, 1500), rng.uniform(-25, 25, 1500)])
    elif fn == 9:                       # positive normal floats
        x = np.concatenate([10.0 ** rng.uniform(-37.9, 38.5, 4000), rng.uniform(2.0 ** -24, 1.0, 4000),
                            [1.17549435e-38, 1.0, 1.0 - 2.0 ** -24, 2.0 ** -24, 3.4e38]])
        return x.astype(F), np.zeros(x.size, F)
    elif fn == 10:                      # [2^-96, 2^126] and signed zeros, dense where Box-Muller lives
        x = np.concatenate([2.0 ** rng.uniform(-96, 126, 4000), rng.uniform(0.0, 34.0, 20000),
                            [0.0, -0.0, 2.0 ** -96, 2.0 ** 126, 1.0, 4.0, 33.27]])
        return x.astype(F), np.zeros(x.size, F)
    elif fn >= 11:                      # max, min, clamp: random pairs and every pair of specials
        gx, gy = (a.ravel() for a in np.meshgrid(SPECIAL, SPECIAL))
        return (np.concatenate([rng.uniform(-3, 3, 2000).astype(F), gx]),
                np.concatenate([rng.uniform(-3, 3, 2000).astype(F), gy]))
    else:                               # fmod: random pairs and every pair of specials
        gx, gy = (a.ravel() for a in np.meshgrid(SPECIAL, SPECIAL))
        yr = rng.uniform(0.05, 60, 2000) * rng.choice([-1, 1], 2000)
        return (np.concatenate([rng.uniform(-1e4, 1e4, 2000).astype(F), gx]),
                np.concatenate([yr.astype(F), gy]))
    x = np.concatenate([x.astype(F), SPECIAL])
    return x, np.zeros_like(x)


def _host(fn, x, y):
    if fn == 8:
        return ol.fmodf(x, y)
    if fn == 9:
        return ol.mathf(1, x)          # mppi_logf: the same core for normal arguments
    if fn == 10:
        return np.sqrt(x)              # IEEE, correctly rounded
    if fn >= 11:
        v = ol.math2f(fn - 11, x, y)
        with np.errstate(divide="ignore"):
            return np.where(v == 0, F(1) / v, v).astype(F)
    return ol.mathf(fn, x)


@pytest.mark.parametrize("fn", range(14), ids=["exp", "log", "sin", "cos", "sinpi", "cospi", "tanh", "sqrt",
                                               "fmod", "log_normal", "sqrt_bounded", "max", "min", "clamp"])
def test_math_bit_for_bit(require_gpu, fn):
    x, y = _inputs(fn, np.random.default_rng(100 + fn))
    K = x.size
    blob = np.zeros(4 + 2 * K, F)
    blob[:4].view(np.uint32)[0] = fn
    blob[4::2], blob[5::2] = x, y
    hip = HipMPPI(1, 1, 1, K, 1.0, [1.0], [-1.0], [1.0])
    hip.link_source(MATH_PLUGIN, "eval_fn", "zero_stage", "value", blob.nbytes, 0, 1, name="math_probe.hip")
    hip.bind_params(blob.tobytes(), None)
    hip.step(0, np.zeros(1, F), np.zeros(1, F))
    got = hip.read_costs()
    want = F(0.0) + _host(fn, x, y)
    nan = np.isnan(want)
    assert np.array_equal(np.isnan(got), nan)
    diff = got[~nan].view(np.uint32) != want[~nan].view(np.uint32)
    assert not diff.any(), f"{diff.sum()} of {K} differ, e.g. x={x[~nan][diff][:4]} gpu={got[~nan][diff][:4]} " \
                           f"host={want[~nan][diff][:4]}"
