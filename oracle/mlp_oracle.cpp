// mlp_oracle.cpp — TEST INFRASTRUCTURE: CPU oracle of the learned-dynamics
// surrogate (an MLP 8 -> 128 -> 128 -> 6 with tanh hidden layers).
//
// No reference counterpart: the reference has no learned model.  The
// GPU kernels (mppi_metal_amd/csrc/surrogate.hip, mppi_mlp.h) compute the
// forward pass with binary16 (FP16) operands and FP32 accumulation, the
// backward pass of training with bfloat16 operands, and train with MSE loss
// and Adam; this file states the same mathematics in FP64:
//   z1 = W1 x + b1, h1 = f(z1), z2 = W2 h1 + b2, h2 = f(z2),
//   y = W3 h2 + b3, loss = mean over batch and outputs of (y - t)^2,
//   f(z) = Q(clamp01(z)), Q(y) = s P(v), s = 2 y - 1, v = s^2 - 1/2,
//   P(v) = c0 + v (c1 + v (c2 + v (c3 + v c4))), f'(z) = 2 (P(v) + 2 s^2 P'(v))
//   inside (0, 1) and 0 outside,
// with the device's binary16 coefficients (mppi_mlp.h: Q is tanh(2.77 s)
// within 4.31e-3 once rounded), and its analytic gradients by
// backpropagation.  With `rounded` set the
// FP64 computation sees the GPU's rounded operands: weights and inputs of
// the forward pass rounded to binary16 (nearest even), the hidden
// activations T(z) of the GPU (below) in place of f, and what the backward
// pass keeps in bfloat16 (nearest even): the transposed weights, dL/dy, dL/dz
// of both hidden layers and the inputs and activations of the
// weight-gradient products; the remaining difference is FP32-versus-FP64
// accumulation.
// The GPU's activation is T(z) = Q16(y), y = clamp01(r16(float(z))): r16
// rounds the FP32 pre-activation to binary16 (nearest even), clamp01 maps
// NaN to 0, and Q16 evaluates Q as seven binary16 operations, each rounded
// to nearest even: s = fma(y, 2, -1), v = fma(s, s, -1/2), p = fma(v, c4,
// c3), p = fma(v, p, c2), p = fma(v, p, c1), p = fma(v, p, c0), T = s * p.
// Each of those is exact in FP64 before its rounding (binary16 operands, whose
// products and sums span far fewer than 53 bits), so one rounding of the FP64
// value is the binary16 operation's result.
// Pinned in tests/test_oracle_mlp.py by central finite differences (FP64
// path) and, for T, against numpy's binary16 conversions over every
// binary16 value and against tanh itself.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

namespace {

constexpr int IN = 8, HID = 128, OUT = 6;

double round_bf16(double v) {
    float f = float(v);
    uint32_t u;
    std::memcpy(&u, &f, 4);
    if ((u & 0x7f800000u) == 0x7f800000u) return v;  // inf / nan unchanged
    uint32_t lsb = (u >> 16) & 1u;
    u += 0x7fffu + lsb;
    u &= 0xffff0000u;
    std::memcpy(&f, &u, 4);
    return double(f);
}

// Binary16 nearest-even of a finite v below 65520 in magnitude (larger ones
// overflow and are handled by the caller): 11 significant bits above 2^-14,
// multiples of 2^-24 below.
double round_f16(double v) {
    if (v == 0.0 || !std::isfinite(v)) return v;
    int e;
    std::frexp(std::fabs(v), &e);                  // |v| in [2^(e-1), 2^e)
    const int q = std::max(e - 11, -24);           // exponent of the last kept bit
    return std::ldexp(std::nearbyint(std::ldexp(v, -q)), q);
}

// y of the GPU's activation: clamp01(r16(z)), r16 nearest even (magnitudes
// from 65520 on overflow to infinity), NaN to 0.
double act_y(float z) {
    const double v = z;
    if (!(std::fabs(v) < 65520.0)) return v > 0.0 ? 1.0 : 0.0;   // infinities, overflow, NaN
    const double h = round_f16(v);
    if (!(h > 0.0)) return 0.0;                    // also -0
    return h < 1.0 ? h : 1.0;
}

// The tanh polynomial's binary16 coefficients (mppi_mlp.h).
constexpr double TC0 = 1.3583984375, TC1 = -1.0908203125, TC2 = 1.298828125, TC3 = -2.537109375,
                 TC4 = 2.80859375;

// Q evaluated in binary16 (seven rounded operations).
double q16(double y) {
    const double s = round_f16(2.0 * y - 1.0);
    const double v = round_f16(s * s - 0.5);
    double p = round_f16(v * TC4 + TC3);
    p = round_f16(v * p + TC2);
    p = round_f16(v * p + TC1);
    p = round_f16(v * p + TC0);
    return round_f16(s * p);
}

// dT/dz at y (the backward pass's factor): 2 (P(v) + 2 s^2 P'(v)) inside
// (0, 1), 0 at the clamp.
double dq(double y) {
    if (!(y > 0.0 && y < 1.0)) return 0.0;
    const double s = 2.0 * y - 1.0, v = s * s - 0.5;
    const double p = TC0 + v * (TC1 + v * (TC2 + v * (TC3 + v * TC4)));
    const double dp = TC1 + v * (2.0 * TC2 + v * (3.0 * TC3 + v * 4.0 * TC4));
    return 2.0 * (p + 2.0 * s * s * dp);
}

// The GPU's activation T of an FP32 pre-activation.
double act(float z) { return q16(act_y(z)); }

// The exact activation f and its derivative (FP64 path).
double f_exact(double z) {
    const double y = z <= 0.0 ? 0.0 : z >= 1.0 ? 1.0 : z;
    const double s = 2.0 * y - 1.0, v = s * s - 0.5;
    return s * (TC0 + v * (TC1 + v * (TC2 + v * (TC3 + v * TC4))));
}
double df_exact(double z) {
    return z <= 0.0 || z >= 1.0 ? 0.0 : dq(z);
}

struct Params {
    const float *w1, *b1, *w2, *b2, *w3, *b3;  // row-major [out][in]
};

}  // namespace

extern "C" {

// Forward pass for n samples.  x: [n][8]; y: [n][6].  rounded != 0 uses
// the GPU's binary16 operands and activation T.
void oracle_mlp_forward(const float* w1, const float* b1, const float* w2, const float* b2,
                        const float* w3, const float* b3, const float* x, int n, int rounded,
                        double* y) {
    auto r = [&](double v) { return rounded ? round_f16(v) : v; };
    auto f = [&](double z) { return rounded ? act(float(z)) : f_exact(z); };
    std::vector<double> h1(HID), h2(HID);
    for (int s = 0; s < n; ++s) {
        for (int o = 0; o < HID; ++o) {
            double z = b1[o];
            for (int i = 0; i < IN; ++i) z += r(w1[o * IN + i]) * r(x[s * IN + i]);
            h1[o] = f(z);
        }
        for (int o = 0; o < HID; ++o) {
            double z = b2[o];
            for (int i = 0; i < HID; ++i) z += r(w2[o * HID + i]) * h1[i];
            h2[o] = f(z);
        }
        for (int o = 0; o < OUT; ++o) {
            double z = b3[o];
            for (int i = 0; i < HID; ++i) z += r(w3[o * HID + i]) * h2[i];
            y[s * OUT + o] = z;
        }
    }
}

// Loss and gradients for a batch: loss = sum_{s,o} (y - t)^2 / (n * 6).
// Gradient outputs have the parameters' shapes.  rounded != 0 rounds as the
// GPU does (header); sums in FP64.
double oracle_mlp_grad(const float* w1, const float* b1, const float* w2, const float* b2,
                       const float* w3, const float* b3, const float* x, const float* t, int n,
                       int rounded, double* gw1, double* gb1, double* gw2, double* gb2, double* gw3,
                       double* gb3) {
    auto r = [&](double v) { return rounded ? round_f16(v) : v; };      // forward operands
    auto rb = [&](double v) { return rounded ? round_bf16(v) : v; };    // backward operands
    std::memset(gw1, 0, sizeof(double) * HID * IN);
    std::memset(gb1, 0, sizeof(double) * HID);
    std::memset(gw2, 0, sizeof(double) * HID * HID);
    std::memset(gb2, 0, sizeof(double) * HID);
    std::memset(gw3, 0, sizeof(double) * OUT * HID);
    std::memset(gb3, 0, sizeof(double) * OUT);
    std::vector<double> xr(IN), h1(HID), h2(HID), d1(HID), d2(HID), y(OUT), dy(OUT), dz2(HID), dz1(HID);
    double loss = 0.0;
    const double scale = 2.0 / (double(n) * OUT);
    for (int s = 0; s < n; ++s) {
        for (int i = 0; i < IN; ++i) xr[i] = r(x[s * IN + i]);
        for (int o = 0; o < HID; ++o) {
            double z = b1[o];
            for (int i = 0; i < IN; ++i) z += r(w1[o * IN + i]) * xr[i];
            if (rounded) {
                const double yv = act_y(float(z));
                h1[o] = q16(yv);
                d1[o] = dq(yv);
            } else {
                h1[o] = f_exact(z);
                d1[o] = df_exact(z);
            }
        }
        for (int o = 0; o < HID; ++o) {
            double z = b2[o];
            for (int i = 0; i < HID; ++i) z += r(w2[o * HID + i]) * h1[i];
            if (rounded) {
                const double yv = act_y(float(z));
                h2[o] = q16(yv);
                d2[o] = dq(yv);
            } else {
                h2[o] = f_exact(z);
                d2[o] = df_exact(z);
            }
        }
        for (int o = 0; o < OUT; ++o) {
            double z = b3[o];
            for (int i = 0; i < HID; ++i) z += r(w3[o * HID + i]) * h2[i];
            y[o] = z;
            const double e = z - t[s * OUT + o];
            loss += e * e;
            dy[o] = rb(scale * e);
        }
        for (int o = 0; o < OUT; ++o) {
            gb3[o] += dy[o];
            for (int i = 0; i < HID; ++i) gw3[o * HID + i] += dy[o] * rb(h2[i]);
        }
        for (int i = 0; i < HID; ++i) {
            double a = 0.0;
            for (int o = 0; o < OUT; ++o) a += rb(w3[o * HID + i]) * dy[o];
            dz2[i] = rb(a * d2[i]);
        }
        for (int o = 0; o < HID; ++o) {
            gb2[o] += dz2[o];
            for (int i = 0; i < HID; ++i) gw2[o * HID + i] += dz2[o] * rb(h1[i]);
        }
        for (int i = 0; i < HID; ++i) {
            double a = 0.0;
            for (int o = 0; o < HID; ++o) a += rb(w2[o * HID + i]) * dz2[o];
            dz1[i] = rb(a * d1[i]);
        }
        for (int o = 0; o < HID; ++o) {
            gb1[o] += dz1[o];
            for (int i = 0; i < IN; ++i) gw1[o * IN + i] += dz1[o] * rb(double(x[s * IN + i]));
        }
    }
    return loss / (double(n) * OUT);
}

// One sample through a network of any input / output width (<= 16) with
// the parameter layout [W1 | b1 | W2 | b2 | W3 | b3] (row-major [out][in]),
// as the GPU rollout evaluates it: binary16 weights and inputs, sums
// rounded to FP32 before the FP32 bias is added (the matrix cores
// accumulate in FP32), hidden activations T (above).
void oracle_mlp_forward1(const float* P, uint32_t in, uint32_t out, const float* x, float* y) {
    const float* w1 = P;
    const float* b1 = w1 + HID * in;
    const float* w2 = b1 + HID;
    const float* b2 = w2 + HID * HID;
    const float* w3 = b2 + HID;
    const float* b3 = w3 + out * HID;
    double h1[HID], h2[HID];
    for (int o = 0; o < HID; ++o) {
        double z = 0.0;
        for (uint32_t i = 0; i < in; ++i) z += round_f16(w1[o * in + i]) * round_f16(x[i]);
        h1[o] = act(float(z) + b1[o]);
    }
    for (int o = 0; o < HID; ++o) {
        double z = 0.0;
        for (int i = 0; i < HID; ++i) z += round_f16(w2[o * HID + i]) * h1[i];
        h2[o] = act(float(z) + b2[o]);
    }
    for (uint32_t o = 0; o < out; ++o) {
        double z = 0.0;
        for (int i = 0; i < HID; ++i) z += round_f16(w3[o * HID + i]) * h2[i];
        y[o] = float(z) + b3[o];
    }
}

// T of n FP32 pre-activations (test access to the activation).
void oracle_mlp_act(const float* z, float* out, size_t n) {
    for (size_t i = 0; i < n; ++i) out[i] = float(act(z[i]));
}

}  // extern "C"
