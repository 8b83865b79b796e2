// mlp_oracle.cpp — TEST INFRASTRUCTURE: CPU oracle of the learned-dynamics
// surrogate (an MLP 8 -> 128 -> 128 -> 6 with tanh hidden layers).
//
// No reference counterpart: the reference has no learned model.  The
// GPU kernels (mppi_metal_amd/csrc/surrogate.hip, mppi_mlp.h) compute the
// forward pass with BF16 operands and FP32 accumulation and train with MSE
// loss and Adam; this file states the same mathematics in FP64:
//   z1 = W1 x + b1, h1 = tanh(z1), z2 = W2 h1 + b2, h2 = tanh(z2),
//   y = W3 h2 + b3, loss = mean over batch and outputs of (y - t)^2,
// and its analytic gradients by backpropagation.  `round_bf16` makes the
// FP64 computation see the same rounded operands as the GPU (weights and
// activations rounded to bfloat16, nearest-even), so the remaining difference
// is FP32-versus-FP64 accumulation.
// With bf16 the hidden activations are T(z) = bf16(tanh(r16(float(z)))), the
// GPU's activation (mppi_metal_amd/csrc/mppi_mlp.h): r16 rounds the FP32
// pre-activation to binary16 toward zero, magnitudes above 4 (and
// infinities, NaN) are capped at 4, bf16 is nearest-even of the FP64 tanh.
// Pinned in tests/test_oracle_mlp.py by central finite differences (FP64
// path) and, for T, against numpy over every binary16 magnitude.
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

// bf16 nearest-even of t >= 0, by scaling to 8 significant bits.
double bf16_rne(double t) {
    if (t == 0.0) return 0.0;
    int e;
    const double m = std::frexp(t, &e);            // t = m 2^e, m in [0.5, 1)
    return std::ldexp(std::nearbyint(std::ldexp(m, 8)), e - 8);
}

// Binary16 bits of a >= 0 rounded toward zero (65504 for larger finite
// values, 0x7c00 for infinity and NaN).
uint32_t f16_rtz_bits(float a) {
    uint32_t u;
    std::memcpy(&u, &a, 4);
    if (u >= 0x7f800000u) return 0x7c00u;
    const int e = int(u >> 23) - 127;
    if (e >= 16) return 0x7bffu;
    if (e >= -14) return uint32_t((e + 15) << 10) | ((u >> 13) & 0x3ffu);
    if (e < -24) return 0u;
    return uint32_t(std::floor(std::ldexp(double(a), 24)));   // subnormal: m 2^-24
}

// The GPU's activation T of an FP32 pre-activation.
double act(float z) {
    uint32_t u;
    std::memcpy(&u, &z, 4);
    float a;
    const uint32_t au = u & 0x7fffffffu;
    std::memcpy(&a, &au, 4);
    uint32_t key = f16_rtz_bits(a);
    if (key > 0x4400u) key = 0x4400u;
    const uint32_t ex = key >> 10, mant = key & 0x3ffu;
    const double v = ex == 0 ? std::ldexp(double(mant), -24) : std::ldexp(double(1024 + mant), int(ex) - 25);
    const double t = bf16_rne(std::tanh(v));
    return (u >> 31) ? -t : t;
}

struct Params {
    const float *w1, *b1, *w2, *b2, *w3, *b3;  // row-major [out][in]
};

}  // namespace

extern "C" {

// Forward pass for n samples.  x: [n][8]; y: [n][6].  bf16 != 0 rounds
// weights, inputs and hidden activations to bfloat16 as the GPU does.
void oracle_mlp_forward(const float* w1, const float* b1, const float* w2, const float* b2,
                        const float* w3, const float* b3, const float* x, int n, int bf16,
                        double* y) {
    auto r = [&](double v) { return bf16 ? round_bf16(v) : v; };
    std::vector<double> h1(HID), h2(HID);
    for (int s = 0; s < n; ++s) {
        for (int o = 0; o < HID; ++o) {
            double z = b1[o];
            for (int i = 0; i < IN; ++i) z += r(w1[o * IN + i]) * r(x[s * IN + i]);
            h1[o] = bf16 ? act(float(z)) : std::tanh(z);
        }
        for (int o = 0; o < HID; ++o) {
            double z = b2[o];
            for (int i = 0; i < HID; ++i) z += r(w2[o * HID + i]) * h1[i];
            h2[o] = bf16 ? act(float(z)) : std::tanh(z);
        }
        for (int o = 0; o < OUT; ++o) {
            double z = b3[o];
            for (int i = 0; i < HID; ++i) z += r(w3[o * HID + i]) * h2[i];
            y[s * OUT + o] = z;
        }
    }
}

// Loss and gradients for a batch: loss = sum_{s,o} (y - t)^2 / (n * 6).
// Gradient outputs have the parameters' shapes.  bf16 != 0 rounds what the
// GPU keeps in bfloat16: weights, inputs and activations as in the forward
// pass, and the backward operands dL/dy and dL/dz of both hidden layers (the
// GPU stores them in bf16 for the weight-gradient products); sums in FP64.
double oracle_mlp_grad(const float* w1, const float* b1, const float* w2, const float* b2,
                       const float* w3, const float* b3, const float* x, const float* t, int n,
                       int bf16, double* gw1, double* gb1, double* gw2, double* gb2, double* gw3,
                       double* gb3) {
    auto r = [&](double v) { return bf16 ? round_bf16(v) : v; };
    std::memset(gw1, 0, sizeof(double) * HID * IN);
    std::memset(gb1, 0, sizeof(double) * HID);
    std::memset(gw2, 0, sizeof(double) * HID * HID);
    std::memset(gb2, 0, sizeof(double) * HID);
    std::memset(gw3, 0, sizeof(double) * OUT * HID);
    std::memset(gb3, 0, sizeof(double) * OUT);
    std::vector<double> xr(IN), h1(HID), h2(HID), y(OUT), dy(OUT), dh2(HID), dz2(HID), dh1(HID), dz1(HID);
    double loss = 0.0;
    const double scale = 2.0 / (double(n) * OUT);
    for (int s = 0; s < n; ++s) {
        for (int i = 0; i < IN; ++i) xr[i] = r(x[s * IN + i]);
        for (int o = 0; o < HID; ++o) {
            double z = b1[o];
            for (int i = 0; i < IN; ++i) z += r(w1[o * IN + i]) * xr[i];
            h1[o] = bf16 ? act(float(z)) : std::tanh(z);
        }
        for (int o = 0; o < HID; ++o) {
            double z = b2[o];
            for (int i = 0; i < HID; ++i) z += r(w2[o * HID + i]) * h1[i];
            h2[o] = bf16 ? act(float(z)) : std::tanh(z);
        }
        for (int o = 0; o < OUT; ++o) {
            double z = b3[o];
            for (int i = 0; i < HID; ++i) z += r(w3[o * HID + i]) * h2[i];
            y[o] = z;
            const double e = z - t[s * OUT + o];
            loss += e * e;
            dy[o] = r(scale * e);
        }
        for (int o = 0; o < OUT; ++o) {
            gb3[o] += dy[o];
            for (int i = 0; i < HID; ++i) gw3[o * HID + i] += dy[o] * h2[i];
        }
        for (int i = 0; i < HID; ++i) {
            double a = 0.0;
            for (int o = 0; o < OUT; ++o) a += r(w3[o * HID + i]) * dy[o];
            dh2[i] = a;
            dz2[i] = r(a * (1.0 - h2[i] * h2[i]));
        }
        for (int o = 0; o < HID; ++o) {
            gb2[o] += dz2[o];
            for (int i = 0; i < HID; ++i) gw2[o * HID + i] += dz2[o] * h1[i];
        }
        for (int i = 0; i < HID; ++i) {
            double a = 0.0;
            for (int o = 0; o < HID; ++o) a += r(w2[o * HID + i]) * dz2[o];
            dh1[i] = a;
            dz1[i] = r(a * (1.0 - h1[i] * h1[i]));
        }
        for (int o = 0; o < HID; ++o) {
            gb1[o] += dz1[o];
            for (int i = 0; i < IN; ++i) gw1[o * IN + i] += dz1[o] * xr[i];
        }
    }
    return loss / (double(n) * OUT);
}

// One sample through a network of any input / output width (<= 16) with
// the parameter layout [W1 | b1 | W2 | b2 | W3 | b3] (row-major [out][in]),
// as the GPU rollout evaluates it: bfloat16 weights and inputs, sums
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
        for (uint32_t i = 0; i < in; ++i) z += round_bf16(w1[o * in + i]) * round_bf16(x[i]);
        h1[o] = act(float(z) + b1[o]);
    }
    for (int o = 0; o < HID; ++o) {
        double z = 0.0;
        for (int i = 0; i < HID; ++i) z += round_bf16(w2[o * HID + i]) * h1[i];
        h2[o] = act(float(z) + b2[o]);
    }
    for (uint32_t o = 0; o < out; ++o) {
        double z = 0.0;
        for (int i = 0; i < HID; ++i) z += round_bf16(w3[o * HID + i]) * h2[i];
        y[o] = float(z) + b3[o];
    }
}

// T of n FP32 pre-activations (test access to the activation).
void oracle_mlp_act(const float* z, float* out, size_t n) {
    for (size_t i = 0; i < n; ++i) out[i] = float(act(z[i]));
}

}  // extern "C"
