// mppi_mlp.h — the learned-dynamics surrogate's network on the device: an MLP
// in -> 128 -> 128 -> out with hidden activation T (below), in = state_dim +
// control_dim <= 16, out = state_dim <= 16.  The forward pass runs binary16
// (FP16) operands on the matrix cores with FP32 accumulation; the backward
// pass of training runs bfloat16 operands (gradients need bf16's range).
//
// Layout: every layer is computed transposed, Z^T = W X^T, so the samples are
// the N dimension of v_mfma_f32_16x16x16_{f16,bf16} and the hidden units are
// M.  In that instruction lane l holds
//   A (16x16, M x K):  A[l % 16][4 * (l / 16) + j]          j = 0..3
//   B (16x16, K x N):  B[4 * (l / 16) + j][l % 16]
//   D (16x16, M x N):  D[4 * (l / 16) + j][l % 16]
// so the result tile of hidden block m is, register for register, the B
// operand of K block m of the next layer: layers chain without moving data
// between lanes.  Weights are kept as ready-made A fragments (64 lanes x 4
// 16-bit values each, "fragment-major"), built from the FP32 master copy by
// the pack kernel after every optimiser step.
//
// Fragment table (indices into the fragment buffer, 256 values per fragment):
//   W1   : 8 fragments, m-block mb:            A[i][k] = W1[16mb + i][k]       (k < in)     FP16
//   W2   : 64 fragments, mb * 8 + kb:          A[i][k] = W2[16mb + i][16kb + k]             FP16
//   W3   : 8 fragments, k-block kb:            A[i][k] = W3[i][16kb + k]       (i < out)    FP16
//   W2T  : 64 fragments, mb * 8 + kb:          A[i][k] = W2[16kb + k][16mb + i]             BF16
//   W3T  : 8 fragments, m-block mb:            A[i][k] = W3[k][16mb + i]       (k < out)    BF16
// The rollout's layers 2 and 3 use gfx950's v_mfma_f32_16x16x32_f16 (twice
// the K per instruction), whose lane l holds A[l % 16][8 (l / 16) + j],
// B[8 (l / 16) + j][l % 16], j = 0..7, and D as above.  Its fragments follow
// the 152 above (512 values each, FP16), with the K order of block kb
//   u(kb, p) = 32 kb + (j < 4 ? 4 g + j : 16 + 4 g + (j - 4)),  g = p / 8, j = p % 8,
// so that the two result tiles of m-blocks 2 kb and 2 kb + 1 that a lane
// holds (rows 4 g + j of each), concatenated, are its B operand:
//   W2X  : 32 fragments, mb * 4 + kb:          A[i][p] = W2[16mb + i][u(kb, p)]
//   W3X  : 4 fragments, kb:                    A[i][p] = W3[i][u(kb, p)]       (i < out)
//
// Activation: tanh, as T(z) = Q(y), y = clamp01(r16(z)), with r16 the rounding
// of the FP32 pre-activation to binary16, nearest even (clamp01 maps NaN to
// 0: the hardware's output clamp), and Q the odd polynomial
//   Q = s P(v),  s = 2 y - 1,  v = s^2 - 1/2,
//   P(v) = c0 + v (c1 + v (c2 + v (c3 + v c4))),
// degree 9 in s, its five coefficients binary16 values (MPPI_MLP_TANH_C0..4)
// fitted (Lawson's iteration toward minimax, then a search over their
// last bits with every binary16 rounding in place) to tanh(A s), A = 2.77.
// So T(z) = tanh(A (2 z - 1)) = tanh(5.54 (z - 1/2)) for z in [0, 1], and
// T saturates at Q(0) = -Q(1) beyond; over every binary16 z, T is within
// 4.31e-3 of tanh(5.54 (r16(z) - 1/2)) (inside and beyond the clamp).  The
// centring on z = 1/2 and the slope 5.54 are a change of variable that the
// first two layers' weights and biases absorb (z_tanh = 5.54 (z - 1/2)): the
// network is an MLP with tanh hidden layers, whose biases start at 1/2.
// Every operation is one binary16 operation rounded to nearest even, in
// this order:
//   s = fma(y, 2, -1),  v = fma(s, s, -1/2),
//   p = fma(v, c4, c3),  p = fma(v, p, c2),  p = fma(v, p, c1),  p = fma(v, p, c0),
//   T = s * p.
// The variable v is centred on its range [-1/2, 1/2], which keeps the
// Horner sums free of the cancellation that monomials of s^2 would have in
// binary16.  On packed binary16 pairs T costs eight vector instructions per
// two values (the conversion with its output clamp, six fused multiply-adds
// and a multiply) and no memory access, and its values are the FP16 B
// operands of the next layer as they stand.  dT/dz = 2 (P(v) + 2 s^2 P'(v))
// inside (0, 1), 0 outside.
#ifndef MPPI_HIP_MLP_H
#define MPPI_HIP_MLP_H

#if !defined(__HIPCC_RTC__)
#include <stdint.h>
#endif

#define MPPI_MLP_HID 128
// tanh polynomial (above): binary16 coefficients, exact as written.
#define MPPI_MLP_TANH_C0 1.3583984375f
#define MPPI_MLP_TANH_C1 (-1.0908203125f)
#define MPPI_MLP_TANH_C2 1.298828125f
#define MPPI_MLP_TANH_C3 (-2.537109375f)
#define MPPI_MLP_TANH_C4 2.80859375f
#define MPPI_MLP_FRAG_W1 0
#define MPPI_MLP_FRAG_W2 8
#define MPPI_MLP_FRAG_W3 72
#define MPPI_MLP_FRAG_W2T 80
#define MPPI_MLP_FRAG_W3T 144
#define MPPI_MLP_NUM_FRAGS 152
#define MPPI_MLP_FRAG_ELEMS 256
#define MPPI_MLP_X32_BASE (MPPI_MLP_NUM_FRAGS * MPPI_MLP_FRAG_ELEMS)   // element offset of the x32 fragments
#define MPPI_MLP_X32_ELEMS 512
#define MPPI_MLP_X32_W2 0
#define MPPI_MLP_X32_W3 32
#define MPPI_MLP_X32_FRAGS 36
#define MPPI_MLP_TOTAL_ELEMS (MPPI_MLP_X32_BASE + MPPI_MLP_X32_FRAGS * MPPI_MLP_X32_ELEMS)

// Offsets of the FP32 parameter vector [W1 | b1 | W2 | b2 | W3 | b3] (row
// major [out][in]) for in / out dimensions.
struct mppi_mlp_dims {
    uint32_t in, out;
    uint32_t w1, b1, w2, b2, w3, b3, n_params;
    uint32_t pad;
};

// Input normalisation x_n = wrap(x - center) * inv_half, where wrap reduces
// into [-period/2, period/2] when period > 0 (inputs the dynamics is periodic
// in), as d - period * rint(d * inv_period) with inv_period = 1 / period
// rounded to FP32 once, and inv_half = 0 drops an input (one the dynamics is
// invariant to); output (state delta) delta = y * out_sd + out_mu.
struct mppi_mlp_norm {
    float in_center[16];
    float in_inv_half[16];
    float in_period[16];
    float out_mu[16];
    float out_sd[16];
};

#if defined(__HIPCC_RTC__) || defined(__HIP_DEVICE_COMPILE__) || defined(__HIPCC__)
typedef short mppi_v4s __attribute__((ext_vector_type(4)));     // four 16-bit values (FP16 or BF16 bits)
typedef short mppi_v8s __attribute__((ext_vector_type(8)));
typedef float mppi_v4f __attribute__((ext_vector_type(4)));
typedef __bf16 mppi_v4bf16 __attribute__((ext_vector_type(4)));
typedef _Float16 mppi_v2h __attribute__((ext_vector_type(2)));
typedef _Float16 mppi_v4h __attribute__((ext_vector_type(4)));
typedef _Float16 mppi_v8h __attribute__((ext_vector_type(8)));

__device__ inline float mppi_mlp_input(float v, float center, float inv_half, float period, float inv_period) {
    float d = v - center;
    if (period > 0.0f) d -= period * __builtin_rintf(d * inv_period);
    return d * inv_half;
}

// float -> bfloat16, round to nearest even (v_cvt_pk_bf16_f32 on gfx950).
__device__ inline short mppi_f2bf(float f) {
    return __builtin_bit_cast(short, (__bf16)f);
}
__device__ inline float mppi_bf2f(short h) {
    return __builtin_bit_cast(float, uint32_t((unsigned short)h) << 16);
}
// float -> binary16, round to nearest even.
__device__ inline short mppi_f2h(float f) {
    return __builtin_bit_cast(short, (_Float16)f);
}
__device__ inline float mppi_h2f(short h) {
    return float(__builtin_bit_cast(_Float16, h));
}
// Four floats -> four binary16 (nearest even).
__device__ inline mppi_v4s mppi_pack_f16(mppi_v4f v) {
    return __builtin_bit_cast(mppi_v4s, __builtin_convertvector(v, mppi_v4h));
}

// y = clamp01(r16(z)) of N pre-activations as binary16 (N even): packed
// conversions to nearest even with their output clamp (v_cvt_pk_f16_f32 ...
// clamp; NaN clamps to 0).  The maximum and minimum below are that clamp:
// they fold into the conversion, and have its meaning where they do not.
template <typename VF, typename VH>
__device__ inline VH mppi_act_y(VF z) {
    const VH h = __builtin_convertvector(z, VH);
    return __builtin_elementwise_min(__builtin_elementwise_max(h, (VH)(_Float16)0), (VH)(_Float16)1);
}
// Q(y) of N binary16 values.  With N = 4 or 8 every step is two or four
// independent packed instructions, so a dependent one never follows right
// after the instruction that writes its operand (which on gfx950 costs a
// wait state per packed binary16 instruction of a single pair's chain).
template <typename VH>
__device__ inline VH mppi_act_q(VH y) {
    const VH s = __builtin_elementwise_fma(y, (VH)(_Float16)2, (VH)(_Float16)-1);
    const VH v = __builtin_elementwise_fma(s, s, (VH)(_Float16)-0.5);
    VH p = __builtin_elementwise_fma(v, (VH)(_Float16)MPPI_MLP_TANH_C4, (VH)(_Float16)MPPI_MLP_TANH_C3);
    p = __builtin_elementwise_fma(v, p, (VH)(_Float16)MPPI_MLP_TANH_C2);
    p = __builtin_elementwise_fma(v, p, (VH)(_Float16)MPPI_MLP_TANH_C1);
    p = __builtin_elementwise_fma(v, p, (VH)(_Float16)MPPI_MLP_TANH_C0);
    return s * p;
}
// dT/dz at a binary16 y (training's backward pass, FP32): 2 (P(v) + 2 s^2 P'(v))
// for y in (0, 1), 0 at the clamp.
__device__ inline float mppi_act_deriv(float y) {
    const float s = 2.0f * y - 1.0f, v = s * s - 0.5f;
    const float p = MPPI_MLP_TANH_C0 + v * (MPPI_MLP_TANH_C1 + v * (MPPI_MLP_TANH_C2 + v * (MPPI_MLP_TANH_C3 + v * MPPI_MLP_TANH_C4)));
    const float dp = MPPI_MLP_TANH_C1 + v * (2.0f * MPPI_MLP_TANH_C2 + v * (3.0f * MPPI_MLP_TANH_C3 + v * (4.0f * MPPI_MLP_TANH_C4)));
    return (y > 0.0f && y < 1.0f) ? 2.0f * (p + 2.0f * s * s * dp) : 0.0f;
}
typedef float mppi_v2f __attribute__((ext_vector_type(2)));
typedef float mppi_v8f __attribute__((ext_vector_type(8)));
// T of two pre-activations, as a pair of binary16 (z0 in the low half).
__device__ inline uint32_t mppi_act2(float z0, float z1) {
    return __builtin_bit_cast(uint32_t, mppi_act_q(mppi_act_y<mppi_v2f, mppi_v2h>((mppi_v2f){z0, z1})));
}

__device__ inline mppi_v4s mppi_act4(mppi_v4f z) {
    return __builtin_bit_cast(mppi_v4s, mppi_act_q(mppi_act_y<mppi_v4f, mppi_v4h>(z)));
}

// T of two tiles of four pre-activations at once (four independent packed
// instructions per step).
__device__ inline void mppi_act8(mppi_v4f za, mppi_v4f zb, mppi_v4s* ta, mppi_v4s* tb) {
    const mppi_v8f z = __builtin_shufflevector(za, zb, 0, 1, 2, 3, 4, 5, 6, 7);
    const mppi_v8h t = mppi_act_q(mppi_act_y<mppi_v8f, mppi_v8h>(z));
    *ta = __builtin_bit_cast(mppi_v4s, __builtin_shufflevector(t, t, 0, 1, 2, 3));
    *tb = __builtin_bit_cast(mppi_v4s, __builtin_shufflevector(t, t, 4, 5, 6, 7));
}

// T of four pre-activations and their y (for the derivative, mppi_act_deriv).
__device__ inline mppi_v4s mppi_act4_y(mppi_v4f z, mppi_v4s* y_out) {
    const mppi_v4h y = mppi_act_y<mppi_v4f, mppi_v4h>(z);
    *y_out = __builtin_bit_cast(mppi_v4s, y);
    return __builtin_bit_cast(mppi_v4s, mppi_act_q(y));
}

// v_mfma_f32_16x16x16_f16 (forward pass).
__device__ inline mppi_v4f mppi_mfma(mppi_v4s a, mppi_v4s b, mppi_v4f c) {
    return __builtin_amdgcn_mfma_f32_16x16x16f16(__builtin_bit_cast(mppi_v4h, a), __builtin_bit_cast(mppi_v4h, b),
                                                 c, 0, 0, 0);
}

// v_mfma_f32_16x16x16_bf16 (backward pass and weight gradients).
__device__ inline mppi_v4f mppi_mfma_bf16(mppi_v4s a, mppi_v4s b, mppi_v4f c) {
    return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
}

// v_mfma_f32_16x16x32_f16 (gfx950; the rollout's layers 2 and 3).
__device__ inline mppi_v4f mppi_mfma32(mppi_v8s a, mppi_v8s b, mppi_v4f c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_f16(__builtin_bit_cast(mppi_v8h, a), __builtin_bit_cast(mppi_v8h, b),
                                                  c, 0, 0, 0);
}

// B operand of x32 K-block kb from the two x16 result tiles 2 kb and 2 kb + 1.
__device__ inline mppi_v8s mppi_cat8(mppi_v4s lo, mppi_v4s hi) {
    return __builtin_shufflevector(lo, hi, 0, 1, 2, 3, 4, 5, 6, 7);
}

__device__ inline mppi_v4s mppi_frag(const short* frags, uint32_t f, uint32_t lane) {
    return *reinterpret_cast<const mppi_v4s*>(frags + size_t(f) * MPPI_MLP_FRAG_ELEMS + lane * 4);
}

// Forward pass of one 16-sample block (x16 fragments; training and
// inference over sample lists).  xb: the B operand of the inputs, FP16
// (rows 4*(lane/16)+j, zero beyond `in`).  Returns the output tile (rows
// 4*(lane/16)+j < out valid, without the output bias); h1_out / h2_out
// receive the FP16 hidden activations and y1_out / y2_out their y when
// non-null (for training).  b1 and b2 must be 16-byte aligned: the biases
// initialise the accumulators (four rows per lane), so the bias add costs
// no instruction.
__device__ inline mppi_v4f mppi_mlp_block(const short* frags, const float* b1, const float* b2, mppi_v4s xb,
                                          uint32_t lane, mppi_v4s* h1_out, mppi_v4s* h2_out,
                                          mppi_v4s* y1_out, mppi_v4s* y2_out) {
    const uint32_t g = lane >> 4;
    mppi_v4s h1b[8], h2b[8], y1b[8], y2b[8];
    const mppi_v4f zero = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
    for (int mb = 0; mb < 8; ++mb) {
        const mppi_v4f bias = *reinterpret_cast<const mppi_v4f*>(b1 + 16 * mb + 4 * g);
        h1b[mb] = mppi_act4_y(mppi_mfma(mppi_frag(frags, MPPI_MLP_FRAG_W1 + mb, lane), xb, bias), &y1b[mb]);
    }
#pragma unroll
    for (int mb = 0; mb < 8; ++mb) {
        mppi_v4f acc = *reinterpret_cast<const mppi_v4f*>(b2 + 16 * mb + 4 * g);
#pragma unroll
        for (int kb = 0; kb < 8; ++kb)
            acc = mppi_mfma(mppi_frag(frags, MPPI_MLP_FRAG_W2 + mb * 8 + kb, lane), h1b[kb], acc);
        h2b[mb] = mppi_act4_y(acc, &y2b[mb]);
    }
    mppi_v4f out = zero;
#pragma unroll
    for (int kb = 0; kb < 8; ++kb) out = mppi_mfma(mppi_frag(frags, MPPI_MLP_FRAG_W3 + kb, lane), h2b[kb], out);
    if (h1_out) for (int i = 0; i < 8; ++i) h1_out[i] = h1b[i];
    if (h2_out) for (int i = 0; i < 8; ++i) h2_out[i] = h2b[i];
    if (y1_out) for (int i = 0; i < 8; ++i) y1_out[i] = y1b[i];
    if (y2_out) for (int i = 0; i < 8; ++i) y2_out[i] = y2b[i];
    return out;
}
#endif

#endif  // MPPI_HIP_MLP_H
