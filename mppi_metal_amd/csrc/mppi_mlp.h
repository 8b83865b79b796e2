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
// The rollout runs all three layers on 32x32 tiles (32 hidden units x 32
// samples): layer 1 as v_mfma_f32_32x32x8_f16 (one instruction per eight
// inputs), layers 2 and 3 as gfx950's v_mfma_f32_32x32x16_f16, the shape that
// issues at the matrix cores' full FP16 rate and leaves the vector unit free
// to run the activations of the previous tile meanwhile.  In these
// instructions lane l, with h = l / 32, holds
//   A (32 x K):  A[l % 32][(K / 2) h + j]          j < K / 2
//   B (K x 32):  B[(K / 2) h + j][l % 32]
//   D (32 x 32): D[8 (i / 4) + 4 h + i % 4][l % 32]   i = 0..15
// A result tile of 32 hidden units holds 16 rows per lane; its first and
// last eight registers, as binary16, are the B operands of two K-blocks of 16
// of the next layer when K-block kb of the next layer's weights takes its
// columns in the order
//   u(kb, p) = 16 kb + 8 ((p % 8) / 4) + 4 (p / 8) + p % 4,   p = 0..15,
// so layers chain without moving data between lanes.  These fragments follow
// the 152 above (FP16, element offsets MPPI_MLP_Y_*):
//   W1Y  : 8 fragments of 256, kb1 * 4 + mb:  A[i][k] = W1[32 mb + i][8 kb1 + k]   (8 kb1 + k < in)
//   W2Y  : 32 fragments of 512, mb * 8 + kb:  A[i][p] = W2[32 mb + i][u(kb, p)]
//   W3Y  : 8 fragments of 512, kb:            A[i][p] = W3[i][u(kb, p)]           (i < out)
// With out <= 8 the last layer runs as v_mfma_f32_16x16x32_f16 instead
// (fewer matrix cycles, a quarter of the accumulator registers): its B
// operand, lane l holding column l % 16 and K rows 8 (l / 16) + j, is the
// same layer-2 registers read as two 16-sample groups (l / 16 odd: samples
// 16-31 of the column block) each with half of the K-block, and the A
// fragment keeps the groups apart, rows 0-7 for samples 0-15 and rows 8-15
// for samples 16-31:
//   W3Z  : 8 fragments of 512 (in place of W3Y), kb, lane l, g = l / 16:
//          A[i][8 g + j] = W3[i % 8][u(kb, 8 (g / 2) + j)] if i / 8 == g % 2 (and i % 8 < out), else 0
// and result register i of lane l is output row 4 (g % 2) + i of sample
// l % 16 + 16 (g / 2) of the column block.
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
#define MPPI_MLP_Y_BASE (MPPI_MLP_NUM_FRAGS * MPPI_MLP_FRAG_ELEMS)   // element offset of the 32x32 fragments
#define MPPI_MLP_Y_W1 0          // offsets within the 32x32 set
#define MPPI_MLP_Y_W2 2048
#define MPPI_MLP_Y_W3 18432
#define MPPI_MLP_Y_ELEMS 22528
#define MPPI_MLP_TOTAL_ELEMS (MPPI_MLP_Y_BASE + MPPI_MLP_Y_ELEMS)

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
// Q(y) of N binary16 values, each step N / 2 independent packed
// instructions (the rollout runs the same steps one by one, mppi_act_step,
// so that two tiles' steps alternate: a packed binary16 instruction that
// reads the result of the one right before it costs a wait state on gfx950).
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
typedef unsigned int mppi_v2u __attribute__((ext_vector_type(2)));
typedef float mppi_v8f __attribute__((ext_vector_type(8)));
typedef float mppi_v16f __attribute__((ext_vector_type(16)));
typedef _Float16 mppi_v16h __attribute__((ext_vector_type(16)));
// T of two pre-activations, as a pair of binary16 (z0 in the low half).
__device__ inline uint32_t mppi_act2(float z0, float z1) {
    return __builtin_bit_cast(uint32_t, mppi_act_q(mppi_act_y<mppi_v2f, mppi_v2h>((mppi_v2f){z0, z1})));
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


// v_mfma_f32_32x32x8_f16 and v_mfma_f32_32x32x16_f16 (the rollout's layers).
__device__ inline mppi_v16f mppi_mfma_32x8(mppi_v4s a, mppi_v4s b, mppi_v16f c) {
    return __builtin_amdgcn_mfma_f32_32x32x8f16(__builtin_bit_cast(mppi_v4h, a), __builtin_bit_cast(mppi_v4h, b),
                                                c, 0, 0, 0);
}
__device__ inline mppi_v4f mppi_mfma_16x32(mppi_v8s a, mppi_v8s b, mppi_v4f c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_f16(__builtin_bit_cast(mppi_v8h, a), __builtin_bit_cast(mppi_v8h, b),
                                                  c, 0, 0, 0);
}
__device__ inline mppi_v16f mppi_mfma_32x16(mppi_v8s a, mppi_v8s b, mppi_v16f c) {
    return __builtin_amdgcn_mfma_f32_32x32x16_f16(__builtin_bit_cast(mppi_v8h, a), __builtin_bit_cast(mppi_v8h, b),
                                                  c, 0, 0, 0);
}

// A 32x32 accumulator tile of a bias vector: lane l, register i holds
// b[8 (i / 4) + 4 (l / 32) + i % 4] (b 16-byte aligned, 32 values).
__device__ inline mppi_v16f mppi_bias_tile32(const float* b, uint32_t lane) {
    const float* p = b + 4 * (lane >> 5);
    const mppi_v4f q0 = *reinterpret_cast<const mppi_v4f*>(p), q1 = *reinterpret_cast<const mppi_v4f*>(p + 8);
    const mppi_v4f q2 = *reinterpret_cast<const mppi_v4f*>(p + 16), q3 = *reinterpret_cast<const mppi_v4f*>(p + 24);
    const mppi_v8f lo = __builtin_shufflevector(q0, q1, 0, 1, 2, 3, 4, 5, 6, 7);
    const mppi_v8f hi = __builtin_shufflevector(q2, q3, 0, 1, 2, 3, 4, 5, 6, 7);
    return __builtin_shufflevector(lo, hi, 0, 1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11, 12, 13, 14, 15);
}

// One 32x32 tile's activation in eight steps of eight independent packed
// instructions (the operations of mppi_act_q, in its order), so that the
// steps of two tiles can be spread between matrix instructions.
struct mppi_act_tile {
    mppi_v16f z;
    mppi_v16h s, v, p;
};
// Step `step` of the eight values of half `half` (0: registers 0-7, 1: 8-15).
__device__ inline void mppi_act_half_step(mppi_act_tile& a, int step, int half) {
    mppi_v8h s = half ? a.s.hi : a.s.lo, v = half ? a.v.hi : a.v.lo, p = half ? a.p.hi : a.p.lo;
    switch (step) {
        case 0: {
            const mppi_v8f z = half ? __builtin_shufflevector(a.z, a.z, 8, 9, 10, 11, 12, 13, 14, 15)
                                    : __builtin_shufflevector(a.z, a.z, 0, 1, 2, 3, 4, 5, 6, 7);
            s = mppi_act_y<mppi_v8f, mppi_v8h>(z);
            break;
        }
        case 1: s = __builtin_elementwise_fma(s, (mppi_v8h)(_Float16)2, (mppi_v8h)(_Float16)-1); break;
        case 2: v = __builtin_elementwise_fma(s, s, (mppi_v8h)(_Float16)-0.5); break;
        case 3: p = __builtin_elementwise_fma(v, (mppi_v8h)(_Float16)MPPI_MLP_TANH_C4, (mppi_v8h)(_Float16)MPPI_MLP_TANH_C3); break;
        case 4: p = __builtin_elementwise_fma(v, p, (mppi_v8h)(_Float16)MPPI_MLP_TANH_C2); break;
        case 5: p = __builtin_elementwise_fma(v, p, (mppi_v8h)(_Float16)MPPI_MLP_TANH_C1); break;
        case 6: p = __builtin_elementwise_fma(v, p, (mppi_v8h)(_Float16)MPPI_MLP_TANH_C0); break;
        default: p = s * p; break;
    }
    if (half) {
        a.s.hi = s; a.v.hi = v; a.p.hi = p;
    } else {
        a.s.lo = s; a.v.lo = v; a.p.lo = p;
    }
}
__device__ inline void mppi_act_step(mppi_act_tile& a, int step) {
    switch (step) {
        case 0: {
            const mppi_v8h ya = mppi_act_y<mppi_v8f, mppi_v8h>(__builtin_shufflevector(a.z, a.z, 0, 1, 2, 3, 4, 5, 6, 7));
            const mppi_v8h yb =
                mppi_act_y<mppi_v8f, mppi_v8h>(__builtin_shufflevector(a.z, a.z, 8, 9, 10, 11, 12, 13, 14, 15));
            a.s = __builtin_shufflevector(ya, yb, 0, 1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11, 12, 13, 14, 15);
            break;
        }
        case 1: a.s = __builtin_elementwise_fma(a.s, (mppi_v16h)(_Float16)2, (mppi_v16h)(_Float16)-1); break;
        case 2: a.v = __builtin_elementwise_fma(a.s, a.s, (mppi_v16h)(_Float16)-0.5); break;
        case 3: a.p = __builtin_elementwise_fma(a.v, (mppi_v16h)(_Float16)MPPI_MLP_TANH_C4, (mppi_v16h)(_Float16)MPPI_MLP_TANH_C3); break;
        case 4: a.p = __builtin_elementwise_fma(a.v, a.p, (mppi_v16h)(_Float16)MPPI_MLP_TANH_C2); break;
        case 5: a.p = __builtin_elementwise_fma(a.v, a.p, (mppi_v16h)(_Float16)MPPI_MLP_TANH_C1); break;
        case 6: a.p = __builtin_elementwise_fma(a.v, a.p, (mppi_v16h)(_Float16)MPPI_MLP_TANH_C0); break;
        default: a.p = a.s * a.p; break;
    }
}
__device__ inline mppi_v8s mppi_act_lo(const mppi_act_tile& a) {
    return __builtin_bit_cast(mppi_v8s, __builtin_shufflevector(a.p, a.p, 0, 1, 2, 3, 4, 5, 6, 7));
}
__device__ inline mppi_v8s mppi_act_hi(const mppi_act_tile& a) {
    return __builtin_bit_cast(mppi_v8s, __builtin_shufflevector(a.p, a.p, 8, 9, 10, 11, 12, 13, 14, 15));
}

// Scheduling fence between the slots of mppi_mlp_net: only scalar
// instructions move across it (the LDS reads are placed ahead of their use
// by hand).
#define MPPI_NET_SLOT() __builtin_amdgcn_sched_barrier(0x0004)

// Compile-time loop: f(mppi_ic<I>{}) for I = B .. E - 1 (the slots below
// index register arrays, which must resolve at compile time).
template <int I>
struct mppi_ic {
    static constexpr int value = I;
};
template <int B, int E, typename F>
__device__ inline void mppi_static_for(F&& f) {
    if constexpr (B < E) {
        f(mppi_ic<B>{});
        mppi_static_for<B + 1, E>(f);
    }
}

// Element offset in the 32x32 set of the i-th layer-2/3 fragment that
// mppi_mlp_net consumes (40 in order; each feeds two products, one per
// column block): layer 2 m-blocks 0 and 1, then m-block 2 and layer 3
// K-blocks 0-1, m-block 3 and K-blocks 2-3, K-blocks 4-7.
__device__ constexpr int mppi_net_frag(int i) {
    return i < 16   ? MPPI_MLP_Y_W2 + i * 512
           : i < 24 ? MPPI_MLP_Y_W2 + i * 512
           : i < 26 ? MPPI_MLP_Y_W3 + (i - 24) * 512
           : i < 34 ? MPPI_MLP_Y_W2 + (i - 2) * 512
           : i < 36 ? MPPI_MLP_Y_W3 + (i - 32) * 512
                    : MPPI_MLP_Y_W3 + (i - 32) * 512;
}

// The rollout's network on one wavefront's 64 samples (two column blocks of
// 32), 32x32 tiles, y = W3 T(W2 T(W1 x + b1) + b2) + b3 as output tiles.
// wy: the 32x32 fragment set in LDS; b: [b1 | b2 | b3 padded to 32] in LDS
// (16-byte aligned); xb[nb][kb1]: the layer-1 B operands.  The matrix
// instructions are spread between the steps of the activations that do not
// depend on them, one or two per slot of about eight packed instructions,
// in phases:
//   0      layer 1 of m-block 0;
//   1-4    layer 1 of m-block p ahead, and layer 2 m-block 0's K-blocks as
//          their inputs come out, with the activations of layer-1 m-block
//          p - 1;
//   5      layer 2 m-block 0's last two K-blocks;
//   6-9    layer 2 m-block q with the activations of m-block q - 1 and
//          layer 3's products of m-block q - 2;
//   10     layer 3 of the last m-block.
// Weight fragments are read from LDS two fragments ahead of their use into
// a ring of four, and bias tiles one phase ahead.
template <bool B, typename T, typename F>
struct mppi_pick {
    typedef T type;
};
template <typename T, typename F>
struct mppi_pick<false, T, F> {
    typedef F type;
};

// NARROW (out <= 8): the last layer as 16x16x32 products on W3Z, y[nb] a
// 16x16 tile; otherwise 32x32x16 on W3Y, y[nb] a 32x32 tile.
template <int KB1, bool NARROW>
__device__ inline void mppi_mlp_net(const short* wy, const float* b, const mppi_v4s (&xb)[2][KB1], uint32_t lane,
                                    typename mppi_pick<NARROW, mppi_v4f, mppi_v16f>::type (&y)[2]) {
    mppi_v8s h1[2][8];
    mppi_v8s h2[2][2];
    mppi_act_tile t1[2], t1n[2], t2[2];
    mppi_v16f a2[2], bias_next;
    mppi_v8s ring[4];
    mppi_v4s w1f[KB1][4];
    auto load = [&](int i) { return *reinterpret_cast<const mppi_v8s*>(wy + mppi_net_frag(i) + lane * 8); };
    // product with the i-th fragment for column block nb; the first of the
    // pair reads fragment i + 2 ahead
    auto use = [&](auto I, int nb, mppi_v8s bop, auto& acc) {
        constexpr int i = decltype(I)::value;
        if (nb == 0 && i + 2 < 40) ring[(i + 2) % 4] = load(i + 2);
        if constexpr (sizeof(acc) == sizeof(mppi_v4f)) acc = mppi_mfma_16x32(ring[i % 4], bop, acc);
        else acc = mppi_mfma_32x16(ring[i % 4], bop, acc);
    };
    mppi_static_for<0, KB1>([&](auto K) {
        mppi_static_for<0, 4>([&](auto MB) {
            w1f[K.value][MB.value] =
                *reinterpret_cast<const mppi_v4s*>(wy + MPPI_MLP_Y_W1 + (K.value * 4 + MB.value) * 256 + lane * 4);
        });
    });
    // layer 1 of m-block mb, column block nb, from bias tile bt
    auto l1 = [&](int mb, int nb, const mppi_v16f& bt) {
        mppi_v16f acc = bt;
        mppi_static_for<0, KB1>([&](auto K) { acc = mppi_mfma_32x8(w1f[K.value][mb], xb[nb][K.value], acc); });
        return acc;
    };
    // phase 0
    {
        const mppi_v16f bt = mppi_bias_tile32(b, lane);
        bias_next = mppi_bias_tile32(b + 32, lane);
        t1[0].z = l1(0, 0, bt);
        t1[1].z = l1(0, 1, bt);
        ring[0] = load(0);
        ring[1] = load(1);
    }
    MPPI_NET_SLOT();
    // phases 1-4: sixteen slots, one activation step each; matrix
    // instruction m of n_m in slot (16 m) / n_m
    mppi_static_for<1, 5>([&](auto P) {
        constexpr int p = P.value;
        constexpr int n_l1 = p < 4 ? 2 : 0, n_l2 = p >= 2 ? 4 : 0, n_m = n_l1 + n_l2;
        const mppi_v16f bt = bias_next;
        if constexpr (p < 3) bias_next = mppi_bias_tile32(b + 32 * (p + 1), lane);
        if constexpr (p == 1) a2[0] = a2[1] = mppi_bias_tile32(b + MPPI_MLP_HID, lane);   // layer 2 m-block 0
        if constexpr (p < 4) {
            mppi_static_for<0, 16>([&](auto J) {
                constexpr int j = J.value;
                mppi_static_for<0, n_m>([&](auto M) {
                    constexpr int m = M.value;
                    if constexpr ((16 * m) / n_m == j) {
                        if constexpr (m < n_l1) {
                            t1n[m].z = l1(p, m, bt);
                        } else {
                            constexpr int q = m - n_l1, kb = 2 * (p - 2) + q / 2, nb = q % 2;
                            use(mppi_ic<kb>{}, nb, h1[nb][kb], a2[nb]);
                        }
                    }
                });
                mppi_act_step(t1[j % 2], j / 2);
                MPPI_NET_SLOT();
            });
            mppi_static_for<0, 2>([&](auto NB) {
                constexpr int nb = NB.value;
                h1[nb][2 * (p - 1)] = mppi_act_lo(t1[nb]);
                h1[nb][2 * (p - 1) + 1] = mppi_act_hi(t1[nb]);
                t1[nb] = t1n[nb];
            });
        } else {
            // the last layer-1 tiles' first halves first, so that K-block 6 of
            // layer 2 runs while their second halves are computed
            mppi_static_for<0, 16>([&](auto J) {
                constexpr int j = J.value;
                if constexpr (j % 2 == 0 && j < 8) use(mppi_ic<4 + j / 4>{}, (j / 2) % 2, h1[(j / 2) % 2][4 + j / 4], a2[(j / 2) % 2]);
                if constexpr (j == 9 || j == 12) use(mppi_ic<6>{}, j == 12, h1[j == 12][6], a2[j == 12]);
                mppi_act_half_step(t1[0], j % 8, j / 8);
                mppi_act_half_step(t1[1], j % 8, j / 8);
                if constexpr (j == 7) {
                    h1[0][6] = mppi_act_lo(t1[0]);
                    h1[1][6] = mppi_act_lo(t1[1]);
                }
                MPPI_NET_SLOT();
            });
            h1[0][7] = mppi_act_hi(t1[0]);
            h1[1][7] = mppi_act_hi(t1[1]);
        }
    });
    // phase 5
    use(mppi_ic<7>{}, 0, h1[0][7], a2[0]);
    use(mppi_ic<7>{}, 1, h1[1][7], a2[1]);
    t2[0].z = a2[0];
    t2[1].z = a2[1];
    if constexpr (NARROW)
        y[0] = y[1] = *reinterpret_cast<const mppi_v4f*>(b + 2 * MPPI_MLP_HID + 4 * ((lane >> 4) & 1));
    else
        y[0] = y[1] = mppi_bias_tile32(b + 2 * MPPI_MLP_HID, lane);
    bias_next = mppi_bias_tile32(b + MPPI_MLP_HID + 32, lane);
    MPPI_NET_SLOT();
    // phases 6-9
    mppi_static_for<1, 5>([&](auto Q) {
        constexpr int q = Q.value;
        constexpr int n_l2 = q < 4 ? 16 : 0, n_l3 = q >= 2 ? 4 : 0, n_m = n_l2 + n_l3;
        constexpr int f0 = q == 1 ? 8 : q == 2 ? 16 : q == 3 ? 26 : 36;     // first fragment of the phase
        if constexpr (q < 4) a2[0] = a2[1] = bias_next;
        if constexpr (q < 3) bias_next = mppi_bias_tile32(b + MPPI_MLP_HID + 32 * (q + 1), lane);
        if constexpr (q == 4) {
            // layer 3 K-blocks 4-5 first; the last layer-2 tiles' first halves
            // first, so that K-block 6 runs while their second halves are computed
            mppi_static_for<0, 16>([&](auto J) {
                constexpr int j = J.value;
                if constexpr (j % 2 == 0 && j < 8) use(mppi_ic<36 + j / 4>{}, (j / 2) % 2, h2[(j / 2) % 2][j / 4], y[(j / 2) % 2]);
                if constexpr (j == 9 || j == 12) use(mppi_ic<38>{}, j == 12, mppi_act_lo(t2[j == 12]), y[j == 12]);
                mppi_act_half_step(t2[0], j % 8, j / 8);
                mppi_act_half_step(t2[1], j % 8, j / 8);
                MPPI_NET_SLOT();
            });
            h2[0][1] = mppi_act_hi(t2[0]);
            h2[1][1] = mppi_act_hi(t2[1]);
            return;
        }
        mppi_static_for<0, 16>([&](auto J) {
            constexpr int j = J.value;
            mppi_static_for<0, n_m>([&](auto M) {
                constexpr int m = M.value;
                if constexpr ((16 * m) / n_m == j) {
                    if constexpr (m < n_l2) {
                        use(mppi_ic<f0 + m / 2>{}, m % 2, h1[m % 2][m / 2], a2[m % 2]);
                    } else {
                        constexpr int r = m - n_l2, e = r / 2, nb = r % 2;
                        use(mppi_ic<f0 + n_l2 / 2 + e>{}, nb, h2[nb][e], y[nb]);
                    }
                }
            });
            mppi_act_step(t2[j % 2], j / 2);
            MPPI_NET_SLOT();
        });
        mppi_static_for<0, 2>([&](auto NB) {
            constexpr int nb = NB.value;
            h2[nb][0] = mppi_act_lo(t2[nb]);
            h2[nb][1] = mppi_act_hi(t2[nb]);
            if constexpr (q < 4) t2[nb].z = a2[nb];
        });
    });
    // phase 10
    use(mppi_ic<39>{}, 0, h2[0][1], y[0]);
    use(mppi_ic<39>{}, 1, h2[1][1], y[1]);
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
