// mppi_mlp.h — the learned-dynamics surrogate's network on the device: an MLP
// in -> 128 -> 128 -> out with tanh hidden layers (in = state_dim +
// control_dim <= 16, out = state_dim <= 16), BF16 operands on the matrix
// cores with FP32 accumulation.
//
// Layout: every layer is computed transposed, Z^T = W X^T, so the samples are
// the N dimension of v_mfma_f32_16x16x16_bf16 and the hidden units are M.  In
// that instruction lane l holds
//   A (16x16, M x K):  A[l % 16][4 * (l / 16) + j]          j = 0..3
//   B (16x16, K x N):  B[4 * (l / 16) + j][l % 16]
//   D (16x16, M x N):  D[4 * (l / 16) + j][l % 16]
// so the result tile of hidden block m is, register for register, the B
// operand of K block m of the next layer: layers chain without moving data
// between lanes.  Weights are kept as ready-made A fragments (64 lanes x 4
// bf16 each, "fragment-major"), built from the FP32 master copy by the pack
// kernel after every optimiser step.
//
// Fragment table (indices into the fragment buffer, 256 bf16 per fragment):
//   W1   : 8 fragments, m-block mb:            A[i][k] = W1[16mb + i][k]       (k < in)
//   W2   : 64 fragments, mb * 8 + kb:          A[i][k] = W2[16mb + i][16kb + k]
//   W3   : 8 fragments, k-block kb:            A[i][k] = W3[i][16kb + k]       (i < out)
//   W2T  : 64 fragments, mb * 8 + kb:          A[i][k] = W2[16kb + k][16mb + i]
//   W3T  : 8 fragments, m-block mb:            A[i][k] = W3[k][16mb + i]       (k < out)
// The rollout's layers 2 and 3 use gfx950's v_mfma_f32_16x16x32_bf16 (twice
// the K per instruction), whose lane l holds A[l % 16][8 (l / 16) + j],
// B[8 (l / 16) + j][l % 16], j = 0..7, and D as above.  Its fragments follow
// the 152 above (512 bf16 each), with the K order of block kb
//   u(kb, p) = 32 kb + (j < 4 ? 4 g + j : 16 + 4 g + (j - 4)),  g = p / 8, j = p % 8,
// so that the two result tiles of m-blocks 2 kb and 2 kb + 1 that a lane
// holds (rows 4 g + j of each), concatenated, are its B operand:
//   W2X  : 32 fragments, mb * 4 + kb:          A[i][p] = W2[16mb + i][u(kb, p)]
//   W3X  : 4 fragments, kb:                    A[i][p] = W3[i][u(kb, p)]       (i < out)
//
// Activation: the hidden layers apply T(z) = bf16(tanh(r16(z))), r16 the
// rounding of the FP32 pre-activation to binary16 toward zero (the
// hardware's packed conversion), bf16 the nearest-even rounding of the exact
// tanh.  T is read from a table indexed by the binary16 magnitude bits
// (32,768 entries, built on the host in FP64): magnitudes above 4 (and
// infinities, NaN) hold the entry of 4, +-1, as bf16(tanh(4)) already is,
// so that no cap is needed before the read.  The activations are bf16
// operands of the next layer anyway, so T is tanh at the precision the
// network keeps, for six vector instructions per pair of values.
#ifndef MPPI_HIP_MLP_H
#define MPPI_HIP_MLP_H

#if !defined(__HIPCC_RTC__)
#include <stdint.h>
#endif

#define MPPI_MLP_HID 128
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
#define MPPI_MLP_TANH_KEYS 0x8000                                      // every binary16 magnitude

// Offsets of the FP32 parameter vector [W1 | b1 | W2 | b2 | W3 | b3] (row
// major [out][in]) for in / out dimensions.
struct mppi_mlp_dims {
    uint32_t in, out;
    uint32_t w1, b1, w2, b2, w3, b3, n_params;
    uint32_t pad;
};

// Input normalisation x_n = wrap(x - center) * inv_half, where wrap reduces
// into [-period/2, period/2] when period > 0 (inputs the dynamics is periodic
// in) and inv_half = 0 drops an input (one the dynamics is invariant to);
// output (state delta) delta = y * out_sd + out_mu.
struct mppi_mlp_norm {
    float in_center[16];
    float in_inv_half[16];
    float in_period[16];
    float out_mu[16];
    float out_sd[16];
};

#if defined(__HIPCC_RTC__) || defined(__HIP_DEVICE_COMPILE__) || defined(__HIPCC__)
typedef short mppi_v4bf __attribute__((ext_vector_type(4)));
typedef float mppi_v4f __attribute__((ext_vector_type(4)));
typedef __bf16 mppi_v4bf16 __attribute__((ext_vector_type(4)));

__device__ inline float mppi_mlp_input(float v, float center, float inv_half, float period) {
    float d = v - center;
    if (period > 0.0f) d -= period * __builtin_rintf(d / period);
    return d * inv_half;
}

// float -> bfloat16, round to nearest even (v_cvt_pk_bf16_f32 on gfx950).
__device__ inline short mppi_f2bf(float f) {
    return __builtin_bit_cast(short, (__bf16)f);
}
// Four floats -> four bfloat16 (two packed conversions).
__device__ inline mppi_v4bf mppi_pack_bf16(mppi_v4f v) {
    return __builtin_bit_cast(mppi_v4bf, __builtin_convertvector(v, mppi_v4bf16));
}
__device__ inline float mppi_bf2f(short h) {
    return __builtin_bit_cast(float, uint32_t((unsigned short)h) << 16);
}

typedef short mppi_v8bf __attribute__((ext_vector_type(8)));
typedef __bf16 mppi_v8bf16 __attribute__((ext_vector_type(8)));
typedef unsigned short mppi_u16x2 __attribute__((ext_vector_type(2)));

// T of two pre-activations, as a pair of bf16 (z0 in the low half): the
// packed conversion to binary16 toward zero, one table read per magnitude
// (the table repeats the entry of 4.0, 0x4400, for every larger magnitude,
// infinities and NaN included, so no cap is applied here), the sign put
// back.  `tab` may be in LDS or global memory.
__device__ inline uint32_t mppi_act2(float z0, float z1, const unsigned short* tab) {
    const uint32_t h = __builtin_bit_cast(uint32_t, __builtin_amdgcn_cvt_pkrtz(z0, z1));
    const uint32_t key = h & 0x7fff7fffu;
    // Byte offsets of the two entries: the low key's by one shift of its
    // 16-bit half (an SDWA operand select, which the compiler does not form
    // here by itself), the high key's by one shift (bit 15 is clear).
    uint32_t off_lo;
    asm("v_lshlrev_b32_sdwa %0, 1, %1 dst_sel:DWORD dst_unused:UNUSED_PAD src0_sel:DWORD src1_sel:WORD_0"
        : "=v"(off_lo) : "v"(key));
    const char* base = reinterpret_cast<const char*>(tab);
    const uint32_t lo = *reinterpret_cast<const unsigned short*>(base + off_lo);
    const uint32_t hi = *reinterpret_cast<const unsigned short*>(base + (key >> 15));
    return __builtin_amdgcn_perm(hi, lo, 0x05040100u) | (h & 0x80008000u);
}

__device__ inline mppi_v4bf mppi_act4(mppi_v4f z, const unsigned short* tab) {
    const uint32_t a = mppi_act2(z[0], z[1], tab), b = mppi_act2(z[2], z[3], tab);
    return __builtin_bit_cast(mppi_v4bf, (uint64_t(b) << 32) | a);
}

__device__ inline mppi_v4f mppi_mfma(mppi_v4bf a, mppi_v4bf b, mppi_v4f c) {
    return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
}

// v_mfma_f32_16x16x32_bf16 (gfx950).
__device__ inline mppi_v4f mppi_mfma32(mppi_v8bf a, mppi_v8bf b, mppi_v4f c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(mppi_v8bf16, a),
                                                    __builtin_bit_cast(mppi_v8bf16, b), c, 0, 0, 0);
}

// B operand of x32 K-block kb from the two x16 result tiles 2 kb and 2 kb + 1.
__device__ inline mppi_v8bf mppi_cat8(mppi_v4bf lo, mppi_v4bf hi) {
    return __builtin_shufflevector(lo, hi, 0, 1, 2, 3, 4, 5, 6, 7);
}

__device__ inline mppi_v4bf mppi_frag(const short* frags, uint32_t f, uint32_t lane) {
    return *reinterpret_cast<const mppi_v4bf*>(frags + size_t(f) * MPPI_MLP_FRAG_ELEMS + lane * 4);
}

// Forward pass of one 16-sample block (x16 fragments; training and
// inference over sample lists).  xb: the B operand of the inputs
// (rows 4*(lane/16)+j, zero beyond `in`).  Returns the output tile (rows
// 4*(lane/16)+j < out valid, without the output bias); h1b / h2b receive the
// bf16 hidden activations when non-null (for training).  b1 and b2 must be
// 16-byte aligned: the biases initialise the accumulators (four rows per
// lane), so the bias add costs no instruction.
__device__ inline mppi_v4f mppi_mlp_block(const short* frags, const float* b1, const float* b2,
                                            const unsigned short* tab, mppi_v4bf xb, uint32_t lane,
                                            mppi_v4bf* h1b_out, mppi_v4bf* h2b_out) {
    const uint32_t g = lane >> 4;
    mppi_v4bf h1b[8], h2b[8];
    const mppi_v4f zero = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
    for (int mb = 0; mb < 8; ++mb) {
        const mppi_v4f bias = *reinterpret_cast<const mppi_v4f*>(b1 + 16 * mb + 4 * g);
        h1b[mb] = mppi_act4(mppi_mfma(mppi_frag(frags, MPPI_MLP_FRAG_W1 + mb, lane), xb, bias), tab);
    }
#pragma unroll
    for (int mb = 0; mb < 8; ++mb) {
        mppi_v4f acc = *reinterpret_cast<const mppi_v4f*>(b2 + 16 * mb + 4 * g);
#pragma unroll
        for (int kb = 0; kb < 8; ++kb)
            acc = mppi_mfma(mppi_frag(frags, MPPI_MLP_FRAG_W2 + mb * 8 + kb, lane), h1b[kb], acc);
        h2b[mb] = mppi_act4(acc, tab);
    }
    mppi_v4f out = zero;
#pragma unroll
    for (int kb = 0; kb < 8; ++kb) out = mppi_mfma(mppi_frag(frags, MPPI_MLP_FRAG_W3 + kb, lane), h2b[kb], out);
    if (h1b_out) for (int i = 0; i < 8; ++i) h1b_out[i] = h1b[i];
    if (h2b_out) for (int i = 0; i < 8; ++i) h2b_out[i] = h2b[i];
    return out;
}
#endif

#endif  // MPPI_HIP_MLP_H
