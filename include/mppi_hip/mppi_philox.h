// mppi_philox.h — counter-based random numbers shared by the GPU kernels and
// the CPU oracle.
//
// Philox4x32-10 and Philox2x32-10 are the generators of Salmon, Moraes, Dror
// and Shaw, "Parallel random numbers: as easy as 1, 2, 3" (SC'11), as
// published with Random123: multipliers 0xD2511F53 / 0xCD9E8D57 (4x32) and
// 0xD256D193 (2x32), Weyl key increments 0x9E3779B9 / 0xBB67AE85, ten rounds.
// The 2x32 form is the reference's generator (shaders/philox.h:19-36), kept
// for compatibility checks; the library's noise stream uses the 4x32 form.
//
// Noise stream contract (replaces the reference's flat u32 counter,
// shaders/mppi_rollout.metal:73-87, which overflows at K*H*P >= 2^31):
//   key     = the 64-bit seed (lo word, hi word)
//   counter = (block, sample, step, tag << 28 | agent)
// Sample k of agent a at MPPI step s draws the normals of its H*C control
// noise values as one sequence n = t*C + d; normal n comes from counter block
// b = n / 4, lane j = n % 4: the block's words (w0,w1) and (w2,w3) are two
// Box-Muller pairs giving (z0, z1) and (z2, z3).
#ifndef MPPI_HIP_PHILOX_H
#define MPPI_HIP_PHILOX_H

#include "mppi_math.h"

struct mppi_u32x4 { uint32_t x, y, z, w; };
struct mppi_f32x4 { float x, y, z, w; };

// Stream tags (top 4 bits of counter word 3).
#define MPPI_TAG_CONTROL_NOISE 0u   // rollout control noise
#define MPPI_TAG_DYNAMICS      1u   // stochastic dynamics inside the rollout
#define MPPI_TAG_PLANT         2u   // closed-loop plant step (simulate)
#define MPPI_TAG_SCENE         3u   // synthetic scene generation
#define MPPI_TAG_SURROGATE     4u   // surrogate training data
#define MPPI_AGENT_MASK        0x0fffffffu

MPPI_HD uint32_t mppi_mulhilo32(uint32_t a, uint32_t b, uint32_t* hi) {
    uint64_t p = uint64_t(a) * uint64_t(b);
    *hi = uint32_t(p >> 32);
    return uint32_t(p);
}

MPPI_HD mppi_u32x4 mppi_philox4x32_10(mppi_u32x4 c, uint32_t k0, uint32_t k1) {
MPPI_UNROLL
    for (int i = 0; i < 10; ++i) {
        uint32_t hi0, hi1;
        uint32_t lo0 = mppi_mulhilo32(0xD2511F53u, c.x, &hi0);
        uint32_t lo1 = mppi_mulhilo32(0xCD9E8D57u, c.z, &hi1);
        mppi_u32x4 n;
        n.x = hi1 ^ c.y ^ k0;
        n.y = lo1;
        n.z = hi0 ^ c.w ^ k1;
        n.w = lo0;
        c = n;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    return c;
}

// Philox4x32-10 split for counters whose words x, z, w are the same in every
// lane of a workgroup and whose word y (the sample) differs.  In rounds 1-4
// the lane words are xored with words that depend only on x, z, w and the
// key: mppi_philox_shared computes four of them per x (a: round 2, b and c:
// round 3, d: round 4), mppi_philox_lane_key the round-1 word e = hi(M1 z) ^
// k0, and mppi_philox_lanes finishes the block from y.  XOR is associative and
// every product is the one mppi_philox4x32_10 forms, so the block is the same
// bit for bit.  The rollout keeps the shared words of its noise blocks in LDS.
struct mppi_philox_shared_words { uint32_t a, b, c, d; };

MPPI_HD mppi_philox_shared_words mppi_philox_shared(uint32_t x, uint32_t z, uint32_t w, uint32_t k0,
                                                    uint32_t k1) {
    uint32_t hx, hz, h2, h3;
    const uint32_t lx = mppi_mulhilo32(0xD2511F53u, x, &hx);           // round 1: M0 x
    const uint32_t lz = mppi_mulhilo32(0xCD9E8D57u, z, &hz);           // round 1: M1 z
    const uint32_t z1 = hx ^ w ^ k1;                                   // round 1 word z
    const uint32_t l2 = mppi_mulhilo32(0xCD9E8D57u, z1, &h2);          // round 2: M1 z1
    const uint32_t x2 = h2 ^ lz ^ (k0 + 0x9E3779B9u);                  // round 2 word x
    const uint32_t l3 = mppi_mulhilo32(0xD2511F53u, x2, &h3);          // round 3: M0 x2
    mppi_philox_shared_words r;
    r.a = lx ^ (k1 + 0xBB67AE85u);                 // round 2: z = hi(M0 x1) ^ a
    r.b = l2 ^ (k0 + 2u * 0x9E3779B9u);            // round 3: x = hi(M1 z2) ^ b
    r.c = h3 ^ (k1 + 2u * 0xBB67AE85u);            // round 3: z = w2 ^ c
    r.d = l3 ^ (k1 + 3u * 0xBB67AE85u);            // round 4: z = hi(M0 x3) ^ d
    return r;
}

MPPI_HD uint32_t mppi_philox_lane_key(uint32_t z, uint32_t k0) {
    uint32_t hz;
    mppi_mulhilo32(0xCD9E8D57u, z, &hz);
    return hz ^ k0;                                // round 1: x = y ^ e
}

MPPI_HD mppi_u32x4 mppi_philox_lanes(uint32_t y, uint32_t e, mppi_philox_shared_words s, uint32_t k0,
                                     uint32_t k1) {
    uint32_t h0, h1;
    const uint32_t x1 = y ^ e;                                         // round 1
    const uint32_t w2 = mppi_mulhilo32(0xD2511F53u, x1, &h0);          // round 2
    const uint32_t z2 = h0 ^ s.a;
    const uint32_t y3 = mppi_mulhilo32(0xCD9E8D57u, z2, &h1);          // round 3
    const uint32_t x3 = h1 ^ s.b;
    const uint32_t z3 = w2 ^ s.c;
    mppi_u32x4 c;                                                      // round 4
    const uint32_t lo0 = mppi_mulhilo32(0xD2511F53u, x3, &h0);
    const uint32_t lo1 = mppi_mulhilo32(0xCD9E8D57u, z3, &h1);
    c.x = h1 ^ y3 ^ (k0 + 3u * 0x9E3779B9u);
    c.y = lo1;
    c.z = h0 ^ s.d;
    c.w = lo0;
    k0 += 4u * 0x9E3779B9u;
    k1 += 4u * 0xBB67AE85u;
MPPI_UNROLL
    for (int i = 4; i < 10; ++i) {                                     // rounds 5-10
        uint32_t hi0, hi1;
        const uint32_t l0 = mppi_mulhilo32(0xD2511F53u, c.x, &hi0);
        const uint32_t l1 = mppi_mulhilo32(0xCD9E8D57u, c.z, &hi1);
        mppi_u32x4 n;
        n.x = hi1 ^ c.y ^ k0;
        n.y = l1;
        n.z = hi0 ^ c.w ^ k1;
        n.w = l0;
        c = n;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    return c;
}

// Reference-compatible Philox2x32-10 (counter (x, y), 32-bit key).
MPPI_HD void mppi_philox2x32_10(uint32_t* x, uint32_t* y, uint32_t key) {
MPPI_UNROLL
    for (int i = 0; i < 10; ++i) {
        uint32_t hi;
        uint32_t lo = mppi_mulhilo32(0xD256D193u, *x, &hi);
        uint32_t nx = hi ^ key ^ *y;
        *y = lo;
        *x = nx;
        key += 0x9E3779B9u;
    }
}

// Uniform in (0, 1): the top 23 bits j of x give (2j + 1) * 2^-24, which is
// exact in binary32, symmetric about 1/2, never 0 and never 1.
MPPI_HD float mppi_u01(uint32_t x) {
    return float((x >> 8) | 1u) * 5.9604644775390625e-8f;  // 2^-24
}

// x rounded to the nearest binary16 value (ties to even), as a float: the
// FP16 numeric mode draws its control noise at half precision.  Integer and
// exact float operations only, so the host oracle reproduces it.
MPPI_HD float mppi_round_f16(float x) {
    uint32_t u = mppi_f2u(x);
    const uint32_t sign = u & 0x80000000u;
    uint32_t a = u & 0x7fffffffu;
    if (a > 0x7f800000u) return x;                       // NaN
    if (a >= 0x477ff000u) return mppi_u2f(sign | 0x7f800000u);  // rounds past 65504
    if (a < 0x38800000u) {                               // half subnormal: quantum 2^-24
        const float q = mppi_rintf(mppi_u2f(a) * 16777216.0f) * 5.9604644775390625e-8f;
        return mppi_u2f(sign | mppi_f2u(q));
    }
    a = (a + 0x0fffu + ((a >> 13) & 1u)) & 0xffffe000u;  // keep 10 fraction bits
    return mppi_u2f(sign | a);
}

// Box-Muller: one pair of uniform words -> two standard normals,
// z0 = r cos(2 pi u2), z1 = r sin(2 pi u2), r = sqrt(-2 log u1), with
// u1 = mppi_u01(a), u2 = mppi_u01(b).
MPPI_HD void mppi_box_muller(uint32_t a, uint32_t b, float* z0, float* z1) {
    float s, c;
#if MPPI_FAST
    const float r = mppi_sqrt_boundedf(-2.0f * mppi_log_normalf(mppi_u01(a)));
    mppi_sincospif(2.0f * mppi_u01(b), &s, &c);
#else
    // u1 = m1 2^-24 with m1 = (a >> 8) | 1: the logarithm core takes m1 and
    // the exponent offset -24, the bits of mppi_log_normalf(u1).  -2 log(u1)
    // lies in [1.1e-7, 33.3], inside the range of the bounded square root.
    const uint32_t m1 = (a >> 8) | 1u;
    const float r = mppi_sqrt_boundedf(-2.0f * mppi_log_core_(mppi_f2u(float(m1)), -24));
    // The angle in integers: u2 = m 2^-24 with m = (b >> 8) | 1 odd, so
    // x = 2 u2 = m 2^-23, its nearest quarter turn n = rint(2x) =
    // (m + 2^21) >> 22 (no ties: m is odd), f = x - n/2 = (m - n 2^22) 2^-23
    // exactly, and f pi rounds once, as float(m - n 2^22) (pi 2^-23) does:
    // the bits of mppi_sincospif(x).
    const uint32_t m = (b >> 8) | 1u;
    const uint32_t n = (m + 0x200000u) >> 22;
    const float fpi = float(int32_t(m) - int32_t(n << 22)) * (3.14159274101257324219f * 0x1p-23f);
    mppi_quadrant(n, mppi_sin_poly(fpi), mppi_cos_poly(fpi), &s, &c);
#endif
    *z0 = r * c;
    *z1 = r * s;
}

// Four standard normals from the words of one Philox4x32-10 block.
MPPI_HD mppi_f32x4 mppi_normal4_words(mppi_u32x4 w) {
    mppi_f32x4 z;
    mppi_box_muller(w.x, w.y, &z.x, &z.y);
    mppi_box_muller(w.z, w.w, &z.z, &z.w);
#if defined(MPPI_NOISE_F16) && MPPI_NOISE_F16
    z.x = mppi_round_f16(z.x);
    z.y = mppi_round_f16(z.y);
    z.z = mppi_round_f16(z.z);
    z.w = mppi_round_f16(z.w);
#endif
    return z;
}

// Four standard normals from one Philox4x32-10 block.
MPPI_HD mppi_f32x4 mppi_normal4(mppi_u32x4 ctr, uint32_t k0, uint32_t k1) {
    return mppi_normal4_words(mppi_philox4x32_10(ctr, k0, k1));
}

MPPI_HD mppi_u32x4 mppi_noise_counter(uint32_t block, uint32_t sample, uint32_t step,
                                      uint32_t tag, uint32_t agent) {
    mppi_u32x4 c;
    c.x = block;
    c.y = sample;
    c.z = step;
    c.w = (tag << 28) | (agent & MPPI_AGENT_MASK);
    return c;
}

#endif  // MPPI_HIP_PHILOX_H
