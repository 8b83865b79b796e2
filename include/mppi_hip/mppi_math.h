// mppi_math.h — single-precision elementary functions used identically by the
// GPU rollout (compiled by hipRTC/hipcc for gfx950) and by the CPU oracle
// (compiled by g++).
//
// Why these exist: the MPPI weights are exp(-(c - min)/lambda), so a cost that
// differs by one ulp between two builds changes the weights.  Vendor libm
// (glibc on the host, OCML on the GPU) give different last-bit results, so
// every transcendental on the rollout path is written here from +, -, *, fma,
// correctly rounded / and sqrt, comparisons and bit manipulation only.  Each of
// those operations is IEEE-754 exact-rounded on both targets, so the same
// source gives the same bits on both, provided floating-point contraction is
// off (-ffp-contract=off on both compilers; the library passes it to hipRTC).
//
// Methods (all public, textbook):
//   exp   : Cody-Waite reduction x = k*ln2 + r with ln2 split in two parts,
//           |r| <= ln2/2, Taylor polynomial of degree 7, scaling by 2^k.
//   log   : x = 2^e * m, m in [sqrt(1/2), sqrt(2)), f = m - 1 (exact);
//           log(1 + f) = f + f^2 * P(f) with P of degree 8 fitted on
//           [sqrt(1/2) - 1, sqrt(2) - 1] (no division); at most 0.92 ulp
//           against a double logarithm over every float in [2^-30, 2^8)
//           when fitted, and every float in [1/2, 2) within one ulp
//           (exhaustive) in the tests.
//   sin/cos: Cody-Waite reduction by pi/2 with a three-part pi/2 (the fdlibm
//           split), Taylor polynomials on [-pi/4, pi/4], quadrant selection.
//   sincospi: exact reduction of x to [-1/4, 1/4] (multiples of 1/2 are exact
//           in binary), then the same polynomials.
//   fmod  : exact long division by repeated subtraction of y*2^j (each step
//           is exact by Sterbenz's lemma), as IEEE fmod requires.
//   tanh  : via exp for |x| >= 0.55 and an odd Taylor series below.
// Accuracy against FP64 libm is measured in tests/test_oracle_math.py.
//
// Fast mode (MPPI_FAST_MATH = 1, set by the library for the MIXED and FP16
// numeric modes; GPU code only): exp, log, sqrt, sin, cos, sincospi and tanh
// map to the hardware's single-precision instructions (v_exp_f32, v_log_f32,
// v_sqrt_f32, v_sin_f32 / v_cos_f32 on revolutions, v_rcp_f32), about 1-2 ulp
// and a few instructions each instead of dozens.  Results are then close to,
// not equal to, the CPU oracle's.  fmod stays exact.
#ifndef MPPI_HIP_MATH_H
#define MPPI_HIP_MATH_H

#if defined(__HIPCC_RTC__) || defined(__HIPCC__) || defined(__HIP__)
#define MPPI_HD __host__ __device__ inline
#else
#define MPPI_HD inline
#endif

#if defined(__HIPCC_RTC__)
// hipRTC has no <stdint.h>; the compiler's builtin type macros give the same types.
typedef __UINT8_TYPE__ uint8_t;
typedef __UINT32_TYPE__ uint32_t;
typedef __UINT64_TYPE__ uint64_t;
typedef __INT32_TYPE__ int32_t;
typedef __SIZE_TYPE__ size_t;
#else
#include <stdint.h>
#endif

#if defined(__HIP_DEVICE_COMPILE__) && defined(MPPI_FAST_MATH) && MPPI_FAST_MATH
#define MPPI_FAST 1
#else
#define MPPI_FAST 0
#endif

#if defined(__clang__)
#define MPPI_UNROLL _Pragma("unroll")
#else
#define MPPI_UNROLL
#endif

MPPI_HD uint32_t mppi_f2u(float f) { return __builtin_bit_cast(uint32_t, f); }
MPPI_HD float mppi_u2f(uint32_t u) { return __builtin_bit_cast(float, u); }

MPPI_HD float mppi_fma(float a, float b, float c) { return __builtin_fmaf(a, b, c); }

// True when c holds in any lane of the wavefront (on the host: c itself).
// Guards a rarely needed slow path so that it costs a uniform branch rather
// than divergent control flow in the common case.
MPPI_HD bool mppi_wave_any_(bool c) {
#if defined(__HIP_DEVICE_COMPILE__)
    return __builtin_amdgcn_ballot_w64(c) != 0;
#else
    return c;
#endif
}
MPPI_HD float mppi_sqrtf(float x) {
#if MPPI_FAST
    return __builtin_amdgcn_sqrtf(x);
#else
    return __builtin_sqrtf(x);
#endif
}
MPPI_HD float mppi_fabsf(float x) { return mppi_u2f(mppi_f2u(x) & 0x7fffffffu); }

// 1/x within one ulp for normal x and results: the hardware estimate on the
// GPU, a division on the host.  The two may differ in the last bit, so it is
// only for bounds whose proof allows that error on each side (conservative
// tests that gate an exact computation), never for a value that reaches a
// result.
MPPI_HD float mppi_rcp_estimatef(float x) {
#if defined(__HIP_DEVICE_COMPILE__)
    return __builtin_amdgcn_rcpf(x);
#else
    return 1.0f / x;
#endif
}

// Correctly rounded square root of x in [2^-96, 2^126] or +-0 (finite,
// no subnormal, no scaling needed): the same result as mppi_sqrtf there.  On
// the GPU, the hardware square root (within 1 ulp) is corrected by the
// residuals of its two neighbours, the compiler's own correctly rounded
// sequence without the range handling that the argument does not need.
MPPI_HD float mppi_sqrt_boundedf(float x) {
#if defined(__HIP_DEVICE_COMPILE__) && !MPPI_FAST
    const float s = __builtin_amdgcn_sqrtf(x);
    const uint32_t si = mppi_f2u(s);
    const float down = mppi_u2f(si - 1u), up = mppi_u2f(si + 1u);
    float r = mppi_fma(-down, s, x) <= 0.0f ? down : s;
    return mppi_fma(-up, s, x) > 0.0f ? up : r;
#else
    return mppi_sqrtf(x);
#endif
}

// Maximum and minimum with the rules of the GPU's v_max_f32 / v_min_f32 in
// IEEE mode, stated for the host: a NaN operand yields the other operand,
// and -0 orders below +0.  On the device each is one instruction.
MPPI_HD float mppi_maxf(float a, float b) {
#if defined(__HIP_DEVICE_COMPILE__)
    return __builtin_fmaxf(a, b);
#else
    if (a != a) return b;
    if (b != b) return a;
    if (a == b) return mppi_f2u(a) >> 31 ? b : a;
    return a > b ? a : b;
#endif
}
MPPI_HD float mppi_minf(float a, float b) {
#if defined(__HIP_DEVICE_COMPILE__)
    return __builtin_fminf(a, b);
#else
    if (a != a) return b;
    if (b != b) return a;
    if (a == b) return mppi_f2u(a) >> 31 ? a : b;
    return a < b ? a : b;
#endif
}

// clamp(x, lo, hi) = min(max(x, lo), hi), the reference's Metal clamp (a NaN
// x gives lo): two instructions on the device.
MPPI_HD float mppi_clampf(float x, float lo, float hi) {
    return mppi_minf(mppi_maxf(x, lo), hi);
}

// Round to nearest integer (ties to even) for |x| < 2^22.
MPPI_HD float mppi_rintf(float x) {
    const float magic = 12582912.0f;  // 1.5 * 2^23
    float t = x + magic;
    return t - magic;
}

// 2^k as a float for k in [-126, 127].
MPPI_HD float mppi_pow2i(int k) { return mppi_u2f(uint32_t(k + 127) << 23); }

// ---------------------------------------------------------------------------
// exp
// ---------------------------------------------------------------------------
MPPI_HD float mppi_expf(float x) {
#if MPPI_FAST
    return __builtin_amdgcn_exp2f(x * 1.44269504088896340736f);
#endif
    if (!(x == x)) return x;                       // NaN
    if (x > 88.72283935546875f) return mppi_u2f(0x7f800000u);  // +inf
    if (x < -103.972084045410156f) return 0.0f;    // below the smallest subnormal
    const float log2e = 1.44269502162933349609375f;
    const float ln2_hi = 0.693145751953125f;       // 0x3f317200: 16 trailing zero bits
    const float ln2_lo = 1.428606765330187045e-06f;
    float kf = mppi_rintf(x * log2e);
    float r = mppi_fma(-kf, ln2_hi, x);
    r = mppi_fma(-kf, ln2_lo, r);
    // Taylor coefficients 1/n!, n = 7..0.
    float p = 1.98412698412698412698e-4f;
    p = mppi_fma(p, r, 1.38888888888888888889e-3f);
    p = mppi_fma(p, r, 8.33333333333333333333e-3f);
    p = mppi_fma(p, r, 4.16666666666666666667e-2f);
    p = mppi_fma(p, r, 1.66666666666666666667e-1f);
    p = mppi_fma(p, r, 0.5f);
    p = mppi_fma(p, r, 1.0f);
    p = mppi_fma(p, r, 1.0f);
    int k = int(kf);
    if (k > 127) return (p * 2.0f) * mppi_pow2i(k - 1);
    if (k < -126) return (p * mppi_pow2i(-126)) * mppi_pow2i(k + 126);
    return p * mppi_pow2i(k);
}

// ---------------------------------------------------------------------------
// log
// ---------------------------------------------------------------------------
// log of 2^e * x for a positive normal float x (bit pattern ix): the core of
// mppi_logf, also called directly where the argument is known to be a
// positive normal number (Box-Muller), with the same result.
MPPI_HD float mppi_log_core_(uint32_t ix, int e) {
    e += int(ix >> 23) - 127;
    uint32_t mi = (ix & 0x007fffffu) | 0x3f800000u;  // m in [1, 2)
    if (mi > 0x3fb504f3u) {                        // m > sqrt(2): halve it
        mi -= 0x00800000u;
        e += 1;
    }
    float m = mppi_u2f(mi);
    float f = m - 1.0f;                            // exact (Sterbenz)
    // log(1 + f) = f + f^2 P(f) on [sqrt(1/2) - 1, sqrt(2) - 1], P of degree 8
    // fitted (weighted least squares, iterated towards minimax) to
    // (log1p(f) - f) / f^2 in the error of the result; no division.  With
    // the exponent term below the error is at most 0.92 ulp over every float
    // in [2^-30, 2^8) against double-precision log (checked exhaustively).
    float p = -0.07622639089822769f;
    p = mppi_fma(p, f, 0.12905363738536835f);
    p = mppi_fma(p, f, -0.13243086636066437f);
    p = mppi_fma(p, f, 0.14181660115718842f);
    p = mppi_fma(p, f, -0.1660904586315155f);
    p = mppi_fma(p, f, 0.20001983642578125f);
    p = mppi_fma(p, f, -0.25001585483551025f);
    p = mppi_fma(p, f, 0.33333325386047363f);
    p = mppi_fma(p, f, -0.49999988079071045f);
    float lm = mppi_fma(f * f, p, f);
    const float ln2_hi = 0.693145751953125f;
    const float ln2_lo = 1.428606765330187045e-06f;
    float ef = float(e);
    return mppi_fma(ef, ln2_hi, mppi_fma(ef, ln2_lo, lm));
}

MPPI_HD float mppi_logf(float x) {
#if MPPI_FAST
    return __builtin_amdgcn_logf(x) * 0.693147180559945309417f;
#endif
    if (!(x == x)) return x;
    if (x < 0.0f) return mppi_u2f(0x7fc00000u);    // NaN
    if (x == 0.0f) return mppi_u2f(0xff800000u);   // -inf
    uint32_t ix = mppi_f2u(x);
    if (ix == 0x7f800000u) return x;               // +inf
    const bool subnormal = ix < 0x00800000u;       // scaled by 2^25
    return mppi_log_core_(subnormal ? mppi_f2u(x * 33554432.0f) : ix, subnormal ? -25 : 0);
}

// mppi_logf for an argument known to be a positive normal float: the same
// result without the special-value tests.
MPPI_HD float mppi_log_normalf(float x) {
#if MPPI_FAST
    return __builtin_amdgcn_logf(x) * 0.693147180559945309417f;
#endif
    return mppi_log_core_(mppi_f2u(x), 0);
}

// ---------------------------------------------------------------------------
// sin / cos kernels on [-pi/4, pi/4]
// ---------------------------------------------------------------------------
MPPI_HD float mppi_sin_poly(float r) {
    float r2 = r * r;
    float p = 2.75573192239858906526e-6f;          // 1/9!
    p = mppi_fma(p, r2, -1.98412698412698412698e-4f);  // -1/7!
    p = mppi_fma(p, r2, 8.33333333333333333333e-3f);   // 1/5!
    p = mppi_fma(p, r2, -1.66666666666666666667e-1f);  // -1/3!
    return mppi_fma(r * r2, p, r);
}

MPPI_HD float mppi_cos_poly(float r) {
    float r2 = r * r;
    float p = -2.75573192239858906526e-7f;         // -1/10!
    p = mppi_fma(p, r2, 2.48015873015873015873e-5f);   // 1/8!
    p = mppi_fma(p, r2, -1.38888888888888888889e-3f);  // -1/6!
    p = mppi_fma(p, r2, 4.16666666666666666667e-2f);   // 1/4!
    p = mppi_fma(p, r2, -0.5f);
    return mppi_fma(r2, p, 1.0f);
}

// Select (sin, cos) of r + q*pi/2 from the kernels' values in bit
// operations only (no comparison, so no condition-mask hazards on the GPU):
// odd quadrants swap sin and cos, quadrants 2 and 3 negate sin, quadrants 1
// and 2 negate cos.  Only the low two bits of q are used.
MPPI_HD void mppi_quadrant(uint32_t q, float s, float c, float* so, float* co) {
    const uint32_t swap = 0u - (q & 1u);                 // all ones in odd quadrants
    const uint32_t su = mppi_f2u(s), cu = mppi_f2u(c);
    const uint32_t sa = (cu & swap) | (su & ~swap);
    const uint32_t ca = (su & swap) | (cu & ~swap);
    const uint32_t t = q << 30;                          // bit 1 of q at the sign bit
    *so = mppi_u2f(sa ^ (t & 0x80000000u));
    *co = mppi_u2f(ca ^ ((t + 0x40000000u) & 0x80000000u));
}

// k = round(v) (ties to even) by adding and subtracting 1.5 * 2^23, with the
// intermediate sum y returned as well: while |v| < 2^22, y lies in
// [2^23, 2^24), where floats are the integers, and its low mantissa bits are
// 2^22 + k, so mppi_f2u(y) holds k mod 4 in its low two bits (2^22 is a
// multiple of 4).  Beyond, k and those bits are what the same operations
// give on every target (no float-to-int conversion, whose out-of-range
// result differs between targets).
MPPI_HD float mppi_rint_bits(float v, uint32_t* bits) {
    const float magic = 12582912.0f;  // 1.5 * 2^23
    const float y = v + magic;
    *bits = mppi_f2u(y);
    return y - magic;
}

// sin(x), cos(x).  Accurate for |x| <= 1e5 (the three-part pi/2 keeps the
// reduced argument exact enough there); the same bits on every target for
// every input.
MPPI_HD void mppi_sincosf(float x, float* s, float* c) {
#if MPPI_FAST
    {
        // v_sin/v_cos take revolutions and are accurate for |r| < 256.
        const float r = x * 0.159154943091895335769f;   // 1 / (2 pi)
        if (mppi_fabsf(r) < 256.0f) {
            *s = __builtin_amdgcn_sinf(r);
            *c = __builtin_amdgcn_cosf(r);
            return;
        }
    }
#endif
    const float two_over_pi = 0.636619772367581343076f;
    const float pio2_1 = 1.57079625129699707031e+00f;  // 0x3fc90fda
    const float pio2_2 = 7.54978941586159635335e-08f;  // 0x33a22168
    const float pio2_3 = 5.39030252995776476554e-15f;  // 0x27c234c4
    uint32_t q;
    const float kf = mppi_rint_bits(x * two_over_pi, &q);
    float r = mppi_fma(-kf, pio2_1, x);
    r = mppi_fma(-kf, pio2_2, r);
    r = mppi_fma(-kf, pio2_3, r);
    mppi_quadrant(q, mppi_sin_poly(r), mppi_cos_poly(r), s, c);
}

MPPI_HD float mppi_sinf(float x) { float s, c; mppi_sincosf(x, &s, &c); return s; }
MPPI_HD float mppi_cosf(float x) { float s, c; mppi_sincosf(x, &s, &c); return c; }

// sin(pi*x), cos(pi*x) for |x| < 2^21.
MPPI_HD void mppi_sincospif(float x, float* s, float* c) {
#if MPPI_FAST
    if (mppi_fabsf(x) < 512.0f) {
        *s = __builtin_amdgcn_sinf(x * 0.5f);            // exact scaling to revolutions
        *c = __builtin_amdgcn_cosf(x * 0.5f);
        return;
    }
#endif
    uint32_t q;
    const float n = mppi_rint_bits(x * 2.0f, &q);  // exact scaling
    float f = x - n * 0.5f;                        // exact: f in [-1/4, 1/4]
    const float pi_f = 3.14159274101257324219f;
    mppi_quadrant(q, mppi_sin_poly(f * pi_f), mppi_cos_poly(f * pi_f), s, c);
}

// ---------------------------------------------------------------------------
// fmod: IEEE remainder toward zero, exact.
// ---------------------------------------------------------------------------
MPPI_HD float mppi_fmodf(float x, float y) {
    const float ax = mppi_fabsf(x), ay = mppi_fabsf(y);
    const bool nan_result = !(ax == ax) || !(ay == ay) || ay == 0.0f || ax == mppi_u2f(0x7f800000u);
    // ax < ay: the remainder is |x|.  ay <= ax < 2 ay: one subtraction, exact
    // by Sterbenz.  Both are selects, so the usual case has no branch.
    float r = ax < ay ? ax : ax - ay;
    const bool long_div = !nan_result && !(ax < ay * 2.0f);
    // Otherwise binary long division: yj runs over ay*2^i from the largest
    // value with ax < 2*yj down to ay.  With yj <= r < 2*yj each r - yj is
    // exact (Sterbenz), and halving yj is exact, so the remainder is exact.
    // On the GPU the loop is entered only by wavefronts that need it, through
    // a uniform branch; lanes of such a wavefront that need no long division
    // run it on (1, 1), which ends after one pass, and keep their r.
    if (mppi_wave_any_(long_div)) {
        const float axl = long_div ? ax : 1.0f, ayl = long_div ? ay : 1.0f;
        float yj = ayl;
        while (!(yj > axl * 0.5f)) yj = yj * 2.0f;
        float rl = axl;
        for (;;) {
            if (yj <= rl) rl = rl - yj;
            if (yj == ayl) break;
            yj = yj * 0.5f;
        }
        r = long_div ? rl : r;
    }
    // The result has the sign of x (x itself when |x| < |y|, -0 included).
    r = mppi_u2f(mppi_f2u(r) | (mppi_f2u(x) & 0x80000000u));
    return nan_result ? mppi_u2f(0x7fc00000u) : r;
}

// ---------------------------------------------------------------------------
// tanh
// ---------------------------------------------------------------------------
MPPI_HD float mppi_tanhf(float x) {
#if MPPI_FAST
    {
        const float t = x > 9.0f ? 9.0f : (x < -9.0f ? -9.0f : x);
        return 1.0f - 2.0f * __builtin_amdgcn_rcpf(__builtin_amdgcn_exp2f(t * 2.88539008177792681472f) + 1.0f);
    }
#endif
    float ax = mppi_fabsf(x);
    float r;
    if (ax < 0.55f) {
        // tanh(x) = x - x^3/3 + 2x^5/15 - 17x^7/315 + 62x^9/2835
        //         - 1382x^11/155925 + 21844x^13/6081075 - 929569x^15/638512875
        float x2 = ax * ax;
        float p = -1.45583438705131826e-3f;
        p = mppi_fma(p, x2, 3.59212803657248102e-3f);
        p = mppi_fma(p, x2, -8.86323552990219657e-3f);
        p = mppi_fma(p, x2, 2.18694885361552028e-2f);
        p = mppi_fma(p, x2, -5.39682539682539683e-2f);
        p = mppi_fma(p, x2, 1.33333333333333333e-1f);
        p = mppi_fma(p, x2, -3.33333333333333333e-1f);
        r = mppi_fma(ax * x2, p, ax);
    } else if (ax < 9.0f) {
        float e = mppi_expf(2.0f * ax);
        r = 1.0f - 2.0f / (e + 1.0f);
    } else {
        r = 1.0f;
    }
    return x < 0.0f ? -r : r;
}

#endif  // MPPI_HIP_MATH_H
