// mppi_core.h — the device contract between the library-owned MPPI kernels
// and user model plugins.  It replaces the reference's shaders/mppi_core.h:11-42
// (callable typedefs, thread-local limits, function-table slots).
//
// A plugin is a source file of plain functions marked MPPI_CALLABLE.  The
// library compiles it with hipRTC in one translation unit with its own kernels
// (direct, inlinable calls; no function tables), and the CPU oracle compiles
// the same file with g++.  Callables (names are chosen in CallableNames):
//
//   dynamics (required; reference MppiDynamicsFn, mppi_core.h:11-15):
//     MPPI_CALLABLE void f(float* state, const float* control,
//                          const uint8_t* model_params, mppi_rng rng);
//     Propagates the state in place by one step.
//   stage cost (optional; reference MppiStageCostFn, mppi_core.h:18-21):
//     MPPI_CALLABLE float f(const float* state, const float* control,
//                           const uint8_t* model_params, const uint8_t* cost_params);
//     Called with the state AFTER the dynamics step and the clamped control.
//   terminal cost (optional; reference MppiTerminalCostFn, mppi_core.h:24-26):
//     MPPI_CALLABLE float f(const float* state, const uint8_t* model_params,
//                           const uint8_t* cost_params);
//   Either cost may add a trailing `mppi_time` argument (new; see below).
//   cost preparation (optional; new, no reference counterpart): a plugin
//     that defines MPPI_COST_PREPARE as the name of a device function
//       void f(const uint8_t* cost_params, uint32_t thread, uint32_t threads);
//     has it called on the GPU by every thread of a workgroup (thread in
//     [0, threads)) with the workgroup's cost parameters before a workgroup
//     barrier that precedes the workgroup's first cost evaluation with them.
//     It can fill workgroup memory (__shared__ arrays declared at file scope
//     under `#if defined(__HIP_DEVICE_COMPILE__)`) that the costs then read;
//     the costs' results must not depend on it (the host oracle has no
//     workgroups and never calls it).
//
// The state argument is passed through mppi_state_ref / mppi_cstate_ref and
// the parameter blobs through mppi_blob_ref, which convert implicitly to
// float*, uint8_t* or void* — the reference's plugins declare the same slot
// as both `thread uint8_t*` and `thread const float*`
// (test_user_dynamics.metal:27 vs mppi_core.h:18), and both forms compile.
//
// Floating-point rule for plugins: use the mppi_* functions of mppi_math.h for
// transcendentals (mppi_sinf, mppi_cosf, mppi_expf, mppi_logf, mppi_sqrtf,
// mppi_fmodf, mppi_tanhf).  Both compilers then produce the same bits.
#ifndef MPPI_HIP_CORE_H
#define MPPI_HIP_CORE_H

#include "mppi_math.h"
#include "mppi_philox.h"

#define MPPI_CALLABLE MPPI_HD

// True when `c` holds for this sample or, on the GPU, for any sample of the
// same wavefront.  A callable can guard work that only some samples need so
// that a wavefront in which no sample needs it skips it (a uniform branch
// instead of computing and discarding the work on every lane):
//   bool need = false;
//   for (i...) need |= cheap_test(i);
//   if (mppi_any(need)) for (i...) if (cheap_test(i)) expensive(i);
// Guard a loop or a large block: the optimiser evaluates a small guarded
// block unconditionally and selects its result.  Results are the same either
// way; on the host it is the identity.
MPPI_HD bool mppi_any(bool c) { return mppi_wave_any_(c); }

// Random-number handle given to the dynamics callable.  Draw j of timestep t
// of sample k uses counter (t << 8 | j, k, step, tag << 28 | agent); a
// callable may make up to 256 draws of four words each per step.
struct mppi_rng {
    uint32_t block_base;   // (t << 8)
    uint32_t sample;
    uint32_t step;
    uint32_t tag_agent;
    uint32_t key0, key1;
};

MPPI_HD mppi_u32x4 mppi_rng_bits(const mppi_rng& r, uint32_t draw) {
    mppi_u32x4 c;
    c.x = r.block_base + (draw & 0xffu);
    c.y = r.sample;
    c.z = r.step;
    c.w = r.tag_agent;
    return mppi_philox4x32_10(c, r.key0, r.key1);
}

// Four uniforms in (0, 1).
MPPI_HD mppi_f32x4 mppi_rng_uniform4(const mppi_rng& r, uint32_t draw) {
    mppi_u32x4 w = mppi_rng_bits(r, draw);
    mppi_f32x4 u;
    u.x = mppi_u01(w.x); u.y = mppi_u01(w.y); u.z = mppi_u01(w.z); u.w = mppi_u01(w.w);
    return u;
}

// Four standard normals.
MPPI_HD mppi_f32x4 mppi_rng_normal4(const mppi_rng& r, uint32_t draw) {
    mppi_u32x4 w = mppi_rng_bits(r, draw);
    mppi_f32x4 z;
    mppi_box_muller(w.x, w.y, &z.x, &z.y);
    mppi_box_muller(w.z, w.w, &z.z, &z.w);
    return z;
}

// Time of a cost evaluation: timestep t within the horizon (0-based; the
// terminal cost gets t = H) and the MPPI step index.  A stage or terminal
// cost may take it as a trailing argument (e.g. to follow a moving target);
// the library detects which form a plugin uses, so plugins without it are
// called as in the reference.
struct mppi_time {
    uint32_t t;
    uint32_t step;
};

// Argument proxies (see the header comment).
struct mppi_state_ref {
    float* p;
    MPPI_HD operator float*() const { return p; }
    MPPI_HD operator const float*() const { return p; }
    MPPI_HD operator uint8_t*() const { return reinterpret_cast<uint8_t*>(p); }
    MPPI_HD operator const uint8_t*() const { return reinterpret_cast<const uint8_t*>(p); }
    MPPI_HD operator void*() const { return p; }
    MPPI_HD operator const void*() const { return p; }
};

struct mppi_cstate_ref {
    const float* p;
    MPPI_HD operator const float*() const { return p; }
    MPPI_HD operator const uint8_t*() const { return reinterpret_cast<const uint8_t*>(p); }
    MPPI_HD operator const void*() const { return p; }
};

struct mppi_blob_ref {
    const uint8_t* p;
    MPPI_HD operator const uint8_t*() const { return p; }
    MPPI_HD operator const void*() const { return p; }
    MPPI_HD operator const float*() const { return reinterpret_cast<const float*>(p); }
};

// Default costs, linked when a plugin names no stage or terminal cost
// (reference mppi_rollout.metal:11-28).
MPPI_CALLABLE float mppi_default_stage_cost(const float*, const float*, const uint8_t*,
                                            const uint8_t*) {
    return 0.0f;
}
MPPI_CALLABLE float mppi_default_terminal_cost(const float*, const uint8_t*, const uint8_t*) {
    return 0.0f;
}

// Size limits.  The reference caps thread-local state at 128 bytes
// (mppi_core.h:32-33) and silently truncates larger states; here the state
// array is sized from the JIT-time state_dim, and these caps are checked by
// the driver before compiling.
#define MPPI_MAX_STATE_DIM 64
#define MPPI_MAX_CONTROL_DIM 16

#endif  // MPPI_HIP_CORE_H
