// Kinematic unicycle ("nonholonomic vehicle"), the model of the reference's
// example plugin (examples/nonholonomic_vehicle/nonholonomic_vehicle.metal:24-91),
// plus a circular-obstacle hinge penalty for the benchmark scenes.
//
// State [x, y, theta, v], control [a, omega].
//   x += v cos(theta) dt; y += v sin(theta) dt; theta += omega dt; v += a dt;
//   theta wrapped to [-pi, pi) with fmod; v clamped to [-v_max, v_max].
// Model params: UnicycleModel {v_max, dt} (8 bytes, as the reference's).
// Cost params:  UnicycleCost, whose first 24 bytes are the reference's
//   CostParams {x_goal, y_goal, goal_radius, pos_weight, heading_weight,
//   control_weight}; the obstacle fields follow.
//
// Callables:
//   mppi_dynamics                       — dynamics
//   mppi_stage_cost / mppi_terminal_cost — the reference's goal costs; they
//       read only the first 24 bytes, so the reference's CostParams works.
//   unicycle_stage_cost_obstacles       — goal cost plus
//       obstacle_weight * sum_i max(0, r_i - |p - c_i|) over the first
//       n_obstacles slots (at most 16); it needs the whole UnicycleCost
//       (224 bytes).
//   unicycle_terminal_cost              — 10 * pos_weight * |p - g|^2
#include "mppi_core.h"

#define UNICYCLE_MAX_OBSTACLES 16

struct UnicycleModel {
    float v_max;
    float dt;
};

struct UnicycleCost {
    float x_goal;
    float y_goal;
    float goal_radius;
    float pos_weight;
    float heading_weight;
    float control_weight;
    float obstacle_weight;
    uint32_t n_obstacles;
    float obstacles[UNICYCLE_MAX_OBSTACLES][3];  // (cx, cy, radius)
};

MPPI_CALLABLE void mppi_dynamics(float* state, const float* control,
                                 const uint8_t* model_params_raw, mppi_rng rng) {
    const UnicycleModel* params = reinterpret_cast<const UnicycleModel*>(model_params_raw);
    float dt = params->dt;
    float theta = state[2];
    float v = state[3];
    float s, c;
    mppi_sincosf(theta, &s, &c);
    state[0] += v * c * dt;
    state[1] += v * s * dt;
    state[2] += control[1] * dt;
    state[3] += control[0] * dt;
    const float pi = 3.1415926535f;
    float th = mppi_fmodf(state[2] + pi, 2.0f * pi);
    if (th < 0.0f) th += 2.0f * pi;
    state[2] = th - pi;
    if (state[3] > params->v_max) state[3] = params->v_max;
    if (state[3] < -params->v_max) state[3] = -params->v_max;
}

MPPI_CALLABLE float mppi_stage_cost(const float* state, const float* control,
                                    const uint8_t* model_params, const uint8_t* cost_params_raw) {
    const UnicycleCost* p = reinterpret_cast<const UnicycleCost*>(cost_params_raw);
    float dx = state[0] - p->x_goal;
    float dy = state[1] - p->y_goal;
    float cost = p->pos_weight * (dx * dx + dy * dy);
    cost += p->control_weight * (control[0] * control[0] + control[1] * control[1]);
    return cost;
}

// The first pass's square test for obstacle (cx, cy, r), as scaled
// coordinates ox = x a + bx, oy = y a + by with a ~ 1 / (r t), bx = -cx a,
// by = -cy a: a sample inside the square has max(|ox|, |oy|) < 1.  Proof that
// every sample the exact pass counts (d2 < r*r, rounded as written there) is
// flagged: d2 >= fl(ox^2) and fl(r^2) <= r^2 (1 + 2^-24) give |x - cx| <
// r (1 + 2^-22) (likewise y; r in [2^-60, 2^60] keeps r^2 normal and finite,
// and an underflowed ox^2 means |x - cx| < 2^-62).  The estimates of a
// (one ulp each for 1/r and 1/(1 + s), and two roundings) and the roundings
// of bx and of the fused multiply-add move ox by less than
// 1.5 * 2^-21 + (|cx| / r) 2^-24 beyond (x - cx) / (r t); the effective t,
// (1 + s) / (1 - 2^-18) with s = (|cx| + |cy|) / r 2^-21, exceeds 1 by
// more than four times that.  Radii outside [2^-60, 2^60] are always sent to
// the exact pass (a = 0); a slot with r <= 0 or NaN is never counted, as its
// hinge max(0, r - |p - c|) is 0 (bx = by = 2 keeps its square empty).  The
// gate only decides where the exact pass runs: the costs are the exact sum
// over the obstacles, which is what the host computes (without the gate).
struct UnicycleGate {
    float a, bx, by;
};

MPPI_HD UnicycleGate unicycle_gate(float cx, float cy, float r) {
    if (!(r > 0.0f)) return {0.0f, 2.0f, 2.0f};
    if (!(r >= 0x1p-60f && r <= 0x1p60f)) return {0.0f, 0.0f, 0.0f};
    const float q = mppi_rcp_estimatef(r);
    const float s = (mppi_fabsf(cx) + mppi_fabsf(cy)) * q * 0x1p-21f;
    const float a = (q * 0.999996185302734375f) * mppi_rcp_estimatef(1.0f + s);   // 1 - 2^-18
    return {a, -cx * a, -cy * a};
}

MPPI_CALLABLE float unicycle_stage_cost_obstacles(const float* state, const float* control,
                                                  const uint8_t* model_params,
                                                  const uint8_t* cost_params_raw) {
    const UnicycleCost* p = reinterpret_cast<const UnicycleCost*>(cost_params_raw);
    float cost = mppi_stage_cost(state, control, model_params, cost_params_raw);
    float pen = 0.0f;
    uint32_t n = p->n_obstacles < UNICYCLE_MAX_OBSTACLES ? p->n_obstacles : UNICYCLE_MAX_OBSTACLES;
#if defined(__HIP_DEVICE_COMPILE__)
    // max(0, r - |p - c|) summed over the obstacles.  The square root is
    // only needed inside an obstacle, so a cheap first pass tests the square
    // |ox|, |oy| < r around each circle, in coordinates scaled by the radius
    // (two fused multiply-adds per obstacle; unicycle_gate).  The exact pass
    // runs only in wavefronts where some sample is in some square.  The first
    // pass tests all 16 slots whatever n is, so the obstacle parameters do
    // not depend on n and stay in registers across the whole horizon.  Slots
    // at or past n may hold anything: they can only send a wavefront to the
    // exact pass, which skips them.
    float sq[UNICYCLE_MAX_OBSTACLES];   // max(|ox|, |oy|) / r, scaled: < 1 inside the square
    float margin = 2.0f;                // their minimum
    MPPI_UNROLL
    for (uint32_t j = 0; j < UNICYCLE_MAX_OBSTACLES; ++j) {
        const UnicycleGate g = unicycle_gate(p->obstacles[j][0], p->obstacles[j][1], p->obstacles[j][2]);
        const float ox = mppi_fma(state[0], g.a, g.bx);
        const float oy = mppi_fma(state[1], g.a, g.by);
        sq[j] = __builtin_fmaxf(__builtin_fabsf(ox), __builtin_fabsf(oy));
        margin = __builtin_fminf(margin, sq[j]);
    }
    // The exact pass keeps slot order and takes the distance only for
    // obstacles whose square holds some sample of the wavefront, and a square
    // root only for samples inside the circle.  Marked unlikely so that it is
    // compiled out of line rather than evaluated on every step.
    if (__builtin_expect(mppi_any(margin < 1.0f), 0)) {
        MPPI_UNROLL
        for (uint32_t i = 0; i < UNICYCLE_MAX_OBSTACLES; ++i) {
            if (__builtin_expect(mppi_any(sq[i] < 1.0f), 0)) {
                float ox = state[0] - p->obstacles[i][0];
                float oy = state[1] - p->obstacles[i][1];
                float r = p->obstacles[i][2];
                float d2 = ox * ox + oy * oy;
                if (i < n && d2 < r * r) pen += r - mppi_sqrtf(d2);
            }
        }
    }
#else
    // The definition, slot by slot in the same order.
    for (uint32_t i = 0; i < n; ++i) {
        const float ox = state[0] - p->obstacles[i][0];
        const float oy = state[1] - p->obstacles[i][1];
        const float r = p->obstacles[i][2];
        const float d2 = ox * ox + oy * oy;
        if (r > 0.0f && d2 < r * r) pen += r - mppi_sqrtf(d2);
    }
#endif
    return cost + p->obstacle_weight * pen;
}

MPPI_CALLABLE float mppi_terminal_cost(const float* state, const uint8_t* model_params,
                                       const uint8_t* cost_params_raw) {
    const UnicycleCost* p = reinterpret_cast<const UnicycleCost*>(cost_params_raw);
    float dx = state[0] - p->x_goal;
    float dy = state[1] - p->y_goal;
    return p->pos_weight * (dx * dx + dy * dy) * 10.0f;
}

MPPI_CALLABLE float unicycle_terminal_cost(const float* state, const uint8_t* model_params,
                                           const uint8_t* cost_params_raw) {
    return mppi_terminal_cost(state, model_params, cost_params_raw);
}
