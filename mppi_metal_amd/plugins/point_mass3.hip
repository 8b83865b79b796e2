// Three-axis point mass with process noise: a three-control model (it
// exercises noise blocks that straddle timesteps, C = 3) whose dynamics draws
// from its random handle.  No reference counterpart: the reference's one
// stochastic plugin is the WEZ pursuit model
// (examples/wez_environment_batch/wez_environment.metal:166-168).
//
// State   [px, py, pz, vx, vy, vz];   Control [ax, ay, az]
//   p += dt v;   v += dt a + process_std * w,   w ~ N(0, I) from draw 0 of
//   the step's random handle (the same counter on the GPU and the oracle)
// Stage cost:    w_pos |p - goal|^2 + w_vel |v|^2 + w_control |a|^2
// Terminal cost: w_terminal |p - goal|^2
#include "mppi_core.h"

struct PointMassModel {
    float dt, process_std;
    float pad[2];
};

struct PointMassCost {
    float goal[3];
    float w_pos, w_vel, w_control, w_terminal;
    float pad;
};

MPPI_CALLABLE void pm3_dynamics(float* s, const float* u, const uint8_t* model_raw, mppi_rng rng) {
    const PointMassModel* p = reinterpret_cast<const PointMassModel*>(model_raw);
    const mppi_f32x4 w = mppi_rng_normal4(rng, 0);
    const float dt = p->dt;
    s[0] += dt * s[3];
    s[1] += dt * s[4];
    s[2] += dt * s[5];
    s[3] += dt * u[0] + p->process_std * w.x;
    s[4] += dt * u[1] + p->process_std * w.y;
    s[5] += dt * u[2] + p->process_std * w.z;
}

MPPI_CALLABLE float pm3_stage_cost(const float* s, const float* u, const uint8_t* model_raw,
                                   const uint8_t* cost_raw) {
    const PointMassCost* c = reinterpret_cast<const PointMassCost*>(cost_raw);
    const float ex = s[0] - c->goal[0], ey = s[1] - c->goal[1], ez = s[2] - c->goal[2];
    return c->w_pos * (ex * ex + ey * ey + ez * ez) + c->w_vel * (s[3] * s[3] + s[4] * s[4] + s[5] * s[5]) +
           c->w_control * (u[0] * u[0] + u[1] * u[1] + u[2] * u[2]);
}

MPPI_CALLABLE float pm3_terminal_cost(const float* s, const uint8_t* model_raw, const uint8_t* cost_raw) {
    const PointMassCost* c = reinterpret_cast<const PointMassCost*>(cost_raw);
    const float ex = s[0] - c->goal[0], ey = s[1] - c->goal[1], ez = s[2] - c->goal[2];
    return c->w_terminal * (ex * ex + ey * ey + ez * ez);
}
