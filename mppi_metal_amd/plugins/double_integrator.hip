// Planar double integrator, the model of the reference's test plugin
// (shaders/test_user_dynamics.metal:8-43).
// State [px, py, vx, vy], control [ax, ay], dt = 0.1.
// Stage and terminal cost: px^2 + py^2.
#include "mppi_core.h"

MPPI_CALLABLE void mppi_dynamics(uint8_t* state_raw, const float* control,
                                 const uint8_t* model_params, mppi_rng rng) {
    float* state = reinterpret_cast<float*>(state_raw);
    const float dt = 0.1f;
    state[0] += state[2] * dt;
    state[1] += state[3] * dt;
    state[2] += control[0] * dt;
    state[3] += control[1] * dt;
}

MPPI_CALLABLE float mppi_stage_cost(const float* state, const float* control,
                                    const uint8_t* model_params, const uint8_t* cost_params) {
    return state[0] * state[0] + state[1] * state[1];
}

MPPI_CALLABLE float mppi_terminal_cost(const float* state, const uint8_t* model_params,
                                       const uint8_t* cost_params) {
    return state[0] * state[0] + state[1] * state[1];
}
