// Quadrotor with thrust and body-rate commands, for the serving workload.
// No reference counterpart (the reference bundles a double integrator, a
// unicycle and a pursuit model).
//
// State   [px, py, pz, vx, vy, vz, roll, pitch, yaw, p, q, r]   (world frame
//         position and velocity, ZYX Euler angles, body rates)
// Control [thrust, p_cmd, q_cmd, r_cmd]  (collective thrust per unit mass
//         [m/s^2], commanded body rates [rad/s])
//
//   p'   = v
//   v'   = thrust * R(roll, pitch, yaw) e3 - g e3 - drag * v
//   eul' = W(roll, pitch) * omega      (ZYX Euler-angle kinematics)
//   omega' = (omega_cmd - omega) / tau  (first-order rate loop)
// Explicit Euler with `substeps` steps of dt / substeps; pitch is kept within
// +-1.4 rad so W stays finite.
//
// Stage cost: w_pos |p - p_goal|^2 + w_vel |v|^2 + w_att (roll^2 + pitch^2)
//             + w_thrust (thrust - g)^2 + w_rate |omega_cmd|^2
// Terminal cost: w_terminal |p - p_goal|^2 + w_vel |v|^2.
#include "mppi_core.h"

struct QuadModel {
    float gravity, drag, tau, dt;
    uint32_t substeps;
    float pad[3];
};

struct QuadCost {
    float goal[3];
    float w_pos, w_vel, w_att, w_thrust, w_rate, w_terminal;
    float pad[3];
};

MPPI_CALLABLE void quad_dynamics(float* s, const float* u, const uint8_t* model_raw, mppi_rng rng) {
    const QuadModel* p = reinterpret_cast<const QuadModel*>(model_raw);
    const uint32_t n = p->substeps > 0 ? p->substeps : 1u;
    const float h = p->dt / float(n);
    for (uint32_t i = 0; i < n; ++i) {
        float sr, cr, sp, cp, sy, cy;
        mppi_sincosf(s[6], &sr, &cr);
        mppi_sincosf(s[7], &sp, &cp);
        mppi_sincosf(s[8], &sy, &cy);
        const float tp = sp / cp;
        // thrust direction R e3
        const float ax = u[0] * (cr * sp * cy + sr * sy) - p->drag * s[3];
        const float ay = u[0] * (cr * sp * sy - sr * cy) - p->drag * s[4];
        const float az = u[0] * (cr * cp) - p->gravity - p->drag * s[5];
        const float wx = s[9], wy = s[10], wz = s[11];
        const float droll = wx + sr * tp * wy + cr * tp * wz;
        const float dpitch = cr * wy - sr * wz;
        const float dyaw = (sr * wy + cr * wz) / cp;
        s[0] += h * s[3];
        s[1] += h * s[4];
        s[2] += h * s[5];
        s[3] += h * ax;
        s[4] += h * ay;
        s[5] += h * az;
        s[6] += h * droll;
        s[7] += h * dpitch;
        s[8] += h * dyaw;
        s[9] += h * (u[1] - wx) / p->tau;
        s[10] += h * (u[2] - wy) / p->tau;
        s[11] += h * (u[3] - wz) / p->tau;
        if (s[7] > 1.4f) s[7] = 1.4f;
        if (s[7] < -1.4f) s[7] = -1.4f;
    }
}

MPPI_CALLABLE float quad_stage_cost(const float* s, const float* u, const uint8_t* model_raw,
                                    const uint8_t* cost_raw) {
    const QuadModel* m = reinterpret_cast<const QuadModel*>(model_raw);
    const QuadCost* c = reinterpret_cast<const QuadCost*>(cost_raw);
    const float ex = s[0] - c->goal[0], ey = s[1] - c->goal[1], ez = s[2] - c->goal[2];
    const float dthr = u[0] - m->gravity;
    float cost = c->w_pos * (ex * ex + ey * ey + ez * ez);
    cost += c->w_vel * (s[3] * s[3] + s[4] * s[4] + s[5] * s[5]);
    cost += c->w_att * (s[6] * s[6] + s[7] * s[7]);
    cost += c->w_thrust * dthr * dthr;
    cost += c->w_rate * (u[1] * u[1] + u[2] * u[2] + u[3] * u[3]);
    return cost;
}

MPPI_CALLABLE float quad_terminal_cost(const float* s, const uint8_t* model_raw, const uint8_t* cost_raw) {
    const QuadCost* c = reinterpret_cast<const QuadCost*>(cost_raw);
    const float ex = s[0] - c->goal[0], ey = s[1] - c->goal[1], ez = s[2] - c->goal[2];
    return c->w_terminal * (ex * ex + ey * ey + ez * ez) +
           c->w_vel * (s[3] * s[3] + s[4] * s[4] + s[5] * s[5]);
}
