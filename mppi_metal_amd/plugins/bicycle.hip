// Dynamic bicycle with a linear tyre model, for the multi-agent pursuit and
// learned-surrogate workloads.  No reference counterpart: the reference's
// batch example uses a kinematic unicycle (examples/nonholonomic_vehicle_batch)
// and a pursuit model with adversaries (examples/wez_environment_batch); this
// model replaces them as the benchmark workload names.
//
// State   [x, y, psi, vx, vy, r]   (position, heading, body velocities, yaw rate)
// Control [steer, throttle]        (front wheel angle [rad], longitudinal accel [m/s^2])
//
// Continuous dynamics, body frame:
//   x'   = vx cos(psi) - vy sin(psi)        y' = vx sin(psi) + vy cos(psi)
//   psi' = r
//   vx'  = throttle - drag * vx + r * vy
//   vy'  = (Fyf + Fyr) / m - r * vx
//   r'   = (lf * Fyf - lr * Fyr) / Iz
//   Fyf  = Cf * (steer - (vy + lf r) / u),  Fyr = Cr * (-(vy - lr r) / u),
//   u    = max(vx, v_min)  (linearised slip angles)
// integrated with `substeps` explicit Euler steps of dt / substeps; vx is
// clamped to [0, v_max].
//
// Stage cost (pursuit of a moving target among circular threat zones):
//   w_pos |p - g(t)|^2 + w_u |u|^2 + w_zone * sum_i max(0, r_i - |p - c_i|)
//   with the target moving at constant velocity, g(t) = g0 + v_g * (step + t + 1) * dt
//   (the stage cost sees the state after the step, hence t + 1).
// Terminal cost: w_term |p - g(H)|^2.
#include "mppi_core.h"

#define BICYCLE_MAX_ZONES 32

struct BicycleModel {
    float mass, yaw_inertia, lf, lr, cf, cr, drag, v_min, v_max, dt;
    uint32_t substeps;
    float pad;
};

struct BicycleCost {
    float target_x, target_y, target_vx, target_vy;   // target at closed-loop step 0
    float w_pos, w_control, w_zone, w_terminal;
    float dt;                                         // time per MPPI step
    uint32_t n_zones;
    float pad[2];
    float zones[BICYCLE_MAX_ZONES][3];                // (cx, cy, radius)
};

MPPI_CALLABLE void bicycle_dynamics(float* s, const float* u, const uint8_t* model_raw,
                                    mppi_rng rng) {
    const BicycleModel* p = reinterpret_cast<const BicycleModel*>(model_raw);
    const uint32_t n = p->substeps > 0 ? p->substeps : 1u;
    const float h = p->dt / float(n);
    const float steer = u[0], acc = u[1];
    // Reciprocals of the parameters: they do not change along a rollout, so
    // the compiler computes them once per kernel rather than dividing twice
    // per substep.
    const float inv_mass = 1.0f / p->mass, inv_iz = 1.0f / p->yaw_inertia;
    for (uint32_t i = 0; i < n; ++i) {
        float sp, cp;
        mppi_sincosf(s[2], &sp, &cp);
        const float vx = s[3], vy = s[4], r = s[5];
        const float uu = vx > p->v_min ? vx : p->v_min;
        const float inv_u = 1.0f / uu;
        const float fyf = p->cf * (steer - (vy + p->lf * r) * inv_u);
        const float fyr = p->cr * (-(vy - p->lr * r) * inv_u);
        const float dx = vx * cp - vy * sp;
        const float dy = vx * sp + vy * cp;
        const float dvx = acc - p->drag * vx + r * vy;
        const float dvy = (fyf + fyr) * inv_mass - r * vx;
        const float dr = (p->lf * fyf - p->lr * fyr) * inv_iz;
        s[0] += h * dx;
        s[1] += h * dy;
        s[2] += h * r;
        s[3] += h * dvx;
        s[4] += h * dvy;
        s[5] += h * dr;
        if (s[3] < 0.0f) s[3] = 0.0f;
        if (s[3] > p->v_max) s[3] = p->v_max;
    }
}

MPPI_CALLABLE float bicycle_pursuit_cost(const float* s, const float* u, const uint8_t* model,
                                         const uint8_t* cost_raw, mppi_time tm) {
    const BicycleCost* c = reinterpret_cast<const BicycleCost*>(cost_raw);
    const float tau = float(tm.step + tm.t + 1u) * c->dt;
    const float ex = s[0] - (c->target_x + c->target_vx * tau);
    const float ey = s[1] - (c->target_y + c->target_vy * tau);
    float cost = c->w_pos * (ex * ex + ey * ey);
    cost += c->w_control * (u[0] * u[0] + u[1] * u[1]);
    float pen = 0.0f;
    const uint32_t nz = c->n_zones < BICYCLE_MAX_ZONES ? c->n_zones : BICYCLE_MAX_ZONES;
    // The square root is only needed inside a zone.  A first pass tests the
    // square |ox|, |oy| < r around each zone (inside the circle in the
    // rounded arithmetic below implies inside the square, rounding being
    // monotone), eight zones per group so that their parameter loads issue
    // together; slots in [nz, nz rounded up to 8) may hold anything and only
    // send a wavefront to the exact pass.  The exact pass, taken only by
    // wavefronts with some sample in some square, visits the zones in order
    // and measures the distance only for zones whose square holds some sample
    // of the wavefront.
    const uint32_t nz8 = (nz + 7u) & ~7u;
    float margin = 1.0f;        // min over zones of max(|ox|, |oy|) - r
    for (uint32_t i = 0; i < nz8; i += 8) {
        MPPI_UNROLL
        for (uint32_t j = 0; j < 8; ++j) {
            const float ox = s[0] - c->zones[i + j][0];
            const float oy = s[1] - c->zones[i + j][1];
            margin = __builtin_fminf(margin, __builtin_fmaxf(__builtin_fabsf(ox), __builtin_fabsf(oy)) -
                                                 c->zones[i + j][2]);
        }
    }
    if (__builtin_expect(mppi_any(margin < 0.0f), 0)) {
        for (uint32_t i = 0; i < nz; ++i) {
            const float ox = s[0] - c->zones[i][0];
            const float oy = s[1] - c->zones[i][1];
            const float rz = c->zones[i][2];
            if (mppi_any(__builtin_fmaxf(__builtin_fabsf(ox), __builtin_fabsf(oy)) < rz)) {
                const float d2 = ox * ox + oy * oy;
                if (d2 < rz * rz) pen += rz - mppi_sqrtf(d2);
            }
        }
    }
    return cost + c->w_zone * pen;
}

MPPI_CALLABLE float bicycle_terminal_cost(const float* s, const uint8_t* model,
                                          const uint8_t* cost_raw, mppi_time tm) {
    const BicycleCost* c = reinterpret_cast<const BicycleCost*>(cost_raw);
    const float tau = float(tm.step + tm.t) * c->dt;
    const float ex = s[0] - (c->target_x + c->target_vx * tau);
    const float ey = s[1] - (c->target_y + c->target_vy * tau);
    return c->w_terminal * (ex * ex + ey * ey);
}
