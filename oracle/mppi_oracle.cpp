// mppi_oracle.cpp — TEST INFRASTRUCTURE: the CPU oracle of the MPPI step.
//
// Only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg load
// this library, and only as the checker or the timed CPU baseline.  The
// product (mppi_metal_amd/libmppi_hip.so) never calls it.
//
// It restates, in scalar C++, the reference's MPPI iteration:
//   rollout        shaders/mppi_rollout.metal:36-96 (one sample: noise, clamp,
//                  dynamics, stage cost on the post-dynamics state, terminal)
//   weights        shaders/mppi_reduce.metal:14-121 (min, exp(-(c-min)/lambda),
//                  normalise, mean cost, ESS = 1/sum w^2)
//   reduce         shaders/mppi_reduce.metal:130-168 (u = clamp(u_nom +
//                  sum_k w_k eps_k), k ascending, unclamped noise)
//   propagate/shift shaders/mppi_reduce.metal:175-231 and the batch variant
//                  :413-488 (history, plant step, hold-last shift)
//   driver loop    src/mppi_driver.cpp:326-339 (host warm start), 343-421,
//                  423-491
// with the noise stream of the new build (Philox4x32-10, counter
// (block, sample, step, tag<<28 | agent), 64-bit key = seed) in place of the
// reference's flat Philox2x32 counter (shaders/mppi_rollout.metal:73-87).
//
// Two precisions:
//   FP32  the plugin callables are the bundled plugin sources compiled for
//         the host, the transcendentals are include/mppi_hip/mppi_math.h;
//         every operation is in the order the reference states it, so sample
//         costs are expected to match the GPU bit for bit.
//   FP64  the same algorithm in double with libm transcendentals and models
//         restated by hand in double.  It shares nothing with the GPU path
//         but the integer Philox stream, and is the yardstick each FP32
//         result's error is stated against.
// The Philox generator here is this directory's own (oracle_philox.h).
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

#include "mppi_core.h"
#include "oracle_philox.h"

namespace plug_di {
#include "../mppi_metal_amd/plugins/double_integrator.hip"
}
namespace plug_uni {
#include "../mppi_metal_amd/plugins/unicycle.hip"
}
namespace plug_bic {
#include "../mppi_metal_amd/plugins/bicycle.hip"
}
namespace plug_quad {
#include "../mppi_metal_amd/plugins/quadrotor.hip"
}
namespace plug_pm3 {
#include "../mppi_metal_amd/plugins/point_mass3.hip"
}

namespace {

// ---------------------------------------------------------------------------
// Numeric traits: the FP32 path uses the shared elementary functions, the FP64
// path libm.
// ---------------------------------------------------------------------------
template <typename T> struct Num;
template <> struct Num<float> {
    static float log(float x) { return mppi_logf(x); }
    static float sqrt(float x) { return mppi_sqrtf(x); }
    static float exp(float x) { return mppi_expf(x); }
    static void sincospi(float x, float* s, float* c) { mppi_sincospif(x, s, c); }
};
template <> struct Num<double> {
    static double log(double x) { return std::log(x); }
    static double sqrt(double x) { return std::sqrt(x); }
    static double exp(double x) { return std::exp(x); }
    static void sincospi(double x, double* s, double* c) {
        *s = std::sin(M_PI * x);
        *c = std::cos(M_PI * x);
    }
};

template <typename T> T u01(uint32_t x) { return T((x >> 8) | 1u) * T(5.9604644775390625e-8); }

// Four standard normals of counter block (block, sample, step, tag<<28|agent).
template <typename T>
void normals4(uint32_t block, uint32_t sample, uint32_t step, uint32_t tag, uint32_t agent,
              uint64_t seed, T z[4]) {
    const uint32_t ctr[4] = {block, sample, step, (tag << 28) | (agent & 0x0fffffffu)};
    const uint32_t key[2] = {uint32_t(seed), uint32_t(seed >> 32)};
    uint32_t w[4];
    oracle_philox4x32_10(ctr, key, w);
    for (int p = 0; p < 2; ++p) {
        T u1 = u01<T>(w[2 * p]), u2 = u01<T>(w[2 * p + 1]);
        T r = Num<T>::sqrt(T(-2) * Num<T>::log(u1));
        T s, c;
        Num<T>::sincospi(T(2) * u2, &s, &c);
        z[2 * p] = r * c;
        z[2 * p + 1] = r * s;
    }
}

// clamp(x, lo, hi) = min(max(x, lo), hi) as Metal's clamp
// (shaders/mppi_rollout.metal:79-86, shaders/mppi_reduce.metal:167): a NaN
// operand of max or min yields the other operand, -0 orders below +0 (the
// GPU's v_max_f32 / v_min_f32; mppi_maxf / mppi_minf).
template <typename T> T max_v(T a, T b) {
    if (a != a) return b;
    if (b != b) return a;
    if (a == b) return std::signbit(a) ? b : a;
    return a > b ? a : b;
}
template <typename T> T min_v(T a, T b) {
    if (a != a) return b;
    if (b != b) return a;
    if (a == b) return std::signbit(a) ? a : b;
    return a < b ? a : b;
}
template <typename T> T clampv(T x, T lo, T hi) { return min_v(max_v(x, lo), hi); }

// x rounded to the nearest binary16 value, ties to even (IEEE 754 binary16:
// 11 significant bits, quantum 2^-24 below 2^-14, overflow at 65520), by
// scaling to an integer multiple of the quantum; independent of the
// library's bit-level mppi_round_f16.
double round_f16(double x) {
    if (!std::isfinite(x) || x == 0.0) return x;
    const double ax = std::fabs(x);
    if (ax >= 65520.0) return std::copysign(INFINITY, x);
    int e;
    std::frexp(ax, &e);                         // ax in [2^(e-1), 2^e)
    const double quantum = std::ldexp(1.0, std::max(e - 11, -24));
    return std::copysign(std::nearbyint(ax / quantum) * quantum, x);
}

// ---------------------------------------------------------------------------
// Models.  FP32: the plugin sources.  FP64: hand restatements of the same
// formulas.
// ---------------------------------------------------------------------------
struct RngCtx {
    uint32_t t, sample, step, tag, agent;
    uint64_t seed;
};

template <typename T>
using DynFn = void (*)(T* x, const T* u, const uint8_t* model, const RngCtx& rng);
template <typename T>
using StageFn = T (*)(const T* x, const T* u, const uint8_t* model, const uint8_t* cost,
                      const mppi_time& tm);
template <typename T>
using TermFn = T (*)(const T* x, const uint8_t* model, const uint8_t* cost, const mppi_time& tm);

mppi_rng to_rng(const RngCtx& r) {
    return mppi_rng{r.t << 8, r.sample, r.step, (r.tag << 28) | (r.agent & 0x0fffffffu),
                    uint32_t(r.seed), uint32_t(r.seed >> 32)};
}

template <typename T>
T zero_stage(const T*, const T*, const uint8_t*, const uint8_t*, const mppi_time&) { return T(0); }
template <typename T>
T zero_term(const T*, const uint8_t*, const uint8_t*, const mppi_time&) { return T(0); }

// ---- double integrator (shaders/test_user_dynamics.metal:8-43) ----
void di_dyn_d(double* s, const double* u, const uint8_t*, const RngCtx&) {
    const double dt = double(0.1f);
    s[0] += s[2] * dt;
    s[1] += s[3] * dt;
    s[2] += u[0] * dt;
    s[3] += u[1] * dt;
}
double di_cost_d(const double* s, const double*, const uint8_t*, const uint8_t*, const mppi_time&) {
    return s[0] * s[0] + s[1] * s[1];
}
double di_term_d(const double* s, const uint8_t*, const uint8_t*, const mppi_time&) { return s[0] * s[0] + s[1] * s[1]; }

// ---- unicycle (examples/nonholonomic_vehicle/nonholonomic_vehicle.metal:24-91) ----
struct UniModel { float v_max, dt; };
struct UniCost {
    float x_goal, y_goal, goal_radius, pos_weight, heading_weight, control_weight;
    float obstacle_weight;
    uint32_t n_obstacles;
    float obstacles[16][3];
};
void uni_dyn_d(double* s, const double* u, const uint8_t* mp, const RngCtx&) {
    const UniModel* p = reinterpret_cast<const UniModel*>(mp);
    const double dt = p->dt, th = s[2], v = s[3];
    s[0] += v * std::cos(th) * dt;
    s[1] += v * std::sin(th) * dt;
    s[2] += u[1] * dt;
    s[3] += u[0] * dt;
    const double pi = double(3.1415926535f);
    double w = std::fmod(s[2] + pi, 2.0 * pi);
    if (w < 0) w += 2.0 * pi;
    s[2] = w - pi;
    if (s[3] > p->v_max) s[3] = p->v_max;
    if (s[3] < -p->v_max) s[3] = -p->v_max;
}
double uni_stage_d(const double* s, const double* u, const uint8_t*, const uint8_t* cp,
                   const mppi_time&) {
    const UniCost* p = reinterpret_cast<const UniCost*>(cp);
    double dx = s[0] - p->x_goal, dy = s[1] - p->y_goal;
    return p->pos_weight * (dx * dx + dy * dy) + p->control_weight * (u[0] * u[0] + u[1] * u[1]);
}
double uni_stage_obs_d(const double* s, const double* u, const uint8_t* mp, const uint8_t* cp,
                       const mppi_time& tm) {
    const UniCost* p = reinterpret_cast<const UniCost*>(cp);
    double c = uni_stage_d(s, u, mp, cp, tm), pen = 0.0;
    uint32_t n = std::min<uint32_t>(p->n_obstacles, 16);
    for (uint32_t i = 0; i < n; ++i) {
        double ox = s[0] - p->obstacles[i][0], oy = s[1] - p->obstacles[i][1];
        pen += std::max(0.0, double(p->obstacles[i][2]) - std::sqrt(ox * ox + oy * oy));
    }
    return c + p->obstacle_weight * pen;
}
double uni_term_d(const double* s, const uint8_t*, const uint8_t* cp, const mppi_time&) {
    const UniCost* p = reinterpret_cast<const UniCost*>(cp);
    double dx = s[0] - p->x_goal, dy = s[1] - p->y_goal;
    return p->pos_weight * (dx * dx + dy * dy) * 10.0;
}

// ---- three-axis point mass (mppi_metal_amd/plugins/point_mass3.hip) ----
struct Pm3Model { float dt, process_std, pad[2]; };
struct Pm3Cost { float goal[3], w_pos, w_vel, w_control, w_terminal, pad; };
void pm3_dyn_d(double* s, const double* u, const uint8_t* mp, const RngCtx& r) {
    const Pm3Model* p = reinterpret_cast<const Pm3Model*>(mp);
    double w[4];
    normals4<double>((r.t << 8) + 0u, r.sample, r.step, r.tag, r.agent, r.seed, w);
    const double dt = p->dt;
    for (int i = 0; i < 3; ++i) s[i] += dt * s[3 + i];
    for (int i = 0; i < 3; ++i) s[3 + i] += dt * u[i] + double(p->process_std) * w[i];
}
double pm3_stage_d(const double* s, const double* u, const uint8_t*, const uint8_t* cp, const mppi_time&) {
    const Pm3Cost* c = reinterpret_cast<const Pm3Cost*>(cp);
    double ep = 0, ev = 0, eu = 0;
    for (int i = 0; i < 3; ++i) {
        const double e = s[i] - c->goal[i];
        ep += e * e;
        ev += s[3 + i] * s[3 + i];
        eu += u[i] * u[i];
    }
    return c->w_pos * ep + c->w_vel * ev + c->w_control * eu;
}
double pm3_term_d(const double* s, const uint8_t*, const uint8_t* cp, const mppi_time&) {
    const Pm3Cost* c = reinterpret_cast<const Pm3Cost*>(cp);
    double ep = 0;
    for (int i = 0; i < 3; ++i) {
        const double e = s[i] - c->goal[i];
        ep += e * e;
    }
    return c->w_terminal * ep;
}

// ---- dynamic bicycle (mppi_metal_amd/plugins/bicycle.hip) ----
struct BicModel {
    float mass, yaw_inertia, lf, lr, cf, cr, drag, v_min, v_max, dt;
    uint32_t substeps;
    float pad;
};
struct BicCost {
    float target_x, target_y, target_vx, target_vy, w_pos, w_control, w_zone, w_terminal, dt;
    uint32_t n_zones;
    float pad[2];
    float zones[32][3];
};
void bic_dyn_d(double* s, const double* u, const uint8_t* mp, const RngCtx&) {
    const BicModel* p = reinterpret_cast<const BicModel*>(mp);
    const uint32_t n = p->substeps > 0 ? p->substeps : 1u;
    const double h = double(p->dt) / n;
    for (uint32_t i = 0; i < n; ++i) {
        const double vx = s[3], vy = s[4], r = s[5], sp = std::sin(s[2]), cp = std::cos(s[2]);
        const double uu = std::max(vx, double(p->v_min));
        const double fyf = p->cf * (u[0] - (vy + p->lf * r) / uu);
        const double fyr = p->cr * (-(vy - p->lr * r) / uu);
        const double d[6] = {vx * cp - vy * sp, vx * sp + vy * cp, r, u[1] - p->drag * vx + r * vy,
                             (fyf + fyr) / p->mass - r * vx, (p->lf * fyf - p->lr * fyr) / p->yaw_inertia};
        for (int j = 0; j < 6; ++j) s[j] += h * d[j];
        s[3] = std::min(std::max(s[3], 0.0), double(p->v_max));
    }
}
double bic_target_err2(const double* s, const BicCost* c, double tau) {
    const double ex = s[0] - (c->target_x + c->target_vx * tau);
    const double ey = s[1] - (c->target_y + c->target_vy * tau);
    return ex * ex + ey * ey;
}
double bic_stage_d(const double* s, const double* u, const uint8_t*, const uint8_t* cp,
                   const mppi_time& tm) {
    const BicCost* c = reinterpret_cast<const BicCost*>(cp);
    double cost = c->w_pos * bic_target_err2(s, c, double(tm.step + tm.t + 1) * c->dt) +
                  c->w_control * (u[0] * u[0] + u[1] * u[1]);
    double pen = 0.0;
    for (uint32_t i = 0; i < std::min<uint32_t>(c->n_zones, 32); ++i) {
        const double ox = s[0] - c->zones[i][0], oy = s[1] - c->zones[i][1];
        pen += std::max(0.0, double(c->zones[i][2]) - std::sqrt(ox * ox + oy * oy));
    }
    return cost + c->w_zone * pen;
}
double bic_term_d(const double* s, const uint8_t*, const uint8_t* cp, const mppi_time& tm) {
    const BicCost* c = reinterpret_cast<const BicCost*>(cp);
    return c->w_terminal * bic_target_err2(s, c, double(tm.step + tm.t) * c->dt);
}

// ---- quadrotor (mppi_metal_amd/plugins/quadrotor.hip) ----
struct QModel { float gravity, drag, tau, dt; uint32_t substeps; float pad[3]; };
struct QCost { float goal[3], w_pos, w_vel, w_att, w_thrust, w_rate, w_terminal, pad[3]; };
void quad_dyn_d(double* s, const double* u, const uint8_t* mp, const RngCtx&) {
    const QModel* p = reinterpret_cast<const QModel*>(mp);
    const uint32_t n = p->substeps > 0 ? p->substeps : 1u;
    const double h = double(p->dt) / n;
    for (uint32_t i = 0; i < n; ++i) {
        const double sr = std::sin(s[6]), cr = std::cos(s[6]), sp = std::sin(s[7]), cp = std::cos(s[7]);
        const double sy = std::sin(s[8]), cy = std::cos(s[8]), tp = sp / cp;
        const double wx = s[9], wy = s[10], wz = s[11];
        const double d[12] = {s[3], s[4], s[5],
                              u[0] * (cr * sp * cy + sr * sy) - p->drag * s[3],
                              u[0] * (cr * sp * sy - sr * cy) - p->drag * s[4],
                              u[0] * (cr * cp) - p->gravity - p->drag * s[5],
                              wx + sr * tp * wy + cr * tp * wz, cr * wy - sr * wz, (sr * wy + cr * wz) / cp,
                              (u[1] - wx) / p->tau, (u[2] - wy) / p->tau, (u[3] - wz) / p->tau};
        for (int j = 0; j < 12; ++j) s[j] += h * d[j];
        s[7] = std::min(std::max(s[7], -1.4), 1.4);
    }
}
double quad_stage_d(const double* s, const double* u, const uint8_t* mp, const uint8_t* cp,
                    const mppi_time&) {
    const QModel* m = reinterpret_cast<const QModel*>(mp);
    const QCost* c = reinterpret_cast<const QCost*>(cp);
    const double ex = s[0] - c->goal[0], ey = s[1] - c->goal[1], ez = s[2] - c->goal[2];
    const double dthr = u[0] - m->gravity;
    return c->w_pos * (ex * ex + ey * ey + ez * ez) + c->w_vel * (s[3] * s[3] + s[4] * s[4] + s[5] * s[5]) +
           c->w_att * (s[6] * s[6] + s[7] * s[7]) + c->w_thrust * dthr * dthr +
           c->w_rate * (u[1] * u[1] + u[2] * u[2] + u[3] * u[3]);
}
double quad_term_d(const double* s, const uint8_t*, const uint8_t* cp, const mppi_time&) {
    const QCost* c = reinterpret_cast<const QCost*>(cp);
    const double ex = s[0] - c->goal[0], ey = s[1] - c->goal[1], ez = s[2] - c->goal[2];
    return c->w_terminal * (ex * ex + ey * ey + ez * ez) + c->w_vel * (s[3] * s[3] + s[4] * s[4] + s[5] * s[5]);
}

struct Callable {
    const char* model;
    const char* name;
    int role;  // 0 dynamics, 1 stage, 2 terminal
    void* f32;
    void* f64;
};

#define FP(x) reinterpret_cast<void*>(x)
const Callable kCallables[] = {
    {"double_integrator", "mppi_dynamics", 0,
     FP(+[](float* x, const float* u, const uint8_t* m, const RngCtx& r) {
         plug_di::mppi_dynamics(mppi_state_ref{x}, u, m, to_rng(r));
     }),
     FP(&di_dyn_d)},
    {"double_integrator", "mppi_stage_cost", 1,
     FP(+[](const float* x, const float* u, const uint8_t* m, const uint8_t* c, const mppi_time&) {
         return plug_di::mppi_stage_cost(x, u, m, c);
     }),
     FP(&di_cost_d)},
    {"double_integrator", "mppi_terminal_cost", 2,
     FP(+[](const float* x, const uint8_t* m, const uint8_t* c, const mppi_time&) {
         return plug_di::mppi_terminal_cost(x, m, c);
     }),
     FP(&di_term_d)},
    {"unicycle", "mppi_dynamics", 0,
     FP(+[](float* x, const float* u, const uint8_t* m, const RngCtx& r) {
         plug_uni::mppi_dynamics(x, u, m, to_rng(r));
     }),
     FP(&uni_dyn_d)},
    {"unicycle", "mppi_stage_cost", 1,
     FP(+[](const float* x, const float* u, const uint8_t* m, const uint8_t* c, const mppi_time&) {
         return plug_uni::mppi_stage_cost(x, u, m, c);
     }),
     FP(&uni_stage_d)},
    {"unicycle", "unicycle_stage_cost_obstacles", 1,
     FP(+[](const float* x, const float* u, const uint8_t* m, const uint8_t* c, const mppi_time&) {
         return plug_uni::unicycle_stage_cost_obstacles(x, u, m, c);
     }),
     FP(&uni_stage_obs_d)},
    {"unicycle", "mppi_terminal_cost", 2,
     FP(+[](const float* x, const uint8_t* m, const uint8_t* c, const mppi_time&) {
         return plug_uni::mppi_terminal_cost(x, m, c);
     }),
     FP(&uni_term_d)},
    {"unicycle", "unicycle_terminal_cost", 2,
     FP(+[](const float* x, const uint8_t* m, const uint8_t* c, const mppi_time&) {
         return plug_uni::unicycle_terminal_cost(x, m, c);
     }),
     FP(&uni_term_d)},
    {"bicycle", "bicycle_dynamics", 0,
     FP(+[](float* x, const float* u, const uint8_t* m, const RngCtx& r) {
         plug_bic::bicycle_dynamics(x, u, m, to_rng(r));
     }),
     FP(&bic_dyn_d)},
    {"bicycle", "bicycle_pursuit_cost", 1,
     FP(+[](const float* x, const float* u, const uint8_t* m, const uint8_t* c, const mppi_time& tm) {
         return plug_bic::bicycle_pursuit_cost(x, u, m, c, tm);
     }),
     FP(&bic_stage_d)},
    {"bicycle", "bicycle_terminal_cost", 2,
     FP(+[](const float* x, const uint8_t* m, const uint8_t* c, const mppi_time& tm) {
         return plug_bic::bicycle_terminal_cost(x, m, c, tm);
     }),
     FP(&bic_term_d)},
    {"point_mass3", "pm3_dynamics", 0,
     FP(+[](float* x, const float* u, const uint8_t* m, const RngCtx& r) {
         plug_pm3::pm3_dynamics(x, u, m, to_rng(r));
     }),
     FP(&pm3_dyn_d)},
    {"point_mass3", "pm3_stage_cost", 1,
     FP(+[](const float* x, const float* u, const uint8_t* m, const uint8_t* c, const mppi_time&) {
         return plug_pm3::pm3_stage_cost(x, u, m, c);
     }),
     FP(&pm3_stage_d)},
    {"point_mass3", "pm3_terminal_cost", 2,
     FP(+[](const float* x, const uint8_t* m, const uint8_t* c, const mppi_time&) {
         return plug_pm3::pm3_terminal_cost(x, m, c);
     }),
     FP(&pm3_term_d)},
    {"quadrotor", "quad_dynamics", 0,
     FP(+[](float* x, const float* u, const uint8_t* m, const RngCtx& r) {
         plug_quad::quad_dynamics(x, u, m, to_rng(r));
     }),
     FP(&quad_dyn_d)},
    {"quadrotor", "quad_stage_cost", 1,
     FP(+[](const float* x, const float* u, const uint8_t* m, const uint8_t* c, const mppi_time&) {
         return plug_quad::quad_stage_cost(x, u, m, c);
     }),
     FP(&quad_stage_d)},
    {"quadrotor", "quad_terminal_cost", 2,
     FP(+[](const float* x, const uint8_t* m, const uint8_t* c, const mppi_time&) {
         return plug_quad::quad_terminal_cost(x, m, c);
     }),
     FP(&quad_term_d)},
};
#undef FP

const Callable* find_callable(const char* model, const char* name, int role) {
    for (const Callable& c : kCallables)
        if (c.role == role && std::strcmp(c.model, model) == 0 && std::strcmp(c.name, name) == 0)
            return &c;
    return nullptr;
}

}  // namespace

extern "C" {

// Problem description shared by all oracle entry points.
struct oracle_problem {
    uint32_t D, C, H, K;
    float lambda;
    uint32_t cov_full;        // 0: sigma = C std devs; 1: sigma = C*C covariance
    const float* sigma;
    const float* u_min;
    const float* u_max;
    uint64_t seed;
    const char* model;        // "double_integrator" | "unicycle"
    const char* dynamics;
    const char* stage_cost;   // NULL or "" -> zero cost
    const char* terminal_cost;
    const void* model_params;
    const void* cost_params;
    uint32_t noise_f16;       // 1: normals rounded to binary16 (the FP16 numeric mode)
};

// mlp_oracle.cpp
void oracle_mlp_forward1(const float* P, uint32_t in, uint32_t out, const float* x, float* y);

}  // extern "C"

namespace {

template <typename T> struct Resolved {
    DynFn<T> dyn = nullptr;
    StageFn<T> stage = nullptr;
    TermFn<T> term = nullptr;
    std::vector<T> scale;     // sigma or lower Cholesky factor
    std::vector<T> lo, hi;
    const uint8_t* model = nullptr;
    const uint8_t* cost = nullptr;
    bool ok = false;
};

// Lower Cholesky factor, computed in double and rounded to float exactly as
// the library's host code does before upload.
bool cholesky_f(const float* a, uint32_t n, std::vector<float>* out) {
    std::vector<double> L(size_t(n) * n, 0.0);
    for (uint32_t i = 0; i < n; ++i)
        for (uint32_t j = 0; j <= i; ++j) {
            double s = double(a[i * n + j]);
            for (uint32_t k = 0; k < j; ++k) s -= L[i * n + k] * L[j * n + k];
            if (i == j) {
                if (!(s > 0.0)) return false;
                L[i * n + i] = std::sqrt(s);
            } else {
                L[i * n + j] = s / L[j * n + j];
            }
        }
    out->assign(size_t(n) * n, 0.0f);
    for (size_t i = 0; i < L.size(); ++i) (*out)[i] = float(L[i]);
    return true;
}

template <typename T> Resolved<T> resolve(const oracle_problem* p) {
    Resolved<T> r;
    const bool f32 = sizeof(T) == 4;
    const Callable* d = find_callable(p->model, p->dynamics, 0);
    if (!d) return r;
    r.dyn = reinterpret_cast<DynFn<T>>(f32 ? d->f32 : d->f64);
    if (p->stage_cost && *p->stage_cost) {
        const Callable* s = find_callable(p->model, p->stage_cost, 1);
        if (!s) return r;
        r.stage = reinterpret_cast<StageFn<T>>(f32 ? s->f32 : s->f64);
    } else {
        r.stage = &zero_stage<T>;
    }
    if (p->terminal_cost && *p->terminal_cost) {
        const Callable* t = find_callable(p->model, p->terminal_cost, 2);
        if (!t) return r;
        r.term = reinterpret_cast<TermFn<T>>(f32 ? t->f32 : t->f64);
    } else {
        r.term = &zero_term<T>;
    }
    if (p->cov_full) {
        std::vector<float> L;
        if (!cholesky_f(p->sigma, p->C, &L)) return r;
        r.scale.assign(L.begin(), L.end());
    } else {
        r.scale.assign(p->sigma, p->sigma + p->C);
    }
    r.lo.assign(p->u_min, p->u_min + p->C);
    r.hi.assign(p->u_max, p->u_max + p->C);
    static const uint8_t zeros[256] = {0};
    r.model = p->model_params ? static_cast<const uint8_t*>(p->model_params) : zeros;
    r.cost = p->cost_params ? static_cast<const uint8_t*>(p->cost_params) : zeros;
    r.ok = true;
    return r;
}

// Noise eps(t, .) of sample k: unclamped, already scaled (sigma z or L z).
template <typename T>
void noise_eps(const oracle_problem* p, const Resolved<T>& r, const T* z_all, uint32_t t, T* eps) {
    const uint32_t C = p->C;
    const T* z = z_all + size_t(t) * C;
    if (p->cov_full) {
        for (uint32_t i = 0; i < C; ++i) {
            T e = T(0);
            for (uint32_t j = 0; j <= i; ++j) {
                if (sizeof(T) == 4) e = T(mppi_fma(float(r.scale[i * C + j]), float(z[j]), float(e)));
                else e = std::fma(r.scale[i * C + j], z[j], e);
            }
            eps[i] = e;
        }
    } else {
        for (uint32_t i = 0; i < C; ++i) eps[i] = z[i] * r.scale[i];
    }
}

// All H*C normals of sample k (normal n = t*C + d from block n/4, lane n%4).
template <typename T>
void sample_normals(const oracle_problem* p, uint32_t k, uint32_t step, uint32_t agent, T* z) {
    const uint32_t HC = p->H * p->C;
    T blk[4];
    for (uint32_t n = 0; n < HC; ++n) {
        if ((n & 3) == 0) normals4<T>(n >> 2, k, step, MPPI_TAG_CONTROL_NOISE, agent, p->seed, blk);
        z[n] = p->noise_f16 ? T(round_f16(double(blk[n & 3]))) : blk[n & 3];
    }
}

// Learned-dynamics surrogate in place of the plugin dynamics (no reference
// counterpart): parameters and the 80-float normalisation block (input
// centre, input 1/half-range, input period, output mean, output sd; 16 each).
struct Surrogate {
    const float* params;
    const float* norm;
};

// x <- x + delta(x, u) as the GPU's surrogate rollout computes it.
void surrogate_dynamics(const Surrogate& s, uint32_t D, uint32_t C, float* x, const float* u) {
    const float* center = s.norm;
    const float* inv_half = s.norm + 16;
    const float* period = s.norm + 32;
    const float* mu = s.norm + 48;
    const float* sd = s.norm + 64;
    float xin[16], y[16];
    for (uint32_t k = 0; k < D + C; ++k) {
        float d = (k < D ? x[k] : u[k - D]) - center[k];
        if (period[k] > 0.0f) d -= period[k] * std::nearbyint(d * (1.0f / period[k]));
        xin[k] = d * inv_half[k];
    }
    oracle_mlp_forward1(s.params, D + C, D, xin, y);
    for (uint32_t k = 0; k < D; ++k) x[k] += y[k] * sd[k] + mu[k];
}

// Rollout of samples [k0, k1) (mppi_rollout.metal:36-96); with `sur` the
// network replaces the plugin dynamics (FP32 only).
template <typename T>
void rollout_range(const oracle_problem* p, const Resolved<T>& r, uint32_t step, uint32_t agent,
                   const float* x0, const T* u_nom, uint32_t k0, uint32_t k1, T* costs,
                   const Surrogate* sur = nullptr) {
    const uint32_t D = p->D, C = p->C, H = p->H;
    std::vector<T> z(size_t(H) * C), x(D), u(C), eps(C);
    for (uint32_t k = k0; k < k1; ++k) {
        sample_normals<T>(p, k, step, agent, z.data());
        for (uint32_t i = 0; i < D; ++i) x[i] = T(x0[i]);
        T total = T(0);
        for (uint32_t t = 0; t < H; ++t) {
            noise_eps<T>(p, r, z.data(), t, eps.data());
            for (uint32_t i = 0; i < C; ++i)
                u[i] = clampv<T>(u_nom[t * C + i] + eps[i], r.lo[i], r.hi[i]);
            if (sur) {
                if constexpr (sizeof(T) == 4) surrogate_dynamics(*sur, D, C, x.data(), u.data());
            } else {
                RngCtx rc{t, k, step, MPPI_TAG_DYNAMICS, agent, p->seed};
                r.dyn(x.data(), u.data(), r.model, rc);
            }
            total += r.stage(x.data(), u.data(), r.model, r.cost, mppi_time{t, step});
        }
        total += r.term(x.data(), r.model, r.cost, mppi_time{H, step});
        costs[k] = total;
    }
}

template <typename T>
void rollout(const oracle_problem* p, const Resolved<T>& r, uint32_t step, uint32_t agent,
             const float* x0, const T* u_nom, T* costs, int threads) {
    const uint32_t K = p->K;
    if (threads <= 1 || K < 256) {
        rollout_range<T>(p, r, step, agent, x0, u_nom, 0, K, costs);
        return;
    }
    std::vector<std::thread> pool;
    const uint32_t per = (K + threads - 1) / threads;
    for (int i = 0; i < threads; ++i) {
        uint32_t a = std::min<uint32_t>(K, i * per), b = std::min<uint32_t>(K, a + per);
        if (a >= b) break;
        pool.emplace_back([=, &r] { rollout_range<T>(p, r, step, agent, x0, u_nom, a, b, costs); });
    }
    for (auto& th : pool) th.join();
}

bool finite(double x) { return std::isfinite(x); }

// Softmax weights and diagnostics (mppi_reduce.metal:14-121).  Non-finite
// costs get weight 0; if every cost is non-finite the weights are uniform.
template <typename T>
void weights(const oracle_problem* p, const T* costs, T* w, T diag[3]) {
    const uint32_t K = p->K;
    T m = T(INFINITY);
    for (uint32_t k = 0; k < K; ++k) {
        T c = finite(double(costs[k])) ? costs[k] : T(INFINITY);
        m = c < m ? c : m;
    }
    T sum_w = T(0), sum_c = T(0);
    const bool uniform = !finite(double(m));
    for (uint32_t k = 0; k < K; ++k) {
        T e;
        if (uniform) e = T(1);
        else e = finite(double(costs[k])) ? Num<T>::exp(-(costs[k] - m) / T(p->lambda)) : T(0);
        w[k] = e;
        sum_w += e;
        sum_c += costs[k];
    }
    T sum_w2 = T(0);
    for (uint32_t k = 0; k < K; ++k) {
        w[k] = w[k] / sum_w;
        sum_w2 += w[k] * w[k];
    }
    diag[0] = m;
    diag[1] = sum_c / T(K);
    diag[2] = T(1) / sum_w2;
}

// Weighted-noise average (mppi_reduce.metal:130-168): for each (t, d),
// u = clamp(u_nom + sum_{k=0..K-1} w_k * eps_k(t, d)), k ascending.  With
// threads > 1 the samples are split into contiguous ranges whose partial sums
// are added in range order (used by the CPU baseline timing).
template <typename T>
void reduce_range(const oracle_problem* p, const Resolved<T>& r, uint32_t step, uint32_t agent,
                  const T* w, uint32_t k0, uint32_t k1, T* acc) {
    const uint32_t C = p->C, H = p->H, HC = H * C;
    std::vector<T> z(HC), eps(C);
    for (uint32_t k = k0; k < k1; ++k) {
        sample_normals<T>(p, k, step, agent, z.data());
        for (uint32_t t = 0; t < H; ++t) {
            noise_eps<T>(p, r, z.data(), t, eps.data());
            for (uint32_t i = 0; i < C; ++i) acc[t * C + i] += w[k] * eps[i];
        }
    }
}

template <typename T>
void reduce(const oracle_problem* p, const Resolved<T>& r, uint32_t step, uint32_t agent,
            const T* w, const T* u_nom, T* u_out, int threads = 1) {
    const uint32_t C = p->C, H = p->H, HC = H * C, K = p->K;
    std::vector<T> acc(HC, T(0));
    if (threads <= 1 || K < 1024) {
        reduce_range<T>(p, r, step, agent, w, 0, K, acc.data());
    } else {
        std::vector<std::vector<T>> part(threads, std::vector<T>(HC, T(0)));
        std::vector<std::thread> pool;
        const uint32_t per = (K + threads - 1) / threads;
        for (int i = 0; i < threads; ++i) {
            uint32_t a = std::min<uint32_t>(K, i * per), b = std::min<uint32_t>(K, a + per);
            pool.emplace_back([=, &r, &part] { reduce_range<T>(p, r, step, agent, w, a, b, part[i].data()); });
        }
        for (auto& th : pool) th.join();
        for (int i = 0; i < threads; ++i)
            for (uint32_t n = 0; n < HC; ++n) acc[n] += part[i][n];
    }
    for (uint32_t n = 0; n < HC; ++n) {
        const uint32_t i = n % C;
        u_out[n] = clampv<T>(u_nom[n] + acc[n], r.lo[i], r.hi[i]);
    }
}

template <typename T>
int step_impl(const oracle_problem* p, uint32_t step, uint32_t agent, const float* x0,
              const T* u_nom, T* u_out, T* diag3, T* costs_out, T* weights_out, int threads) {
    Resolved<T> r = resolve<T>(p);
    if (!r.ok) return -1;
    std::vector<T> costs(p->K), w(p->K);
    T diag[3];
    rollout<T>(p, r, step, agent, x0, u_nom, costs.data(), threads);
    weights<T>(p, costs.data(), w.data(), diag);
    reduce<T>(p, r, step, agent, w.data(), u_nom, u_out, threads);
    if (diag3) std::copy(diag, diag + 3, diag3);
    if (costs_out) std::copy(costs.begin(), costs.end(), costs_out);
    if (weights_out) std::copy(w.begin(), w.end(), weights_out);
    return 0;
}

// Closed loop for n_agents agents over T steps: MPPI iteration, history,
// plant step with the dynamics (counter (0, 0xffffffff, step, PLANT|agent)),
// hold-last shift (mppi_reduce.metal:175-231, :413-488).
template <typename T>
int simulate_impl(const oracle_problem* p, uint32_t T_steps, uint32_t start_step, uint32_t n_agents,
                  const float* x0_packed, const float* u_nom_packed, float* states_out,
                  float* controls_out, float* costs_out, float* terminal_out, int threads) {
    Resolved<T> r = resolve<T>(p);
    if (!r.ok) return -1;
    const uint32_t D = p->D, C = p->C, H = p->H, HC = H * C;
    for (uint32_t a = 0; a < n_agents; ++a) {
        std::vector<float> x(x0_packed + size_t(a) * D, x0_packed + size_t(a + 1) * D);
        std::vector<T> un(HC, T(0)), uo(HC), costs(p->K), w(p->K), xs(D);
        if (u_nom_packed)
            for (uint32_t n = 0; n < HC; ++n) un[n] = T(u_nom_packed[size_t(a) * HC + n]);
        for (uint32_t s = 0; s < T_steps; ++s) {
            const uint32_t step = start_step + s;
            T diag[3];
            rollout<T>(p, r, step, a, x.data(), un.data(), costs.data(), threads);
            weights<T>(p, costs.data(), w.data(), diag);
            reduce<T>(p, r, step, a, w.data(), un.data(), uo.data(), threads);
            const size_t row = size_t(s) * n_agents + a;
            if (states_out) for (uint32_t i = 0; i < D; ++i) states_out[row * D + i] = x[i];
            if (controls_out) for (uint32_t i = 0; i < C; ++i) controls_out[row * C + i] = float(uo[i]);
            if (costs_out) costs_out[row] = float(diag[0]);
            for (uint32_t i = 0; i < D; ++i) xs[i] = T(x[i]);
            RngCtx rc{0, 0xffffffffu, step, MPPI_TAG_PLANT, a, p->seed};
            r.dyn(xs.data(), uo.data(), r.model, rc);
            for (uint32_t i = 0; i < D; ++i) x[i] = float(xs[i]);
            for (uint32_t n = 0; n < HC; ++n) {
                const uint32_t t = n / C, i = n % C;
                un[n] = uo[(t + 1 < H ? t + 1 : H - 1) * C + i];
            }
        }
        if (terminal_out) for (uint32_t i = 0; i < D; ++i) terminal_out[size_t(a) * D + i] = x[i];
    }
    return 0;
}

}  // namespace

extern "C" {

void oracle_philox4x32(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
    oracle_philox4x32_10(ctr, key, out);
}
void oracle_philox2x32(const uint32_t ctr[2], uint32_t key, uint32_t out[2]) {
    oracle_philox2x32_10(ctr, key, out);
}
void oracle_normals4(uint32_t block, uint32_t sample, uint32_t step, uint32_t tag, uint32_t agent,
                     uint64_t seed, float out[4]) {
    normals4<float>(block, sample, step, tag, agent, seed, out);
}
void oracle_normals4_64(uint32_t block, uint32_t sample, uint32_t step, uint32_t tag,
                        uint32_t agent, uint64_t seed, double out[4]) {
    normals4<double>(block, sample, step, tag, agent, seed, out);
}

// The shared single-precision elementary functions, compiled for the host,
// so tests can bound their error against FP64 libm.
// fn: 0 exp, 1 log, 2 sin, 3 cos, 4 sin(pi x), 5 cos(pi x), 6 tanh, 7 sqrt
void oracle_mathf(int fn, const float* x, float* y, size_t n) {
    for (size_t i = 0; i < n; ++i) {
        float s, c;
        switch (fn) {
            case 0: y[i] = mppi_expf(x[i]); break;
            case 1: y[i] = mppi_logf(x[i]); break;
            case 2: y[i] = mppi_sinf(x[i]); break;
            case 3: y[i] = mppi_cosf(x[i]); break;
            case 4: mppi_sincospif(x[i], &s, &c); y[i] = s; break;
            case 5: mppi_sincospif(x[i], &s, &c); y[i] = c; break;
            case 6: y[i] = mppi_tanhf(x[i]); break;
            default: y[i] = mppi_sqrtf(x[i]); break;
        }
    }
}
void oracle_fmodf(const float* x, const float* y, float* out, size_t n) {
    for (size_t i = 0; i < n; ++i) out[i] = mppi_fmodf(x[i], y[i]);
}
// The library's two-argument helpers: 0 mppi_maxf, 1 mppi_minf, 2
// mppi_clampf(x, -1/2, y).
void oracle_math2f(int fn, const float* x, const float* y, float* out, size_t n) {
    for (size_t i = 0; i < n; ++i)
        out[i] = fn == 0 ? mppi_maxf(x[i], y[i]) : fn == 1 ? mppi_minf(x[i], y[i]) : mppi_clampf(x[i], -0.5f, y[i]);
}

// The library header's Philox4x32-10 (include/mppi_hip/mppi_philox.h)
// compiled for the host, so tests can compare it with this oracle's own.
void oracle_library_philox4x32(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
    mppi_u32x4 r = mppi_philox4x32_10(mppi_u32x4{ctr[0], ctr[1], ctr[2], ctr[3]}, key[0], key[1]);
    out[0] = r.x; out[1] = r.y; out[2] = r.z; out[3] = r.w;
}

// The same block through the header's split form (the shared words of x, z,
// w, then the lane part from y), as the GPU rollout computes it.
void oracle_library_philox4x32_split(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
    const mppi_philox_shared_words s = mppi_philox_shared(ctr[0], ctr[2], ctr[3], key[0], key[1]);
    mppi_u32x4 r = mppi_philox_lanes(ctr[1], mppi_philox_lane_key(ctr[2], key[0]), s, key[0], key[1]);
    out[0] = r.x; out[1] = r.y; out[2] = r.z; out[3] = r.w;
}

// Costs of samples [k0, k1) only (costs is indexed by absolute sample).
int oracle_rollout_range(const oracle_problem* p, uint32_t step, uint32_t agent, const float* x0,
                         const float* u_nom, uint32_t k0, uint32_t k1, float* costs) {
    Resolved<float> r = resolve<float>(p);
    if (!r.ok || k1 > p->K || k0 > k1) return -1;
    rollout_range<float>(p, r, step, agent, x0, u_nom, k0, k1, costs);
    return 0;
}

// Realised cost of a closed loop: out[i] = stage(x[i], u[i], time {0,
// steps[i]}) for n (post-step state, applied control) pairs, the cost the
// rollout's first timestep assigns to what the plant actually did.
// The GPU pre-filters of the obstacle and zone costs, evaluated on the host
// from the plugins' own (host-callable) functions, so that tests can check
// that they flag every obstacle or zone a state counts.
// unicycle: bit j of flags[i] is set when the square gate of slot j (scaled
// coordinates, unicycle_gate) holds state i = (xy[2i], xy[2i+1]).
int oracle_unicycle_gate_flags(const uint8_t* cost, uint32_t n, const float* xy, uint32_t* flags) {
    const auto* p = reinterpret_cast<const plug_uni::UnicycleCost*>(cost);
    for (uint32_t i = 0; i < n; ++i) {
        uint32_t f = 0;
        for (uint32_t j = 0; j < UNICYCLE_MAX_OBSTACLES; ++j) {
            const plug_uni::UnicycleGate g =
                plug_uni::unicycle_gate(p->obstacles[j][0], p->obstacles[j][1], p->obstacles[j][2]);
            const float ox = std::fma(xy[2 * i], g.a, g.bx), oy = std::fma(xy[2 * i + 1], g.a, g.by);
            if (std::fmax(std::fabs(ox), std::fabs(oy)) < 1.0f) f |= 1u << j;
        }
        flags[i] = f;
    }
    return 0;
}

// bicycle: cand[i] = the zones the grid lists for state i (its column's and
// row's sets and the zones every wavefront tests), as bicycle_candidates
// forms them on the GPU from the same functions.
int oracle_bicycle_candidates(const uint8_t* cost, uint32_t n, const float* xy, uint32_t* cand) {
    const auto* c = reinterpret_cast<const plug_bic::BicycleCost*>(cost);
    const plug_bic::BicycleFrame f = plug_bic::bicycle_frame(c);
    for (uint32_t i = 0; i < n; ++i) {
        const uint32_t cu = plug_bic::bicycle_band(xy[2 * i], f.x0, f.inv);
        const uint32_t cv = plug_bic::bicycle_band(xy[2 * i + 1], f.y0, f.inv);
        cand[i] = (plug_bic::bicycle_band_zones(c, f, true, cu) & plug_bic::bicycle_band_zones(c, f, false, cv)) |
                  f.always;
    }
    return 0;
}

int oracle_stage_costs(const oracle_problem* p, uint32_t n, const float* states, const float* controls,
                       const uint32_t* steps, float* out) {
    Resolved<float> r = resolve<float>(p);
    if (!r.ok) return -1;
    for (uint32_t i = 0; i < n; ++i)
        out[i] = r.stage(states + size_t(i) * p->D, controls + size_t(i) * p->C, r.model, r.cost,
                         mppi_time{0u, steps[i]});
    return 0;
}

// The binary16 rounding of the FP16 mode: this oracle's restatement and the
// library header's (include/mppi_hip/mppi_philox.h) compiled for the host.
void oracle_round_f16(const float* x, float* out, size_t n) {
    for (size_t i = 0; i < n; ++i) out[i] = float(round_f16(double(x[i])));
}
void oracle_library_round_f16(const float* x, float* out, size_t n) {
    for (size_t i = 0; i < n; ++i) out[i] = mppi_round_f16(x[i]);
}

// One step x <- x + delta(x, u) of the surrogate network.
void oracle_surrogate_dynamics(uint32_t D, uint32_t C, float* x, const float* u, const float* params,
                               const float* norm80) {
    surrogate_dynamics(Surrogate{params, norm80}, D, C, x, u);
}

// Costs of samples [k0, k1) with the surrogate network as dynamics.
int oracle_rollout_surrogate_range(const oracle_problem* p, uint32_t step, uint32_t agent,
                                   const float* x0, const float* u_nom, uint32_t k0, uint32_t k1,
                                   const float* params, const float* norm80, float* costs) {
    Resolved<float> r = resolve<float>(p);
    if (!r.ok || k1 > p->K || k0 > k1 || p->D + p->C > 16) return -1;
    const Surrogate s{params, norm80};
    rollout_range<float>(p, r, step, agent, x0, u_nom, k0, k1, costs, &s);
    return 0;
}

int oracle_rollout(const oracle_problem* p, uint32_t step, uint32_t agent, const float* x0,
                   const float* u_nom, float* costs, int threads) {
    Resolved<float> r = resolve<float>(p);
    if (!r.ok) return -1;
    rollout<float>(p, r, step, agent, x0, u_nom, costs, threads);
    return 0;
}

int oracle_rollout64(const oracle_problem* p, uint32_t step, uint32_t agent, const float* x0,
                     const double* u_nom, double* costs, int threads) {
    Resolved<double> r = resolve<double>(p);
    if (!r.ok) return -1;
    rollout<double>(p, r, step, agent, x0, u_nom, costs, threads);
    return 0;
}

int oracle_weights(const oracle_problem* p, const float* costs, float* w, float* diag3) {
    weights<float>(p, costs, w, diag3);
    return 0;
}

int oracle_weights64(const oracle_problem* p, const double* costs, double* w, double* diag3) {
    weights<double>(p, costs, w, diag3);
    return 0;
}

int oracle_reduce(const oracle_problem* p, uint32_t step, uint32_t agent, const float* w,
                  const float* u_nom, float* u_out) {
    Resolved<float> r = resolve<float>(p);
    if (!r.ok) return -1;
    reduce<float>(p, r, step, agent, w, u_nom, u_out);
    return 0;
}

int oracle_reduce64(const oracle_problem* p, uint32_t step, uint32_t agent, const double* w,
                    const double* u_nom, double* u_out) {
    Resolved<double> r = resolve<double>(p);
    if (!r.ok) return -1;
    reduce<double>(p, r, step, agent, w, u_nom, u_out);
    return 0;
}

// The magnitude of the update's sum, per control: sum_k w_k |eps_k| in FP64
// (the conditioning of u_nom + sum_k w_k eps_k, against which the tests
// state the update's elementwise error).
int oracle_reduce_abs(const oracle_problem* p, uint32_t step, uint32_t agent, const double* w,
                      double* out) {
    Resolved<double> r = resolve<double>(p);
    if (!r.ok) return -1;
    const uint32_t C = p->C, H = p->H, HC = H * C;
    std::vector<double> z(HC), eps(C);
    std::fill(out, out + HC, 0.0);
    for (uint32_t k = 0; k < p->K; ++k) {
        if (w[k] == 0.0) continue;
        sample_normals<double>(p, k, step, agent, z.data());
        for (uint32_t t = 0; t < H; ++t) {
            noise_eps<double>(p, r, z.data(), t, eps.data());
            for (uint32_t i = 0; i < C; ++i) out[t * C + i] += w[k] * std::fabs(eps[i]);
        }
    }
    return 0;
}

int oracle_step(const oracle_problem* p, uint32_t step, uint32_t agent, const float* x0,
                const float* u_nom, float* u_out, float* diag3, float* costs_out,
                float* weights_out, int threads) {
    return step_impl<float>(p, step, agent, x0, u_nom, u_out, diag3, costs_out, weights_out, threads);
}

int oracle_step64(const oracle_problem* p, uint32_t step, uint32_t agent, const float* x0,
                  const double* u_nom, double* u_out, double* diag3, double* costs_out,
                  double* weights_out, int threads) {
    return step_impl<double>(p, step, agent, x0, u_nom, u_out, diag3, costs_out, weights_out, threads);
}

int oracle_simulate(const oracle_problem* p, uint32_t T, uint32_t start_step, uint32_t n_agents,
                    const float* x0_packed, const float* u_nom_packed, float* states_out,
                    float* controls_out, float* costs_out, float* terminal_out, int threads) {
    return simulate_impl<float>(p, T, start_step, n_agents, x0_packed, u_nom_packed, states_out,
                                controls_out, costs_out, terminal_out, threads);
}

int oracle_simulate64(const oracle_problem* p, uint32_t T, uint32_t start_step, uint32_t n_agents,
                      const float* x0_packed, const float* u_nom_packed, float* states_out,
                      float* controls_out, float* costs_out, float* terminal_out, int threads) {
    return simulate_impl<double>(p, T, start_step, n_agents, x0_packed, u_nom_packed, states_out,
                                 controls_out, costs_out, terminal_out, threads);
}

}  // extern "C"
