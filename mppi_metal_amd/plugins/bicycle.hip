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
// clamped to [0, v_max] (min(max(vx, 0), v_max): a NaN gives 0).
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
        const float uu = mppi_maxf(vx, p->v_min);
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
        s[3] = mppi_clampf(s[3], 0.0f, p->v_max);
    }
}

// Zone pre-filter (GPU).  A sample counts zone i (i < n_zones) when
// d2 < r*r with d2 = fl(fl(ox^2) + fl(oy^2)), ox = fl(x - cx), oy = fl(y - cy);
// that implies |x - cx| < r (1 + 2^-22) and likewise for y (fl(ox^2) <=
// fl(d2) < fl(r^2), each rounding within 2^-24), so the sample lies in the
// square of half-side r' = r (1 + 2^-20) around the centre.  Each workgroup
// maps the plane to a BICYCLE_GRID x BICYCLE_GRID grid of square cells whose
// frame (bicycle_frame) leaves two empty cells around the union of those
// squares, and keeps in LDS, for every column and every row of cells, the
// set of zones whose square, widened by 1/8 cell, meets it
// (bicycle_prepare).  A sample's set, its column's and row's sets
// intersected, holds every zone it can count.  The frame's arithmetic
// (relative error a few 2^-24 of coordinates within G + 2 cells, and of the
// squares' bounds) moves a sample or a bound by far less than the 1/8 cell
// margin; a frame whose coordinates cannot resolve 1/64 of a cell is not
// used, and its zones are tested by every wavefront instead, as are radii
// outside [2^-60, 2^60].  Zones with r <= 0 or NaN, or a centre that is not
// finite, never count and are in no set.  Only the exact pass decides what
// counts, and it runs for a zone in every wavefront where some sample's set
// holds it, so the cost is the exact sum over the zones, which is what the
// host computes.
#define BICYCLE_GRID 64

struct BicycleFrame {
    float x0, y0, inv;   // cell coordinates u = (x - x0) inv, v = (y - y0) inv
    uint32_t grid;       // zones on the grid
    uint32_t always;     // zones every wavefront tests
};

MPPI_HD bool bicycle_finite(float v) { return (mppi_f2u(v) & 0x7f800000u) != 0x7f800000u; }

MPPI_HD float bicycle_square_half(float r) { return r * 1.00000095367431640625f; }   // r (1 + 2^-20)

MPPI_HD BicycleFrame bicycle_frame(const BicycleCost* c) {
    const uint32_t nz = c->n_zones < BICYCLE_MAX_ZONES ? c->n_zones : BICYCLE_MAX_ZONES;
    float xlo = 3.0e38f, xhi = -3.0e38f, ylo = 3.0e38f, yhi = -3.0e38f;
    uint32_t grid = 0, always = 0;
    for (uint32_t j = 0; j < nz; ++j) {
        const float cx = c->zones[j][0], cy = c->zones[j][1], r = c->zones[j][2];
        if (!(r > 0.0f) || !bicycle_finite(cx) || !bicycle_finite(cy)) continue;
        if (!(r >= 0x1p-60f && r <= 0x1p60f)) {
            always |= 1u << j;
            continue;
        }
        const float h = bicycle_square_half(r);
        xlo = __builtin_fminf(xlo, cx - h);
        xhi = __builtin_fmaxf(xhi, cx + h);
        ylo = __builtin_fminf(ylo, cy - h);
        yhi = __builtin_fmaxf(yhi, cy + h);
        grid |= 1u << j;
    }
    BicycleFrame f{0.0f, 0.0f, 0.0f, grid, always};
    if (grid) {
        const float e = __builtin_fmaxf(xhi - xlo, yhi - ylo);
        const float w = e * (1.0f / float(BICYCLE_GRID - 4));   // two cells of margin each side
        const float inv = 1.0f / w;
        const float m = __builtin_fmaxf(__builtin_fmaxf(mppi_fabsf(xlo), mppi_fabsf(xhi)),
                                        __builtin_fmaxf(mppi_fabsf(ylo), mppi_fabsf(yhi)));
        if (bicycle_finite(e) && w > 0.0f && bicycle_finite(inv) && m * 0x1p-17f <= w) {
            f.x0 = xlo - 2.0f * w;
            f.y0 = ylo - 2.0f * w;
            f.inv = inv;
        } else {
            f.always |= grid;
            f.grid = 0;
        }
    }
    return f;
}

// The set of zones whose widened square meets column band `band` (col) or
// row band `band` of the frame's grid; band BICYCLE_GRID is empty.
MPPI_HD uint32_t bicycle_band_zones(const BicycleCost* c, const BicycleFrame& f, bool col, uint32_t band) {
    const float o = col ? f.x0 : f.y0;
    uint32_t m = 0;
    if (band < BICYCLE_GRID) {
        for (uint32_t j = 0; j < BICYCLE_MAX_ZONES; ++j) {
            if (!((f.grid >> j) & 1u)) continue;
            const float uc = (c->zones[j][col ? 0 : 1] - o) * f.inv;
            const float ur = bicycle_square_half(c->zones[j][2]) * f.inv;
            if (uc + ur >= float(band) - 0.125f && uc - ur <= float(band) + 1.125f) m |= 1u << j;
        }
    }
    return m;
}

// The band of coordinate v (x with the frame's x0, y with its y0): the
// cell index clamped to [0, BICYCLE_GRID]; NaN maps to 0.
MPPI_HD uint32_t bicycle_band(float v, float o, float inv) {
    return uint32_t(__builtin_fminf(__builtin_fmaxf((v - o) * inv, 0.0f), float(BICYCLE_GRID)));
}

#if defined(__HIP_DEVICE_COMPILE__)
// Zone sets of the columns (x) and rows (y) of cells; entry BICYCLE_GRID,
// empty, takes every coordinate at or past the last cell.  The frame (x0,
// y0, inv, always) is in LDS too, read by the costs with one broadcast load.
__shared__ uint32_t bicycle_cols[BICYCLE_GRID + 1];
__shared__ uint32_t bicycle_rows[BICYCLE_GRID + 1];
__shared__ __attribute__((aligned(16))) float bicycle_frame_lds[4];

// The workgroup's preparation of its costs (MPPI_COST_PREPARE, mppi_core.h).
__device__ void bicycle_prepare(const uint8_t* cost_raw, uint32_t thread, uint32_t threads) {
    const BicycleCost* c = reinterpret_cast<const BicycleCost*>(cost_raw);
    const BicycleFrame f = bicycle_frame(c);
    if (thread == 0) {
        bicycle_frame_lds[0] = f.x0;
        bicycle_frame_lds[1] = f.y0;
        bicycle_frame_lds[2] = f.inv;
        bicycle_frame_lds[3] = mppi_u2f(f.always);
    }
    for (uint32_t i = thread; i < 2 * (BICYCLE_GRID + 1); i += threads) {
        const bool col = i <= BICYCLE_GRID;
        const uint32_t band = col ? i : i - (BICYCLE_GRID + 1);
        (col ? bicycle_cols : bicycle_rows)[band] = bicycle_band_zones(c, f, col, band);
    }
}
#define MPPI_COST_PREPARE bicycle_prepare

// The zones sample (x, y) can count: its column's and row's sets, and the
// zones every wavefront tests.
__device__ uint32_t bicycle_candidates(float x, float y) {
    const float4 f = *reinterpret_cast<const float4*>(bicycle_frame_lds);
    return (bicycle_cols[bicycle_band(x, f.x, f.z)] & bicycle_rows[bicycle_band(y, f.y, f.z)]) | mppi_f2u(f.w);
}
#endif

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
#if defined(__HIP_DEVICE_COMPILE__)
    // The exact pass keeps zone order and runs only in wavefronts where some
    // sample can count some zone, and there only for those zones.
    const uint32_t cand = bicycle_candidates(s[0], s[1]);
    if (__builtin_expect(mppi_any(cand != 0u), 0)) {
        for (uint32_t i = 0; i < nz; ++i) {
            if (mppi_any(((cand >> i) & 1u) != 0u)) {
                const float ox = s[0] - c->zones[i][0];
                const float oy = s[1] - c->zones[i][1];
                const float rz = c->zones[i][2];
                const float d2 = ox * ox + oy * oy;
                if (d2 < rz * rz) pen += rz - mppi_sqrtf(d2);
            }
        }
    }
#else
    // The definition, zone by zone in the same order.
    for (uint32_t i = 0; i < nz; ++i) {
        const float ox = s[0] - c->zones[i][0];
        const float oy = s[1] - c->zones[i][1];
        const float rz = c->zones[i][2];
        const float d2 = ox * ox + oy * oy;
        if (rz > 0.0f && d2 < rz * rz) pen += rz - mppi_sqrtf(d2);
    }
#endif
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
