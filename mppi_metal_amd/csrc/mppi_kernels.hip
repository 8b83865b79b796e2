// mppi_kernels.hip — library-owned MPPI kernels for gfx950.
//
// This file is not compiled on its own: the library hands it to hipRTC
// together with the user's plugin source (as the in-memory header
// "mppi_user_plugin.hip") and these JIT-time definitions:
//   MPPI_SDIM, MPPI_CDIM            state and control dimensions
//   MPPI_DYNAMICS                   name of the dynamics callable
//   MPPI_STAGE_COST, MPPI_TERMINAL_COST  cost callables (or the defaults)
//   MPPI_COV_FULL                   0: diagonal sigma, 1: lower Cholesky factor
//   MPPI_BATCH_KERNELS              1: also emit the tile-table and persistent
//                                   kernels of the uneven multi-agent step
//   MPPI_SURROGATE_KERNELS          1: also emit the learned-surrogate kernels
//                                   (training data from the plugin dynamics,
//                                   rollout with the MLP as dynamics)
// so the plugin's callables are direct, inlined calls (the reference reaches
// them through a visible function table at every timestep,
// shaders/mppi_rollout.metal:90-91).
//
// The two-kernel MPPI step is three launches:
//   rollout  one sample per lane; control noise from Philox4x32-10 in
//            registers (never stored); clamp; dynamics; stage cost on the
//            post-dynamics state; terminal cost; FP32 cost accumulation.
//            Each workgroup also reduces its costs to (min, sum exp, sum cost,
//            sum exp^2) with wave shuffles and LDS.  Reference: mppi_rollout
//            (mppi_rollout.metal:36-96, batch :103-170) and the min / sum
//            passes of mppi_compute_weights (mppi_reduce.metal:14-121, :242-356).
//   reduce   every workgroup first combines its agent's rollout partials into
//            the softmax normaliser (log-sum-exp form), then the weighted-noise
//            average, parallel over samples: every lane regenerates its noise
//            blocks from the same counters and accumulates w~_k * z_k into
//            workgroup partials.  Reference: mppi_reduce
//            (mppi_reduce.metal:130-168, :363-406).
//   finalize one workgroup per agent sums the partials, writes u_next, the
//            diagnostics and, in closed-loop mode, steps the plant and shifts
//            the warm start.  Reference: mppi_propagate_and_shift
//            (mppi_reduce.metal:175-231, :413-488).
// Work is cut into tiles: rollout (agent, sample block), reduce (agent,
// sample chunk, noise-block group), and the one-launch step's tile (agent,
// sample block: rollout, then the tile's weighted noise sums, the last tile
// of the agent forming the update).  Three dispatch modes share the tile
// bodies: grid-mapped (one workgroup per tile in grid order), tile table
// (workgroup i runs tile i of a host-built, longest-first table) and
// persistent (a fixed grid takes tiles from a ticket counter).
// Cross-block sums (normaliser and weighted noise) run in FP64 when K >= 2^18.
#if !defined(__HIPCC_RTC__)
#include <hip/hip_runtime.h>   // hipRTC provides these builtins itself
#endif
#include "mppi_core.h"
#include "mppi_launch.h"
#include "mppi_user_plugin.hip"

#ifndef MPPI_STAGE_COST
#define MPPI_STAGE_COST mppi_default_stage_cost
#endif
#ifndef MPPI_TERMINAL_COST
#define MPPI_TERMINAL_COST mppi_default_terminal_cost
#endif
#ifndef MPPI_COV_FULL
#define MPPI_COV_FULL 0
#endif
#ifndef MPPI_BATCH_KERNELS
#define MPPI_BATCH_KERNELS 0
#endif
#ifndef MPPI_SURROGATE_KERNELS
#define MPPI_SURROGATE_KERNELS 0
#endif

#define SDIM MPPI_SDIM
#define CDIM MPPI_CDIM
#define MPPI_REDUCE_BPG 8                 // noise blocks per reduce tile
#define MPPI_NV (4 * MPPI_REDUCE_BPG)     // accumulators per reduce thread
#define MPPI_FP64_MIN_K (1u << 18)
#define MPPI_MAX_HC 4096
#define MPPI_FUSED_MAX_HC 128             // horizon * control_dim limit of the one-pass step
#define MPPI_FUSED_ZS (MPPI_FUSED_MAX_HC + 4)  // LDS row of a lane's normals (16-byte multiple)

constexpr int mppi_gcd(int a, int b) { return b == 0 ? a : mppi_gcd(b, a % b); }
// Timesteps whose C*G noise values fill whole Philox blocks.
constexpr int MPPI_G = 4 / mppi_gcd(CDIM, 4);
constexpr int MPPI_NG = MPPI_G * CDIM;     // normals per group (multiple of 4)
constexpr int MPPI_BG = MPPI_NG / 4;       // Philox blocks per group
// The one-pass step writes each group's normals whole when a group that
// starts inside the horizon (t0 <= H - 1) and ends past it still fits the
// padding of a lane's row: (G - 1) C <= ZS - MAX_HC.
constexpr bool MPPI_FUSED_OVERHANG_OK = CDIM * (MPPI_G - 1) <= MPPI_FUSED_ZS - MPPI_FUSED_MAX_HC;

// Cost calls with or without the optional trailing mppi_time argument: the
// first overload is viable only when the plugin's function takes it.
struct mppi_form_plain {};
struct mppi_form_timed : mppi_form_plain {};
template <class X, class U, class M, class C>
__device__ inline auto mppi_stage(mppi_form_timed, X x, U u, M m, C c, mppi_time tm)
    -> decltype(MPPI_STAGE_COST(x, u, m, c, tm)) { return MPPI_STAGE_COST(x, u, m, c, tm); }
template <class X, class U, class M, class C>
__device__ inline auto mppi_stage(mppi_form_plain, X x, U u, M m, C c, mppi_time)
    -> decltype(MPPI_STAGE_COST(x, u, m, c)) { return MPPI_STAGE_COST(x, u, m, c); }
template <class X, class M, class C>
__device__ inline auto mppi_terminal(mppi_form_timed, X x, M m, C c, mppi_time tm)
    -> decltype(MPPI_TERMINAL_COST(x, m, c, tm)) { return MPPI_TERMINAL_COST(x, m, c, tm); }
template <class X, class M, class C>
__device__ inline auto mppi_terminal(mppi_form_plain, X x, M m, C c, mppi_time)
    -> decltype(MPPI_TERMINAL_COST(x, m, c)) { return MPPI_TERMINAL_COST(x, m, c); }

// The plugin's optional preparation of its costs for a workgroup
// (MPPI_COST_PREPARE, mppi_core.h): every thread of the workgroup calls it
// with the workgroup's cost parameters before the barrier that precedes the
// time loop; the previous user of the workgroup's memory finished before an
// earlier barrier.
__device__ inline void mppi_prepare_costs(const uint8_t* cost, uint32_t threads) {
#ifdef MPPI_COST_PREPARE
    MPPI_COST_PREPARE(mppi_blob_ref{cost}, threadIdx.x, threads);
#else
    (void)cost;
    (void)threads;
#endif
}

// Wavefront reductions (every lane of the wavefront active).  Within each
// row of 16 lanes the values are combined through DPP lane patterns, which
// cost one VALU instruction each instead of an LDS permute and its wait:
// quad_perm [1,0,3,2] (lane ^ 1), quad_perm [2,3,0,1] (lane ^ 2), then the
// half-row mirror (lane i of 8 takes lane 7 - i, which lies in the other
// quad) and the row mirror (lane i of 16 takes lane 15 - i, in the other
// half).  After these four steps every lane of a row holds the row's value;
// the four rows are read out to scalar registers and combined as
// (r0 op r1) op (r2 op r3), so every lane returns the same value.
template <int CTRL>
__device__ inline uint32_t mppi_dpp_(uint32_t v) {
    return uint32_t(__builtin_amdgcn_mov_dpp(int(v), CTRL, 0xf, 0xf, false));
}
template <int CTRL>
__device__ inline float mppi_dpp_(float v) {
    return __builtin_bit_cast(float, mppi_dpp_<CTRL>(__builtin_bit_cast(uint32_t, v)));
}
template <int CTRL>
__device__ inline double mppi_dpp_(double v) {
    const uint64_t b = __builtin_bit_cast(uint64_t, v);
    const uint32_t lo = mppi_dpp_<CTRL>(uint32_t(b)), hi = mppi_dpp_<CTRL>(uint32_t(b >> 32));
    return __builtin_bit_cast(double, (uint64_t(hi) << 32) | lo);
}
__device__ inline float mppi_readlane_(float v, int lane) {
    return __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, v), lane));
}
__device__ inline double mppi_readlane_(double v, int lane) {
    const uint64_t b = __builtin_bit_cast(uint64_t, v);
    const uint32_t lo = uint32_t(__builtin_amdgcn_readlane(int(uint32_t(b)), lane));
    const uint32_t hi = uint32_t(__builtin_amdgcn_readlane(int(uint32_t(b >> 32)), lane));
    return __builtin_bit_cast(double, (uint64_t(hi) << 32) | lo);
}

template <typename T, typename Op>
__device__ inline T mppi_wave_reduce_(T v, Op op) {
    v = op(v, mppi_dpp_<0xB1>(v));
    v = op(v, mppi_dpp_<0x4E>(v));
    v = op(v, mppi_dpp_<0x141>(v));
    v = op(v, mppi_dpp_<0x140>(v));
    return op(op(mppi_readlane_(v, 0), mppi_readlane_(v, 16)), op(mppi_readlane_(v, 32), mppi_readlane_(v, 48)));
}

__device__ inline float mppi_wave_min(float v) {
    return mppi_wave_reduce_(v, [](float x, float w) { return w < x ? w : x; });
}

template <typename T>
__device__ inline T mppi_wave_sum(T v) {
    return mppi_wave_reduce_(v, [](T x, T w) { return x + w; });
}

__device__ inline bool mppi_isfinite(float x) {
    return (mppi_f2u(x) & 0x7f800000u) != 0x7f800000u;
}

// Control for one timestep from the normals z[0..CDIM): u = clamp(u_nom + eps).
__device__ inline void mppi_make_control(const float* un, const float* z, const float* scale,
                                         const float* lo, const float* hi, float* u) {
#if MPPI_COV_FULL
#pragma unroll
    for (int i = 0; i < CDIM; ++i) {
        float e = 0.0f;
#pragma unroll
        for (int j = 0; j <= i; ++j) e = mppi_fma(scale[i * CDIM + j], z[j], e);
        u[i] = mppi_clampf(un[i] + e, lo[i], hi[i]);
    }
#else
#pragma unroll
    for (int i = 0; i < CDIM; ++i) {
        float e = z[i] * scale[i];
        u[i] = mppi_clampf(un[i] + e, lo[i], hi[i]);
    }
#endif
}

// ---------------------------------------------------------------------------
// Softmax normaliser: combine the per-workgroup (min, sum e, sum c, sum e^2)
// partials of one agent: m = min m_g, S = sum_g S_g exp(-(m_g - m)/lambda),
// Q likewise with the squared factor, ESS = S^2 / Q (= 1 / sum w^2).
// T is the accumulation type: float, or double when K >= 2^18.  Every
// workgroup that needs the normaliser forms it from the rows itself, in the
// same order, so all get the same values: a workgroup that wrote rows would
// otherwise have to publish them at agent scope before a last-arriving
// workgroup combined them, and that release (a write-back of the XCD's L2,
// whose caches are not coherent with the other XCDs') costs more, summed
// over thousands of workgroups, than the rows' few loads per workgroup.
// ---------------------------------------------------------------------------
template <typename T>
struct mppi_norm {
    float m;          // min m_g (not finite: no finite cost; then S = Q = K)
    T S, Cs, Q;
};

template <int BLOCK, typename T>
__device__ mppi_norm<T> mppi_normaliser(const mppi_step_args& a, const mppi_agent_desc& d) {
    constexpr int WAVES = BLOCK / 64;
    __shared__ float s_min[WAVES];
    __shared__ T s_sum[3][WAVES];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const float* st = a.blk_stats + d.blk_off * 4;
    float mg = __builtin_inff();
    for (uint32_t i = threadIdx.x; i < d.n_blk; i += BLOCK) {
        float v = st[i * 4];
        mg = v < mg ? v : mg;
    }
    mg = mppi_wave_min(mg);
    if (lane == 0) s_min[wave] = mg;
    __syncthreads();
    mg = s_min[0];
#pragma unroll
    for (int w = 1; w < WAVES; ++w) mg = s_min[w] < mg ? s_min[w] : mg;
    T S = T(0), Cs = T(0), Q = T(0);
    for (uint32_t i = threadIdx.x; i < d.n_blk; i += BLOCK) {
        float bm = st[i * 4 + 0];
        float f = (mppi_isfinite(bm) && mppi_isfinite(mg)) ? mppi_expf(-(bm - mg) / a.lambda) : 0.0f;
        S += T(st[i * 4 + 1]) * T(f);
        Cs += T(st[i * 4 + 2]);
        Q += T(st[i * 4 + 3]) * (T(f) * T(f));
    }
    S = mppi_wave_sum(S);
    Cs = mppi_wave_sum(Cs);
    Q = mppi_wave_sum(Q);
    if (lane == 0) { s_sum[0][wave] = S; s_sum[1][wave] = Cs; s_sum[2][wave] = Q; }
    __syncthreads();
    S = s_sum[0][0]; Cs = s_sum[1][0]; Q = s_sum[2][0];
#pragma unroll
    for (int w = 1; w < WAVES; ++w) { S += s_sum[0][w]; Cs += s_sum[1][w]; Q += s_sum[2][w]; }
    if (!mppi_isfinite(mg)) {
        // Every cost is NaN or infinite: fall back to uniform weights.
        S = T(d.K);
        Q = T(d.K);
    }
    __syncthreads();                        // s_min and s_sum are free again
    return mppi_norm<T>{mg, S, Cs, Q};
}

// The normaliser into the agent's state (thread 0), as the update and the
// per-sample weights read it; also re-arms the agent's tile arrival counter.
template <typename T>
__device__ void mppi_store_normaliser(const mppi_step_args& a, const mppi_agent_desc& d, const mppi_norm<T>& n) {
    if (threadIdx.x == 0) {
        mppi_agent_state* ag = &a.agents[d.slot];
        ag->min_cost = n.m;
        ag->weight_sum = double(n.S);
        ag->mean_cost = double(n.Cs / T(d.K));
        ag->ess = double((n.S * n.S) / n.Q);
        ag->rollout_arrivals = 0u;
    }
}

// Workgroup softmax partials over one rollout tile's costs (min, sum e,
// sum c, sum e^2), one statistics row per SUB samples (SUB = BLOCK: one row
// per tile; a tile of BLOCK = 2 SUB samples writes rows 2 blk and 2 blk + 1,
// the second only if it is inside the agent's n_blk rows).
template <int BLOCK, int SUB = BLOCK>
__device__ void mppi_block_stats(const mppi_step_args& a, const mppi_agent_desc& d, uint32_t blk, bool active,
                                 float total) {
    constexpr int WAVES = BLOCK / 64, WPR = SUB / 64, ROWS = BLOCK / SUB;
    static_assert(BLOCK % SUB == 0 && SUB % 64 == 0, "rows of whole wavefronts");
    __shared__ float s_red[4][WAVES];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, r0 = (wave / WPR) * WPR;
    float c = (active && mppi_isfinite(total)) ? total : __builtin_inff();
    float m = mppi_wave_min(c);
    if (lane == 0) s_red[0][wave] = m;
    __syncthreads();
    m = s_red[0][r0];
#pragma unroll
    for (int w = 1; w < WPR; ++w) m = s_red[0][r0 + w] < m ? s_red[0][r0 + w] : m;
    float e = (mppi_isfinite(c) && mppi_isfinite(m)) ? mppi_expf(-(c - m) / a.lambda) : 0.0f;
    float se = mppi_wave_sum(e);
    float sq = mppi_wave_sum(e * e);
    float sc = mppi_wave_sum(active ? total : 0.0f);
    if (lane == 0) { s_red[1][wave] = se; s_red[2][wave] = sc; s_red[3][wave] = sq; }
    __syncthreads();
    if (threadIdx.x < uint32_t(ROWS)) {
        const int rr = int(threadIdx.x) * WPR;
        float mr = s_red[0][rr], S = 0.0f, Cs = 0.0f, Q = 0.0f;
#pragma unroll
        for (int w = 0; w < WPR; ++w) {
            mr = s_red[0][rr + w] < mr ? s_red[0][rr + w] : mr;
            S += s_red[1][rr + w]; Cs += s_red[2][rr + w]; Q += s_red[3][rr + w];
        }
        const uint32_t row = blk * ROWS + threadIdx.x;
        if (row < d.n_blk) {
            float* bs = a.blk_stats + (d.blk_off + row) * 4;
            bs[0] = mr; bs[1] = S; bs[2] = Cs; bs[3] = Q;
        }
    }
}

template <int SAMPLES, int THREADS>
__device__ void mppi_fused_stats(const mppi_step_args& a, uint32_t agent, const mppi_agent_desc& d,
                                 uint32_t blk, bool active, float total, float* s_z);

// ---------------------------------------------------------------------------
// Rollout of samples [blk*BLOCK, blk*BLOCK + BLOCK) of one agent: returns
// the sample's cost (0 for lanes past K) and stores it.  s_un: LDS of at
// least MPPI_UN_LDS floats, free again once the function returns (after a
// barrier).
// ---------------------------------------------------------------------------
constexpr int MPPI_UN_LDS = MPPI_MAX_HC + 2 * MPPI_NG;

template <int BLOCK>
__device__ float mppi_rollout_costs(const mppi_step_args& a, const mppi_agent_desc& d, uint32_t blk,
                                    float* s_un) {
    const uint32_t k = blk * BLOCK + threadIdx.x;
    const uint32_t H = d.H;
    const bool active = k < d.K;
    const uint32_t slot = d.slot;
    const uint32_t step = a.agents[slot].step;
    const uint32_t tag_agent = (MPPI_TAG_CONTROL_NOISE << 28) | (slot & MPPI_AGENT_MASK);
    const uint32_t dyn_tag = (MPPI_TAG_DYNAMICS << 28) | (slot & MPPI_AGENT_MASK);
    const uint8_t* model = a.model + slot * a.model_stride;
    const uint8_t* cost = a.cost + slot * a.cost_stride;

    // The nominal sequence goes to LDS and the noise scale and bounds to
    // registers, so the time loop issues no memory waits of its own.
    // 2 * MPPI_NG slots of padding: the time loop reads one group ahead.
    for (uint32_t i = threadIdx.x; i < H * CDIM; i += BLOCK) s_un[i] = a.u_nom[d.seq_off + i];
    float scale[MPPI_COV_FULL ? CDIM * CDIM : CDIM], lo[CDIM], hi[CDIM];
#pragma unroll
    for (int i = 0; i < (MPPI_COV_FULL ? CDIM * CDIM : CDIM); ++i) scale[i] = a.noise_scale[i];
#pragma unroll
    for (int i = 0; i < CDIM; ++i) {
        lo[i] = a.u_min[i];
        hi[i] = a.u_max[i];
    }
    mppi_prepare_costs(cost, BLOCK);
    __syncthreads();

    float total = 0.0f;
    if (active) {
        float x[SDIM];
#pragma unroll
        for (int i = 0; i < SDIM; ++i) x[i] = a.x[size_t(slot) * SDIM + i];
        // The nominal controls of group t0 + G are read while group t0 runs,
        // so the LDS latency hides behind a group's work.  The read past the
        // horizon lands in the padding (t0 < H, so (t0 + G) C + NG <= H C + 2 NG).
        float un_next[MPPI_NG];
#pragma unroll
        for (int i = 0; i < MPPI_NG; ++i) un_next[i] = s_un[i];
        for (uint32_t t0 = 0; t0 < H; t0 += MPPI_G) {
            float un[MPPI_NG];
#pragma unroll
            for (int i = 0; i < MPPI_NG; ++i) {
                un[i] = un_next[i];
                un_next[i] = s_un[(t0 + MPPI_G) * CDIM + i];
            }
            const uint32_t b0 = (t0 * CDIM) >> 2;
            float z[MPPI_NG];
#pragma unroll
            for (int q = 0; q < MPPI_BG; ++q) {
                mppi_f32x4 v = mppi_normal4(mppi_u32x4{b0 + q, k, step, tag_agent}, a.key0, a.key1);
                z[4 * q + 0] = v.x; z[4 * q + 1] = v.y;
                z[4 * q + 2] = v.z; z[4 * q + 3] = v.w;
            }
#pragma unroll
            for (int g = 0; g < MPPI_G; ++g) {
                const uint32_t t = t0 + g;
                if (t < H) {
                    float u[CDIM];
                    mppi_make_control(un + g * CDIM, z + g * CDIM, scale, lo, hi, u);
                    mppi_rng rng{t << 8, k, step, dyn_tag, a.key0, a.key1};
                    MPPI_DYNAMICS(mppi_state_ref{x}, u, mppi_blob_ref{model}, rng);
                    total += mppi_stage(mppi_form_timed{}, mppi_cstate_ref{x}, (const float*)u,
                                        mppi_blob_ref{model}, mppi_blob_ref{cost},
                                        mppi_time{t, step});
                }
            }
        }
        total += mppi_terminal(mppi_form_timed{}, mppi_cstate_ref{x}, mppi_blob_ref{model},
                               mppi_blob_ref{cost}, mppi_time{H, step});
        a.costs[d.cost_off + k] = total;
    }
    return total;
}

// Rollout tile of the two-kernel step: the costs, then the tile's softmax
// partials (mppi_block_stats; the last tile of an agent forms the normaliser).
template <int BLOCK>
__device__ void mppi_rollout_tile(const mppi_step_args& a, uint32_t agent, uint32_t blk) {
    __shared__ __attribute__((aligned(16))) float s_un[MPPI_UN_LDS];
    const mppi_agent_desc d = a.desc[agent];
    const float total = mppi_rollout_costs<BLOCK>(a, d, blk, s_un);
    mppi_block_stats<BLOCK>(a, d, blk, blk * BLOCK + threadIdx.x < d.K, total);
}

// ---------------------------------------------------------------------------
// Rollout tile of the one-pass step with two wavefronts per SIMD: SAMPLES
// samples and 2 * SAMPLES threads.  Threads [SAMPLES, 2 SAMPLES) produce the
// noise (Philox and Box-Muller) into the LDS rows that the tile keeps for
// its noise sums anyway; threads [0, SAMPLES) run the rollout and read each
// group's normals from their row.  Every SIMD holds one wavefront of each
// kind, so the scalar instructions, branches, hazard waits and quarter-rate
// multiplies of one issue beside the other's vector work.  A producer
// publishes how many groups it has finished (after each of the first
// MPPI_PHASE_G groups, then every MPPI_PHASE_G groups): a consumer runs group
// g once the count is past g, while the producer goes on with the next ones.
// ---------------------------------------------------------------------------
constexpr uint32_t MPPI_PHASE_G = 4;    // groups of timesteps per published count

template <int SAMPLES, bool ONE>
__device__ void mppi_fused_tile(const mppi_step_args& a, uint32_t agent, uint32_t blk) {
    constexpr int THREADS = 2 * SAMPLES;
    constexpr int ZS = MPPI_FUSED_ZS;
    const mppi_agent_desc d = ONE ? a.desc0 : a.desc[agent];
    const bool producer = threadIdx.x >= uint32_t(SAMPLES);
    const uint32_t sid = producer ? threadIdx.x - SAMPLES : threadIdx.x;
    const uint32_t k = blk * SAMPLES + sid;
    const uint32_t H = d.H;
    const bool active = k < d.K;
    const uint32_t slot = d.slot;
    const uint32_t step = ONE ? a.step0 : a.agents[slot].step;
    const uint32_t tag_agent = (MPPI_TAG_CONTROL_NOISE << 28) | (slot & MPPI_AGENT_MASK);
    const uint32_t dyn_tag = (MPPI_TAG_DYNAMICS << 28) | (slot & MPPI_AGENT_MASK);
    const uint8_t* model = a.model + slot * a.model_stride;
    const uint8_t* cost = a.cost + slot * a.cost_stride;
    const uint32_t n_groups = (H + MPPI_G - 1) / MPPI_G;

    // The nominal sequence (read by the consumers with each group's
    // normals), the shared Philox words of
    // every noise block (mppi_philox_shared) and the normals, one row per
    // sample ([sample][ZS], rows 16-byte aligned).
    __shared__ __attribute__((aligned(16))) float s_un[MPPI_FUSED_MAX_HC + MPPI_NG];
    __shared__ __attribute__((aligned(16))) mppi_philox_shared_words s_ph[MPPI_FUSED_MAX_HC / 4 + MPPI_BG];
    __shared__ __attribute__((aligned(16))) float s_z[SAMPLES * ZS];
    for (uint32_t i = threadIdx.x; i < H * CDIM; i += THREADS) s_un[i] = a.u_nom[d.seq_off + i];
    for (uint32_t b = threadIdx.x; b < n_groups * MPPI_BG; b += THREADS)
        s_ph[b] = mppi_philox_shared(b, step, tag_agent, a.key0, a.key1);
    // Progress of producer wavefront w (the rows of samples [64 w, 64 w + 64),
    // which consumer wavefront w reads): groups finished.
    __shared__ uint32_t s_prog[SAMPLES / 64];
    if (threadIdx.x < uint32_t(SAMPLES / 64)) s_prog[threadIdx.x] = 0;
    float* row = s_z + sid * ZS;
    float total = 0.0f;
    mppi_prepare_costs(cost, THREADS);
    __syncthreads();                         // s_un, s_ph, s_prog and the costs' preparation written

    // Producer wavefront w publishes with a workgroup release fence (its LDS
    // writes are complete) and a relaxed store of the count of finished
    // groups to s_prog[w]; consumer wavefront w reads s_prog[w] with relaxed
    // loads before a group past the count it last saw, until the group is
    // there, and adds a zero taken from that value (bit 31, never set) to its
    // row address, so its reads of the group's normals are issued only after
    // the count arrived.
    // No barrier or fence sits in the consumers' time loop: both keep the
    // compiler from hoisting the plugin's loop-invariant parameter loads and
    // set-up out of it.  The producers never wait, so the wait ends; it is
    // bounded all the same, and a consumer that reaches the bound sets
    // MPPI_STATUS_WAIT_TIMEOUT in its agent's status, which the update copies
    // to the host, so the call fails rather than return results computed
    // from normals that were not there.
    const uint32_t spin_max = a.wait_limit ? a.wait_limit : (1u << 20);
    const uint32_t wave = (threadIdx.x & (SAMPLES - 1)) >> 6;
    // The consumers are the critical path (their timesteps depend on each
    // other; the producers run ahead and have slack), so they take the
    // higher wavefront priority: the SIMD issues a consumer instruction
    // whenever one is ready and fills the gaps with producer work.
    if (producer) {
        const uint32_t ph_e = mppi_philox_lane_key(step, a.key0);
#pragma unroll 1
        for (uint32_t g = 0; g < n_groups; ++g) {
#pragma unroll
            for (int q = 0; q < MPPI_BG; ++q) {
                const uint32_t b = g * MPPI_BG + q;
                mppi_f32x4 v = mppi_normal4_words(mppi_philox_lanes(k, ph_e, s_ph[b], a.key0, a.key1));
                // samples past K keep zero noise, so the tile's noise sums
                // need no select
                if (!active) v = mppi_f32x4{0.0f, 0.0f, 0.0f, 0.0f};
                if (MPPI_FUSED_OVERHANG_OK || (b + 1) * 4 <= H * CDIM) {
                    reinterpret_cast<float4*>(row)[b] = float4{v.x, v.y, v.z, v.w};
                } else {
                    const float vv[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
                    for (int j = 0; j < 4; ++j)
                        if (b * 4 + j < H * CDIM) row[b * 4 + j] = vv[j];
                }
            }
            // The count of finished groups is published after each of the
            // first MPPI_PHASE_G groups, so that the consumers start after
            // one group, then every MPPI_PHASE_G groups and after the last.
            const uint32_t done = g + 1;
            if (done <= MPPI_PHASE_G || done % MPPI_PHASE_G == 0 || done == n_groups) {
                __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
                if ((threadIdx.x & 63) == 0)
                    __hip_atomic_store(&s_prog[wave], done, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
            }
        }
    } else {
        __builtin_amdgcn_s_setprio(3);
        float scale[MPPI_COV_FULL ? CDIM * CDIM : CDIM], lo[CDIM], hi[CDIM];
#pragma unroll
        for (int i = 0; i < (MPPI_COV_FULL ? CDIM * CDIM : CDIM); ++i) scale[i] = a.noise_scale[i];
#pragma unroll
        for (int i = 0; i < CDIM; ++i) {
            lo[i] = a.u_min[i];
            hi[i] = a.u_max[i];
        }
        float x[SDIM];
#pragma unroll
        for (int i = 0; i < SDIM; ++i) x[i] = a.x[size_t(slot) * SDIM + i];
        // Samples past K run the loop too (their noise is zero and their
        // cost is not stored), so the loop holds no per-lane branch around
        // the barriers.  Blocks of four values past the horizon read the
        // last block inside it (their timesteps are skipped).
        const uint32_t last_b = (H * CDIM - 1) / 4;
        uint32_t ready = 0, zero = 0;
        bool gave_up = false;
        for (uint32_t t0 = 0; t0 < H; t0 += MPPI_G) {
            const uint32_t g = t0 / MPPI_G;
            const uint32_t need = g + 1 + a.wait_bias;
            if (ready < need) {
                uint32_t got = __hip_atomic_load(&s_prog[wave], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
                uint32_t spin = 0;
                for (; got < need && spin < spin_max; ++spin)
                    got = __hip_atomic_load(&s_prog[wave], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
                if (got < need) {
                    // Gave up: every later group counts as ready (the rows
                    // are read as they are); the agent is flagged after the
                    // loop (a global atomic in the loop could alias the
                    // plugin's parameter blob and keep its loads in the loop).
                    gave_up = true;
                    got = 0x7fffffffu;
                }
                ready = got;
                zero = got & 0x80000000u;
            }
            float z[MPPI_NG], un[MPPI_NG];
#pragma unroll
            for (int q = 0; q < MPPI_BG; ++q) {
                const uint32_t b0 = g * MPPI_BG + q, b = b0 < last_b ? b0 : last_b;
                const float4 zq = reinterpret_cast<const float4*>(row + zero)[b];
                const float4 uq = reinterpret_cast<const float4*>(s_un)[b];
                z[4 * q + 0] = zq.x; z[4 * q + 1] = zq.y;
                z[4 * q + 2] = zq.z; z[4 * q + 3] = zq.w;
                un[4 * q + 0] = uq.x; un[4 * q + 1] = uq.y;
                un[4 * q + 2] = uq.z; un[4 * q + 3] = uq.w;
            }
#pragma unroll
            for (int gg = 0; gg < MPPI_G; ++gg) {
                const uint32_t t = t0 + gg;
                if (t < H) {
                    float u[CDIM];
                    mppi_make_control(un + gg * CDIM, z + gg * CDIM, scale, lo, hi, u);
                    mppi_rng rng{t << 8, k, step, dyn_tag, a.key0, a.key1};
                    MPPI_DYNAMICS(mppi_state_ref{x}, u, mppi_blob_ref{model}, rng);
                    total += mppi_stage(mppi_form_timed{}, mppi_cstate_ref{x}, (const float*)u,
                                        mppi_blob_ref{model}, mppi_blob_ref{cost},
                                        mppi_time{t, step});
                }
            }
        }
        total += mppi_terminal(mppi_form_timed{}, mppi_cstate_ref{x}, mppi_blob_ref{model},
                               mppi_blob_ref{cost}, mppi_time{H, step});
        if (active) a.costs[d.cost_off + k] = total;
        if (gave_up && (threadIdx.x & 63) == 0)
            __hip_atomic_fetch_or(&a.agents[slot].status, MPPI_STATUS_WAIT_TIMEOUT, __ATOMIC_RELAXED,
                                  __HIP_MEMORY_SCOPE_AGENT);
    }
    __builtin_amdgcn_s_setprio(0);
    mppi_fused_stats<SAMPLES, THREADS>(a, agent, d, blk, active, total, s_z);
}

// ---------------------------------------------------------------------------
// Finalize (mppi_finalize_agents, one workgroup per agent): sum the
// weighted-noise partials over sample chunks, divide by the normaliser, write
// u_next and the diagnostics; in closed loop also log history, step the plant
// and shift.
// ---------------------------------------------------------------------------
// zs[col] = (sum_r scale_r * rows[r * W + col]) / S for col < W (W a
// multiple of 4), summed in a fixed order; scale_r = scale[r * sstride], or
// 1 when scale is null.  Rows were
// written by other workgroups, possibly on other XCDs: each thread owns a
// float4 of columns and a stride of rows and keeps 16 row loads in flight,
// so the sum costs a few memory latencies rather than one per row.
template <int BLOCK, int U = 16>
__device__ void mppi_sum_rows(const float* rows, uint32_t n_rows, uint32_t W, const float* scale,
                              uint32_t sstride, float S, float* zs) {
    __shared__ float4 s_part[BLOCK];
    const uint32_t Q = W / 4;                            // column quads
    const uint32_t T = Q < BLOCK ? BLOCK / Q : 1;        // threads per quad
    const uint32_t quad0 = threadIdx.x / T, tq = threadIdx.x % T;
    for (uint32_t qb = 0; qb < Q; qb += BLOCK / T) {
        const uint32_t q = qb + quad0;
        float4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
        if (q < Q && quad0 < BLOCK / T) {
            const float4* base = reinterpret_cast<const float4*>(rows) + q;
            uint32_t r = tq;
            for (; r + (U - 1) * T < n_rows; r += U * T) {
                float4 v[U];
#pragma unroll
                for (int u = 0; u < U; ++u) v[u] = base[size_t(r + u * T) * Q];
#pragma unroll
                for (int u = 0; u < U; ++u) {
                    const float f = scale ? scale[(r + u * T) * sstride] : 1.0f;
                    acc.x = mppi_fma(f, v[u].x, acc.x);
                    acc.y = mppi_fma(f, v[u].y, acc.y);
                    acc.z = mppi_fma(f, v[u].z, acc.z);
                    acc.w = mppi_fma(f, v[u].w, acc.w);
                }
            }
            for (; r < n_rows; r += T) {
                const float4 v = base[size_t(r) * Q];
                const float f = scale ? scale[r * sstride] : 1.0f;
                acc.x = mppi_fma(f, v.x, acc.x);
                acc.y = mppi_fma(f, v.y, acc.y);
                acc.z = mppi_fma(f, v.z, acc.z);
                acc.w = mppi_fma(f, v.w, acc.w);
            }
        }
        s_part[threadIdx.x] = acc;
        __syncthreads();
        if (tq == 0 && q < Q && quad0 < BLOCK / T) {
            float4 t = s_part[threadIdx.x];
            for (uint32_t i = 1; i < T; ++i) {
                const float4 p = s_part[threadIdx.x + i];
                t.x += p.x; t.y += p.y; t.z += p.z; t.w += p.w;
            }
            zs[4 * q + 0] = t.x / S;
            zs[4 * q + 1] = t.y / S;
            zs[4 * q + 2] = t.z / S;
            zs[4 * q + 3] = t.w / S;
        }
        __syncthreads();
    }
}

template <int BLOCK, typename T>
__device__ void mppi_sum_partials(const mppi_step_args& a, const mppi_agent_desc& d,
                                  double weight_sum, float* zs) {
    __shared__ T s_col[BLOCK];
    const uint32_t W = d.NB * 4;
    const T S = T(weight_sum);
    const float* pb = a.partials + d.part_off;
    // P columns per pass, R threads per column; each thread keeps eight
    // independent accumulators (combined in a fixed order) so that eight
    // loads are in flight at a time.
    const uint32_t P = W < BLOCK ? W : BLOCK;
    const uint32_t R = BLOCK / P;
    const uint32_t c = threadIdx.x % P, r = threadIdx.x / P;
    for (uint32_t base = 0; base < W; base += P) {
        const uint32_t col = base + c;
        if (r < R && col < W) {
            T acc[8];
#pragma unroll
            for (int u = 0; u < 8; ++u) acc[u] = T(0);
            uint32_t ch = r;
            for (; ch + 7 * R < d.n_chunks; ch += 8 * R) {
#pragma unroll
                for (int u = 0; u < 8; ++u) acc[u] += T(pb[size_t(ch + u * R) * W + col]);
            }
            for (; ch < d.n_chunks; ch += R) acc[0] += T(pb[size_t(ch) * W + col]);
            s_col[r * P + c] = ((acc[0] + acc[1]) + (acc[2] + acc[3])) + ((acc[4] + acc[5]) + (acc[6] + acc[7]));
        }
        __syncthreads();
        if (threadIdx.x < P && base + threadIdx.x < W) {
            T sum = T(0);
            for (uint32_t q = 0; q < R; ++q) sum += s_col[q * P + threadIdx.x];
            zs[base + threadIdx.x] = float(sum / S);
        }
        __syncthreads();
    }
}

// Closed loop, one thread: log the history row (state, first control, best
// cost), step the plant with the plugin dynamics and the first control, and
// advance the agent's step and history counters.  The inputs are read by
// mppi_plant_load, which a caller may issue early to hide their latency.
struct mppi_plant_in {
    float x[SDIM];
    uint32_t step, row;
};

__device__ inline mppi_plant_in mppi_plant_load(const mppi_step_args& a, const mppi_agent_desc& d) {
    mppi_plant_in in;
    const float* xs = a.x + size_t(d.slot) * SDIM;
#pragma unroll
    for (int i = 0; i < SDIM; ++i) in.x[i] = xs[i];
    in.step = a.agents[d.slot].step;
    in.row = a.agents[d.slot].hist;
    return in;
}

__device__ inline void mppi_plant_step(const mppi_step_args& a, uint32_t agent, const mppi_agent_desc& d,
                                       const float* u0, const mppi_plant_in& in, float best) {
    const uint32_t slot = d.slot, row = in.row;
    float x[SDIM];
    float u[CDIM];
#pragma unroll
    for (int i = 0; i < SDIM; ++i) x[i] = in.x[i];
#pragma unroll
    for (int i = 0; i < CDIM; ++i) u[i] = u0[i];
    if (a.hist_states)
        for (int i = 0; i < SDIM; ++i) a.hist_states[(size_t(row) * a.N + agent) * SDIM + i] = x[i];
    if (a.hist_controls)
        for (int i = 0; i < CDIM; ++i) a.hist_controls[(size_t(row) * a.N + agent) * CDIM + i] = u[i];
    if (a.hist_costs) a.hist_costs[size_t(row) * a.N + agent] = best;
    mppi_rng rng{0u, 0xffffffffu, in.step, (MPPI_TAG_PLANT << 28) | (slot & MPPI_AGENT_MASK), a.key0, a.key1};
    MPPI_DYNAMICS(mppi_state_ref{x}, u, mppi_blob_ref{a.model + slot * a.model_stride}, rng);
    float* xs = a.x + size_t(slot) * SDIM;
#pragma unroll
    for (int i = 0; i < SDIM; ++i) xs[i] = x[i];
    if (a.terminal)
        for (int i = 0; i < SDIM; ++i) a.terminal[size_t(agent) * SDIM + i] = x[i];
    mppi_agent_state* ag = &a.agents[slot];
    ag->hist = row + 1;
    ag->step = in.step + 1;
}

// u = clamp(u_nom + sigma z) (diagonal) or clamp(u_nom + L z) (full) for
// control n = t*C + i of a sequence, z the averaged noise of timestep t.
__device__ inline float mppi_update_control(const mppi_step_args& a, float un, const float* zt, uint32_t i) {
#if MPPI_COV_FULL
    float e = 0.0f;
    for (uint32_t j = 0; j <= i; ++j) e = mppi_fma(a.noise_scale[i * CDIM + j], zt[j], e);
#else
    const float e = a.noise_scale[i] * zt[i];
#endif
    return mppi_clampf(un + e, a.u_min[i], a.u_max[i]);
}

// Apply the averaged noise zs (H*C values, already divided by the
// normaliser): u_next, diagnostics, and in closed loop the history, the plant
// step and the warm-start shift.
template <int BLOCK>
__device__ void mppi_apply_update(const mppi_step_args& a, uint32_t agent, const mppi_agent_desc& d,
                                  const float* zs) {
    const uint32_t H = d.H, HC = H * CDIM;
    mppi_agent_state* ag = &a.agents[d.slot];
    float* uo = a.u_out + d.out_off;
    float* un = a.u_nom + d.seq_off;
    for (uint32_t n = threadIdx.x; n < HC; n += BLOCK) {
        const uint32_t t = n / CDIM, i = n % CDIM;
        uo[n] = mppi_update_control(a, un[n], zs + t * CDIM, i);
    }
    __syncthreads();
    if (threadIdx.x == 0 && a.diag) {
        a.diag[agent * 3 + 0] = float(ag->min_cost);
        a.diag[agent * 3 + 1] = float(ag->mean_cost);
        a.diag[agent * 3 + 2] = float(ag->ess);
    }
    if (threadIdx.x == 0 && a.status_out) a.status_out[agent] = ag->status;
    if (a.first_out && threadIdx.x < CDIM) a.first_out[agent * CDIM + threadIdx.x] = uo[threadIdx.x];
    if (a.closed_loop) {
        if (threadIdx.x == 0) mppi_plant_step(a, agent, d, uo, mppi_plant_load(a, d), float(ag->min_cost));
        float* next = a.u_nom_next ? a.u_nom_next + d.seq_off : un;
        for (uint32_t n = threadIdx.x; n < HC; n += BLOCK) {
            const uint32_t t = n / CDIM, i = n % CDIM;
            next[n] = uo[(t + 1 < H ? t + 1 : H - 1) * CDIM + i];
        }
    }
}

// The update of one agent after its reduce tiles (mppi_finalize_agents, one
// workgroup per agent): the normaliser into the agent's state, the reduce
// partials summed over sample chunks and divided by it, the update applied.
template <int BLOCK, typename T>
__device__ void mppi_finalize(const mppi_step_args& a, uint32_t agent, const mppi_agent_desc& d, float* zs) {
    const mppi_norm<T> n = mppi_normaliser<BLOCK, T>(a, d);
    mppi_store_normaliser(a, d, n);
    if (sizeof(T) == 8) mppi_sum_partials<BLOCK, double>(a, d, double(n.S), zs);
    else mppi_sum_rows<BLOCK>(a.partials + d.part_off, d.n_chunks, d.NB * 4, nullptr, 1, float(n.S), zs);
    __syncthreads();
    mppi_apply_update<BLOCK>(a, agent, d, zs);
}

// ---------------------------------------------------------------------------
// One-pass step (FUSED rollout tiles, then mppi_fused_combine).  A tile's
// weights are taken against its own minimum m_g, e_k = exp(-(c_k - m_g)/lambda),
// and its weighted noise sums V_g = sum_k e_k z_k are formed from the normals
// kept in LDS; the tile writes (m_g, S_g, C_g, Q_g) and V_g.  A second, small
// launch combines the tiles: with m = min m_g and f_g = exp(-(m_g - m)/lambda),
// S = sum f_g S_g and the update's average is sum f_g V_g / S.  A tile with no
// finite cost keeps the plain sum of its noise and gets f = 0, which matters
// only when no cost is finite (then f = 1 for every tile and S = K: uniform
// weights, as in the two-kernel step).  The combine is spread over several
// workgroups, one per slice of timesteps, so reading every tile's sums costs
// a few memory latencies on many compute units instead of one compute unit
// reading them all after the slowest tile.
// ---------------------------------------------------------------------------
// SAMPLES samples, THREADS >= SAMPLES threads: thread i < SAMPLES holds
// sample i's cost (`active`, `total`); the other threads add nothing to the
// statistics and share the work of the noise sums.
// Softmax partials of one tile against the tile's own minimum m_g: writes
// (m_g, S_g, C_g, Q_g) to the tile's statistics row and returns the thread's
// weight factor e = exp(-(c - m_g)/lambda), 0 for a lane without a finite
// cost, or, in a tile with no finite cost, 1 for every active lane (the tile
// then keeps the plain sum of its noise).  When s_e is not null the factor is
// also stored at s_e[threadIdx.x] before the function's last barrier.
template <int THREADS>
__device__ float mppi_tile_stats(const mppi_step_args& a, const mppi_agent_desc& d, uint32_t blk, bool active,
                                 float total, float* s_e) {
    constexpr int WAVES = THREADS / 64;
    __shared__ float s_red[4][WAVES];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    float c = (active && mppi_isfinite(total)) ? total : __builtin_inff();
    float m = mppi_wave_min(c);
    if (lane == 0) s_red[0][wave] = m;
    __syncthreads();
    m = s_red[0][0];
#pragma unroll
    for (int w = 1; w < WAVES; ++w) m = s_red[0][w] < m ? s_red[0][w] : m;
    const float e = (mppi_isfinite(c) && mppi_isfinite(m)) ? mppi_expf(-(c - m) / a.lambda) : 0.0f;
    const float ew = mppi_isfinite(m) ? e : (active ? 1.0f : 0.0f);
    if (s_e) s_e[threadIdx.x] = ew;
    const float se = mppi_wave_sum(e);
    const float sq = mppi_wave_sum(e * e);
    const float sc = mppi_wave_sum(active ? total : 0.0f);
    if (lane == 0) { s_red[1][wave] = se; s_red[2][wave] = sc; s_red[3][wave] = sq; }
    __syncthreads();
    if (threadIdx.x == 0) {
        float S = 0.0f, Cs = 0.0f, Q = 0.0f;
#pragma unroll
        for (int w = 0; w < WAVES; ++w) { S += s_red[1][w]; Cs += s_red[2][w]; Q += s_red[3][w]; }
        float* bs = a.blk_stats + (d.blk_off + blk) * 4;
        bs[0] = m; bs[1] = S; bs[2] = Cs; bs[3] = Q;
    }
    return ew;
}

template <int SAMPLES, int THREADS>
__device__ void mppi_fused_stats(const mppi_step_args& a, uint32_t agent, const mppi_agent_desc& d,
                                 uint32_t blk, bool active, float total, float* s_z) {
    constexpr int ZS = MPPI_FUSED_ZS;
    __shared__ __attribute__((aligned(16))) float s_e[SAMPLES];
    const uint32_t W = d.NB * 4;
    active = active && threadIdx.x < uint32_t(SAMPLES);
    (void)mppi_tile_stats<THREADS>(a, d, blk, active, total, threadIdx.x < uint32_t(SAMPLES) ? s_e : nullptr);
    // V_g[v] = sum over the tile's lanes of e_l z_l[v].  Thread (quad, sg)
    // sums four columns over the 32 lanes of sample group sg in order (two
    // interleaved accumulators), reading four normals and four weights per
    // LDS access; the eight groups are then added in order.  Lanes past K
    // have e = 0 and zero noise, and a lane with a non-finite cost has e = 0
    // and finite noise, so every term is finite and no select is needed.
    {
        constexpr uint32_t QUADS = MPPI_FUSED_MAX_HC / 4, GROUPS = THREADS / QUADS, PER = SAMPLES / GROUPS;
        static_assert(PER % 4 == 0, "sample groups of whole quads");
        __shared__ float4 s_q[GROUPS][QUADS];
        const uint32_t quad = threadIdx.x % QUADS, sg = threadIdx.x / QUADS, Q = W / 4;
        float4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
        if (quad < Q) {
            float4 a0 = {0.0f, 0.0f, 0.0f, 0.0f}, a1 = {0.0f, 0.0f, 0.0f, 0.0f};
            const float4* e4 = reinterpret_cast<const float4*>(s_e);
            const uint32_t l0 = sg * PER;
            for (uint32_t l = l0; l < l0 + PER; l += 4) {
                const float4 e = e4[l / 4];
                const float ev[4] = {e.x, e.y, e.z, e.w};
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    const float4 z = reinterpret_cast<const float4*>(s_z + (l + j) * ZS)[quad];
                    float4& t = (j & 1) ? a1 : a0;
                    t.x = mppi_fma(ev[j], z.x, t.x);
                    t.y = mppi_fma(ev[j], z.y, t.y);
                    t.z = mppi_fma(ev[j], z.z, t.z);
                    t.w = mppi_fma(ev[j], z.w, t.w);
                }
            }
            acc = float4{a0.x + a1.x, a0.y + a1.y, a0.z + a1.z, a0.w + a1.w};
        }
        s_q[sg][quad] = acc;
        __syncthreads();
        if (sg == 0 && quad < Q) {
            float4 t = s_q[0][quad];
            for (uint32_t g = 1; g < GROUPS; ++g) {
                const float4 u = s_q[g][quad];
                t.x += u.x; t.y += u.y; t.z += u.z; t.w += u.w;
            }
            // Slice-major rows (mppi_combine_width): the quad's columns lie
            // in one slice, as a slice spans a multiple of 32 columns or the
            // whole sequence.  Columns past H*C hold normals of the next
            // group or zeros; the combine never reads them.
            constexpr uint32_t SPAN = mppi_combine_tpw(CDIM) * CDIM, WIDTH = mppi_combine_width(CDIM);
            const uint32_t c = 4 * quad, sl = c / SPAN;
            reinterpret_cast<float4*>(a.partials + d.part_off + (size_t(sl) * d.n_blk + blk) * WIDTH +
                                      (c - sl * SPAN))[0] = t;
        }
    }
    // The rows are read by mppi_fused_combine, a later launch.
}

// Combine of the one-pass step for one agent: workgroup `slice` handles the
// timesteps [t0, t1) = [slice*TPW, min((slice+1)*TPW, H)), whose columns of
// the tiles' sums are one block of whole 128-byte lines (slice-major rows,
// mppi_combine_width).  Every workgroup reads all tile statistics (one row
// per thread, 16 bytes per tile) and forms the same normaliser in the same
// order; its columns' sums over the tiles are requested before that, so the
// two reads overlap.  Slice 0 also writes the diagnostics and agent state and
// steps the plant.  In closed loop the warm start shifted by one timestep
// goes to u_nom_next: each slice writes its own timesteps t >= 1 to t - 1,
// and the slice holding H - 1 also repeats it at H - 1, so no slice reads
// another's timesteps.
constexpr uint32_t MPPI_COMBINE_TPW = mppi_combine_tpw(CDIM);

template <int BLOCK, bool ONE>
__device__ void mppi_fused_combine_slice(const mppi_step_args& a, uint32_t agent, uint32_t slice) {
    constexpr int WAVES = BLOCK / 64;
    constexpr uint32_t TPW = MPPI_COMBINE_TPW;
    constexpr uint32_t MAXCOL = mppi_combine_width(CDIM);    // row width, >= TPW * CDIM
    constexpr uint32_t RG = BLOCK / MAXCOL;                   // row groups
    static_assert(RG >= 1 && MAXCOL <= MPPI_FUSED_MAX_HC, "a slice fits the workgroup");
    // rows (tiles) per thread: at most BLOCK tiles over RG row groups,
    // loaded in batches of at most 32
    constexpr int RPT = int((BLOCK + RG - 1) / RG);
    constexpr int BATCH = RPT < 32 ? RPT : 32;
    __shared__ float s_f[BLOCK];
    __shared__ float s_min[WAVES];
    __shared__ float s_sum[3][WAVES];
    __shared__ float s_part[BLOCK];
    __shared__ float s_z[MAXCOL];
    __shared__ float s_u[MAXCOL];
    const mppi_agent_desc d = ONE ? a.desc0 : a.desc[agent];
    const uint32_t H = d.H, t0 = slice * TPW;
    if (t0 >= H) return;                               // the whole workgroup
    const uint32_t t1 = t0 + TPW < H ? t0 + TPW : H;   // own timesteps [t0, t1)
    const uint32_t c0 = t0 * CDIM, ncol = (t1 - t0) * CDIM;
    const uint32_t n_rows = d.n_blk;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;

    // Every global read is issued here, before any reduction: the tiles'
    // sums of this slice's columns (the first batch), the tile statistics,
    // this thread's nominal control and its bounds and noise factors, and
    // (slice 0) the plant's state, so that the kernel waits on memory once.
    const uint32_t seq = d.seq_off;
    const uint32_t ci = threadIdx.x % CDIM;
    float un = 0.0f, lo = 0.0f, hi = 0.0f;
    float sc[MPPI_COV_FULL ? CDIM : 1];
    if (threadIdx.x < ncol) {
        un = a.u_nom[seq + c0 + threadIdx.x];
        lo = a.u_min[ci];
        hi = a.u_max[ci];
#if MPPI_COV_FULL
#pragma unroll
        for (uint32_t j = 0; j < CDIM; ++j) sc[j] = j <= ci ? a.noise_scale[ci * CDIM + j] : 0.0f;
#else
        sc[0] = a.noise_scale[ci];
#endif
    }
    // Columns are spread over MAXCOL threads whatever ncol is (threads of
    // columns past ncol load nothing), so the layout is fixed at compile time.
    const uint32_t col = threadIdx.x % MAXCOL, rg = threadIdx.x / MAXCOL;
    const bool col_ok = col < ncol && rg < RG;
    const float* rows = a.partials + d.part_off + size_t(slice) * n_rows * MAXCOL + col;
    float v[BATCH];
#pragma unroll
    for (int q = 0; q < BATCH; ++q) {
        const uint32_t r = rg + uint32_t(q) * RG;
        v[q] = (col_ok && r < n_rows) ? rows[size_t(r) * MAXCOL] : 0.0f;
    }
    float4 st = {__builtin_inff(), 0.0f, 0.0f, 0.0f};
    if (threadIdx.x < n_rows) st = reinterpret_cast<const float4*>(a.blk_stats + d.blk_off * 4)[threadIdx.x];
    mppi_plant_in plant{};
    if (slice == 0 && threadIdx.x == 0 && a.closed_loop) plant = mppi_plant_load(a, d);
    float mg = mppi_wave_min(st.x);
    if (lane == 0) s_min[wave] = mg;
    __syncthreads();
    mg = s_min[0];
#pragma unroll
    for (int w = 1; w < WAVES; ++w) mg = s_min[w] < mg ? s_min[w] : mg;
    const float f = (mppi_isfinite(st.x) && mppi_isfinite(mg)) ? mppi_expf(-(st.x - mg) / a.lambda) : 0.0f;
    const bool uniform = !mppi_isfinite(mg);
    s_f[threadIdx.x] = uniform ? 1.0f : f;
    float S = mppi_wave_sum(st.y * f);
    float Cs = mppi_wave_sum(st.z);
    float Q = mppi_wave_sum(st.w * (f * f));
    if (lane == 0) { s_sum[0][wave] = S; s_sum[1][wave] = Cs; s_sum[2][wave] = Q; }
    __syncthreads();
    S = s_sum[0][0]; Cs = s_sum[1][0]; Q = s_sum[2][0];
#pragma unroll
    for (int w = 1; w < WAVES; ++w) { S += s_sum[0][w]; Cs += s_sum[1][w]; Q += s_sum[2][w]; }
    if (uniform) {
        S = float(d.K);
        Q = float(d.K);
    }
    // the scaled sums of this workgroup's columns: rows in order per thread,
    // then row groups in order
    float acc = 0.0f;
#pragma unroll
    for (int q = 0; q < BATCH; ++q) {
        const uint32_t r = rg + uint32_t(q) * RG;
        if (col_ok && r < n_rows) acc = mppi_fma(s_f[r], v[q], acc);
    }
    for (int q0 = BATCH; q0 < RPT && rg + uint32_t(q0) * RG < n_rows; q0 += BATCH) {
#pragma unroll
        for (int q = 0; q < BATCH; ++q) {
            const uint32_t r = rg + uint32_t(q0 + q) * RG;
            v[q] = (col_ok && q0 + q < RPT && r < n_rows) ? rows[size_t(r) * MAXCOL] : 0.0f;
        }
#pragma unroll
        for (int q = 0; q < BATCH; ++q) {
            const uint32_t r = rg + uint32_t(q0 + q) * RG;
            if (col_ok && q0 + q < RPT && r < n_rows) acc = mppi_fma(s_f[r], v[q], acc);
        }
    }
    s_part[threadIdx.x] = acc;
    __syncthreads();
    if (threadIdx.x < ncol) {
        float t = 0.0f;
#pragma unroll
        for (uint32_t g = 0; g < RG; ++g) t += s_part[g * MAXCOL + threadIdx.x];
        s_z[threadIdx.x] = t / S;
    }
    __syncthreads();
    if (threadIdx.x < ncol) {
        const uint32_t n = threadIdx.x;
        const float* zt = s_z + (n - ci);
#if MPPI_COV_FULL
        float e = 0.0f;
#pragma unroll
        for (uint32_t j = 0; j < CDIM; ++j)
            if (j <= ci) e = mppi_fma(sc[j], zt[j], e);
#else
        const float e = sc[0] * zt[ci];
#endif
        const float u = mppi_clampf(un + e, lo, hi);
        s_u[n] = u;
        a.u_out[d.out_off + c0 + n] = u;
        if (a.closed_loop && a.u_nom_next) {
            // the shifted warm start: timestep t to t - 1, and H - 1 kept
            const uint32_t t = t0 + n / CDIM;
            if (t >= 1) a.u_nom_next[seq + c0 + n - CDIM] = u;
            if (t + 1 == H) a.u_nom_next[seq + c0 + n] = u;
        }
    }
    __syncthreads();
    if (slice == 0) {
        mppi_agent_state* ag = &a.agents[d.slot];
        if (threadIdx.x == 0) {
            ag->min_cost = mg;
            ag->weight_sum = double(S);
            ag->mean_cost = double(Cs / float(d.K));
            ag->ess = double((S * S) / Q);
            if (a.diag) {
                a.diag[agent * 3 + 0] = mg;
                a.diag[agent * 3 + 1] = float(double(Cs / float(d.K)));
                a.diag[agent * 3 + 2] = float(double((S * S) / Q));
            }
            if (a.status_out) a.status_out[agent] = ag->status;
            if (a.closed_loop) mppi_plant_step(a, agent, d, s_u, plant, mg);
        }
        if (a.first_out && threadIdx.x < CDIM) a.first_out[agent * CDIM + threadIdx.x] = s_u[threadIdx.x];
    }
}

// ---------------------------------------------------------------------------
// Reduce tile: samples of chunk `chunk` times noise blocks of group `grp`.
// ---------------------------------------------------------------------------
template <int BLOCK>
__device__ void mppi_reduce_tile(const mppi_step_args& a, uint32_t agent, uint32_t chunk,
                                 uint32_t grp) {
    const mppi_agent_desc d = a.desc[agent];
    const uint32_t K = d.K, NB = d.NB;
    const mppi_agent_state* ag = &a.agents[d.slot];
    const uint32_t step = ag->step;
    // The agent's normaliser, formed here from the rollout's rows.
    float m, Sf;
    if (d.K >= MPPI_FP64_MIN_K) {
        const mppi_norm<double> n = mppi_normaliser<BLOCK, double>(a, d);
        m = n.m;
        Sf = float(n.S);
    } else {
        const mppi_norm<float> n = mppi_normaliser<BLOCK, float>(a, d);
        m = n.m;
        Sf = n.S;
    }
    const bool uniform = !mppi_isfinite(m);
    const uint32_t tag_agent = (MPPI_TAG_CONTROL_NOISE << 28) | (d.slot & MPPI_AGENT_MASK);
    const uint32_t b_begin = grp * MPPI_REDUCE_BPG;
    const float* costs = a.costs + d.cost_off;

    float acc[MPPI_REDUCE_BPG][4];
#pragma unroll
    for (int q = 0; q < MPPI_REDUCE_BPG; ++q)
#pragma unroll
        for (int j = 0; j < 4; ++j) acc[q][j] = 0.0f;

    const uint32_t k_base = chunk * BLOCK * d.spt + threadIdx.x;
    for (uint32_t s = 0; s < d.spt; ++s) {
        const uint32_t k = k_base + s * BLOCK;
        if (k >= K) break;
        const float c = costs[k];
        float w;
        if (uniform) w = 1.0f;
        else w = mppi_isfinite(c) ? mppi_expf(-(c - m) / a.lambda) : 0.0f;
        if (grp == 0 && a.weights) a.weights[d.w_off + k] = w / Sf;
        // A weight that underflowed to exactly 0 adds exactly nothing: a
        // wavefront whose samples all have zero weight skips regenerating
        // their noise (most samples, when lambda is small against the
        // spread of the costs).
        if (!mppi_any(w != 0.0f)) continue;
#pragma unroll
        for (int q = 0; q < MPPI_REDUCE_BPG; ++q) {
            if (b_begin + q < NB) {
                mppi_f32x4 z = mppi_normal4(mppi_u32x4{b_begin + q, k, step, tag_agent},
                                            a.key0, a.key1);
                acc[q][0] = mppi_fma(w, z.x, acc[q][0]);
                acc[q][1] = mppi_fma(w, z.y, acc[q][1]);
                acc[q][2] = mppi_fma(w, z.z, acc[q][2]);
                acc[q][3] = mppi_fma(w, z.w, acc[q][3]);
            }
        }
    }

    // ---- workgroup sum of the accumulators, transposed through LDS ----
    __shared__ float s_acc[BLOCK * (MPPI_NV + 1)];
#pragma unroll
    for (int q = 0; q < MPPI_REDUCE_BPG; ++q)
#pragma unroll
        for (int j = 0; j < 4; ++j) s_acc[threadIdx.x * (MPPI_NV + 1) + q * 4 + j] = acc[q][j];
    __syncthreads();
    {
        constexpr int NSEG = BLOCK / MPPI_NV;     // row segments
        constexpr int ROWS = BLOCK / NSEG;        // rows per segment
        const int v = threadIdx.x % MPPI_NV, seg = threadIdx.x / MPPI_NV;
        float sum = 0.0f;
        for (int r = 0; r < ROWS; ++r) sum += s_acc[(seg * ROWS + r) * (MPPI_NV + 1) + v];
        __syncthreads();
        s_acc[seg * MPPI_NV + v] = sum;
        __syncthreads();
        if (threadIdx.x < MPPI_NV) {
            float tot = 0.0f;
#pragma unroll
            for (int g = 0; g < NSEG; ++g) tot += s_acc[g * MPPI_NV + threadIdx.x];
            const uint32_t b = b_begin + threadIdx.x / 4;
            if (b < NB) a.partials[d.part_off + (size_t(chunk) * NB + b) * 4 + (threadIdx.x & 3)] = tot;
        }
    }
    // The rows are summed by mppi_finalize_agents, a later launch.
}

// ---------------------------------------------------------------------------
// One-launch step tile (the uneven batch's one-launch form): samples
// [blk*BLOCK, blk*BLOCK + BLOCK) of one agent.  The tile rolls out its
// samples, takes their weights against its own minimum m_g, regenerates
// their noise and forms its weighted noise sums V_g (one row of 4 NB
// floats), as the one-pass step's tiles do with the noise kept in LDS; the
// last tile of the agent to finish (an arrival counter; nothing waits) then
// combines the tiles: m = min m_g, f_g = exp(-(m_g - m)/lambda),
// S = sum f_g S_g, and the update's average is sum f_g V_g / S.  Every tile
// of the agent has read the agent's state, nominal and step counter before
// it arrives, so the last one updates them in place.  K < 2^18 (FP32
// cross-tile sums).
// ---------------------------------------------------------------------------
// V_g[n] = sum over the tile's lanes of e_l z_l[n], n < 4 NB, the noise
// regenerated from its counters: per group of MPPI_REDUCE_BPG noise blocks
// every lane forms e z for its 32 values and the workgroup sums them through
// an LDS transpose (the reduce tile's scheme, one sample per lane).
// s_acc: BLOCK * (MPPI_NV + 1) floats of LDS.
template <int BLOCK>
__device__ void mppi_tile_noise_sums(const mppi_step_args& a, const mppi_agent_desc& d, uint32_t blk, float e,
                                     float* s_acc) {
    const uint32_t k = blk * BLOCK + threadIdx.x;
    const uint32_t NB = d.NB;
    const uint32_t step = a.agents[d.slot].step;
    const uint32_t tag_agent = (MPPI_TAG_CONTROL_NOISE << 28) | (d.slot & MPPI_AGENT_MASK);
    float* row = a.partials + d.part_off + size_t(blk) * NB * 4;
    // A wavefront whose samples all have weight 0 adds nothing.
    const bool any = mppi_any(e != 0.0f);
    for (uint32_t b_begin = 0; b_begin < NB; b_begin += MPPI_REDUCE_BPG) {
        float acc[MPPI_REDUCE_BPG][4];
#pragma unroll
        for (int q = 0; q < MPPI_REDUCE_BPG; ++q)
#pragma unroll
            for (int j = 0; j < 4; ++j) acc[q][j] = 0.0f;
        if (any) {
#pragma unroll
            for (int q = 0; q < MPPI_REDUCE_BPG; ++q) {
                if (b_begin + q < NB) {
                    const mppi_f32x4 z = mppi_normal4(mppi_u32x4{b_begin + q, k, step, tag_agent}, a.key0, a.key1);
                    acc[q][0] = e * z.x;
                    acc[q][1] = e * z.y;
                    acc[q][2] = e * z.z;
                    acc[q][3] = e * z.w;
                }
            }
        }
#pragma unroll
        for (int q = 0; q < MPPI_REDUCE_BPG; ++q)
#pragma unroll
            for (int j = 0; j < 4; ++j) s_acc[threadIdx.x * (MPPI_NV + 1) + q * 4 + j] = acc[q][j];
        __syncthreads();
        constexpr int NSEG = BLOCK / MPPI_NV;     // row segments
        constexpr int ROWS = BLOCK / NSEG;        // rows per segment
        const int v = threadIdx.x % MPPI_NV, seg = threadIdx.x / MPPI_NV;
        float sum = 0.0f;
        for (int r = 0; r < ROWS; ++r) sum += s_acc[(seg * ROWS + r) * (MPPI_NV + 1) + v];
        __syncthreads();
        s_acc[seg * MPPI_NV + v] = sum;
        __syncthreads();
        if (threadIdx.x < MPPI_NV) {
            float tot = 0.0f;
#pragma unroll
            for (int g = 0; g < NSEG; ++g) tot += s_acc[g * MPPI_NV + threadIdx.x];
            const uint32_t b = b_begin + threadIdx.x / 4;
            if (b < NB) row[b * 4 + (threadIdx.x & 3)] = tot;
        }
        __syncthreads();                          // s_acc is reused by the next group
    }
}

// The last tile of an agent: normaliser, scaled row sums, update.
template <int BLOCK>
__device__ void mppi_onelaunch_combine(const mppi_step_args& a, uint32_t agent, const mppi_agent_desc& d,
                                       float* zs) {
    const mppi_norm<float> n = mppi_normaliser<BLOCK, float>(a, d);
    mppi_store_normaliser(a, d, n);
    const float m = n.m;
    const bool uniform = !mppi_isfinite(m);
    // Each tile row's factor f_g (the one mppi_normaliser used for S)
    // replaces the row's m_g, which is no longer needed.
    float* st = a.blk_stats + d.blk_off * 4;
    for (uint32_t i = threadIdx.x; i < d.n_blk; i += BLOCK) {
        const float bm = st[i * 4];
        st[i * 4] = uniform ? 1.0f : (mppi_isfinite(bm) ? mppi_expf(-(bm - m) / a.lambda) : 0.0f);
    }
    __syncthreads();
    mppi_sum_rows<BLOCK>(a.partials + d.part_off, d.n_blk, d.NB * 4, st, 4, n.S, zs);
    __syncthreads();
    mppi_apply_update<BLOCK>(a, agent, d, zs);
}

template <int BLOCK>
__device__ void mppi_onelaunch_tile(const mppi_step_args& a, uint32_t agent, uint32_t blk) {
    constexpr int LDS = MPPI_UN_LDS > BLOCK * (MPPI_NV + 1) ? MPPI_UN_LDS : BLOCK * (MPPI_NV + 1);
    __shared__ __attribute__((aligned(16))) float s_lds[LDS];
    __shared__ int s_last;
    const mppi_agent_desc d = a.desc[agent];
    const bool active = blk * BLOCK + threadIdx.x < d.K;
    const float total = mppi_rollout_costs<BLOCK>(a, d, blk, s_lds);
    const float e = mppi_tile_stats<BLOCK>(a, d, blk, active, total, nullptr);   // its barriers free s_lds
    mppi_tile_noise_sums<BLOCK>(a, d, blk, e, s_lds);
    if (threadIdx.x == 0) {
        __threadfence();                          // the row and the statistics before the arrival
        const uint32_t prev = atomicAdd(&a.agents[d.slot].rollout_arrivals, 1u);
        s_last = (prev == d.n_blk - 1);
    }
    __syncthreads();
    if (!__builtin_amdgcn_readfirstlane(s_last)) return;
    __threadfence();
    mppi_onelaunch_combine<BLOCK>(a, agent, d, s_lds);
}

// ---------------------------------------------------------------------------
// Entry points.  Grid-mapped (the single-agent and uniform batch path):
//   rollout grid (max n_blk, N), reduce grid (max n_chunks, max n_groups, N).
// ---------------------------------------------------------------------------
extern "C" __global__ void __launch_bounds__(256) mppi_rollout(mppi_step_args a) {
    const uint32_t agent = blockIdx.y;
    if (blockIdx.x >= a.desc[agent].n_blk) return;
    mppi_rollout_tile<256>(a, agent, blockIdx.x);
}

extern "C" __global__ void __launch_bounds__(256) mppi_reduce(mppi_step_args a) {
    const uint32_t agent = blockIdx.z;
    if (blockIdx.x >= a.desc[agent].n_chunks || blockIdx.y >= a.desc[agent].n_groups) return;
    mppi_reduce_tile<256>(a, agent, blockIdx.x, blockIdx.y);
}

// The update of every agent after the reduce (grid N: one workgroup per
// agent, whatever workgroup size the rollout and reduce tiles had).
extern "C" __global__ void __launch_bounds__(256) mppi_finalize_agents(mppi_step_args a) {
    __shared__ float s_zs[MPPI_MAX_HC + 4];
    const uint32_t agent = blockIdx.x;
    const mppi_agent_desc d = a.desc[agent];
    if (d.K >= MPPI_FP64_MIN_K) mppi_finalize<256, double>(a, agent, d, s_zs);
    else mppi_finalize<256, float>(a, agent, d, s_zs);
}

// Rollout of the one-pass step (grid as mppi_rollout, 512 threads per tile of
// 256 samples; H*C <= MPPI_FUSED_MAX_HC, K < 2^18).
extern "C" __global__ void __launch_bounds__(512) mppi_rollout_fused(mppi_step_args a) {
    const uint32_t agent = blockIdx.y;
    if (blockIdx.x >= a.desc[agent].n_blk) return;
    mppi_fused_tile<256, false>(a, agent, blockIdx.x);
}

// The same for one agent whose descriptor and step counter are in the
// arguments (desc0, step0; grid n_blk x 1).
extern "C" __global__ void __launch_bounds__(512) mppi_rollout_fused_one(mppi_step_args a) {
    mppi_fused_tile<256, true>(a, 0, blockIdx.x);
}

// Per-sample weights of the last iteration, w_k = exp(-(c_k - m)/lambda) / S
// (the reduce kernel's formula), for a one-pass step, which does not form
// them; grid (ceil(K / 256), N).
// Combine of the one-pass step (grid: timestep slices x agents).
extern "C" __global__ void __launch_bounds__(256) mppi_fused_combine(mppi_step_args a) {
    mppi_fused_combine_slice<256, false>(a, blockIdx.y, blockIdx.x);
}

extern "C" __global__ void __launch_bounds__(256) mppi_fused_combine_one(mppi_step_args a) {
    mppi_fused_combine_slice<256, true>(a, 0, blockIdx.x);
}

extern "C" __global__ void __launch_bounds__(256) mppi_weights(mppi_step_args a) {
    const uint32_t agent = blockIdx.y;
    const mppi_agent_desc d = a.desc[agent];
    const uint32_t k = blockIdx.x * 256 + threadIdx.x;
    if (k >= d.K) return;
    const mppi_agent_state* ag = &a.agents[d.slot];
    const float m = float(ag->min_cost);
    const float Sf = float(ag->weight_sum);
    const float c = a.costs[d.cost_off + k];
    float w;
    if (!mppi_isfinite(m)) w = 1.0f;
    else w = mppi_isfinite(c) ? mppi_expf(-(c - m) / a.lambda) : 0.0f;
    a.weights[d.w_off + k] = w / Sf;
}

#if MPPI_BATCH_KERNELS
// Tile-table and persistent kernels of the uneven multi-agent step.  Tiles
// are (agent, x, y) triples: rollout and one-launch step (agent, block, 0),
// reduce (agent, chunk, group).
//   table       workgroup i runs tile i;
//   persistent  a grid of G workgroups takes tiles from a ticket counter
//               (one 64-bit counter per kernel, never reset): ticket t names
//               tile t mod P, P = n_tiles + G, and a workgroup leaves at its
//               first ticket past the last tile.  Tickets are handed out in
//               order, so a launch's first n_tiles tickets are the tiles and
//               the next G are one exit each; each workgroup takes exactly
//               one exit, so a launch takes exactly P tickets and the next
//               launch starts at a multiple of P again (the host zeroes the
//               counters whenever n_tiles or G changes).  Nothing waits on
//               another workgroup: every tile runs to its end on its own,
//               and every workgroup leaves after at most n_tiles + G + P
//               tickets whatever the counter holds.
enum { MPPI_KIND_ROLLOUT = 0, MPPI_KIND_REDUCE = 1, MPPI_KIND_STEP = 2 };

template <int BLOCK, int KIND>
__device__ inline void mppi_run_tile(const mppi_step_args& a, uint32_t i) {
    const uint32_t* t = a.tiles + size_t(i) * 3;
    if (KIND == MPPI_KIND_ROLLOUT) mppi_rollout_tile<BLOCK>(a, t[0], t[1]);
    else if (KIND == MPPI_KIND_REDUCE) mppi_reduce_tile<BLOCK>(a, t[0], t[1], t[2]);
    else mppi_onelaunch_tile<BLOCK>(a, t[0], t[1]);
}

// Every branch of the loop is uniform to the compiler, and a barrier closes
// each iteration.  Both matter: the tile index comes from LDS, which the
// compiler must treat as varying across lanes, and the tiles end in regions
// run by thread 0 alone; with a varying exit and no closing barrier, the
// compiler may give thread 0's path back to the loop head (through the ticket
// fetch) its own outer loop and let the other lanes re-enter the loop's
// barriers without it, in the same wavefront -- they then re-run the tile
// with a stale index and, once thread 0 is masked off, never leave.  The
// tile index is read into a scalar register (readfirstlane), as is every
// "last tile" flag the tiles branch on.
template <int BLOCK, int KIND>
__device__ inline void mppi_persistent(const mppi_step_args& a) {
    __shared__ uint32_t s_tile;
    const uint64_t P = uint64_t(a.n_tiles) + gridDim.x;
    for (;;) {
        if (threadIdx.x == 0) {
            const uint64_t t = __hip_atomic_fetch_add(a.ticket, uint64_t(1), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            s_tile = uint32_t(t % P);
        }
        __syncthreads();
        const uint32_t i = __builtin_amdgcn_readfirstlane(s_tile);
        if (i >= a.n_tiles) break;
        mppi_run_tile<BLOCK, KIND>(a, i);
        __syncthreads();       // the tile, thread 0's last writes included, ends here for every thread
    }
}

#define MPPI_BATCH_ENTRIES(B)                                                                    \
    extern "C" __global__ void __launch_bounds__(B) mppi_rollout_table_b##B(mppi_step_args a) {    \
        mppi_run_tile<B, MPPI_KIND_ROLLOUT>(a, blockIdx.x);                                      \
    }                                                                                            \
    extern "C" __global__ void __launch_bounds__(B) mppi_reduce_table_b##B(mppi_step_args a) {     \
        mppi_run_tile<B, MPPI_KIND_REDUCE>(a, blockIdx.x);                                       \
    }                                                                                            \
    extern "C" __global__ void __launch_bounds__(B) mppi_step_table_b##B(mppi_step_args a) {       \
        mppi_run_tile<B, MPPI_KIND_STEP>(a, blockIdx.x);                                         \
    }                                                                                            \
    extern "C" __global__ void __launch_bounds__(B) mppi_rollout_persist_b##B(mppi_step_args a) {  \
        mppi_persistent<B, MPPI_KIND_ROLLOUT>(a);                                                \
    }                                                                                            \
    extern "C" __global__ void __launch_bounds__(B) mppi_reduce_persist_b##B(mppi_step_args a) {   \
        mppi_persistent<B, MPPI_KIND_REDUCE>(a);                                                 \
    }                                                                                            \
    extern "C" __global__ void __launch_bounds__(B) mppi_step_persist_b##B(mppi_step_args a) {     \
        mppi_persistent<B, MPPI_KIND_STEP>(a);                                                   \
    }

MPPI_BATCH_ENTRIES(64)
MPPI_BATCH_ENTRIES(128)
MPPI_BATCH_ENTRIES(256)
#endif

#if MPPI_SURROGATE_KERNELS
#include "mppi_mlp.h"

// Training transitions: (state, control) drawn uniformly from the sampling
// box with Philox (tag SURROGATE, counter (block, i, stream, tag)), labelled
// with one step of the plugin's dynamics: t = f(x, u) - x.
extern "C" __global__ void __launch_bounds__(256) mppi_surrogate_data(mppi_surrogate_data_args a) {
    const uint32_t i = blockIdx.x * 256 + threadIdx.x;
    if (i >= a.M) return;
    constexpr int IN = SDIM + CDIM;
    float v[IN];
#pragma unroll
    for (int q = 0; q < (IN + 3) / 4; ++q) {
        mppi_u32x4 w = mppi_philox4x32_10(
            mppi_u32x4{uint32_t(q), i, a.stream, MPPI_TAG_SURROGATE << 28}, a.key0, a.key1);
        const uint32_t ws[4] = {w.x, w.y, w.z, w.w};
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int k = 4 * q + j;
            if (k < IN) v[k] = a.lo[k] + (a.hi[k] - a.lo[k]) * mppi_u01(ws[j]);
        }
    }
#pragma unroll
    for (int k = 0; k < IN; ++k) a.x[size_t(i) * IN + k] = mppi_mlp_input(v[k], a.center[k], a.inv_half[k], a.period[k], a.inv_period[k]);
    float xs[SDIM];
#pragma unroll
    for (int k = 0; k < SDIM; ++k) xs[k] = v[k];
    mppi_rng rng{0u, i, a.stream, (MPPI_TAG_SURROGATE << 28) | 1u, a.key0, a.key1};
    MPPI_DYNAMICS(mppi_state_ref{xs}, v + SDIM, mppi_blob_ref{a.model}, rng);
#pragma unroll
    for (int k = 0; k < SDIM; ++k) a.t[size_t(i) * SDIM + k] = xs[k] - v[k];
}

// Rollout with the MLP as dynamics (one sample per lane, MPPI_SUR_BLOCK-thread
// workgroups, grid (ceil(K / MPPI_SUR_BLOCK), N); the reduce kernel is shared).  Per
// timestep every wavefront evaluates the network for its 64 samples as two
// 32-sample column blocks on 32x32 tiles (mppi_mlp_net in mppi_mlp.h):
//   layer 1   v_mfma_f32_32x32x8_f16, one per (m-block of 32 units, column
//             block, eight inputs), its B operands the lanes' own inputs
//             after one exchange of lane halves;
//   layer 2   v_mfma_f32_32x32x16_f16: each A fragment is read from LDS once
//             and applied to both column blocks; its B operands are the
//             layer-1 results a lane already holds;
//   layer 3   the same (v_mfma_f32_16x16x32_f16 with at most eight outputs),
//             accumulated as each m-block of layer 2 comes out;
// activations on packed binary16 values in registers, which the vector unit
// runs while the matrix core works through the products they do not feed.
// The state deltas go through the wavefront's own LDS rows (s_dx), whose
// operations complete in order, so a timestep needs wavefront-level ordering
// only (no workgroup barrier).

extern "C" __global__ void __launch_bounds__(MPPI_SUR_BLOCK) mppi_rollout_surrogate(mppi_surrogate_rollout_args sa) {
    constexpr int IN = SDIM + CDIM;
    constexpr int WAVES = MPPI_SUR_BLOCK / 64;
    constexpr int KB1 = IN <= 8 ? 1 : 2;          // layer-1 K-blocks of eight inputs
    static_assert(IN <= 16 && SDIM <= 16, "surrogate network: inputs and outputs of at most 16");
    __shared__ __attribute__((aligned(16))) short s_wy[MPPI_MLP_Y_ELEMS];
    __shared__ __attribute__((aligned(16))) float s_b[2 * MPPI_MLP_HID + 32];
    __shared__ float s_dx[WAVES][64][SDIM + 1];
    const mppi_step_args& a = sa.s;
    const uint32_t agent = blockIdx.y;
    const mppi_agent_desc d = a.desc[agent];
    if (blockIdx.x * MPPI_SUR_BLOCK >= d.K) return;
    for (uint32_t e = threadIdx.x; e < MPPI_MLP_Y_ELEMS / 8; e += MPPI_SUR_BLOCK)
        reinterpret_cast<uint4*>(s_wy)[e] = reinterpret_cast<const uint4*>(sa.frags + MPPI_MLP_Y_BASE)[e];
    for (uint32_t e = threadIdx.x; e < MPPI_MLP_HID; e += MPPI_SUR_BLOCK) {
        s_b[e] = sa.params[sa.b1 + e];
        s_b[MPPI_MLP_HID + e] = sa.params[sa.b2 + e];
    }
    if (threadIdx.x < 32) s_b[2 * MPPI_MLP_HID + threadIdx.x] = threadIdx.x < SDIM ? sa.params[sa.b3 + threadIdx.x] : 0.0f;
    mppi_prepare_costs(a.cost + d.slot * a.cost_stride, MPPI_SUR_BLOCK);
    __syncthreads();

    const uint32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6, h = lane >> 5, c = lane & 31;
    const uint32_t blk = blockIdx.x;
    const uint32_t k = blk * MPPI_SUR_BLOCK + threadIdx.x;
    const uint32_t H = d.H;
    const bool active = k < d.K;
    const uint32_t slot = d.slot;
    const uint32_t step = a.agents[slot].step;
    const uint32_t tag_agent = (MPPI_TAG_CONTROL_NOISE << 28) | (slot & MPPI_AGENT_MASK);
    const float* __restrict__ un = a.u_nom + d.seq_off;
    const uint8_t* model = a.model + slot * a.model_stride;
    const uint8_t* cost = a.cost + slot * a.cost_stride;

    float x[SDIM];
#pragma unroll
    for (int i = 0; i < SDIM; ++i) x[i] = a.x[size_t(slot) * SDIM + i];
    float total = 0.0f;
    for (uint32_t t0 = 0; t0 < H; t0 += MPPI_G) {
        float z[MPPI_NG];
        const uint32_t b0 = (t0 * CDIM) >> 2;
#pragma unroll
        for (int q = 0; q < MPPI_BG; ++q) {
            mppi_f32x4 v = mppi_normal4(mppi_u32x4{b0 + q, k, step, tag_agent}, a.key0, a.key1);
            z[4 * q + 0] = v.x; z[4 * q + 1] = v.y;
            z[4 * q + 2] = v.z; z[4 * q + 3] = v.w;
        }
#pragma unroll
        for (int gi = 0; gi < MPPI_G; ++gi) {
            const uint32_t t = t0 + gi;
            if (t < H) {
                float u[CDIM];
                mppi_make_control(un + t * CDIM, z + gi * CDIM, a.noise_scale, a.u_min, a.u_max, u);
                // this lane's sample: normalised inputs as FP16, zeros past IN
                mppi_v4f xin[2 * KB1];
#pragma unroll
                for (int q = 0; q < 2 * KB1; ++q)
#pragma unroll
                    for (int j = 0; j < 4; ++j) {
                        const int kk = 4 * q + j;
                        xin[q][j] = kk < SDIM ? mppi_mlp_input(x[kk < SDIM ? kk : 0], sa.in_center[kk], sa.in_inv_half[kk],
                                                               sa.in_period[kk], sa.in_inv_period[kk])
                                  : kk < IN ? mppi_mlp_input(u[kk - SDIM < CDIM ? kk - SDIM : 0], sa.in_center[kk],
                                                             sa.in_inv_half[kk], sa.in_period[kk], sa.in_inv_period[kk])
                                            : 0.0f;
                    }
                // layer-1 B operands: lane l holds inputs 8 kb1 + 4 h .. + 3 of sample
                // 32 nb + c, so lanes 32-63 of column block 0 take the second four
                // inputs of samples 0-31, and lanes 0-31 of column block 1 the first
                // four of samples 32-63: one exchange of lane halves per register
                // (v_permlane32_swap: the upper half of the first operand with the
                // lower half of the second)
                mppi_v4s xb[2][KB1];
#pragma unroll
                for (int kb1 = 0; kb1 < KB1; ++kb1) {
                    const mppi_v2u lo = __builtin_bit_cast(mppi_v2u, mppi_pack_f16(xin[2 * kb1]));
                    const mppi_v2u hi = __builtin_bit_cast(mppi_v2u, mppi_pack_f16(xin[2 * kb1 + 1]));
                    const auto r0 = __builtin_amdgcn_permlane32_swap(lo.x, hi.x, false, false);
                    const auto r1 = __builtin_amdgcn_permlane32_swap(lo.y, hi.y, false, false);
                    xb[0][kb1] = __builtin_bit_cast(mppi_v4s, mppi_v2u{r0[0], r1[0]});
                    xb[1][kb1] = __builtin_bit_cast(mppi_v4s, mppi_v2u{r0[1], r1[1]});
                }
                if constexpr (SDIM <= 8) {
                    // the state deltas: register i holds output row 4 (g % 2) + i of
                    // sample 32 nb + lane % 16 + 16 (g / 2), g = lane / 16
                    mppi_v4f y[2];
                    mppi_mlp_net<KB1, true>(s_wy, s_b, xb, lane, y);
                    const uint32_t g = lane >> 4;
#pragma unroll
                    for (int nb = 0; nb < 2; ++nb)
#pragma unroll
                        for (int i = 0; i < 4; ++i) {
                            const uint32_t r = 4 * (g & 1) + i;
                            if (r < SDIM) s_dx[wave][32 * nb + (lane & 15) + 16 * (g >> 1)][r] = y[nb][i];
                        }
                } else {
                    // register i holds output row 8 (i / 4) + 4 h + i % 4 of sample 32 nb + c
                    mppi_v16f y[2];
                    mppi_mlp_net<KB1, false>(s_wy, s_b, xb, lane, y);
#pragma unroll
                    for (int nb = 0; nb < 2; ++nb)
#pragma unroll
                        for (int i = 0; i < 16; ++i) {
                            if (8 * (i / 4) >= SDIM) continue;
                            const uint32_t r = 8 * (i / 4) + 4 * h + (i % 4);
                            if (r < SDIM) s_dx[wave][32 * nb + c][r] = y[nb][i];
                        }
                }
                __builtin_amdgcn_wave_barrier();
#pragma unroll
                for (int i = 0; i < SDIM; ++i) x[i] += s_dx[wave][lane][i] * sa.out_sd[i] + sa.out_mu[i];
                if (active)
                    total += mppi_stage(mppi_form_timed{}, mppi_cstate_ref{x}, (const float*)u,
                                        mppi_blob_ref{model}, mppi_blob_ref{cost}, mppi_time{t, step});
            }
        }
    }
    if (active) {
        total += mppi_terminal(mppi_form_timed{}, mppi_cstate_ref{x}, mppi_blob_ref{model},
                               mppi_blob_ref{cost}, mppi_time{H, step});
        a.costs[d.cost_off + k] = total;
    }
    // softmax partials, one row per 256 samples (the rows the reduce reads)
    mppi_block_stats<MPPI_SUR_BLOCK, 256>(a, d, blk, active, total);
}
#endif
