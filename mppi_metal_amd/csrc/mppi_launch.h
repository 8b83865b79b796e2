// mppi_launch.h — argument blocks of the MPPI kernels, shared by the kernel
// source (compiled by hipRTC) and the host backend that fills them, so both
// sides agree on the layout.  Plain C layout: pointers, 32/64-bit words,
// doubles.
#ifndef MPPI_HIP_LAUNCH_H
#define MPPI_HIP_LAUNCH_H

#if !defined(__HIPCC_RTC__)
#include <stdint.h>
#endif

// Per-agent mutable state on the device.
struct mppi_agent_state {
    uint32_t step;               // low 32 bits of the MPPI step index
    uint32_t hist;               // history row (closed loop)
    uint32_t rollout_arrivals;   // last-workgroup counters, self-resetting
    uint32_t reduce_arrivals;
    double min_cost;
    double weight_sum;           // sum_k exp(-(c_k - min)/lambda)
    double mean_cost;
    double ess;
    uint32_t status;             // MPPI_STATUS_* bits, sticky for a call
    uint32_t reserved;
};

// mppi_agent_state::status bits.
#define MPPI_STATUS_WAIT_TIMEOUT 1u   // a one-pass rollout consumer gave up waiting for its noise

// Per-agent shape and buffer offsets (immutable for a batch shape).  All
// offsets are in elements of the buffer they index.  An agent has two
// indices: its position in the launch (the index of its descriptor; per-call
// outputs diag, histories, terminal and first_out are indexed by it) and its
// slot (its persistent data: state x, mppi_agent_state, parameter blobs, and
// the noise stream's agent word).  The one-call paths use slot = position.
struct mppi_agent_desc {
    uint32_t K;                  // samples
    uint32_t H;                  // horizon
    uint32_t n_blk;              // rollout workgroups: ceil(K / block)
    uint32_t NB;                 // noise blocks per sample: ceil(H*C / 4)
    uint32_t n_chunks;           // reduce sample chunks: ceil(K / (block * spt))
    uint32_t n_groups;           // reduce noise-block groups: ceil(NB / 8)
    uint32_t spt;                // reduce samples per thread
    uint32_t slot;               // persistent agent index (see above)
    uint64_t cost_off;           // costs (floats)
    uint64_t seq_off;            // u_nom (floats, H*C per agent)
    uint64_t blk_off;            // blk_stats (4 floats per rollout workgroup)
    uint64_t part_off;           // partials (floats)
    uint64_t w_off;              // weights (floats)
    uint64_t out_off;            // u_out (floats, H*C per agent)
};

// Slices of the one-pass step's combine.  The combine splits the horizon
// into slices of mppi_combine_tpw(C) timesteps, one workgroup each; the
// rollout tiles store their weighted noise sums slice-major (slice s, tile g:
// row s * n_blk + g), each row mppi_combine_width(C) floats, a multiple of 32
// floats (128 bytes), so that a slice reads whole cache lines.  The slice
// spans lcm(C, 32) columns of the sequence where that is at most 128 (the
// one-pass step's H * C limit), else floor(128 / C) timesteps, which then
// cover the whole horizon.
constexpr uint32_t mppi_pow2_factor(uint32_t c) { return (c & (0u - c)) > 32u ? 32u : (c & (0u - c)); }
constexpr uint32_t mppi_combine_tpw(uint32_t C) {
    return (32u / mppi_pow2_factor(C)) * C <= 128u ? 32u / mppi_pow2_factor(C) : 128u / C;
}
constexpr uint32_t mppi_combine_width(uint32_t C) { return (mppi_combine_tpw(C) * C + 31u) / 32u * 32u; }
constexpr uint32_t mppi_combine_slices(uint32_t H, uint32_t C) {
    return (H + mppi_combine_tpw(C) - 1u) / mppi_combine_tpw(C);
}

struct mppi_step_args {
    const mppi_agent_desc* desc; // [N]
    mppi_agent_state* agents;    // [N]
    const uint32_t* tiles;       // table / persistent modes: [n_tiles][3]
    uint64_t* ticket;            // persistent mode: the launch's ticket counter
    uint32_t n_tiles;
    uint32_t N;
    float* x;                    // [N][D] state (x0; plant state in closed loop)
    float* u_nom;                // packed by desc.seq_off
    float* u_nom_next;           // closed loop: the shifted warm start goes here (null: into u_nom)
    float* u_out;                // packed by desc.seq_off
    const uint8_t* model;        // model blob of agent 0
    const uint8_t* cost;         // cost blob of agent 0
    uint64_t model_stride;       // bytes between agents' model blobs (0: shared)
    uint64_t cost_stride;        // bytes between agents' cost blobs (0: shared)
    const float* noise_scale;    // [C] sigma or [C][C] lower factor
    const float* u_min;          // [C]
    const float* u_max;          // [C]
    float* costs;                // packed by desc.cost_off
    float* weights;              // packed by desc.cost_off, or null
    float* blk_stats;            // packed by desc.blk_off
    float* partials;             // packed by desc.part_off
    float* diag;                 // [N][3]
    float* hist_states;          // [T][N][D] or null
    float* hist_controls;        // [T][N][C] or null
    float* hist_costs;           // [T][N] or null
    float* terminal;             // [N][D] or null
    float* first_out;            // [N][C] first control of u_out, or null
    uint32_t closed_loop;        // 1: plant step + warm-start shift after the reduce
    uint32_t key0, key1;
    float lambda;
    uint32_t* status_out;        // [N] copy of each agent's status after the update, or null
    uint32_t wait_limit;         // one-pass rollout: polls before a consumer gives up (0: default)
    uint32_t wait_bias;          // added to the group count a consumer waits for; nonzero only to test
                                 // the give-up path (the count then never arrives)
    // One-agent launches (N == 1, one != 0) of the one-pass step: agent 0's
    // descriptor and step counter, as the host knows them, so that the
    // kernels start without first loading desc[0] and agents[slot].step.
    mppi_agent_desc desc0;
    uint32_t one;
    uint32_t step0;
};

// Learned-dynamics surrogate (mppi_mlp.h): training data and rollout.
struct mppi_surrogate_data_args {
    float* x;                    // [M][in]  normalised (state, control) inputs
    float* t;                    // [M][out] state deltas of the plugin dynamics
    const uint8_t* model;        // model blob of the plugin dynamics
    uint32_t M;
    uint32_t key0, key1;
    uint32_t stream;             // counter word 2: data stream id
    float lo[16], hi[16];        // sampling box of (state, control)
    float center[16], inv_half[16], period[16], inv_period[16];   // input normalisation (mppi_mlp.h)
};

// Samples (threads) per workgroup of the surrogate rollout: eight
// wavefronts, two per SIMD.  (Twelve, with registers cut to three per SIMD,
// ran slower at K = 2^18: 342 workgroups take two rounds of 256 CUs.)
constexpr uint32_t MPPI_SUR_BLOCK = 512;

struct mppi_surrogate_rollout_args {
    mppi_step_args s;
    const short* frags;          // A fragments (mppi_mlp.h)
    const float* params;         // FP32 parameters (biases are read from here)
    uint32_t b1, b2, b3;         // bias offsets into params
    uint32_t pad;
    float in_center[16], in_inv_half[16], in_period[16], in_inv_period[16], out_mu[16], out_sd[16];
};

#endif  // MPPI_HIP_LAUNCH_H
