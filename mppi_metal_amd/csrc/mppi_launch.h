// mppi_launch.h — argument blocks of the MPPI kernels, shared by the kernel
// source (compiled by hipRTC) and the host backend that fills them, so both
// sides agree on the layout.  Plain C layout: pointers, 32-bit words, doubles.
#ifndef MPPI_HIP_LAUNCH_H
#define MPPI_HIP_LAUNCH_H

#if !defined(__HIPCC_RTC__)
#include <stdint.h>
#endif

struct mppi_agent_state {
    uint32_t step;               // low 32 bits of the MPPI step index
    uint32_t hist;               // history row (closed loop)
    uint32_t rollout_arrivals;   // last-workgroup counters, self-resetting
    uint32_t reduce_arrivals;
    double min_cost;
    double weight_sum;           // sum_k exp(-(c_k - min)/lambda)
    double mean_cost;
    double ess;
};

struct mppi_rollout_args {
    const float* x0;             // [N][D]
    const float* u_nom;          // [N][H][C]
    const uint8_t* model;
    const uint8_t* cost;
    const float* noise_scale;    // [C] sigma or [C][C] lower factor
    const float* u_min;          // [C]
    const float* u_max;          // [C]
    float* costs;                // [N][K]
    float* blk_stats;            // [N][n_blk][4]
    mppi_agent_state* agents;    // [N]
    uint32_t K, H, N, n_blk;
    uint32_t key0, key1;
    float lambda;
    uint32_t pad;
};

struct mppi_reduce_args {
    const float* costs;          // [N][K]
    mppi_agent_state* agents;    // [N]
    float* u_nom;                // [N][H][C]; shifted in place in closed loop
    float* u_out;                // [N][H][C]
    const float* noise_scale;
    const float* u_min;
    const float* u_max;
    float* weights;              // [N][K] or null
    float* partials;             // [N][n_chunks][NB][4]
    float* diag;                 // [N][3]
    float* x_state;              // [N][D] plant state (closed loop)
    float* hist_states;          // [T][N][D] or null
    float* hist_controls;        // [T][N][C] or null
    float* hist_costs;           // [T][N] or null
    float* terminal;             // [N][D] or null
    const uint8_t* model;
    uint32_t K, H, N, NB;
    uint32_t n_chunks, n_groups, spt, closed_loop;
    uint32_t key0, key1;
    float lambda;
    uint32_t pad;
};

#endif  // MPPI_HIP_LAUNCH_H
