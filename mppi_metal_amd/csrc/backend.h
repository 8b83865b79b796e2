// backend.h — internal declarations shared by the C-ABI implementation files
// (hip_backend.cpp: contexts, plugin linking, single-agent and uniform-batch
// steps; batch_backend.cpp: the uneven multi-agent step).  Not installed.
#pragma once

#include "mppi_hip.h"
#include "mppi_launch.h"

#include <hip/hip_runtime.h>

#include <memory>
#include <string>
#include <utility>
#include <vector>

namespace mppi_hip_impl {

constexpr uint32_t kBlock = 256;      // workgroup size of the grid-mapped kernels
constexpr uint32_t kReduceBPG = 8;    // noise blocks per reduce tile
constexpr uint32_t kMaxHC = 4096;     // horizon * control_dim limit
constexpr uint32_t kFusedMaxHC = 128; // horizon * control_dim limit of the one-pass step
// Size of the per-device scratch array of mppi_hip_step_noalloc (noalloc.hip).
constexpr size_t kNoallocScratchBytes = size_t(8) << 20;

// One loaded code object and the kernels looked up in it.
struct LoadedModule {
    hipModule_t module = nullptr;
    hipFunction_t rollout = nullptr;  // grid-mapped kernels
    hipFunction_t reduce = nullptr;
    hipFunction_t rollout_fused = nullptr;  // one-pass step: tiles
    hipFunction_t fused_combine = nullptr;  // one-pass step: combine and update
    hipFunction_t rollout_fused_one = nullptr;   // the same for one agent described in the arguments
    hipFunction_t fused_combine_one = nullptr;
    hipFunction_t finalize = nullptr;       // two-kernel step: the update, one workgroup per agent
    hipFunction_t weights = nullptr;        // per-sample weights after a one-pass step
    // batch kernels, indexed [block 64/128/256 -> 0/1/2][table 0 / persistent 1]
    hipFunction_t rollout_b[3][2] = {};
    hipFunction_t reduce_b[3][2] = {};
    hipFunction_t step_b[3][2] = {};        // one-launch step
    // surrogate kernels (MPPI_SURROGATE_KERNELS)
    hipFunction_t surrogate_data = nullptr;
    hipFunction_t rollout_surrogate = nullptr;
};

struct SurrogateState;  // surrogate.hip

struct DeviceBuffer {
    void* ptr = nullptr;
    size_t bytes = 0;
    void release();
};

struct PinnedBuffer {
    void* ptr = nullptr;
    size_t bytes = 0;
    void release();
};

size_t align_up(size_t x, size_t a);

}  // namespace mppi_hip_impl

struct mppi_hip_ctx {
    // configuration
    uint32_t D = 0, C = 0, H = 0, K = 0;
    float lambda = 1.0f;
    uint32_t cov_form = MPPI_HIP_COV_DIAGONAL;
    std::vector<float> noise_scale;  // sigma (diag) or lower Cholesky factor (full)
    std::vector<float> u_min, u_max;
    uint32_t numeric_mode = MPPI_HIP_FP32;
    int device = 0;
    std::string arch;
    int num_cus = 256;

    // plugin (kept so the batch kernels can be compiled later)
    bool linked = false;
    std::shared_ptr<mppi_hip_impl::LoadedModule> mod;        // plain kernels
    std::shared_ptr<mppi_hip_impl::LoadedModule> batch_mod;  // + batch kernels, lazily
    std::string plugin_src, plugin_name, plugin_dir;
    std::string dyn_name, stage_name, term_name;
    size_t model_bytes = 0, cost_bytes = 0;
    uint32_t key0 = 0, key1 = 0;

    // runtime
    hipStream_t stream = nullptr;
    std::string error;
    uint32_t spt_override = 0;
    // bounded wait of the one-pass rollout (mppi_hip_set_wait_limit)
    uint32_t wait_limit = 0, wait_bias = 0;

    // device buffers of the single-agent / uniform-batch path
    mppi_hip_impl::DeviceBuffer consts;  // noise_scale | u_min | u_max | zero blob
    mppi_hip_impl::DeviceBuffer model_bound, cost_bound;
    bool model_is_bound = false, cost_is_bound = false;
    mppi_hip_impl::DeviceBuffer up;      // desc | agents | x | u_nom | model ovr | cost ovr
    mppi_hip_impl::DeviceBuffer down;    // u_out | diag
    mppi_hip_impl::DeviceBuffer costs, blk_stats, weights, partials;
    // closed-loop results: states | controls | best costs | end states |
    // status words, in one device block copied back in one transfer
    mppi_hip_impl::DeviceBuffer hist;
    mppi_hip_impl::PinnedBuffer up_host, down_host, hist_host;
    uint32_t last_agents = 0;
    // where read_costs / read_weights find the last iteration's data
    const float* last_costs = nullptr;
    const float* last_weights = nullptr;
    std::vector<mppi_agent_desc> last_desc;

    // step path (mppi_hip_set_step_path); a one-pass step leaves the
    // per-sample weights to be formed when they are read
    uint32_t step_path = MPPI_HIP_PATH_AUTO;
    bool weights_pending = false;
    mppi_step_args weights_args{};
    uint32_t weights_grid_x = 0, weights_grid_y = 0;

    // learned-dynamics surrogate (surrogate.hip)
    std::shared_ptr<mppi_hip_impl::SurrogateState> sur;

    // profiling
    bool profiling = false;
    std::vector<hipEvent_t> ev_pool;
    std::vector<std::pair<hipEvent_t, hipEvent_t>> pending_rollout, pending_reduce;
    double prof_rollout_ms = 0.0, prof_reduce_ms = 0.0;
    uint64_t prof_launches = 0;
};

namespace mppi_hip_impl {

int fail(mppi_hip_ctx* ctx, int code, const std::string& msg);
int ensure_device_buffer(mppi_hip_ctx* ctx, DeviceBuffer* b, size_t bytes);
int ensure_pinned_buffer(mppi_hip_ctx* ctx, PinnedBuffer* b, size_t bytes);
int check_ctx(mppi_hip_ctx* ctx);
// Compile (or fetch from the caches) and load the plugin module; `batch`
// adds the tile-table and persistent kernels, `surrogate` the learned-dynamics
// data and rollout kernels.
int load_module(mppi_hip_ctx* ctx, bool batch, bool surrogate, std::shared_ptr<LoadedModule>* out);
// Rollout kernel and argument block of the surrogate when it is enabled on
// ctx (else returns false).
bool surrogate_rollout(mppi_hip_ctx* ctx, const mppi_step_args& a, hipFunction_t* fn,
                       void* args_storage, size_t storage_bytes);
hipEvent_t take_event(mppi_hip_ctx* ctx);
void collect_profile(mppi_hip_ctx* ctx);
// Common kernel arguments (constants, bound blobs, work buffers) of a step.
void fill_common_args(mppi_hip_ctx* ctx, mppi_step_args* a);
// MPPI_HIP_E_DEVICE when any of the n status words reports a failed device wait.
int check_status(mppi_hip_ctx* ctx, const uint32_t* status, uint32_t n);
// noalloc.hip: take the device address of the scratch array of `device` (the
// device must be current; once per device), read it back (null with the
// reason in *why when it could not be taken), and enqueue the kernel that
// initialises a call's descriptor, agent state and status word.
void noalloc_setup(int device);
void* noalloc_scratch(int device, std::string* why);
hipError_t noalloc_init(hipStream_t stream, const mppi_agent_desc& d, uint32_t step, mppi_agent_desc* desc_out,
                        mppi_agent_state* agent_out, uint32_t* status_out);

}  // namespace mppi_hip_impl

#define MPPI_HIP_TRY(ctx, expr)                                                          \
    do {                                                                                 \
        hipError_t e_ = (expr);                                                          \
        if (e_ != hipSuccess)                                                            \
            return ::mppi_hip_impl::fail(ctx, MPPI_HIP_E_DEVICE,                          \
                                         std::string(#expr) + " failed: " + hipGetErrorString(e_)); \
    } while (0)
