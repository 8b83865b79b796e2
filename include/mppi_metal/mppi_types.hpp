// mppi_types.hpp — plain data types of the MPPIDriver API.
//
// Drop-in counterpart of the reference's include/mppi_metal/mppi_types.hpp
// (:15-188): every struct, field name, type and default is kept so code
// written against the reference compiles unchanged.  Differences in meaning:
//   - ModelPluginSpec::user_metallib_path names the plugin SOURCE file
//     (a .hip file of MPPI_CALLABLE functions, see include/mppi_hip/mppi_core.h)
//     that the library compiles with hipRTC; plugin_source_path() is an alias.
//   - RngConfig::seed is used in full (64 bits; the reference truncates it to
//     32, src/metal_backend.mm:565).
//   - ControlNoiseConfig with CovarianceForm::kFull is implemented: sigma is
//     the control covariance (row-major, SPD) and the noise is L z with L its
//     Cholesky factor.  With kDiagonal, sigma holds standard deviations
//     (noise sigma_d z_d), as in the reference's kernels.
#pragma once

#include <cstddef>
#include <cstdint>
#include <filesystem>
#include <string>
#include <string_view>

namespace mppi_metal {

struct ControlNoiseConfig {
    enum class CovarianceForm { kDiagonal, kFull };

    CovarianceForm covariance_form = CovarianceForm::kDiagonal;
    // kDiagonal: control_dim standard deviations; kFull: control_dim^2
    // covariance entries, row-major.
    const float* sigma = nullptr;
    // Same shape as sigma.  Required non-null (validated) but not used by the
    // update law, as in the reference.
    const float* sigma_inv = nullptr;
    uint32_t control_dim = 0;
};

struct ControlBounds {
    const float* u_min = nullptr;  // control_dim
    const float* u_max = nullptr;  // control_dim
    uint32_t control_dim = 0;
};

struct DriverConfig {
    uint32_t state_dim = 0;
    uint32_t control_dim = 0;
    uint32_t horizon = 0;
    uint32_t sample_count = 0;
    float lambda = 1.0f;
    ControlNoiseConfig noise;
    ControlBounds bounds;

    enum class NumericMode { kFp32, kFp16, kMixed };
    NumericMode numeric_mode = NumericMode::kFp32;
};

struct RngConfig {
    // Key of the counter-based Philox4x32-10 noise stream.
    uint64_t seed = 0;
};

struct CallableNames {
    // Required: the dynamics callable.
    std::string_view dynamics = "mppi_dynamics";
    // Optional: empty means the library's zero cost.
    std::string_view stage_cost = "";
    std::string_view terminal_cost = "";
};

struct ParamLayout {
    size_t model_params_size = 0;
    size_t model_params_alignment = 1;
    size_t cost_params_size = 0;
    size_t cost_params_alignment = 1;
    std::string schema_hash;
};

struct ModelPluginSpec {
    // Path of the plugin source file (kept under the reference's field name).
    std::filesystem::path user_metallib_path;
    CallableNames callables;
    ParamLayout param_layout;
    RngConfig rng;

    std::filesystem::path& plugin_source_path() { return user_metallib_path; }
    const std::filesystem::path& plugin_source_path() const { return user_metallib_path; }
};

struct StateView {
    const float* data = nullptr;
    uint32_t state_dim = 0;
};

struct ControlSequenceView {
    const float* data = nullptr;
    uint32_t horizon = 0;
    uint32_t control_dim = 0;
};

struct MutableControlSequenceView {
    float* data = nullptr;
    uint32_t horizon = 0;
    uint32_t control_dim = 0;
};

struct ByteView {
    const void* data = nullptr;
    size_t bytes = 0;
};

struct ProblemParams {
    ByteView model_params;
    ByteView cost_params;
};

struct StepInputs {
    StateView x0;
    // Null data: use the driver's warm-start sequence.
    ControlSequenceView u_nominal_override;
    // bytes == 0: use the bound ProblemParams.
    ByteView model_params_override;
    ByteView cost_params_override;
};

struct StepDiagnostics {
    float best_cost = 0.0f;
    float mean_cost = 0.0f;
    float effective_sample_size = 0.0f;
    bool converged = false;
};

struct SimulationResults {
    float* states_out = nullptr;    // num_steps * state_dim
    float* controls_out = nullptr;  // num_steps * control_dim
    float* costs_out = nullptr;     // num_steps (best cost per step)
    uint32_t num_steps = 0;
};

// Outputs of simulate_batch, indexed [step * num_agents * dim + agent * dim + e].
struct BatchSimulationResults {
    float* states_out = nullptr;           // num_agents * num_steps * state_dim
    float* controls_out = nullptr;         // num_agents * num_steps * control_dim
    float* costs_out = nullptr;            // num_agents * num_steps
    float* terminal_states_out = nullptr;  // num_agents * state_dim
    uint32_t num_agents = 0;
    uint32_t num_steps = 0;
};

}  // namespace mppi_metal
