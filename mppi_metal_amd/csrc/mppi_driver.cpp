// mppi_driver.cpp — MPPIDriver facade over the HIP C ABI.
//
// Keeps the reference driver's contract (src/mppi_driver.cpp): the same
// validation rules in the same order and with the same key words in the
// messages (state_dim, control_dim, horizon, sample_count, sigma, sigma_inv,
// u_min, u_max, "not initialized", "null", the missing callable's name), the
// host warm start after step() (shift left, hold last, :326-337), the step
// index that selects the noise stream (:339, :419, :489) and the reset
// semantics (:493-537).  The device work is done by include/mppi_hip.h.
#include "mppi_metal/mppi_driver.hpp"

#include "mppi_hip.h"

#include <algorithm>
#include <string>
#include <vector>

namespace mppi_metal {

namespace {

struct ErrorSink {
    std::string* out;
    bool operator()(const std::string& msg) const {
        if (out) *out = msg;
        return false;
    }
};

size_t noise_len(const ControlNoiseConfig& n, uint32_t cdim) {
    return n.covariance_form == ControlNoiseConfig::CovarianceForm::kDiagonal ? cdim
                                                                              : size_t(cdim) * cdim;
}

// Validation of the configuration, in the reference's order
// (src/mppi_driver.cpp:100-159).  Returns an empty string when valid.
std::string validate_config(const DriverConfig& c) {
    if (c.state_dim == 0) return "DriverConfig::state_dim must be > 0.";
    if (c.control_dim == 0) return "DriverConfig::control_dim must be > 0.";
    if (c.horizon == 0) return "DriverConfig::horizon must be > 0.";
    if (c.sample_count == 0) return "DriverConfig::sample_count must be > 0.";
    if (c.noise.control_dim != c.control_dim)
        return "ControlNoiseConfig::control_dim (" + std::to_string(c.noise.control_dim) +
               ") must match DriverConfig::control_dim (" + std::to_string(c.control_dim) + ").";
    if (c.noise.sigma == nullptr) return "ControlNoiseConfig::sigma must not be null.";
    if (c.noise.sigma_inv == nullptr) return "ControlNoiseConfig::sigma_inv must not be null.";
    if (c.bounds.control_dim != c.control_dim)
        return "ControlBounds::control_dim (" + std::to_string(c.bounds.control_dim) +
               ") must match DriverConfig::control_dim (" + std::to_string(c.control_dim) + ").";
    if (c.bounds.u_min == nullptr) return "ControlBounds::u_min must not be null when control_dim > 0.";
    if (c.bounds.u_max == nullptr) return "ControlBounds::u_max must not be null when control_dim > 0.";
    for (uint32_t i = 0; i < c.control_dim; ++i)
        if (c.bounds.u_min[i] > c.bounds.u_max[i])
            return "ControlBounds::u_min[" + std::to_string(i) + "] (" +
                   std::to_string(c.bounds.u_min[i]) + ") > u_max[" + std::to_string(i) + "] (" +
                   std::to_string(c.bounds.u_max[i]) + ").";
    return {};
}

std::string validate_x0(const StateView& x0, uint32_t state_dim) {
    if (x0.data == nullptr) return "StepInputs::x0.data must not be null.";
    if (x0.state_dim != state_dim)
        return "StepInputs::x0.state_dim (" + std::to_string(x0.state_dim) +
               ") must match DriverConfig::state_dim (" + std::to_string(state_dim) + ").";
    return {};
}

}  // namespace

class MPPIDriver::Impl {
public:
    explicit Impl(const DriverConfig& cfg) : config(cfg) {
        const uint32_t C = cfg.control_dim;
        if (cfg.noise.sigma && C) {
            sigma.assign(cfg.noise.sigma, cfg.noise.sigma + noise_len(cfg.noise, C));
            config.noise.sigma = sigma.data();
        }
        if (cfg.noise.sigma_inv && C) {
            sigma_inv.assign(cfg.noise.sigma_inv, cfg.noise.sigma_inv + noise_len(cfg.noise, C));
            config.noise.sigma_inv = sigma_inv.data();
        }
        if (cfg.bounds.u_min && C) {
            u_min.assign(cfg.bounds.u_min, cfg.bounds.u_min + C);
            config.bounds.u_min = u_min.data();
        }
        if (cfg.bounds.u_max && C) {
            u_max.assign(cfg.bounds.u_max, cfg.bounds.u_max + C);
            config.bounds.u_max = u_max.data();
        }
        u_nominal.assign(size_t(cfg.horizon) * cfg.control_dim, 0.0f);
    }
    ~Impl() { release(); }

    void release() {
        if (ctx) mppi_hip_destroy(ctx);
        ctx = nullptr;
        initialized = false;
    }

    size_t seq_len() const { return size_t(config.horizon) * config.control_dim; }

    DriverConfig config;
    std::vector<float> sigma, sigma_inv, u_min, u_max;
    std::vector<float> u_nominal;      // warm start, horizon * control_dim
    std::vector<float> u_scratch;
    ParamLayout layout;
    mppi_hip_ctx* ctx = nullptr;
    bool initialized = false;
    uint64_t step_index = 0;
};

MPPIDriver::MPPIDriver(const DriverConfig& config) : impl_(new Impl(config)) {}
MPPIDriver::~MPPIDriver() = default;
MPPIDriver::MPPIDriver(MPPIDriver&&) noexcept = default;
MPPIDriver& MPPIDriver::operator=(MPPIDriver&&) noexcept = default;

bool MPPIDriver::initialize(const ModelPluginSpec& model, std::string* error) {
    ErrorSink err{error};
    const DriverConfig& cfg = impl_->config;
    std::string msg = validate_config(cfg);
    if (!msg.empty()) return err(msg);

    impl_->release();
    mppi_hip_config hc{};
    hc.state_dim = cfg.state_dim;
    hc.control_dim = cfg.control_dim;
    hc.horizon = cfg.horizon;
    hc.sample_count = cfg.sample_count;
    hc.lambda = cfg.lambda;
    hc.covariance_form = cfg.noise.covariance_form == ControlNoiseConfig::CovarianceForm::kFull
                             ? MPPI_HIP_COV_FULL
                             : MPPI_HIP_COV_DIAGONAL;
    hc.sigma = cfg.noise.sigma;
    hc.u_min = cfg.bounds.u_min;
    hc.u_max = cfg.bounds.u_max;
    hc.numeric_mode = cfg.numeric_mode == DriverConfig::NumericMode::kFp32   ? MPPI_HIP_FP32
                      : cfg.numeric_mode == DriverConfig::NumericMode::kFp16 ? MPPI_HIP_FP16
                                                                             : MPPI_HIP_MIXED;
    hc.device = -1;

    // The plugin path is checked before any device work, so these errors
    // do not depend on a GPU being present.
    const std::string path = model.user_metallib_path.string();
    if (path.empty()) return err("user_metallib_path is empty.");

    mppi_hip_ctx* ctx = nullptr;
    if (mppi_hip_create(&hc, &ctx) != MPPI_HIP_OK) return err(mppi_hip_last_error(nullptr));
    const std::string dyn(model.callables.dynamics);
    const std::string sc(model.callables.stage_cost);
    const std::string tc(model.callables.terminal_cost);
    int rc = mppi_hip_link_plugin(ctx, path.c_str(), dyn.c_str(), sc.c_str(), tc.c_str(),
                                  model.param_layout.model_params_size,
                                  model.param_layout.cost_params_size, model.rng.seed);
    if (rc != MPPI_HIP_OK) {
        std::string m = mppi_hip_last_error(ctx);
        mppi_hip_destroy(ctx);
        return err(m);
    }
    impl_->ctx = ctx;
    impl_->layout = model.param_layout;
    impl_->initialized = true;
    return true;
}

bool MPPIDriver::set_problem_params(const ProblemParams& params, std::string* error) {
    ErrorSink err{error};
    if (!impl_->initialized) return err("Driver is not initialized. Call initialize() first.");
    const ParamLayout& L = impl_->layout;
    // Size rules of src/mppi_driver.cpp:186-224.
    if (L.model_params_size > 0 && params.model_params.bytes == 0)
        return err("ParamLayout declares model_params_size " + std::to_string(L.model_params_size) +
                   " but provided model_params.bytes is 0.");
    if (params.model_params.bytes > 0 && params.model_params.bytes != L.model_params_size)
        return err("model_params.bytes (" + std::to_string(params.model_params.bytes) +
                   ") does not match ParamLayout::model_params_size (" +
                   std::to_string(L.model_params_size) + ").");
    if (params.model_params.bytes > 0 && params.model_params.data == nullptr)
        return err("model_params.data is null but bytes > 0.");
    if (L.cost_params_size > 0 && params.cost_params.bytes == 0)
        return err("ParamLayout declares cost_params_size " + std::to_string(L.cost_params_size) +
                   " but provided cost_params.bytes is 0.");
    if (params.cost_params.bytes > 0 && params.cost_params.bytes != L.cost_params_size)
        return err("cost_params.bytes (" + std::to_string(params.cost_params.bytes) +
                   ") does not match ParamLayout::cost_params_size (" +
                   std::to_string(L.cost_params_size) + ").");
    if (params.cost_params.bytes > 0 && params.cost_params.data == nullptr)
        return err("cost_params.data is null but bytes > 0.");
    if (mppi_hip_bind_params(impl_->ctx, params.model_params.data, params.model_params.bytes,
                             params.cost_params.data, params.cost_params.bytes) != MPPI_HIP_OK)
        return err(mppi_hip_last_error(impl_->ctx));
    return true;
}

bool MPPIDriver::step(const StepInputs& inputs, const MutableControlSequenceView& u_next_out,
                      StepDiagnostics* diagnostics, std::string* error) {
    ErrorSink err{error};
    if (!impl_->initialized) return err("Driver is not initialized. Call initialize() first.");
    const DriverConfig& cfg = impl_->config;
    std::string msg = validate_x0(inputs.x0, cfg.state_dim);
    if (!msg.empty()) return err(msg);
    if (u_next_out.data == nullptr) return err("u_next_out.data must not be null.");
    if (u_next_out.horizon != cfg.horizon || u_next_out.control_dim != cfg.control_dim)
        return err("u_next_out shape mismatch.");
    const float* nominal = impl_->u_nominal.data();
    if (inputs.u_nominal_override.data != nullptr) {
        if (inputs.u_nominal_override.horizon != cfg.horizon ||
            inputs.u_nominal_override.control_dim != cfg.control_dim)
            return err("u_nominal_override shape mismatch.");
        nominal = inputs.u_nominal_override.data;
    }
    const size_t n = impl_->seq_len();
    impl_->u_scratch.resize(n);
    float diag[3] = {0.0f, 0.0f, 0.0f};
    if (mppi_hip_step(impl_->ctx, impl_->step_index, inputs.x0.data, nominal,
                      inputs.model_params_override.data, inputs.model_params_override.bytes,
                      inputs.cost_params_override.data, inputs.cost_params_override.bytes,
                      impl_->u_scratch.data(), diag) != MPPI_HIP_OK)
        return err(mppi_hip_last_error(impl_->ctx));
    std::copy(impl_->u_scratch.begin(), impl_->u_scratch.end(), u_next_out.data);
    if (diagnostics) {
        diagnostics->best_cost = diag[0];
        diagnostics->mean_cost = diag[1];
        diagnostics->effective_sample_size = diag[2];
        diagnostics->converged = false;
    }
    // Warm start: u_nom[t] = u[t+1], last step holds u[H-1].
    const uint32_t C = cfg.control_dim, H = cfg.horizon;
    for (uint32_t t = 0; t < H; ++t) {
        const uint32_t src = t + 1 < H ? t + 1 : H - 1;
        for (uint32_t d = 0; d < C; ++d)
            impl_->u_nominal[size_t(t) * C + d] = impl_->u_scratch[size_t(src) * C + d];
    }
    ++impl_->step_index;
    return true;
}

bool MPPIDriver::simulate(uint32_t num_steps, const StepInputs& inputs,
                          const SimulationResults& results, std::string* error) {
    ErrorSink err{error};
    if (!impl_->initialized) return err("Driver is not initialized. Call initialize() first.");
    if (num_steps == 0) return err("num_steps must be > 0.");
    const DriverConfig& cfg = impl_->config;
    std::string msg = validate_x0(inputs.x0, cfg.state_dim);
    if (!msg.empty()) return err(msg);
    if (results.num_steps != num_steps) return err("SimulationResults::num_steps mismatch.");
    const float* nominal = impl_->u_nominal.data();
    if (inputs.u_nominal_override.data != nullptr) {
        if (inputs.u_nominal_override.horizon != cfg.horizon ||
            inputs.u_nominal_override.control_dim != cfg.control_dim)
            return err("u_nominal_override shape mismatch.");
        nominal = inputs.u_nominal_override.data;
    }
    if (mppi_hip_simulate(impl_->ctx, num_steps, impl_->step_index, inputs.x0.data, nominal,
                          inputs.model_params_override.data, inputs.model_params_override.bytes,
                          inputs.cost_params_override.data, inputs.cost_params_override.bytes,
                          results.states_out, results.controls_out,
                          results.costs_out) != MPPI_HIP_OK)
        return err(mppi_hip_last_error(impl_->ctx));
    // As in the reference, the host warm start is left as it was.
    impl_->step_index += num_steps;
    return true;
}

bool MPPIDriver::simulate_batch(uint32_t num_agents, uint32_t num_steps, const float* x0_packed,
                                const BatchSimulationResults& results, std::string* error) {
    ErrorSink err{error};
    if (!impl_->initialized) return err("Driver is not initialized. Call initialize() first.");
    if (num_agents == 0) return err("num_agents must be > 0.");
    if (num_steps == 0) return err("num_steps must be > 0.");
    if (x0_packed == nullptr) return err("x0_packed must not be null.");
    if (results.num_agents != num_agents) return err("BatchSimulationResults::num_agents mismatch.");
    if (results.num_steps != num_steps) return err("BatchSimulationResults::num_steps mismatch.");
    // Every agent starts from a zero nominal and the bound parameters.
    if (mppi_hip_simulate_batch(impl_->ctx, num_agents, num_steps, impl_->step_index, x0_packed,
                                nullptr, results.states_out, results.controls_out,
                                results.costs_out, results.terminal_states_out) != MPPI_HIP_OK)
        return err(mppi_hip_last_error(impl_->ctx));
    impl_->step_index += num_steps;
    return true;
}

bool MPPIDriver::reset(std::string* /*error*/) {
    std::fill(impl_->u_nominal.begin(), impl_->u_nominal.end(), 0.0f);
    impl_->step_index = 0;
    return true;
}

bool MPPIDriver::reset(const ControlSequenceView& u_nominal, std::string* error) {
    ErrorSink err{error};
    const DriverConfig& cfg = impl_->config;
    if (u_nominal.horizon != cfg.horizon)
        return err("reset: u_nominal.horizon (" + std::to_string(u_nominal.horizon) +
                   ") must match DriverConfig::horizon (" + std::to_string(cfg.horizon) + ").");
    if (u_nominal.control_dim != cfg.control_dim)
        return err("reset: u_nominal.control_dim (" + std::to_string(u_nominal.control_dim) +
                   ") must match DriverConfig::control_dim (" + std::to_string(cfg.control_dim) +
                   ").");
    const size_t n = impl_->seq_len();
    if (n > 0 && u_nominal.data == nullptr)
        return err("reset: u_nominal.data must not be null when horizon * control_dim > 0.");
    std::copy(u_nominal.data, u_nominal.data + n, impl_->u_nominal.begin());
    impl_->step_index = 0;
    return true;
}

}  // namespace mppi_metal
