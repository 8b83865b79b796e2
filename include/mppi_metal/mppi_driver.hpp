// mppi_driver.hpp — the MPPIDriver facade, a drop-in for the reference's
// include/mppi_metal/mppi_driver.hpp:18-96 backed by the HIP C ABI
// (include/mppi_hip.h) instead of the Metal backend.
//
// Lifecycle (unchanged from the reference):
//   1. construct with a DriverConfig (pointer data is deep-copied);
//   2. initialize() with a ModelPluginSpec: validates, compiles and links the
//      plugin with the library kernels (hipRTC);
//   3. optionally set_problem_params() (blobs are copied to the device);
//   4. step() in a loop (blocking), or simulate() / simulate_batch();
//   5. reset() to clear the warm start.
// One driver owns one HIP stream; a driver is not thread-safe, use one per
// host thread.
#pragma once

#include "mppi_metal/mppi_types.hpp"

#include <memory>
#include <string>

namespace mppi_metal {

class MPPIDriver {
public:
    explicit MPPIDriver(const DriverConfig& config);
    ~MPPIDriver();

    MPPIDriver(const MPPIDriver&) = delete;
    MPPIDriver& operator=(const MPPIDriver&) = delete;
    MPPIDriver(MPPIDriver&&) noexcept;
    MPPIDriver& operator=(MPPIDriver&&) noexcept;

    bool initialize(const ModelPluginSpec& model, std::string* error = nullptr);
    bool set_problem_params(const ProblemParams& params, std::string* error = nullptr);

    // Blocking MPPI iteration; writes the updated control sequence and warm
    // starts the internal nominal (shift left, hold last).
    bool step(const StepInputs& inputs, const MutableControlSequenceView& u_next_out,
              StepDiagnostics* diagnostics, std::string* error = nullptr);

    // Zero the warm start; step index returns to 0 (as the reference's code
    // does, src/mppi_driver.cpp:493-500).
    bool reset(std::string* error = nullptr);
    bool reset(const ControlSequenceView& u_nominal, std::string* error = nullptr);

    // num_steps closed-loop steps on the device (plant = plugin dynamics).
    bool simulate(uint32_t num_steps, const StepInputs& initial_inputs,
                  const SimulationResults& results, std::string* error = nullptr);

    // num_agents closed loops from x0_packed, zero initial nominal.
    bool simulate_batch(uint32_t num_agents, uint32_t num_steps, const float* x0_packed,
                        const BatchSimulationResults& results, std::string* error = nullptr);

private:
    class Impl;
    std::unique_ptr<Impl> impl_;
};

}  // namespace mppi_metal

// The same API under the new library's own name.
namespace mppi_hip_api = mppi_metal;
