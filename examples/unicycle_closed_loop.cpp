// Closed-loop unicycle runs through the MPPIDriver API, logged in the
// trajectory CSV of the reference's examples:
//   sim_id,step,x,y,theta,v,a_cmd,omega_cmd,best_cost
// (examples/nonholonomic_vehicle/main.cpp:88-186 writes it from a host loop,
// examples/nonholonomic_vehicle_concurrent/main.cpp:98-177 from simulate(),
// examples/nonholonomic_vehicle_batch/main.cpp:136-173 from simulate_batch()).
//
// Modes:
//   step      each simulation runs on the host: step(), then the plant (the
//             unicycle with Gaussian actuation noise), then the warm start
//             shifted one step with a zero last control, set with reset(u);
//   simulate  each simulation is one simulate() call (the loop on the GPU);
//   batch     all simulations are agents of one simulate_batch() call.
// In the step and simulate modes every host thread owns one driver (one
// context, one stream) and takes simulations from a shared counter, which
// is how the reference's examples use several drivers on one device.  Each
// simulation starts with reset(), so its rows depend only on its index,
// not on the thread that ran it or on the number of threads.
//
// Usage: unicycle_closed_loop --plugin PATH [--mode step|simulate|batch]
//        [--sims N] [--threads T] [--steps S] [--samples K] [--horizon H]
//        [--out FILE]
#include "mppi_metal/mppi_driver.hpp"

#include <atomic>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iostream>
#include <mutex>
#include <random>
#include <sstream>
#include <string>
#include <thread>
#include <vector>

namespace {

// The plugin's model and cost blobs (mppi_metal_amd/plugins/unicycle.hip:
// UnicycleModel, and the first 24 bytes of UnicycleCost, which the
// mppi_stage_cost / mppi_terminal_cost callables read).
struct Model {
    float v_max, dt;
};
struct Goal {
    float x, y, radius, pos_weight, heading_weight, control_weight;
};

struct Options {
    std::string mode = "step", plugin, out = "trajectory.csv";
    int sims = 16, threads = 4, steps = 300;
    uint32_t samples = 256, horizon = 20;
};

bool parse(int argc, char** argv, Options* o) {
    for (int i = 1; i < argc; ++i) {
        const std::string a = argv[i];
        if (i + 1 >= argc) return false;
        const char* v = argv[++i];
        if (a == "--mode") o->mode = v;
        else if (a == "--plugin") o->plugin = v;
        else if (a == "--out") o->out = v;
        else if (a == "--sims") o->sims = std::atoi(v);
        else if (a == "--threads") o->threads = std::atoi(v);
        else if (a == "--steps") o->steps = std::atoi(v);
        else if (a == "--samples") o->samples = uint32_t(std::atoi(v));
        else if (a == "--horizon") o->horizon = uint32_t(std::atoi(v));
        else return false;
    }
    return !o->plugin.empty() && o->sims > 0 && o->threads > 0 && o->steps > 0 &&
           (o->mode == "step" || o->mode == "simulate" || o->mode == "batch");
}

const float kSigma[2] = {2.0f, 1.0f};
const float kSigmaInv[2] = {0.5f, 1.0f};
const float kUMin[2] = {-3.0f, -1.5f};
const float kUMax[2] = {3.0f, 1.5f};
const Model kModel = {2.0f, 0.05f};
const Goal kGoal = {5.0f, 5.0f, 0.2f, 10.0f, 0.0f, 0.1f};

mppi_metal::DriverConfig make_config(const Options& o) {
    mppi_metal::DriverConfig c;
    c.state_dim = 4;
    c.control_dim = 2;
    c.horizon = o.horizon;
    c.sample_count = o.samples;
    c.lambda = 1.0f;
    c.noise.covariance_form = mppi_metal::ControlNoiseConfig::CovarianceForm::kDiagonal;
    c.noise.sigma = kSigma;
    c.noise.sigma_inv = kSigmaInv;
    c.noise.control_dim = 2;
    c.bounds.u_min = kUMin;
    c.bounds.u_max = kUMax;
    c.bounds.control_dim = 2;
    return c;
}

bool make_driver(const Options& o, mppi_metal::MPPIDriver* d, std::string* err) {
    mppi_metal::ModelPluginSpec spec;
    spec.user_metallib_path = o.plugin;
    spec.callables.dynamics = "mppi_dynamics";
    spec.callables.stage_cost = "mppi_stage_cost";
    spec.callables.terminal_cost = "mppi_terminal_cost";
    spec.param_layout.model_params_size = sizeof(Model);
    spec.param_layout.cost_params_size = sizeof(Goal);
    if (!d->initialize(spec, err)) return false;
    mppi_metal::ProblemParams p;
    p.model_params = {reinterpret_cast<const uint8_t*>(&kModel), sizeof(Model)};
    p.cost_params = {reinterpret_cast<const uint8_t*>(&kGoal), sizeof(Goal)};
    return d->set_problem_params(p, err);
}

void row(std::ostream& os, int sim, int step, const float* x, const float* u, float best) {
    os << sim << ',' << step << ',' << x[0] << ',' << x[1] << ',' << x[2] << ',' << x[3] << ',' << u[0]
       << ',' << u[1] << ',' << best << '\n';
}

bool at_goal(const float* x) {
    const float dx = x[0] - kGoal.x, dy = x[1] - kGoal.y;
    return std::sqrt(dx * dx + dy * dy) < kGoal.radius;
}

// The plant of the host loop: the plugin's unicycle equations in single
// precision, driven by the commanded controls plus actuation noise.
void plant(float* x, float a, float omega) {
    const float dt = kModel.dt, pi = 3.14159265358979f;
    const float th = x[2], v = x[3];
    x[0] += v * std::cos(th) * dt;
    x[1] += v * std::sin(th) * dt;
    x[2] += omega * dt;
    x[3] += a * dt;
    float w = std::fmod(x[2] + pi, 2.0f * pi);
    if (w < 0.0f) w += 2.0f * pi;
    x[2] = w - pi;
    x[3] = std::min(std::max(x[3], -kModel.v_max), kModel.v_max);
}

// One simulation in step mode; returns the steps taken, or -1 on error.
int run_step_mode(mppi_metal::MPPIDriver& d, const Options& o, int sim, std::ostream& os, std::string* err) {
    const uint32_t H = o.horizon;
    std::vector<float> x = {0.0f, 0.0f, 0.0f, 0.0f}, u(H * 2, 0.0f);
    std::mt19937 gen(uint32_t(1000 + sim));
    std::normal_distribution<float> na(0.0f, 0.2f), nw(0.0f, 0.1f);
    if (!d.reset(err)) return -1;
    mppi_metal::MutableControlSequenceView out{u.data(), H, 2};
    for (int t = 0; t < o.steps; ++t) {
        mppi_metal::StepInputs in;
        in.x0 = {x.data(), 4};
        mppi_metal::StepDiagnostics diag;
        if (!d.step(in, out, &diag, err)) return -1;
        row(os, sim, t, x.data(), u.data(), diag.best_cost);
        plant(x.data(), u[0] + na(gen), u[1] + nw(gen));
        if (at_goal(x.data())) return t + 1;
        for (uint32_t k = 0; k + 1 < H; ++k) {
            u[2 * k] = u[2 * k + 2];
            u[2 * k + 1] = u[2 * k + 3];
        }
        u[2 * H - 2] = 0.0f;
        u[2 * H - 1] = 0.0f;
        if (!d.reset(mppi_metal::ControlSequenceView{u.data(), H, 2}, err)) return -1;
    }
    return o.steps;
}

int run_simulate_mode(mppi_metal::MPPIDriver& d, const Options& o, int sim, std::ostream& os, std::string* err) {
    const uint32_t T = uint32_t(o.steps);
    std::vector<float> x0 = {0.0f, 0.0f, 0.0f, 0.0f};
    std::vector<float> xs(T * 4), us(T * 2), cs(T);
    if (!d.reset(err)) return -1;
    mppi_metal::StepInputs in;
    in.x0 = {x0.data(), 4};
    mppi_metal::SimulationResults r;
    r.states_out = xs.data();
    r.controls_out = us.data();
    r.costs_out = cs.data();
    r.num_steps = T;
    if (!d.simulate(T, in, r, err)) return -1;
    for (uint32_t t = 0; t < T; ++t) {
        row(os, sim, int(t), &xs[4 * t], &us[2 * t], cs[t]);
        if (at_goal(&xs[4 * t])) return int(t) + 1;
    }
    return int(T);
}

}  // namespace

int main(int argc, char** argv) {
    Options o;
    if (!parse(argc, argv, &o)) {
        std::fprintf(stderr,
                     "usage: %s --plugin PATH [--mode step|simulate|batch] [--sims N] [--threads T] "
                     "[--steps S] [--samples K] [--horizon H] [--out FILE]\n",
                     argv[0]);
        return 2;
    }
    const mppi_metal::DriverConfig config = make_config(o);
    std::ofstream csv(o.out);
    csv << "sim_id,step,x,y,theta,v,a_cmd,omega_cmd,best_cost\n";
    std::atomic<int> reached{0}, failed{0};

    if (o.mode == "batch") {
        mppi_metal::MPPIDriver d(config);
        std::string err;
        if (!make_driver(o, &d, &err)) {
            std::cerr << "driver: " << err << "\n";
            return 1;
        }
        const uint32_t N = uint32_t(o.sims), T = uint32_t(o.steps);
        std::vector<float> x0(N * 4, 0.0f), xs(size_t(N) * T * 4), us(size_t(N) * T * 2), cs(size_t(N) * T);
        mppi_metal::BatchSimulationResults r;
        r.states_out = xs.data();
        r.controls_out = us.data();
        r.costs_out = cs.data();
        r.num_agents = N;
        r.num_steps = T;
        if (!d.simulate_batch(N, T, x0.data(), r, &err)) {
            std::cerr << "simulate_batch: " << err << "\n";
            return 1;
        }
        for (uint32_t a = 0; a < N; ++a)   // layout [step][agent][dim]
            for (uint32_t t = 0; t < T; ++t) {
                const float* x = &xs[(size_t(t) * N + a) * 4];
                row(csv, int(a), int(t), x, &us[(size_t(t) * N + a) * 2], cs[size_t(t) * N + a]);
                if (at_goal(x)) {
                    ++reached;
                    break;
                }
            }
    } else {
        std::atomic<int> next{0};
        std::mutex csv_mutex;
        std::vector<std::thread> pool;
        for (int i = 0; i < o.threads; ++i)
            pool.emplace_back([&]() {
                mppi_metal::MPPIDriver d(config);
                std::string err;
                if (!make_driver(o, &d, &err)) {
                    std::lock_guard<std::mutex> lock(csv_mutex);
                    std::cerr << "driver: " << err << "\n";
                    ++failed;
                    return;
                }
                for (int sim = next++; sim < o.sims; sim = next++) {
                    std::ostringstream rows;
                    const int taken = o.mode == "step" ? run_step_mode(d, o, sim, rows, &err)
                                                       : run_simulate_mode(d, o, sim, rows, &err);
                    std::lock_guard<std::mutex> lock(csv_mutex);
                    if (taken < 0) {
                        std::cerr << "simulation " << sim << ": " << err << "\n";
                        ++failed;
                        return;
                    }
                    csv << rows.str();
                    if (taken < o.steps) ++reached;
                }
            });
        for (auto& t : pool) t.join();
    }
    std::cout << reached.load() << "/" << o.sims << " reached the goal; " << failed.load() << " failed\n";
    return failed.load() ? 1 : 0;
}
