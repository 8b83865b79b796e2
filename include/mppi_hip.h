/* mppi_hip.h — C ABI of the MI355X MPPI backend.
 *
 * This is the seam that replaces the reference's MetalBackend class
 * (src/metal_backend.hpp:22-108), one entry point per method, with plain
 * pointers and sizes so it can be bound from C, C++ (the MPPIDriver facade in
 * include/mppi_metal/mppi_driver.hpp), Python ctypes or any other FFI.
 *
 * Conventions
 *   - Every function returns an int: MPPI_HIP_OK (0) or a negative
 *     mppi_hip_status.  The message of the last failure on a context is
 *     returned by mppi_hip_last_error(ctx); failures before a context exists
 *     (mppi_hip_create) are reported by mppi_hip_last_error(NULL), which is
 *     per calling thread.
 *   - All host pointers are caller-owned; nothing is retained after a call
 *     returns (parameter blobs are copied at bind time).
 *   - A context owns one HIP stream on one device.  Calls on one context are
 *     not thread-safe; distinct contexts may be used from distinct threads
 *     (the reference's concurrency model: one driver per host thread).
 *   - Calls that return results (step, simulate, simulate_batch) block until
 *     the device work has finished, like the reference's waitUntilCompleted
 *     (src/metal_backend.mm:719-720).
 */
#ifndef MPPI_HIP_H
#define MPPI_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum mppi_hip_status {
    MPPI_HIP_OK = 0,
    MPPI_HIP_E_INVALID = -1,          /* bad argument, shape or size */
    MPPI_HIP_E_NOT_INITIALIZED = -2,  /* no plugin linked yet */
    MPPI_HIP_E_PLUGIN = -3,           /* plugin source missing or callable not found */
    MPPI_HIP_E_COMPILE = -4,          /* hipRTC compilation or module load failed */
    MPPI_HIP_E_DEVICE = -5,           /* HIP runtime error (no device, launch failure) */
    MPPI_HIP_E_NO_MEMORY = -6         /* device or pinned allocation failed */
};

enum mppi_hip_covariance { MPPI_HIP_COV_DIAGONAL = 0, MPPI_HIP_COV_FULL = 1 };
enum mppi_hip_numeric_mode { MPPI_HIP_FP32 = 0, MPPI_HIP_FP16 = 1, MPPI_HIP_MIXED = 2 };

typedef struct mppi_hip_ctx mppi_hip_ctx;

/* Problem dimensions and noise model; the fields of the reference's
 * DriverConfig (include/mppi_metal/mppi_types.hpp:46-62). */
typedef struct mppi_hip_config {
    uint32_t state_dim;
    uint32_t control_dim;
    uint32_t horizon;
    uint32_t sample_count;
    float lambda;
    uint32_t covariance_form;   /* mppi_hip_covariance */
    /* DIAGONAL: control_dim standard deviations (epsilon_d = sigma_d * z_d).
     * FULL: control_dim x control_dim covariance, row-major, symmetric
     * positive definite; epsilon = L z with L the lower Cholesky factor. */
    const float* sigma;
    const float* u_min;         /* control_dim */
    const float* u_max;         /* control_dim */
    uint32_t numeric_mode;      /* mppi_hip_numeric_mode (the reference declares the modes,
                                   DriverConfig::NumericMode, and uses none):
                                   FP32  every operation exactly rounded as written; sample
                                         costs bit-identical to the CPU oracle
                                   MIXED FP32 data, fast hardware transcendentals, fused
                                         multiply-adds and approximate division in the
                                         library's noise path and in plugin mppi_* math
                                   FP16  MIXED with the control noise rounded to binary16 */
    int32_t device;             /* HIP device ordinal; -1 = current device */
} mppi_hip_config;

/* Create a context: copies the configuration, selects the device and creates
 * the stream.  Replaces MetalBackend construction plus the device part of
 * MetalBackend::initialize (src/metal_backend.mm:225-231). */
int mppi_hip_create(const mppi_hip_config* config, mppi_hip_ctx** out_ctx);

/* Release all device resources of a context.  Replaces ~MetalBackend and
 * MetalBackend::reset (src/metal_backend.mm:1130). */
void mppi_hip_destroy(mppi_hip_ctx* ctx);

/* Message of the last failure on ctx (or of the calling thread if ctx is
 * NULL).  Valid until the next call on the same context / thread. */
const char* mppi_hip_last_error(const mppi_hip_ctx* ctx);

/* Compile the plugin at plugin_path with the library kernels through hipRTC
 * and load the code object.  dynamics is required; stage_cost and
 * terminal_cost may be NULL or "" (the library's zero costs are used).
 * model_bytes / cost_bytes are the ParamLayout sizes; seed is the 64-bit
 * RngConfig::seed (all 64 bits are used).
 * Replaces MetalBackend::initialize (src/metal_backend.mm:225-568): library
 * and user metallib loading, callable lookup, MTLLinkedFunctions and the
 * visible function tables. */
int mppi_hip_link_plugin(mppi_hip_ctx* ctx, const char* plugin_path, const char* dynamics,
                         const char* stage_cost, const char* terminal_cost,
                         size_t model_bytes, size_t cost_bytes, uint64_t seed);

/* Same as mppi_hip_link_plugin, with the plugin given as source text;
 * source_name is used in diagnostics.  No reference counterpart (the
 * reference only loads plugins from files); used by callers that generate
 * plugins. */
int mppi_hip_link_plugin_source(mppi_hip_ctx* ctx, const char* source, const char* source_name,
                                const char* dynamics, const char* stage_cost,
                                const char* terminal_cost, size_t model_bytes,
                                size_t cost_bytes, uint64_t seed);

/* Copy the model / cost parameter blobs to the device for later steps.
 * Sizes must equal the linked layout (0 means "none").  Replaces
 * MetalBackend::bind_params (src/metal_backend.mm:574-598). */
int mppi_hip_bind_params(mppi_hip_ctx* ctx, const void* model, size_t model_bytes,
                         const void* cost, size_t cost_bytes);

/* One blocking MPPI iteration: rollout, softmax weights, weighted-noise
 * reduce.  x0: state_dim floats; u_nominal: horizon*control_dim floats;
 * overrides with bytes == 0 use the bound blobs; u_out: horizon*control_dim
 * floats; diag3 (may be NULL): best cost, mean cost, effective sample size.
 * step_index selects the noise stream (low 32 bits are used).
 * Replaces MetalBackend::dispatch_rollout (src/metal_backend.mm:604-741). */
int mppi_hip_step(mppi_hip_ctx* ctx, uint64_t step_index, const float* x0,
                  const float* u_nominal, const void* model_override, size_t model_override_bytes,
                  const void* cost_override, size_t cost_override_bytes, float* u_out,
                  float* diag3);

/* mppi_hip_step without allocation: for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid it).  No counterpart in the reference.  Inputs and outputs are
 * device pointers owned by the caller: x0_dev (state_dim floats),
 * u_nominal_dev (horizon*control_dim floats, read only), u_out_dev
 * (horizon*control_dim floats), diag3_dev (3 floats, may be NULL); the
 * bound parameter blobs are used (no overrides).  The step's work buffers
 * (costs, workgroup statistics, partial sums, descriptor, agent state) live
 * in a file-scope device array of the library, one per device, at most
 * 8 MiB (K = 65536 at horizon*control_dim = 128 needs about 0.4 MiB);
 * sizes beyond it fail with MPPI_HIP_E_INVALID.  The call issues no
 * hipMalloc, hipFree or other allocating routine.  stream: a hipStream_t
 * to enqueue on (the call then returns without waiting and without the
 * wait check of mppi_hip_set_wait_limit), or NULL for the context's stream,
 * waiting for the result like mppi_hip_step.  Every call on one device
 * shares that device's array and overwrites the previous call's contents:
 * calls that overlap on one device (concurrent streams or threads) race.
 * Afterwards mppi_hip_read_costs / read_weights report no iteration. */
int mppi_hip_step_noalloc(mppi_hip_ctx* ctx, uint64_t step_index, const float* x0_dev,
                          const float* u_nominal_dev, float* u_out_dev, float* diag3_dev, void* stream);

/* num_steps closed-loop steps on the device: each step runs the MPPI
 * iteration, logs (state, first optimal control, best cost), steps the plant
 * with the plugin dynamics and shifts the warm start (hold last).  The noise
 * stream of step i is start_step + i.  Output pointers may be NULL.
 * states_out: num_steps*state_dim, controls_out: num_steps*control_dim,
 * costs_out: num_steps.
 * Replaces MetalBackend::dispatch_simulation (src/metal_backend.mm:747-929). */
int mppi_hip_simulate(mppi_hip_ctx* ctx, uint32_t num_steps, uint64_t start_step,
                      const float* x0, const float* u_nominal, const void* model_override,
                      size_t model_override_bytes, const void* cost_override,
                      size_t cost_override_bytes, float* states_out, float* controls_out,
                      float* costs_out);

/* num_agents independent closed loops sharing the linked plugin, the bound
 * parameters, K and H.  x0_packed: num_agents*state_dim; u_nominal_packed:
 * num_agents*horizon*control_dim, or NULL for all zeros.  Outputs (each may be
 * NULL) are indexed [step][agent][element]:
 * states_out num_steps*num_agents*state_dim, controls_out
 * num_steps*num_agents*control_dim, costs_out num_steps*num_agents;
 * terminal_out num_agents*state_dim is the state after the last plant step.
 * Agent a draws its noise from counter word 3 = agent a (the key is the seed
 * for every agent).
 * Replaces MetalBackend::dispatch_batch_simulation
 * (src/metal_backend.mm:935-1124). */
int mppi_hip_simulate_batch(mppi_hip_ctx* ctx, uint32_t num_agents, uint32_t num_steps,
                            uint64_t start_step, const float* x0_packed,
                            const float* u_nominal_packed, float* states_out,
                            float* controls_out, float* costs_out, float* terminal_out);

/* Inspection of the last iteration for agent `agent` (0 for step/simulate):
 * its K sample costs, or its K normalised weights.  No reference counterpart
 * (the reference keeps these in private Metal buffers); used by parity tests. */
int mppi_hip_read_costs(mppi_hip_ctx* ctx, uint32_t agent, float* out, size_t count);
int mppi_hip_read_weights(mppi_hip_ctx* ctx, uint32_t agent, float* out, size_t count);

/* Which kernels run an MPPI iteration of mppi_hip_step / simulate /
 * simulate_batch.  No reference counterpart.
 *   TWO_KERNEL  rollout, then a reduce that regenerates the noise from its
 *               counters and forms the weighted average in parallel over
 *               samples;
 *   FUSED       one pass over the noise: the rollout keeps each tile's
 *               normals in LDS (producer wavefronts write them, consumer
 *               wavefronts roll out from them) and forms the tile's weighted
 *               sums, then a small combine launch forms the update (needs
 *               horizon * control_dim <= 128 and at most 256 tiles of 256
 *               samples, K <= 65536, else the two kernels run); per-sample
 *               weights are then formed when mppi_hip_read_weights asks;
 *   AUTO        FUSED where it applies and the rollout has at most one
 *               workgroup per compute unit, else TWO_KERNEL (the default). */
enum mppi_hip_step_path { MPPI_HIP_PATH_AUTO = 0, MPPI_HIP_PATH_TWO_KERNEL = 1, MPPI_HIP_PATH_FUSED = 2 };
int mppi_hip_set_step_path(mppi_hip_ctx* ctx, uint32_t path);

/* Launch-shape knobs of the reduce kernel (samples per thread); 0 restores
 * the built-in choice.  No reference counterpart; used by the autotuner. */
int mppi_hip_set_reduce_samples_per_thread(mppi_hip_ctx* ctx, uint32_t spt);

/* The one-pass step's rollout wavefronts wait (bounded) for the noise that
 * their producer wavefronts write to LDS.  A wavefront that reaches the bound
 * flags its agent, and the call (step, simulate, simulate_batch) then fails
 * with MPPI_HIP_E_DEVICE instead of returning results computed from missing
 * noise.  polls: loads of the producer's progress word before giving up (0:
 * the default, 2^20).  force_timeout != 0 makes every wait target unreachable
 * so that each wait runs to the bound and gives up (every kernel still ends
 * normally); it exists to test that error path.  No reference counterpart. */
int mppi_hip_set_wait_limit(mppi_hip_ctx* ctx, uint32_t polls, int force_timeout);

/* Device timing: when enabled, every rollout and reduce launch is bracketed by
 * HIP events on the context's stream and their durations are summed.  This
 * includes the launches of a batch on the context (mppi_hip_batch_run): its
 * one-launch step counts as a rollout, and while profiling a batch launches
 * each step directly instead of replaying its captured graph.
 * mppi_hip_profile_read returns the sums in milliseconds and the launch count
 * and clears them.  No reference counterpart. */
int mppi_hip_profile_enable(mppi_hip_ctx* ctx, int enable);
int mppi_hip_profile_read(mppi_hip_ctx* ctx, double* rollout_ms, double* reduce_ms,
                          uint64_t* launches);

/* ---------------------------------------------------------------------------
 * Batched multi-agent step with uneven sample counts and horizons.
 * No reference counterpart as such: the reference's simulate_batch
 * (src/metal_backend.mm:935-1124) requires one K and one H for all agents.
 * A batch is bound to a context (plugin, dimensions, noise model, lambda,
 * bounds, seed) and holds N agents, agent a with K_a samples and horizon
 * H_a (H_a * control_dim <= 4096).  An MPPI step of all agents runs over
 * tile lists sorted longest first, in one of two forms:
 *   two launches  a rollout launch (tiles: agent x block of samples) and a
 *                 reduce launch that regenerates the noise (tiles: agent x
 *                 sample chunk x noise-block group);
 *   one launch    every tile (agent x block of samples) rolls out its samples,
 *                 weights them against its own best cost and sums its
 *                 weighted noise; the last tile of each agent to finish forms
 *                 the agent's update (needs K_a < 2^18 for every agent).
 * Each launch runs one workgroup per tile (tile table) or a persistent grid
 * of compute units x wgs_per_cu workgroups taking tiles from a ticket
 * counter.  The step is captured once in a HIP graph that is replayed for
 * every step.  The launch shape is chosen by mppi_hip_batch_autotune or set
 * explicitly.
 * Agent a draws its noise from counter word 3 = a, so a batch agent and a
 * single-agent context see the same noise when their indices agree.
 * ------------------------------------------------------------------------- */
typedef struct mppi_hip_batch mppi_hip_batch;

typedef struct mppi_hip_batch_tuning {
    uint32_t block;        /* workgroup size: 64, 128 or 256 */
    uint32_t persistent;   /* 0: one workgroup per tile; 1: persistent grid taking tiles
                              from a ticket counter */
    uint32_t wgs_per_cu;   /* persistent grid size = compute units * wgs_per_cu (1..16) */
    uint32_t reduce_spt;   /* samples per reduce thread (1..64); 0: built-in heuristic
                              (two-launch form only) */
    uint32_t order;        /* 0: longest tiles first; 1: agent order */
    uint32_t launches;     /* kernel launches per step: 2 (rollout, reduce) or 1 */
} mppi_hip_batch_tuning;

/* sample_counts, horizons: num_agents entries each. */
int mppi_hip_batch_create(mppi_hip_ctx* ctx, uint32_t num_agents, const uint32_t* sample_counts,
                          const uint32_t* horizons, mppi_hip_batch** out_batch);
void mppi_hip_batch_destroy(mppi_hip_batch* batch);

/* Parameter blobs of the linked layout.  stride 0: one blob shared by all
 * agents; otherwise num_agents blobs, stride bytes apart (stride >= size). */
int mppi_hip_batch_set_params(mppi_hip_batch* batch, const void* model, size_t model_stride,
                              const void* cost, size_t cost_stride);

int mppi_hip_batch_set_tuning(mppi_hip_batch* batch, const mppi_hip_batch_tuning* tuning);
int mppi_hip_batch_get_tuning(const mppi_hip_batch* batch, mppi_hip_batch_tuning* tuning);

/* Time the candidate launch shapes over trial_steps open-loop steps each
 * (state is left unchanged) and keep the fastest; its device time per step
 * in milliseconds goes to best_ms (may be NULL). */
int mppi_hip_batch_autotune(mppi_hip_batch* batch, uint32_t trial_steps, double* best_ms);

/* Set every agent's state, warm start and step counter.  x0_packed:
 * num_agents * state_dim; u_nominal_packed: sum_a H_a * control_dim floats,
 * agent after agent, or NULL for zeros. */
int mppi_hip_batch_reset(mppi_hip_batch* batch, uint64_t start_step, const float* x0_packed,
                         const float* u_nominal_packed);

/* num_steps MPPI steps of all agents.  closed_loop 1: after each step every
 * agent logs (state, first control, best cost), steps its plant with the
 * dynamics and shifts its warm start; closed_loop 0: open-loop solves from
 * the current state and warm start.  History outputs (may be NULL) are
 * [step][agent][element].  Blocks until done. */
int mppi_hip_batch_run(mppi_hip_batch* batch, uint32_t num_steps, int closed_loop,
                       float* states_out, float* controls_out, float* costs_out);

/* Current state (num_agents * state_dim), last u_next (packed like the warm
 * start) and diagnostics (num_agents * 3); each may be NULL. */
int mppi_hip_batch_read(mppi_hip_batch* batch, float* x_out, float* u_out_packed, float* diag_out);

/* ---------------------------------------------------------------------------
 * Learned-dynamics surrogate.  No reference counterpart.  An MLP
 * (state_dim + control_dim) -> 128 -> 128 -> state_dim with a saturating
 * hidden activation T (mppi_hip_surrogate_activation), FP16 weights on the
 * matrix cores and FP32 accumulation (BF16 in training's backward pass), trained on
 * the GPU on transitions that the linked plugin's own dynamics generates
 * (states drawn uniformly from [state_lo, state_hi], controls from the
 * control bounds), predicting the state change of one step.  When enabled,
 * the rollout of mppi_hip_step / mppi_hip_simulate uses the network in place
 * of the plugin dynamics; the plugin stays the plant of the closed loop and
 * the judge of the network's error.  Requires state_dim + control_dim <= 16.
 * ------------------------------------------------------------------------- */
/* How the network sees a state component. */
#define MPPI_HIP_SURROGATE_INPUT_BOX 0       /* normalised over [lo, hi] */
#define MPPI_HIP_SURROGATE_INPUT_PERIODIC 1  /* the dynamics is periodic in it with period
                                                hi - lo: wrapped into [lo, hi] first */
#define MPPI_HIP_SURROGATE_INPUT_IGNORED 2   /* the dynamics is invariant to it: not an input */

/* Floats in the normalisation block of export / import: input centre, input
 * 1 / half-range (0 for an ignored input), input period (0 if not periodic),
 * output mean, output standard deviation; 16 each. */
#define MPPI_HIP_SURROGATE_NORM_FLOATS 80

typedef struct mppi_hip_surrogate_config {
    const float* state_lo;      /* state_dim: sampling box of the training states */
    const float* state_hi;
    const uint32_t* state_input; /* state_dim MPPI_HIP_SURROGATE_INPUT_* (NULL: all BOX) */
    uint32_t num_transitions;   /* generated transitions (rounded up to whole batches) */
    uint32_t batch;             /* samples per optimiser step, a multiple of 128 */
    uint32_t steps;             /* Adam steps */
    float learning_rate;        /* peak step size (cosine decay to 10%) */
    uint64_t seed;              /* data stream and initialisation */
} mppi_hip_surrogate_config;

/* Generate the data and train; loss_history (steps floats, may be NULL)
 * receives the batch MSE (normalised units) of every step. */
int mppi_hip_surrogate_train(mppi_hip_ctx* ctx, const mppi_hip_surrogate_config* config,
                             float* loss_history);

/* Error of one predicted step on num_samples fresh transitions (a data stream
 * the training never saw): per state component, the RMSE of the predicted
 * state change and that RMSE divided by the standard deviation of the true
 * change (each array state_dim floats, either may be NULL). */
int mppi_hip_surrogate_evaluate(mppi_hip_ctx* ctx, uint32_t num_samples, uint64_t seed,
                                float* rmse, float* relative_rmse);

/* Next states predicted by the network for n (state, control) pairs. */
int mppi_hip_surrogate_predict(mppi_hip_ctx* ctx, uint32_t n, const float* states,
                               const float* controls, float* next_states);

/* 1: rollouts use the network; 0: the plugin dynamics (the default). */
int mppi_hip_surrogate_use(mppi_hip_ctx* ctx, int enable);

/* Checkpoint: FP32 parameters ([W1 | b1 | W2 | b2 | W3 | b3], row-major
 * [out][in]) and the normalisation block (MPPI_HIP_SURROGATE_NORM_FLOATS). */
int mppi_hip_surrogate_export(mppi_hip_ctx* ctx, float* params, size_t capacity, size_t* n_params,
                              float* norm);
int mppi_hip_surrogate_import(mppi_hip_ctx* ctx, const float* params, size_t n_params,
                              const float* norm);

/* One forward and backward pass on n samples (a multiple of 128) of
 * normalised inputs x (n * (state_dim + control_dim)) and raw state-change
 * targets (n * state_dim), without updating: the summed gradients of the
 * batch MSE in grads (n_params floats) and the loss. */
int mppi_hip_surrogate_gradients(mppi_hip_ctx* ctx, uint32_t n, const float* x, const float* targets,
                                 float* grads, float* loss);

/* The network's hidden activation T of n pre-activations z, evaluated on the
 * device: tanh as T(z) = Q(y), y = clamp01(r16(z)) (z rounded to binary16,
 * nearest even, clamped to [0, 1], NaN to 0), Q = s P(v), s = 2 y - 1,
 * v = s^2 - 1/2, P of degree 4 with binary16 coefficients, evaluated as
 * seven binary16 operations each rounded to nearest even (s = fma(y, 2, -1),
 * v = fma(s, s, -1/2), four Horner fused multiply-adds, T = s * p); T is
 * within 4.31e-3 of tanh(5.54 (r16(z) - 1/2)) for every z.  out receives the
 * n binary16 results as floats.  No
 * reference counterpart: it lets an application reproduce the network on
 * the CPU bit for bit in its activations. */
int mppi_hip_surrogate_activation(mppi_hip_ctx* ctx, uint32_t n, const float* z, float* out);

/* ---------------------------------------------------------------------------
 * Serving: agents admitted against an HBM budget and solved in batches of any
 * composition.  No reference counterpart (the reference serves concurrent
 * request streams with one blocking driver per host thread,
 * examples/nonholonomic_vehicle_concurrent/main.cpp:98-177).
 * A server is bound to a context (plugin, noise model, numeric mode, seed,
 * bound model blob) and owns one device arena of hbm_budget_bytes.  Admitting
 * an agent reserves its workspace for its K and H (costs, weights, workgroup
 * statistics, reduce partials, warm start, update) in the arena and stores its
 * state, warm start and cost blob on the device; when no free range of the
 * arena fits, admission fails with MPPI_HIP_E_NO_MEMORY and can be retried
 * after releases.  A solve runs one closed-loop MPPI step for any set of
 * admitted agents in one pair of kernel launches: the plant step with the
 * plugin dynamics and the warm-start shift happen on the device, and only the
 * first control and the diagnostics come back.  An agent's noise stream is
 * keyed by its id, so its results do not depend on its batch.
 * ------------------------------------------------------------------------- */
typedef struct mppi_hip_server mppi_hip_server;

int mppi_hip_server_create(mppi_hip_ctx* ctx, uint32_t max_agents, size_t hbm_budget_bytes,
                           mppi_hip_server** out_server);
void mppi_hip_server_destroy(mppi_hip_server* server);

/* Arena bytes an agent with K samples and horizon H reserves. */
int mppi_hip_server_workspace_bytes(const mppi_hip_server* server, uint32_t K, uint32_t H,
                                    size_t* bytes);

/* x0: state_dim floats; u_nominal: H*control_dim floats or NULL (zeros);
 * cost_params: cost_bytes of the linked layout (its per-agent cost blob; the
 * model blob is the context's bound one).  Writes the agent id. */
int mppi_hip_server_admit(mppi_hip_server* server, uint32_t K, uint32_t H, const float* x0,
                          const float* u_nominal, const void* cost_params, size_t cost_bytes,
                          uint32_t* agent_id);
int mppi_hip_server_release(mppi_hip_server* server, uint32_t agent_id);

/* One closed-loop step of n distinct admitted agents.  first_controls:
 * n*control_dim (the control applied by the plant step), diag: n*3 (best
 * cost, mean cost, effective sample size); either may be NULL.  Blocking.
 * Afterwards mppi_hip_read_costs / read_weights(ctx, i, ...) return agent
 * agent_ids[i]'s sample costs and weights of this solve. */
int mppi_hip_server_solve(mppi_hip_server* server, uint32_t n, const uint32_t* agent_ids,
                          float* first_controls, float* diag);

/* An agent's device state (state_dim) and warm start (H*control_dim). */
int mppi_hip_server_read(mppi_hip_server* server, uint32_t agent_id, float* x_out,
                         float* u_nominal_out);

int mppi_hip_server_usage(const mppi_hip_server* server, size_t* used_bytes, size_t* budget_bytes,
                          uint32_t* agents);

/* Serving loop over a request trace in real time: requests arrive at
 * arrival_s (seconds after the call; non-decreasing); arrived requests are
 * admitted in arrival order while they fit the budget (the oldest waits when
 * it does not), every admitted request is in the next batch (at most
 * max_batch agents, 0 = no limit), and a request completes after `solves`
 * closed-loop solves, when its agent is released.  latency_s[i] = completion
 * minus arrival of request i (n doubles). */
typedef struct mppi_hip_request {
    double arrival_s;
    uint32_t K, H;
    uint32_t solves;
    uint32_t reserved;
    const float* x0;            /* state_dim */
    const float* u_nominal;     /* H*control_dim, or NULL */
    const void* cost_params;    /* the linked cost layout's size, or NULL when it is 0 */
} mppi_hip_request;

typedef struct mppi_hip_replay_stats {
    double seconds;             /* wall time of the whole replay */
    double sample_steps;        /* sum of K*H over every solve */
    double max_workspace_bytes; /* largest arena use after a batch */
    uint64_t batches;
    uint32_t max_batch;
    uint32_t admission_waits;   /* loop passes in which the oldest request did not fit */
} mppi_hip_replay_stats;

int mppi_hip_server_replay(mppi_hip_server* server, uint32_t n, const mppi_hip_request* requests,
                           uint32_t max_batch, double* latency_s, mppi_hip_replay_stats* stats);

/* Library version string. */
const char* mppi_hip_version(void);

/* Measured issue ceilings of `device` for roofline figures: the rate at
 * which its SIMDs retire plain (unpacked) FP32 vector instructions, in lane
 * operations per second (an FMA counts once), and 32x32x16 FP16 matrix
 * instructions (the surrogate rollout's products), in flops per second, each
 * from long streams of independent instructions of that kind at eight
 * wavefronts per SIMD; and, in mixed_over_sum (may be NULL), the time of a
 * stream with one matrix instruction after every sixteen FP32 FMAs over the
 * sum of the times of its two parts alone (1: the SIMD does not overlap
 * them; 0.5: fully).  About 40 ms.  No reference counterpart (measurement
 * only).  The published FP32 vector peak counts packed FP32 instructions,
 * which this library does not emit. */
int mppi_hip_calibrate(int device, double* valu_lane_ops_per_s, double* mfma_f16_flops_per_s,
                       double* mixed_over_sum);

/* Compile a plugin with the library kernels for gfx950 without a device
 * (build hosts, CI): the same sources and options as linking.  No reference
 * counterpart (the reference compiles its shaders ahead of time with xcrun,
 * CMakeLists.txt:21-45).  kernel_sets: bit 0 adds the uneven-batch kernels, bit 1
 * the surrogate kernels.  On MPPI_HIP_E_COMPILE the compiler log is copied
 * into log (log_capacity bytes, may be NULL). */
int mppi_hip_compile_check(const char* plugin_source, const char* include_dir, const char* dynamics_fn,
                           const char* stage_cost_fn, const char* terminal_cost_fn, uint32_t state_dim,
                           uint32_t control_dim, int covariance_full, uint32_t numeric_mode,
                           uint32_t kernel_sets, char* log, size_t log_capacity);

#ifdef __cplusplus
}
#endif

#endif /* MPPI_HIP_H */
