// batch_backend.cpp — the uneven multi-agent MPPI step (mppi_hip_batch_* in
// include/mppi_hip.h).
//
// A batch keeps every agent's state on the device: descriptors (K_a, H_a,
// buffer offsets), step counters, states, warm starts, parameter blobs.  One
// MPPI step of all agents is either a rollout launch and a reduce launch over
// tile tables, or one launch whose tiles roll out, sum their weighted noise
// and, in the last tile of each agent, form the update; the step is captured
// once in a HIP graph and the graph is replayed for every step (the step
// counters advance on the device, so the captured arguments never change).
// The launch shape — one or two launches, workgroup size, one workgroup per
// tile or a persistent grid taking tiles from a ticket counter, reduce
// samples per thread, tile order — is a mppi_hip_batch_tuning, chosen by the
// autotuner below by timing candidates with HIP events.
//
// Reference: MetalBackend::dispatch_batch_simulation
// (src/metal_backend.mm:935-1124) encodes the batched rollout, weights,
// reduce and propagate kernels (shaders/mppi_rollout.metal:103-170,
// shaders/mppi_reduce.metal:242-488) for agents that share K and H.
#include "backend.h"

#include <hip/hip_ext.h>

#include <algorithm>
#include <cstring>
#include <numeric>

using namespace mppi_hip_impl;

struct mppi_hip_batch {
    mppi_hip_ctx* ctx = nullptr;
    uint32_t N = 0;
    std::vector<uint32_t> K, H;
    mppi_hip_batch_tuning tuning{256, 0, 4, 0, 0, 2};
    std::shared_ptr<LoadedModule> mod;

    // layout of the current tuning
    std::vector<mppi_agent_desc> desc;
    std::vector<uint32_t> rollout_tiles, reduce_tiles;   // triples
    size_t seq_total = 0;                                // sum H_a * C

    // device memory
    DeviceBuffer d_desc, d_agents, d_x, d_unom, d_uout, d_diag, d_rtiles, d_dtiles, d_ticket;
    DeviceBuffer d_costs, d_weights, d_blk, d_part, d_model, d_cost;
    DeviceBuffer d_hist_states, d_hist_controls, d_hist_costs;
    size_t hist_rows = 0;
    size_t model_stride = 0, cost_stride = 0;
    bool params_set = false;

    // captured step graphs, keyed by (closed_loop, history pointers)
    struct Graph {
        hipGraphExec_t exec = nullptr;
        int closed_loop = -1;
        const void* hs = nullptr;
        const void* hc = nullptr;
        const void* hk = nullptr;
    };
    Graph graph[2];
};

namespace {

int bfail(mppi_hip_batch* b, int code, const std::string& msg) {
    return fail(b ? b->ctx : nullptr, code, msg);
}

void drop_graphs(mppi_hip_batch* b) {
    for (auto& g : b->graph) {
        if (g.exec) (void)hipGraphExecDestroy(g.exec);
        g = mppi_hip_batch::Graph{};
    }
}

int block_index(uint32_t block) { return block == 64 ? 0 : block == 128 ? 1 : 2; }

// The one-launch form sums across tiles in FP32, which the library keeps
// below 2^18 samples per agent.
bool one_launch_ok(const mppi_hip_batch* b) {
    for (uint32_t k : b->K)
        if (k >= (1u << 18)) return false;
    return true;
}

// Descriptors and tile tables for the current tuning; (re)allocates the
// tuning-dependent work buffers.
int build_layout(mppi_hip_batch* b) {
    mppi_hip_ctx* ctx = b->ctx;
    const uint32_t B = b->tuning.block, C = ctx->C;
    b->desc.assign(b->N, mppi_agent_desc{});
    uint64_t cost_off = 0, seq_off = 0, blk_off = 0, part_off = 0;
    uint64_t total_work = 0;
    for (uint32_t a = 0; a < b->N; ++a) total_work += uint64_t(b->K[a]) * ((b->H[a] * C + 3) / 4 + 7) / 8;
    for (uint32_t a = 0; a < b->N; ++a) {
        mppi_agent_desc& d = b->desc[a];
        d.K = b->K[a];
        d.H = b->H[a];
        d.n_blk = (d.K + B - 1) / B;
        d.NB = (d.H * C + 3) / 4;
        d.n_groups = (d.NB + kReduceBPG - 1) / kReduceBPG;
        uint32_t spt = b->tuning.reduce_spt;
        if (spt == 0) {
            // About 2048 reduce workgroups over the whole batch.
            uint64_t s = (total_work + uint64_t(B) * 2048 - 1) / (uint64_t(B) * 2048);
            spt = uint32_t(std::max<uint64_t>(1, std::min<uint64_t>(64, s)));
        }
        d.spt = spt;
        d.n_chunks = (d.K + B * spt - 1) / (B * spt);
        d.slot = a;
        d.cost_off = cost_off;
        d.w_off = cost_off;
        d.seq_off = seq_off;
        d.out_off = seq_off;
        d.blk_off = blk_off;
        d.part_off = part_off;
        cost_off += d.K;
        seq_off += uint64_t(d.H) * C;
        blk_off += d.n_blk;
        // rows of partials: the reduce's sample chunks, or one per tile in
        // the one-launch form
        part_off += uint64_t(b->tuning.launches == 1 ? d.n_blk : d.n_chunks) * d.NB * 4;
    }
    b->seq_total = seq_off;
    // Tile tables.  Rollout tile cost ~ H_a; reduce tile cost ~ spt * noise blocks.
    std::vector<uint32_t> order(b->N);
    std::iota(order.begin(), order.end(), 0u);
    if (b->tuning.order == 0)
        std::stable_sort(order.begin(), order.end(),
                         [&](uint32_t x, uint32_t y) { return b->H[x] > b->H[y]; });
    b->rollout_tiles.clear();
    b->reduce_tiles.clear();
    for (uint32_t a : order)
        for (uint32_t i = 0; i < b->desc[a].n_blk; ++i) b->rollout_tiles.insert(b->rollout_tiles.end(), {a, i, 0u});
    if (b->tuning.order == 0)
        std::stable_sort(order.begin(), order.end(), [&](uint32_t x, uint32_t y) {
            return b->desc[x].spt * std::min<uint32_t>(b->desc[x].NB, kReduceBPG) >
                   b->desc[y].spt * std::min<uint32_t>(b->desc[y].NB, kReduceBPG);
        });
    for (uint32_t a : order)
        for (uint32_t g = 0; g < b->desc[a].n_groups; ++g)
            for (uint32_t c = 0; c < b->desc[a].n_chunks; ++c)
                b->reduce_tiles.insert(b->reduce_tiles.end(), {a, c, g});

    int rc;
    if ((rc = ensure_device_buffer(ctx, &b->d_desc, sizeof(mppi_agent_desc) * b->N))) return rc;
    if ((rc = ensure_device_buffer(ctx, &b->d_rtiles, sizeof(uint32_t) * b->rollout_tiles.size()))) return rc;
    if ((rc = ensure_device_buffer(ctx, &b->d_dtiles, sizeof(uint32_t) * b->reduce_tiles.size()))) return rc;
    if ((rc = ensure_device_buffer(ctx, &b->d_costs, sizeof(float) * cost_off))) return rc;
    if ((rc = ensure_device_buffer(ctx, &b->d_weights, sizeof(float) * cost_off))) return rc;
    if ((rc = ensure_device_buffer(ctx, &b->d_blk, sizeof(float) * 4 * blk_off))) return rc;
    if ((rc = ensure_device_buffer(ctx, &b->d_part, sizeof(float) * part_off))) return rc;
    MPPI_HIP_TRY(ctx, hipMemcpyAsync(b->d_desc.ptr, b->desc.data(), sizeof(mppi_agent_desc) * b->N,
                                     hipMemcpyHostToDevice, ctx->stream));
    MPPI_HIP_TRY(ctx, hipMemcpyAsync(b->d_rtiles.ptr, b->rollout_tiles.data(),
                                     sizeof(uint32_t) * b->rollout_tiles.size(), hipMemcpyHostToDevice,
                                     ctx->stream));
    MPPI_HIP_TRY(ctx, hipMemcpyAsync(b->d_dtiles.ptr, b->reduce_tiles.data(),
                                     sizeof(uint32_t) * b->reduce_tiles.size(), hipMemcpyHostToDevice,
                                     ctx->stream));
    // The persistent kernels' ticket counters start every launch at a
    // multiple of n_tiles + grid; both may have changed.
    MPPI_HIP_TRY(ctx, hipMemsetAsync(b->d_ticket.ptr, 0, sizeof(uint64_t) * 3, ctx->stream));
    MPPI_HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    drop_graphs(b);
    return MPPI_HIP_OK;
}

mppi_step_args batch_args(mppi_hip_batch* b, int closed_loop) {
    mppi_hip_ctx* ctx = b->ctx;
    mppi_step_args a{};
    fill_common_args(ctx, &a);
    a.desc = static_cast<const mppi_agent_desc*>(b->d_desc.ptr);
    a.agents = static_cast<mppi_agent_state*>(b->d_agents.ptr);
    a.N = b->N;
    a.x = static_cast<float*>(b->d_x.ptr);
    a.u_nom = static_cast<float*>(b->d_unom.ptr);
    a.u_out = static_cast<float*>(b->d_uout.ptr);
    a.model = static_cast<const uint8_t*>(b->d_model.ptr);
    a.cost = static_cast<const uint8_t*>(b->d_cost.ptr);
    a.model_stride = b->model_stride;
    a.cost_stride = b->cost_stride;
    a.costs = static_cast<float*>(b->d_costs.ptr);
    a.weights = static_cast<float*>(b->d_weights.ptr);
    a.blk_stats = static_cast<float*>(b->d_blk.ptr);
    a.partials = static_cast<float*>(b->d_part.ptr);
    a.diag = static_cast<float*>(b->d_diag.ptr);
    a.closed_loop = closed_loop ? 1u : 0u;
    return a;
}

// Launch f on the context's stream; with `prof`, a start and stop event
// attached to the dispatch itself time the kernel alone and are queued on
// `pending` for collect_profile.
int launch_timed(mppi_hip_ctx* ctx, hipFunction_t f, uint32_t grid, uint32_t block, void** params, bool prof,
                 std::vector<std::pair<hipEvent_t, hipEvent_t>>& pending) {
    hipEvent_t e0 = prof ? take_event(ctx) : nullptr, e1 = prof ? take_event(ctx) : nullptr;
    if (e0 && e1) {
        MPPI_HIP_TRY(ctx, hipExtModuleLaunchKernel(f, grid * block, 1, 1, block, 1, 1, 0, ctx->stream, params,
                                                   nullptr, e0, e1, 0));
        pending.push_back({e0, e1});
        return MPPI_HIP_OK;
    }
    MPPI_HIP_TRY(ctx, hipModuleLaunchKernel(f, grid, 1, 1, block, 1, 1, 0, ctx->stream, params, nullptr));
    return MPPI_HIP_OK;
}

// Enqueue one step on the context's stream: the rollout and reduce launches,
// or the one-launch step.  With `prof` (never under capture) each launch is
// timed by its own events: the rollout, or the one-launch step, as the
// profile's rollout time, the reduce as its reduce time.
int enqueue_step(mppi_hip_batch* b, const mppi_step_args& base, bool prof = false) {
    mppi_hip_ctx* ctx = b->ctx;
    const int bi = block_index(b->tuning.block);
    const int pm = b->tuning.persistent ? 1 : 0;
    const uint32_t B = b->tuning.block;
    const uint32_t nr = uint32_t(b->rollout_tiles.size() / 3), nd = uint32_t(b->reduce_tiles.size() / 3);
    const uint32_t pgrid = uint32_t(ctx->num_cus) * std::max(1u, b->tuning.wgs_per_cu);
    uint64_t* tickets = static_cast<uint64_t*>(b->d_ticket.ptr);
    mppi_step_args ra = base, da = base;
    ra.tiles = static_cast<const uint32_t*>(b->d_rtiles.ptr);
    ra.n_tiles = nr;
    da.tiles = static_cast<const uint32_t*>(b->d_dtiles.ptr);
    da.n_tiles = nd;
    void* rp[] = {&ra};
    void* dp[] = {&da};
    const uint32_t rgrid = pm ? std::min(nr, pgrid) : nr;
    const uint32_t dgrid = pm ? std::min(nd, pgrid) : nd;
    int rc;
    if (b->tuning.launches == 1) {
        ra.ticket = tickets + 2;
        if ((rc = launch_timed(ctx, b->mod->step_b[bi][pm], rgrid, B, rp, prof, ctx->pending_rollout))) return rc;
    } else {
        ra.ticket = tickets + 0;
        da.ticket = tickets + 1;
        if ((rc = launch_timed(ctx, b->mod->rollout_b[bi][pm], rgrid, B, rp, prof, ctx->pending_rollout))) return rc;
        if ((rc = launch_timed(ctx, b->mod->reduce_b[bi][pm], dgrid, B, dp, prof, ctx->pending_reduce))) return rc;
        // the update: one workgroup per agent (timed with the reduce)
        if ((rc = launch_timed(ctx, b->mod->finalize, b->N, 256, dp, prof, ctx->pending_reduce))) return rc;
    }
    if (prof) ctx->prof_launches += 1;
    return MPPI_HIP_OK;
}

// The graph of one step for (closed_loop, history pointers); captured on
// first use.
int step_graph(mppi_hip_batch* b, const mppi_step_args& a, hipGraphExec_t* out) {
    mppi_hip_ctx* ctx = b->ctx;
    mppi_hip_batch::Graph& g = b->graph[a.closed_loop ? 1 : 0];
    if (g.exec && g.hs == a.hist_states && g.hc == a.hist_controls && g.hk == a.hist_costs) {
        *out = g.exec;
        return MPPI_HIP_OK;
    }
    if (g.exec) (void)hipGraphExecDestroy(g.exec);
    g = mppi_hip_batch::Graph{};
    MPPI_HIP_TRY(ctx, hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
    int rc = enqueue_step(b, a);
    hipGraph_t graph = nullptr;
    hipError_t e = hipStreamEndCapture(ctx->stream, &graph);
    if (rc) {
        if (graph) (void)hipGraphDestroy(graph);
        return rc;
    }
    if (e != hipSuccess) return fail(ctx, MPPI_HIP_E_DEVICE, std::string("hipStreamEndCapture failed: ") + hipGetErrorString(e));
    e = hipGraphInstantiate(&g.exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    if (e != hipSuccess) return fail(ctx, MPPI_HIP_E_DEVICE, std::string("hipGraphInstantiate failed: ") + hipGetErrorString(e));
    g.closed_loop = int(a.closed_loop);
    g.hs = a.hist_states;
    g.hc = a.hist_controls;
    g.hk = a.hist_costs;
    *out = g.exec;
    return MPPI_HIP_OK;
}

int run_steps(mppi_hip_batch* b, uint32_t T, int closed_loop, float* hs, float* hc, float* hk) {
    mppi_step_args a = batch_args(b, closed_loop);
    a.hist_states = hs;
    a.hist_controls = hc;
    a.hist_costs = hk;
    if (b->ctx->profiling) {
        // Timed launches, one step at a time, outside any graph.
        for (uint32_t s = 0; s < T; ++s) {
            int rc = enqueue_step(b, a, true);
            if (rc) return rc;
        }
        return MPPI_HIP_OK;
    }
    hipGraphExec_t exec = nullptr;
    int rc = step_graph(b, a, &exec);
    if (rc) return rc;
    for (uint32_t s = 0; s < T; ++s) MPPI_HIP_TRY(b->ctx, hipGraphLaunch(exec, b->ctx->stream));
    return MPPI_HIP_OK;
}

}  // namespace

extern "C" {

int mppi_hip_batch_create(mppi_hip_ctx* ctx, uint32_t num_agents, const uint32_t* sample_counts,
                          const uint32_t* horizons, mppi_hip_batch** out) {
    if (!out) return fail(ctx, MPPI_HIP_E_INVALID, "out_batch is null.");
    *out = nullptr;
    int rc = check_ctx(ctx);
    if (rc) return rc;
    if (num_agents == 0 || num_agents > (1u << 24))
        return fail(ctx, MPPI_HIP_E_INVALID, "num_agents must be in [1, 2^24].");
    if (!sample_counts || !horizons) return fail(ctx, MPPI_HIP_E_INVALID, "sample_counts and horizons must not be null.");
    for (uint32_t a = 0; a < num_agents; ++a) {
        if (sample_counts[a] == 0) return fail(ctx, MPPI_HIP_E_INVALID, "sample_counts[" + std::to_string(a) + "] must be > 0.");
        if (horizons[a] == 0 || uint64_t(horizons[a]) * ctx->C > kMaxHC)
            return fail(ctx, MPPI_HIP_E_INVALID,
                        "horizons[" + std::to_string(a) + "] must be in [1, 4096 / control_dim].");
    }
    std::unique_ptr<mppi_hip_batch> b(new mppi_hip_batch());
    b->ctx = ctx;
    b->N = num_agents;
    b->K.assign(sample_counts, sample_counts + num_agents);
    b->H.assign(horizons, horizons + num_agents);
    if (!ctx->batch_mod && (rc = load_module(ctx, true, false, &ctx->batch_mod))) return rc;
    b->mod = ctx->batch_mod;
    size_t seq = 0;
    for (uint32_t a = 0; a < num_agents; ++a) seq += size_t(b->H[a]) * ctx->C;
    if ((rc = ensure_device_buffer(ctx, &b->d_agents, sizeof(mppi_agent_state) * num_agents))) return rc;
    if ((rc = ensure_device_buffer(ctx, &b->d_x, sizeof(float) * num_agents * ctx->D))) return rc;
    if ((rc = ensure_device_buffer(ctx, &b->d_unom, sizeof(float) * seq))) return rc;
    if ((rc = ensure_device_buffer(ctx, &b->d_uout, sizeof(float) * seq))) return rc;
    if ((rc = ensure_device_buffer(ctx, &b->d_diag, sizeof(float) * 3 * num_agents))) return rc;
    if ((rc = ensure_device_buffer(ctx, &b->d_ticket, sizeof(uint64_t) * 3))) return rc;
    MPPI_HIP_TRY(ctx, hipMemsetAsync(b->d_agents.ptr, 0, sizeof(mppi_agent_state) * num_agents, ctx->stream));
    MPPI_HIP_TRY(ctx, hipMemsetAsync(b->d_x.ptr, 0, sizeof(float) * num_agents * ctx->D, ctx->stream));
    MPPI_HIP_TRY(ctx, hipMemsetAsync(b->d_unom.ptr, 0, sizeof(float) * seq, ctx->stream));
    // Until set_params: the context's bound blobs, shared by every agent.
    const size_t mb = std::max<size_t>(ctx->model_bytes, 16), cb = std::max<size_t>(ctx->cost_bytes, 16);
    if ((rc = ensure_device_buffer(ctx, &b->d_model, mb))) return rc;
    if ((rc = ensure_device_buffer(ctx, &b->d_cost, cb))) return rc;
    MPPI_HIP_TRY(ctx, hipMemsetAsync(b->d_model.ptr, 0, mb, ctx->stream));
    MPPI_HIP_TRY(ctx, hipMemsetAsync(b->d_cost.ptr, 0, cb, ctx->stream));
    if (ctx->model_is_bound)
        MPPI_HIP_TRY(ctx, hipMemcpyAsync(b->d_model.ptr, ctx->model_bound.ptr, ctx->model_bytes,
                                         hipMemcpyDeviceToDevice, ctx->stream));
    if (ctx->cost_is_bound)
        MPPI_HIP_TRY(ctx, hipMemcpyAsync(b->d_cost.ptr, ctx->cost_bound.ptr, ctx->cost_bytes,
                                         hipMemcpyDeviceToDevice, ctx->stream));
    if ((rc = build_layout(b.get()))) return rc;
    *out = b.release();
    return MPPI_HIP_OK;
}

void mppi_hip_batch_destroy(mppi_hip_batch* b) {
    if (!b) return;
    mppi_hip_ctx* ctx = b->ctx;
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    drop_graphs(b);
    if (ctx->last_costs == b->d_costs.ptr) {
        // read_costs / read_weights must not reach this batch's freed buffers.
        ctx->last_agents = 0;
        ctx->last_costs = nullptr;
        ctx->last_weights = nullptr;
        ctx->weights_pending = false;
        ctx->last_desc.clear();
    }
    for (DeviceBuffer* d : {&b->d_desc, &b->d_agents, &b->d_x, &b->d_unom, &b->d_uout, &b->d_diag,
                            &b->d_rtiles, &b->d_dtiles, &b->d_ticket, &b->d_costs, &b->d_weights,
                            &b->d_blk, &b->d_part, &b->d_model, &b->d_cost, &b->d_hist_states,
                            &b->d_hist_controls, &b->d_hist_costs})
        d->release();
    delete b;
}

int mppi_hip_batch_set_params(mppi_hip_batch* b, const void* model, size_t model_stride,
                              const void* cost, size_t cost_stride) {
    if (!b) return fail(nullptr, MPPI_HIP_E_INVALID, "batch is null.");
    mppi_hip_ctx* ctx = b->ctx;
    int rc = check_ctx(ctx);
    if (rc) return rc;
    if (ctx->model_bytes && !model) return bfail(b, MPPI_HIP_E_INVALID, "model params are null but the layout has model_params_size > 0.");
    if (ctx->cost_bytes && !cost) return bfail(b, MPPI_HIP_E_INVALID, "cost params are null but the layout has cost_params_size > 0.");
    if (model_stride && model_stride < ctx->model_bytes) return bfail(b, MPPI_HIP_E_INVALID, "model_stride is smaller than the model blob.");
    if (cost_stride && cost_stride < ctx->cost_bytes) return bfail(b, MPPI_HIP_E_INVALID, "cost_stride is smaller than the cost blob.");
    const size_t mbytes = model_stride ? model_stride * b->N : ctx->model_bytes;
    const size_t cbytes = cost_stride ? cost_stride * b->N : ctx->cost_bytes;
    if ((rc = ensure_device_buffer(ctx, &b->d_model, std::max<size_t>(mbytes, 16)))) return rc;
    if ((rc = ensure_device_buffer(ctx, &b->d_cost, std::max<size_t>(cbytes, 16)))) return rc;
    if (model && mbytes) MPPI_HIP_TRY(ctx, hipMemcpyAsync(b->d_model.ptr, model, mbytes, hipMemcpyHostToDevice, ctx->stream));
    if (cost && cbytes) MPPI_HIP_TRY(ctx, hipMemcpyAsync(b->d_cost.ptr, cost, cbytes, hipMemcpyHostToDevice, ctx->stream));
    MPPI_HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    b->model_stride = model_stride;
    b->cost_stride = cost_stride;
    b->params_set = true;
    drop_graphs(b);
    return MPPI_HIP_OK;
}

int mppi_hip_batch_set_tuning(mppi_hip_batch* b, const mppi_hip_batch_tuning* t) {
    if (!b || !t) return fail(b ? b->ctx : nullptr, MPPI_HIP_E_INVALID, "batch and tuning must not be null.");
    if (t->block != 64 && t->block != 128 && t->block != 256)
        return bfail(b, MPPI_HIP_E_INVALID, "tuning.block must be 64, 128 or 256.");
    if (t->persistent > 1 || t->order > 1) return bfail(b, MPPI_HIP_E_INVALID, "tuning.persistent and tuning.order must be 0 or 1.");
    if (t->launches != 1 && t->launches != 2) return bfail(b, MPPI_HIP_E_INVALID, "tuning.launches must be 1 or 2.");
    if (t->launches == 1 && !one_launch_ok(b))
        return bfail(b, MPPI_HIP_E_INVALID, "tuning.launches = 1 needs K < 2^18 for every agent.");
    if (t->persistent && (t->wgs_per_cu < 1 || t->wgs_per_cu > 16))
        return bfail(b, MPPI_HIP_E_INVALID, "tuning.wgs_per_cu must be in [1, 16].");
    if (t->reduce_spt > 64) return bfail(b, MPPI_HIP_E_INVALID, "tuning.reduce_spt must be <= 64.");
    int rc = check_ctx(b->ctx);
    if (rc) return rc;
    b->tuning = *t;
    return build_layout(b);
}

int mppi_hip_batch_get_tuning(const mppi_hip_batch* b, mppi_hip_batch_tuning* t) {
    if (!b || !t) return fail(nullptr, MPPI_HIP_E_INVALID, "batch and tuning must not be null.");
    *t = b->tuning;
    return MPPI_HIP_OK;
}

int mppi_hip_batch_reset(mppi_hip_batch* b, uint64_t start_step, const float* x0_packed,
                         const float* u_nominal_packed) {
    if (!b) return fail(nullptr, MPPI_HIP_E_INVALID, "batch is null.");
    mppi_hip_ctx* ctx = b->ctx;
    int rc = check_ctx(ctx);
    if (rc) return rc;
    if (!x0_packed) return bfail(b, MPPI_HIP_E_INVALID, "x0_packed must not be null.");
    std::vector<mppi_agent_state> ag(b->N);
    for (auto& s : ag) s.step = uint32_t(start_step);
    MPPI_HIP_TRY(ctx, hipMemcpyAsync(b->d_agents.ptr, ag.data(), sizeof(mppi_agent_state) * b->N,
                                     hipMemcpyHostToDevice, ctx->stream));
    MPPI_HIP_TRY(ctx, hipMemcpyAsync(b->d_x.ptr, x0_packed, sizeof(float) * b->N * ctx->D,
                                     hipMemcpyHostToDevice, ctx->stream));
    if (u_nominal_packed)
        MPPI_HIP_TRY(ctx, hipMemcpyAsync(b->d_unom.ptr, u_nominal_packed, sizeof(float) * b->seq_total,
                                         hipMemcpyHostToDevice, ctx->stream));
    else
        MPPI_HIP_TRY(ctx, hipMemsetAsync(b->d_unom.ptr, 0, sizeof(float) * b->seq_total, ctx->stream));
    MPPI_HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    return MPPI_HIP_OK;
}

int mppi_hip_batch_run(mppi_hip_batch* b, uint32_t num_steps, int closed_loop, float* states_out,
                       float* controls_out, float* costs_out) {
    if (!b) return fail(nullptr, MPPI_HIP_E_INVALID, "batch is null.");
    mppi_hip_ctx* ctx = b->ctx;
    int rc = check_ctx(ctx);
    if (rc) return rc;
    if (num_steps == 0) return bfail(b, MPPI_HIP_E_INVALID, "num_steps must be > 0.");
    const bool want_hist = closed_loop && (states_out || controls_out || costs_out);
    if ((states_out || controls_out || costs_out) && !closed_loop)
        return bfail(b, MPPI_HIP_E_INVALID, "history outputs need closed_loop = 1.");
    const size_t N = b->N, D = ctx->D, C = ctx->C;
    if (want_hist && b->hist_rows < num_steps) {
        drop_graphs(b);
        if ((rc = ensure_device_buffer(ctx, &b->d_hist_states, sizeof(float) * num_steps * N * D))) return rc;
        if ((rc = ensure_device_buffer(ctx, &b->d_hist_controls, sizeof(float) * num_steps * N * C))) return rc;
        if ((rc = ensure_device_buffer(ctx, &b->d_hist_costs, sizeof(float) * num_steps * N))) return rc;
        b->hist_rows = num_steps;
    }
    if (want_hist) {
        // History rows restart at 0 for this call.
        std::vector<mppi_agent_state> ag(b->N);
        MPPI_HIP_TRY(ctx, hipMemcpyAsync(ag.data(), b->d_agents.ptr, sizeof(mppi_agent_state) * N,
                                         hipMemcpyDeviceToHost, ctx->stream));
        MPPI_HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
        for (auto& s : ag) s.hist = 0;
        MPPI_HIP_TRY(ctx, hipMemcpyAsync(b->d_agents.ptr, ag.data(), sizeof(mppi_agent_state) * N,
                                         hipMemcpyHostToDevice, ctx->stream));
    }
    rc = run_steps(b, num_steps, closed_loop,
                   states_out ? static_cast<float*>(b->d_hist_states.ptr) : nullptr,
                   controls_out ? static_cast<float*>(b->d_hist_controls.ptr) : nullptr,
                   costs_out ? static_cast<float*>(b->d_hist_costs.ptr) : nullptr);
    if (rc) return rc;
    if (states_out)
        MPPI_HIP_TRY(ctx, hipMemcpyAsync(states_out, b->d_hist_states.ptr, sizeof(float) * num_steps * N * D,
                                         hipMemcpyDeviceToHost, ctx->stream));
    if (controls_out)
        MPPI_HIP_TRY(ctx, hipMemcpyAsync(controls_out, b->d_hist_controls.ptr, sizeof(float) * num_steps * N * C,
                                         hipMemcpyDeviceToHost, ctx->stream));
    if (costs_out)
        MPPI_HIP_TRY(ctx, hipMemcpyAsync(costs_out, b->d_hist_costs.ptr, sizeof(float) * num_steps * N,
                                         hipMemcpyDeviceToHost, ctx->stream));
    MPPI_HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    ctx->last_agents = b->N;
    ctx->last_costs = static_cast<const float*>(b->d_costs.ptr);
    ctx->last_weights = static_cast<const float*>(b->d_weights.ptr);
    ctx->weights_pending = false;
    if (b->tuning.launches == 1) {
        // The one-launch form does not form per-sample weights; they are
        // formed from the costs and the agents' normalisers when read.
        ctx->weights_pending = true;
        ctx->weights_args = batch_args(b, closed_loop);
        ctx->weights_grid_x = (*std::max_element(b->K.begin(), b->K.end()) + kBlock - 1) / kBlock;
        ctx->weights_grid_y = b->N;
    }
    ctx->last_desc = b->desc;
    return MPPI_HIP_OK;
}

int mppi_hip_batch_read(mppi_hip_batch* b, float* x_out, float* u_out_packed, float* diag_out) {
    if (!b) return fail(nullptr, MPPI_HIP_E_INVALID, "batch is null.");
    mppi_hip_ctx* ctx = b->ctx;
    int rc = check_ctx(ctx);
    if (rc) return rc;
    if (x_out)
        MPPI_HIP_TRY(ctx, hipMemcpyAsync(x_out, b->d_x.ptr, sizeof(float) * b->N * ctx->D,
                                         hipMemcpyDeviceToHost, ctx->stream));
    if (u_out_packed)
        MPPI_HIP_TRY(ctx, hipMemcpyAsync(u_out_packed, b->d_uout.ptr, sizeof(float) * b->seq_total,
                                         hipMemcpyDeviceToHost, ctx->stream));
    if (diag_out)
        MPPI_HIP_TRY(ctx, hipMemcpyAsync(diag_out, b->d_diag.ptr, sizeof(float) * 3 * b->N,
                                         hipMemcpyDeviceToHost, ctx->stream));
    MPPI_HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    return MPPI_HIP_OK;
}

int mppi_hip_batch_autotune(mppi_hip_batch* b, uint32_t trial_steps, double* best_ms) {
    if (!b) return fail(nullptr, MPPI_HIP_E_INVALID, "batch is null.");
    mppi_hip_ctx* ctx = b->ctx;
    int rc = check_ctx(ctx);
    if (rc) return rc;
    if (trial_steps == 0) trial_steps = 3;
    // Two launches: tile tables over block x reduce spt x order, and the
    // persistent grid at two sizes.  One launch (when it applies): tile
    // table or persistent grid, block x order.
    std::vector<mppi_hip_batch_tuning> cands;
    for (uint32_t block : {64u, 128u, 256u}) {
        for (uint32_t spt : {1u, 2u, 4u, 8u, 16u, 32u})
            for (uint32_t order : {0u, 1u}) cands.push_back({block, 0, 0, spt, order, 2});
        for (uint32_t w : {2u, 4u}) cands.push_back({block, 1, w, 0, 0, 2});
        if (one_launch_ok(b))
            for (uint32_t order : {0u, 1u}) {
                cands.push_back({block, 0, 0, 0, order, 1});
                for (uint32_t w : {2u, 4u}) cands.push_back({block, 1, w, 0, order, 1});
            }
    }
    hipEvent_t e0 = take_event(ctx), e1 = take_event(ctx);
    if (!e0 || !e1) return fail(ctx, MPPI_HIP_E_DEVICE, "hipEventCreate failed.");
    double best = 1e300;
    mppi_hip_batch_tuning best_t = b->tuning;
    for (const auto& t : cands) {
        b->tuning = t;
        if ((rc = build_layout(b))) return rc;
        if ((rc = run_steps(b, 1, 0, nullptr, nullptr, nullptr))) return rc;  // warm-up + capture
        MPPI_HIP_TRY(ctx, hipEventRecord(e0, ctx->stream));
        if ((rc = run_steps(b, trial_steps, 0, nullptr, nullptr, nullptr))) return rc;
        MPPI_HIP_TRY(ctx, hipEventRecord(e1, ctx->stream));
        MPPI_HIP_TRY(ctx, hipEventSynchronize(e1));
        float ms = 0.0f;
        MPPI_HIP_TRY(ctx, hipEventElapsedTime(&ms, e0, e1));
        const double per = double(ms) / trial_steps;
        if (per < best) {
            best = per;
            best_t = t;
        }
    }
    ctx->ev_pool.push_back(e0);
    ctx->ev_pool.push_back(e1);
    b->tuning = best_t;
    if ((rc = build_layout(b))) return rc;
    if (best_ms) *best_ms = best;
    return MPPI_HIP_OK;
}

}  // extern "C"
