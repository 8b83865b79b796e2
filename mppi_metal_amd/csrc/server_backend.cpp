// server_backend.cpp — serving: agents admitted against an HBM budget and
// solved in batches of any composition (mppi_hip_server_* in
// include/mppi_hip.h).  No reference counterpart: the reference runs one
// driver per request stream (examples/nonholonomic_vehicle_concurrent).
//
// Device memory of a server:
//   arena      one allocation of the HBM budget; each admitted agent holds one
//              contiguous workspace in it: costs[K] | weights[K] |
//              block statistics[4 n_blk] | reduce partials[4 NB n_chunks] |
//              warm start[H C] | update[H C] (floats, 256-byte aligned),
//              placed first-fit; admission fails when no free range fits
//   per slot   state x[D], mppi_agent_state, cost blob (max_agents slots,
//              allocated at creation, outside the budget)
// A solve of n agents uploads their descriptors (slot and arena offsets) and
// the tile tables in one copy, runs the tile-table rollout and reduce
// kernels of the uneven batch in closed-loop mode (plant step and warm-start
// shift on the device), and downloads n x (3 diagnostics + first control).
// The noise stream of an agent is keyed by its slot, so its results do not
// depend on which agents it is batched with.
#include "backend.h"

#include <algorithm>
#include <chrono>
#include <cstring>
#include <map>
#include <thread>

using namespace mppi_hip_impl;

namespace {

constexpr size_t kAlignFloats = 64;   // 256-byte workspace alignment
constexpr uint32_t kServerBlock = 256;

size_t round_up(size_t x, size_t a) { return (x + a - 1) / a * a; }

struct Workspace {
    uint32_t K = 0, H = 0, spt = 1, n_blk = 0, NB = 0, n_chunks = 0, n_groups = 0;
    size_t off = 0, floats = 0;   // range in the arena
};

// Shape and arena footprint (floats) of an agent with K samples and horizon H.
Workspace shape(uint32_t K, uint32_t H, uint32_t C) {
    Workspace w;
    w.K = K;
    w.H = H;
    // About 16 reduce sample chunks per agent at the largest K.
    w.spt = std::max<uint32_t>(1, std::min<uint32_t>(8, K / 4096));
    w.n_blk = (K + kServerBlock - 1) / kServerBlock;
    w.NB = (H * C + 3) / 4;
    w.n_chunks = (K + kServerBlock * w.spt - 1) / (kServerBlock * w.spt);
    w.n_groups = (w.NB + kReduceBPG - 1) / kReduceBPG;
    const size_t HC = size_t(H) * C;
    w.floats = round_up(K, kAlignFloats) * 2 + round_up(4 * size_t(w.n_blk), kAlignFloats) +
               round_up(4 * size_t(w.NB) * w.n_chunks, kAlignFloats) + round_up(HC, kAlignFloats) * 2;
    return w;
}

}  // namespace

struct mppi_hip_server {
    mppi_hip_ctx* ctx = nullptr;
    std::shared_ptr<LoadedModule> mod;
    uint32_t max_agents = 0;
    size_t arena_floats = 0;
    size_t cost_stride = 0;
    DeviceBuffer arena, d_x, d_agents, d_cost, d_up, d_down;
    PinnedBuffer up_host, down_host;
    std::map<size_t, size_t> free_ranges;   // offset -> length (floats), coalesced
    std::vector<Workspace> ws;              // per slot
    std::vector<uint8_t> in_use;            // per slot
    std::vector<uint32_t> free_slots;
    size_t used_floats = 0;
    uint32_t active = 0;
};

namespace {

int sfail(mppi_hip_server* s, int code, const std::string& msg) {
    return fail(s ? s->ctx : nullptr, code, msg);
}

bool arena_alloc(mppi_hip_server* s, size_t n, size_t* off) {
    for (auto it = s->free_ranges.begin(); it != s->free_ranges.end(); ++it) {
        if (it->second < n) continue;
        *off = it->first;
        const size_t rest = it->second - n;
        const size_t next = it->first + n;
        s->free_ranges.erase(it);
        if (rest) s->free_ranges[next] = rest;
        s->used_floats += n;
        return true;
    }
    return false;
}

void arena_free(mppi_hip_server* s, size_t off, size_t n) {
    auto it = s->free_ranges.emplace(off, n).first;
    if (it != s->free_ranges.begin()) {   // merge with the previous range
        auto prev = std::prev(it);
        if (prev->first + prev->second == off) {
            prev->second += it->second;
            s->free_ranges.erase(it);
            it = prev;
        }
    }
    auto next = std::next(it);
    if (next != s->free_ranges.end() && it->first + it->second == next->first) {
        it->second += next->second;
        s->free_ranges.erase(next);
    }
    s->used_floats -= n;
}

mppi_agent_desc desc_of(const Workspace& w, uint32_t slot, uint32_t C) {
    mppi_agent_desc d{};
    d.K = w.K;
    d.H = w.H;
    d.n_blk = w.n_blk;
    d.NB = w.NB;
    d.n_chunks = w.n_chunks;
    d.n_groups = w.n_groups;
    d.spt = w.spt;
    d.slot = slot;
    const size_t HC = size_t(w.H) * C;
    size_t o = w.off;
    d.cost_off = o;
    o += round_up(w.K, kAlignFloats);
    d.w_off = o;
    o += round_up(w.K, kAlignFloats);
    d.blk_off = o / 4;   // blk_stats is indexed in workgroups of 4 floats
    o += round_up(4 * size_t(w.n_blk), kAlignFloats);
    d.part_off = o;
    o += round_up(4 * size_t(w.NB) * w.n_chunks, kAlignFloats);
    d.seq_off = o;
    o += round_up(HC, kAlignFloats);
    d.out_off = o;
    return d;
}

int check_server(mppi_hip_server* s) {
    if (!s) return fail(nullptr, MPPI_HIP_E_INVALID, "server is null.");
    return check_ctx(s->ctx);
}

int check_agent(mppi_hip_server* s, uint32_t id) {
    if (id >= s->max_agents || !s->in_use[id])
        return sfail(s, MPPI_HIP_E_INVALID, "agent " + std::to_string(id) + " is not admitted.");
    return MPPI_HIP_OK;
}

int solve_impl(mppi_hip_server* s, uint32_t n, const uint32_t* ids, float* first_controls, float* diag) {
    mppi_hip_ctx* ctx = s->ctx;
    const uint32_t C = ctx->C;
    std::vector<mppi_agent_desc> desc(n);
    std::vector<uint32_t> order(n);
    for (uint32_t i = 0; i < n; ++i) {
        desc[i] = desc_of(s->ws[ids[i]], ids[i], C);
        order[i] = i;
    }
    // Longest rollouts first, so the long tiles start early.
    std::stable_sort(order.begin(), order.end(), [&](uint32_t a, uint32_t b) { return desc[a].H > desc[b].H; });
    std::vector<uint32_t> rt, dt;
    for (uint32_t i : order)
        for (uint32_t b = 0; b < desc[i].n_blk; ++b) rt.insert(rt.end(), {i, b, 0u});
    for (uint32_t i : order)
        for (uint32_t g = 0; g < desc[i].n_groups; ++g)
            for (uint32_t c = 0; c < desc[i].n_chunks; ++c) dt.insert(dt.end(), {i, c, g});
    const size_t b_desc = sizeof(mppi_agent_desc) * n;
    const size_t b_rt = sizeof(uint32_t) * rt.size(), b_dt = sizeof(uint32_t) * dt.size();
    const size_t up = align_up(b_desc, 256) + align_up(b_rt, 256) + b_dt;
    const size_t down = sizeof(float) * size_t(n) * (3 + C);
    int rc;
    if ((rc = ensure_device_buffer(ctx, &s->d_up, up))) return rc;
    if ((rc = ensure_pinned_buffer(ctx, &s->up_host, up))) return rc;
    if ((rc = ensure_device_buffer(ctx, &s->d_down, down))) return rc;
    if ((rc = ensure_pinned_buffer(ctx, &s->down_host, down))) return rc;
    uint8_t* h = static_cast<uint8_t*>(s->up_host.ptr);
    std::memcpy(h, desc.data(), b_desc);
    std::memcpy(h + align_up(b_desc, 256), rt.data(), b_rt);
    std::memcpy(h + align_up(b_desc, 256) + align_up(b_rt, 256), dt.data(), b_dt);
    MPPI_HIP_TRY(ctx, hipMemcpyAsync(s->d_up.ptr, h, up, hipMemcpyHostToDevice, ctx->stream));

    uint8_t* dbase = static_cast<uint8_t*>(s->d_up.ptr);
    float* arena = static_cast<float*>(s->arena.ptr);
    float* out = static_cast<float*>(s->d_down.ptr);
    mppi_step_args a{};
    fill_common_args(ctx, &a);
    a.desc = reinterpret_cast<const mppi_agent_desc*>(dbase);
    a.agents = static_cast<mppi_agent_state*>(s->d_agents.ptr);
    a.N = n;
    a.x = static_cast<float*>(s->d_x.ptr);
    a.u_nom = arena;
    a.u_out = arena;
    a.costs = arena;
    a.weights = arena;
    a.blk_stats = arena;
    a.partials = arena;
    a.model = ctx->model_is_bound ? static_cast<const uint8_t*>(ctx->model_bound.ptr)
                                  : reinterpret_cast<const uint8_t*>(static_cast<const float*>(ctx->consts.ptr) +
                                                                     ctx->noise_scale.size() + 2 * C);
    a.model_stride = 0;
    a.cost = static_cast<const uint8_t*>(s->d_cost.ptr);
    a.cost_stride = s->cost_stride;
    a.diag = out;
    a.first_out = out + size_t(n) * 3;
    a.closed_loop = 1;
    mppi_step_args ra = a, da = a;
    ra.tiles = reinterpret_cast<const uint32_t*>(dbase + align_up(b_desc, 256));
    ra.n_tiles = uint32_t(rt.size() / 3);
    da.tiles = reinterpret_cast<const uint32_t*>(dbase + align_up(b_desc, 256) + align_up(b_rt, 256));
    da.n_tiles = uint32_t(dt.size() / 3);
    void* rp[] = {&ra};
    void* dp[] = {&da};
    MPPI_HIP_TRY(ctx, hipModuleLaunchKernel(s->mod->rollout_b[2][0], ra.n_tiles, 1, 1, kServerBlock, 1, 1, 0,
                                            ctx->stream, rp, nullptr));
    MPPI_HIP_TRY(ctx, hipModuleLaunchKernel(s->mod->reduce_b[2][0], da.n_tiles, 1, 1, kServerBlock, 1, 1, 0,
                                            ctx->stream, dp, nullptr));
    // the update: one workgroup of 256 threads per agent
    MPPI_HIP_TRY(ctx, hipModuleLaunchKernel(s->mod->finalize, uint32_t(n), 1, 1, 256, 1, 1, 0, ctx->stream, dp,
                                            nullptr));
    MPPI_HIP_TRY(ctx, hipMemcpyAsync(s->down_host.ptr, out, down, hipMemcpyDeviceToHost, ctx->stream));
    MPPI_HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    const float* r = static_cast<const float*>(s->down_host.ptr);
    if (diag) std::memcpy(diag, r, sizeof(float) * 3 * n);
    if (first_controls) std::memcpy(first_controls, r + size_t(n) * 3, sizeof(float) * C * n);
    // read_costs / read_weights see this solve (agent = position in ids).
    ctx->last_agents = n;
    ctx->last_costs = arena;
    ctx->last_weights = arena;
    ctx->weights_pending = false;
    ctx->last_desc = desc;
    return MPPI_HIP_OK;
}

}  // namespace

extern "C" {

int mppi_hip_server_create(mppi_hip_ctx* ctx, uint32_t max_agents, size_t hbm_budget_bytes,
                           mppi_hip_server** out) {
    if (!out) return fail(ctx, MPPI_HIP_E_INVALID, "out_server is null.");
    *out = nullptr;
    int rc = check_ctx(ctx);
    if (rc) return rc;
    if (max_agents == 0 || max_agents > (1u << 24))
        return fail(ctx, MPPI_HIP_E_INVALID, "max_agents must be in [1, 2^24].");
    if (hbm_budget_bytes < 4096) return fail(ctx, MPPI_HIP_E_INVALID, "hbm_budget_bytes must be >= 4096.");
    std::unique_ptr<mppi_hip_server> s(new mppi_hip_server());
    s->ctx = ctx;
    s->max_agents = max_agents;
    if (!ctx->linked) return fail(ctx, MPPI_HIP_E_NOT_INITIALIZED, "server: no plugin linked.");
    if (!ctx->batch_mod && (rc = load_module(ctx, true, false, &ctx->batch_mod))) return rc;
    s->mod = ctx->batch_mod;
    s->arena_floats = hbm_budget_bytes / sizeof(float) / kAlignFloats * kAlignFloats;
    s->cost_stride = align_up(std::max<size_t>(ctx->cost_bytes, 16), 16);
    if ((rc = ensure_device_buffer(ctx, &s->arena, sizeof(float) * s->arena_floats))) return rc;
    if ((rc = ensure_device_buffer(ctx, &s->d_x, sizeof(float) * max_agents * ctx->D))) return rc;
    if ((rc = ensure_device_buffer(ctx, &s->d_agents, sizeof(mppi_agent_state) * max_agents))) return rc;
    if ((rc = ensure_device_buffer(ctx, &s->d_cost, s->cost_stride * max_agents))) return rc;
    MPPI_HIP_TRY(ctx, hipMemsetAsync(s->d_agents.ptr, 0, sizeof(mppi_agent_state) * max_agents, ctx->stream));
    MPPI_HIP_TRY(ctx, hipMemsetAsync(s->d_cost.ptr, 0, s->cost_stride * max_agents, ctx->stream));
    MPPI_HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    s->free_ranges[0] = s->arena_floats;
    s->ws.assign(max_agents, Workspace{});
    s->in_use.assign(max_agents, 0);
    s->free_slots.resize(max_agents);
    for (uint32_t i = 0; i < max_agents; ++i) s->free_slots[i] = max_agents - 1 - i;   // pop_back gives 0 first
    *out = s.release();
    return MPPI_HIP_OK;
}

void mppi_hip_server_destroy(mppi_hip_server* s) {
    if (!s) return;
    if (s->ctx && s->ctx->stream) (void)hipStreamSynchronize(s->ctx->stream);
    if (s->ctx && s->ctx->last_costs == s->arena.ptr) {
        s->ctx->last_costs = nullptr;
        s->ctx->last_weights = nullptr;
        s->ctx->last_agents = 0;
    }
    for (DeviceBuffer* b : {&s->arena, &s->d_x, &s->d_agents, &s->d_cost, &s->d_up, &s->d_down}) b->release();
    s->up_host.release();
    s->down_host.release();
    delete s;
}

int mppi_hip_server_workspace_bytes(const mppi_hip_server* s, uint32_t K, uint32_t H, size_t* bytes) {
    if (!s || !bytes) return fail(nullptr, MPPI_HIP_E_INVALID, "server and bytes must not be null.");
    *bytes = sizeof(float) * shape(K, H, s->ctx->C).floats;
    return MPPI_HIP_OK;
}

int mppi_hip_server_admit(mppi_hip_server* s, uint32_t K, uint32_t H, const float* x0,
                          const float* u_nominal, const void* cost_params, size_t cost_bytes,
                          uint32_t* agent_id) {
    int rc = check_server(s);
    if (rc) return rc;
    mppi_hip_ctx* ctx = s->ctx;
    if (!agent_id || !x0) return sfail(s, MPPI_HIP_E_INVALID, "x0 and agent_id must not be null.");
    if (K == 0 || H == 0 || uint64_t(H) * ctx->C > kMaxHC)
        return sfail(s, MPPI_HIP_E_INVALID, "K must be > 0 and H in [1, 4096 / control_dim].");
    if (cost_bytes != ctx->cost_bytes || (cost_bytes && !cost_params))
        return sfail(s, MPPI_HIP_E_INVALID,
                     "cost_params must have the linked layout's size (" + std::to_string(ctx->cost_bytes) + ").");
    if (s->free_slots.empty())
        return sfail(s, MPPI_HIP_E_NO_MEMORY, "no free agent slot (max_agents = " + std::to_string(s->max_agents) + ").");
    Workspace w = shape(K, H, ctx->C);
    if (!arena_alloc(s, w.floats, &w.off))
        return sfail(s, MPPI_HIP_E_NO_MEMORY,
                     "HBM budget: " + std::to_string(sizeof(float) * w.floats) + " bytes needed, " +
                         std::to_string(sizeof(float) * (s->arena_floats - s->used_floats)) + " free.");
    const uint32_t slot = s->free_slots.back();
    const mppi_agent_desc d = desc_of(w, slot, ctx->C);
    float* arena = static_cast<float*>(s->arena.ptr);
    const size_t HC = size_t(H) * ctx->C;
    hipError_t e = hipMemcpyAsync(static_cast<float*>(s->d_x.ptr) + size_t(slot) * ctx->D, x0,
                                  sizeof(float) * ctx->D, hipMemcpyHostToDevice, ctx->stream);
    if (e == hipSuccess)
        e = u_nominal ? hipMemcpyAsync(arena + d.seq_off, u_nominal, sizeof(float) * HC, hipMemcpyHostToDevice,
                                       ctx->stream)
                      : hipMemsetAsync(arena + d.seq_off, 0, sizeof(float) * HC, ctx->stream);
    if (e == hipSuccess && cost_bytes)
        e = hipMemcpyAsync(static_cast<uint8_t*>(s->d_cost.ptr) + size_t(slot) * s->cost_stride, cost_params,
                           cost_bytes, hipMemcpyHostToDevice, ctx->stream);
    if (e == hipSuccess)
        e = hipMemsetAsync(static_cast<mppi_agent_state*>(s->d_agents.ptr) + slot, 0, sizeof(mppi_agent_state),
                           ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess) {
        arena_free(s, w.off, w.floats);
        return sfail(s, MPPI_HIP_E_DEVICE, std::string("server admit: ") + hipGetErrorString(e));
    }
    s->free_slots.pop_back();
    s->ws[slot] = w;
    s->in_use[slot] = 1;
    s->active += 1;
    *agent_id = slot;
    return MPPI_HIP_OK;
}

int mppi_hip_server_release(mppi_hip_server* s, uint32_t agent_id) {
    int rc = check_server(s);
    if (rc) return rc;
    if ((rc = check_agent(s, agent_id))) return rc;
    arena_free(s, s->ws[agent_id].off, s->ws[agent_id].floats);
    s->in_use[agent_id] = 0;
    s->free_slots.push_back(agent_id);
    s->active -= 1;
    return MPPI_HIP_OK;
}

int mppi_hip_server_solve(mppi_hip_server* s, uint32_t n, const uint32_t* agent_ids, float* first_controls,
                          float* diag) {
    int rc = check_server(s);
    if (rc) return rc;
    if (n == 0 || !agent_ids) return sfail(s, MPPI_HIP_E_INVALID, "n must be > 0 and agent_ids non-null.");
    std::vector<uint8_t> seen(s->max_agents, 0);
    for (uint32_t i = 0; i < n; ++i) {
        if ((rc = check_agent(s, agent_ids[i]))) return rc;
        if (seen[agent_ids[i]]++) return sfail(s, MPPI_HIP_E_INVALID, "agent_ids must be distinct.");
    }
    return solve_impl(s, n, agent_ids, first_controls, diag);
}

int mppi_hip_server_read(mppi_hip_server* s, uint32_t agent_id, float* x_out, float* u_nominal_out) {
    int rc = check_server(s);
    if (rc) return rc;
    if ((rc = check_agent(s, agent_id))) return rc;
    mppi_hip_ctx* ctx = s->ctx;
    if (x_out)
        MPPI_HIP_TRY(ctx, hipMemcpyAsync(x_out, static_cast<float*>(s->d_x.ptr) + size_t(agent_id) * ctx->D,
                                         sizeof(float) * ctx->D, hipMemcpyDeviceToHost, ctx->stream));
    if (u_nominal_out) {
        const mppi_agent_desc d = desc_of(s->ws[agent_id], agent_id, ctx->C);
        MPPI_HIP_TRY(ctx, hipMemcpyAsync(u_nominal_out, static_cast<float*>(s->arena.ptr) + d.seq_off,
                                         sizeof(float) * s->ws[agent_id].H * ctx->C, hipMemcpyDeviceToHost,
                                         ctx->stream));
    }
    MPPI_HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    return MPPI_HIP_OK;
}

int mppi_hip_server_usage(const mppi_hip_server* s, size_t* used_bytes, size_t* budget_bytes,
                          uint32_t* agents) {
    if (!s) return fail(nullptr, MPPI_HIP_E_INVALID, "server is null.");
    if (used_bytes) *used_bytes = sizeof(float) * s->used_floats;
    if (budget_bytes) *budget_bytes = sizeof(float) * s->arena_floats;
    if (agents) *agents = s->active;
    return MPPI_HIP_OK;
}

int mppi_hip_server_replay(mppi_hip_server* s, uint32_t n, const mppi_hip_request* requests,
                           uint32_t max_batch, double* latency_s, mppi_hip_replay_stats* stats) {
    int rc = check_server(s);
    if (rc) return rc;
    mppi_hip_ctx* ctx = s->ctx;
    if (n == 0 || !requests || !latency_s) return sfail(s, MPPI_HIP_E_INVALID, "requests and latency_s are required.");
    for (uint32_t i = 0; i < n; ++i) {
        const mppi_hip_request& r = requests[i];
        if (!r.x0 || r.solves == 0 || (i && r.arrival_s < requests[i - 1].arrival_s))
            return sfail(s, MPPI_HIP_E_INVALID, "request " + std::to_string(i) +
                                                    ": x0 required, solves > 0, arrivals non-decreasing.");
        size_t bytes = 0;
        mppi_hip_server_workspace_bytes(s, r.K, r.H, &bytes);
        if (bytes > sizeof(float) * s->arena_floats)
            return sfail(s, MPPI_HIP_E_INVALID, "request " + std::to_string(i) + " can never fit the HBM budget.");
    }
    if (max_batch == 0) max_batch = s->max_agents;
    struct Active {
        uint32_t req, id, left;
    };
    std::vector<Active> active;
    std::vector<uint32_t> pending;   // FIFO of arrived, not admitted
    size_t head = 0;
    uint32_t next = 0, done = 0;
    mppi_hip_replay_stats st{};
    using clock = std::chrono::steady_clock;
    const auto t0 = clock::now();
    auto now_s = [&]() { return std::chrono::duration<double>(clock::now() - t0).count(); };
    std::vector<uint32_t> ids;
    while (done < n) {
        const double now = now_s();
        while (next < n && requests[next].arrival_s <= now) pending.push_back(next++);
        // Admission in arrival order; the head waits while it does not fit.
        while (head < pending.size() && active.size() < max_batch) {
            const mppi_hip_request& r = requests[pending[head]];
            uint32_t id = 0;
            rc = mppi_hip_server_admit(s, r.K, r.H, r.x0, r.u_nominal, r.cost_params, ctx->cost_bytes, &id);
            if (rc == MPPI_HIP_E_NO_MEMORY) {
                st.admission_waits += 1;
                break;
            }
            if (rc) return rc;
            active.push_back({pending[head], id, r.solves});
            ++head;
        }
        if (active.empty()) {
            if (next < n) {
                const double wait = requests[next].arrival_s - now_s();
                if (wait > 2e-4) std::this_thread::sleep_for(std::chrono::duration<double>(wait - 1e-4));
            }
            continue;
        }
        ids.clear();
        for (const Active& a : active) ids.push_back(a.id);
        if ((rc = solve_impl(s, uint32_t(ids.size()), ids.data(), nullptr, nullptr))) return rc;
        const double t = now_s();
        st.batches += 1;
        st.max_batch = std::max<uint32_t>(st.max_batch, uint32_t(ids.size()));
        size_t used = sizeof(float) * s->used_floats;
        st.max_workspace_bytes = std::max<double>(st.max_workspace_bytes, double(used));
        std::vector<Active> keep;
        for (Active& a : active) {
            st.sample_steps += double(s->ws[a.id].K) * s->ws[a.id].H;
            if (--a.left == 0) {
                latency_s[a.req] = t - requests[a.req].arrival_s;
                mppi_hip_server_release(s, a.id);
                ++done;
            } else {
                keep.push_back(a);
            }
        }
        active.swap(keep);
    }
    st.seconds = now_s();
    if (stats) *stats = st;
    return MPPI_HIP_OK;
}

}  // extern "C"
