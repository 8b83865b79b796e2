// hip_backend.cpp — implementation of the C ABI in include/mppi_hip.h.
//
// Replaces the reference's Objective-C++ MetalBackend (src/metal_backend.mm).
// What the reference does with metallibs, MTLLinkedFunctions and visible
// function tables, this does with one hipRTC translation unit per
// (plugin source, callable names, dimensions): the library kernels
// (mppi_kernels.hip) and the plugin are compiled together for the device's
// architecture, the code object is loaded with hipModuleLoadData and cached
// in-process and on disk by a hash of everything that went into it.
//
// Per MPPI step the backend issues one host->device copy (agent state, x0,
// nominal controls, parameter overrides, packed in pinned memory), the two
// kernels (mppi_rollout, mppi_reduce) and one device->host copy (u_next and
// the diagnostics), then synchronises the context's stream.
#include "mppi_hip.h"
#include "mppi_launch.h"

#include <hip/hip_runtime.h>
#include <hip/hiprtc.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <map>
#include <memory>
#include <mutex>
#include <sstream>
#include <string>
#include <sys/stat.h>
#include <vector>

namespace {

// Sources embedded at build time (generated by mppi_metal_amd/build.py).
#include "_jit_sources.inc"

thread_local std::string g_thread_error;

constexpr uint32_t kBlock = 256;
constexpr uint32_t kReduceBPG = 8;
constexpr uint32_t kMaxHC = 4096;

struct HipError {};

size_t align_up(size_t x, size_t a) { return (x + a - 1) / a * a; }

uint64_t fnv1a(const std::string& s, uint64_t h) {
    for (unsigned char c : s) {
        h ^= c;
        h *= 0x100000001b3ULL;
    }
    return h;
}

// ---------------------------------------------------------------------------
// Code-object cache
// ---------------------------------------------------------------------------
struct LoadedModule {
    hipModule_t module = nullptr;
    hipFunction_t rollout = nullptr;
    hipFunction_t reduce = nullptr;
};

std::mutex g_cache_mutex;
std::map<std::string, std::vector<char>> g_code_objects;               // key -> code object
std::map<std::pair<std::string, int>, std::shared_ptr<LoadedModule>> g_modules;  // (key, device)

std::string cache_dir() {
    const char* d = std::getenv("MPPI_HIP_CACHE_DIR");
    if (d && *d) return d;
    const char* home = std::getenv("HOME");
    if (home && *home) return std::string(home) + "/.cache/mppi_hip";
    return "";
}

bool read_file(const std::string& path, std::string* out) {
    std::ifstream f(path, std::ios::binary);
    if (!f) return false;
    std::ostringstream ss;
    ss << f.rdbuf();
    *out = ss.str();
    return true;
}

void mkdirs(const std::string& path) {
    std::string cur;
    for (size_t i = 0; i < path.size(); ++i) {
        cur += path[i];
        if (path[i] == '/' && cur.size() > 1) mkdir(cur.c_str(), 0755);
    }
    mkdir(path.c_str(), 0755);
}

// True when `name` occurs in `src` as a whole identifier.
bool has_identifier(const std::string& src, const std::string& name) {
    auto is_id = [](char c) { return (c >= 'a' && c <= 'z') || (c >= 'A' && c <= 'Z') ||
                                     (c >= '0' && c <= '9') || c == '_'; };
    size_t pos = 0;
    while ((pos = src.find(name, pos)) != std::string::npos) {
        bool left = pos == 0 || !is_id(src[pos - 1]);
        size_t end = pos + name.size();
        bool right = end >= src.size() || !is_id(src[end]);
        if (left && right) return true;
        pos = end;
    }
    return false;
}

bool valid_identifier(const std::string& s) {
    if (s.empty()) return false;
    for (size_t i = 0; i < s.size(); ++i) {
        char c = s[i];
        bool ok = (c >= 'a' && c <= 'z') || (c >= 'A' && c <= 'Z') || c == '_' ||
                  (i > 0 && c >= '0' && c <= '9');
        if (!ok) return false;
    }
    return true;
}

// Lower Cholesky factor of a symmetric positive definite row-major matrix.
bool cholesky(const std::vector<float>& a, uint32_t n, std::vector<float>* l) {
    std::vector<double> L(size_t(n) * n, 0.0);
    for (uint32_t i = 0; i < n; ++i) {
        for (uint32_t j = 0; j <= i; ++j) {
            double s = double(a[size_t(i) * n + j]);
            for (uint32_t k = 0; k < j; ++k) s -= L[size_t(i) * n + k] * L[size_t(j) * n + k];
            if (i == j) {
                if (!(s > 0.0)) return false;
                L[size_t(i) * n + i] = std::sqrt(s);
            } else {
                L[size_t(i) * n + j] = s / L[size_t(j) * n + j];
            }
        }
    }
    l->assign(size_t(n) * n, 0.0f);
    for (size_t i = 0; i < L.size(); ++i) (*l)[i] = float(L[i]);
    return true;
}

struct DeviceBuffer {
    void* ptr = nullptr;
    size_t bytes = 0;
    void release() {
        if (ptr) (void)hipFree(ptr);
        ptr = nullptr;
        bytes = 0;
    }
};

struct PinnedBuffer {
    void* ptr = nullptr;
    size_t bytes = 0;
    void release() {
        if (ptr) (void)hipHostFree(ptr);
        ptr = nullptr;
        bytes = 0;
    }
};

}  // namespace

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

    // plugin
    bool linked = false;
    std::shared_ptr<LoadedModule> mod;
    size_t model_bytes = 0, cost_bytes = 0;
    uint32_t key0 = 0, key1 = 0;

    // runtime
    hipStream_t stream = nullptr;
    std::string error;
    uint32_t spt_override = 0;

    // device buffers
    DeviceBuffer consts;          // noise_scale | u_min | u_max | zero blob
    DeviceBuffer model_bound, cost_bound;
    bool model_is_bound = false, cost_is_bound = false;
    DeviceBuffer up;              // agents | x0 | u_nom | model override | cost override
    DeviceBuffer down;            // u_out | diag
    DeviceBuffer costs, blk_stats, weights, partials;
    DeviceBuffer hist_states, hist_controls, hist_costs, terminal;
    PinnedBuffer up_host, down_host;
    uint32_t last_agents = 0;

    // profiling
    bool profiling = false;
    std::vector<hipEvent_t> ev_pool;
    std::vector<std::pair<hipEvent_t, hipEvent_t>> pending_rollout, pending_reduce;
    double prof_rollout_ms = 0.0, prof_reduce_ms = 0.0;
    uint64_t prof_launches = 0;
};

namespace {

int fail(mppi_hip_ctx* ctx, int code, const std::string& msg) {
    if (ctx) ctx->error = msg;
    else g_thread_error = msg;
    return code;
}

#define HIP_TRY(ctx, expr)                                                              \
    do {                                                                                \
        hipError_t e_ = (expr);                                                         \
        if (e_ != hipSuccess)                                                           \
            return fail(ctx, MPPI_HIP_E_DEVICE,                                         \
                        std::string(#expr) + " failed: " + hipGetErrorString(e_));      \
    } while (0)

int ensure_device_buffer(mppi_hip_ctx* ctx, DeviceBuffer* b, size_t bytes) {
    if (bytes == 0) bytes = 16;
    if (b->bytes >= bytes) return MPPI_HIP_OK;
    b->release();
    hipError_t e = hipMalloc(&b->ptr, bytes);
    if (e != hipSuccess) {
        b->ptr = nullptr;
        return fail(ctx, MPPI_HIP_E_NO_MEMORY,
                    "hipMalloc(" + std::to_string(bytes) + ") failed: " + hipGetErrorString(e));
    }
    b->bytes = bytes;
    return MPPI_HIP_OK;
}

int ensure_pinned_buffer(mppi_hip_ctx* ctx, PinnedBuffer* b, size_t bytes) {
    if (bytes == 0) bytes = 16;
    if (b->bytes >= bytes) return MPPI_HIP_OK;
    b->release();
    hipError_t e = hipHostMalloc(&b->ptr, bytes, hipHostMallocDefault);
    if (e != hipSuccess) {
        b->ptr = nullptr;
        return fail(ctx, MPPI_HIP_E_NO_MEMORY,
                    "hipHostMalloc(" + std::to_string(bytes) + ") failed: " + hipGetErrorString(e));
    }
    b->bytes = bytes;
    return MPPI_HIP_OK;
}

// Layout of the per-call upload block (identical on host and device).
struct UploadPlan {
    size_t agents = 0, x0 = 0, u_nom = 0, model = 0, cost = 0, total = 0;
};

UploadPlan plan_upload(const mppi_hip_ctx* ctx, uint32_t N, size_t model_bytes, size_t cost_bytes) {
    UploadPlan p;
    size_t off = 0;
    p.agents = off;
    off = align_up(off + sizeof(mppi_agent_state) * N, 256);
    p.x0 = off;
    off = align_up(off + sizeof(float) * size_t(N) * ctx->D, 256);
    p.u_nom = off;
    off = align_up(off + sizeof(float) * size_t(N) * ctx->H * ctx->C, 256);
    p.model = off;
    off = align_up(off + model_bytes, 256);
    p.cost = off;
    off = align_up(off + cost_bytes, 256);
    p.total = off;
    return p;
}

struct ReduceShape {
    uint32_t NB, n_groups, spt, n_chunks;
};

ReduceShape reduce_shape(const mppi_hip_ctx* ctx, uint32_t N) {
    ReduceShape r;
    r.NB = (ctx->H * ctx->C + 3) / 4;
    r.n_groups = (r.NB + kReduceBPG - 1) / kReduceBPG;
    if (ctx->spt_override) {
        r.spt = ctx->spt_override;
    } else {
        // About 2048 workgroups over the whole launch, at most 64 samples per thread.
        uint64_t work = uint64_t(ctx->K) * r.n_groups * N;
        uint64_t spt = (work + uint64_t(kBlock) * 2048 - 1) / (uint64_t(kBlock) * 2048);
        r.spt = uint32_t(std::max<uint64_t>(1, std::min<uint64_t>(64, spt)));
    }
    r.n_chunks = (ctx->K + kBlock * r.spt - 1) / (kBlock * r.spt);
    return r;
}

int ensure_work_buffers(mppi_hip_ctx* ctx, uint32_t N) {
    const uint32_t n_blk = (ctx->K + kBlock - 1) / kBlock;
    ReduceShape rs = reduce_shape(ctx, N);
    int rc;
    if ((rc = ensure_device_buffer(ctx, &ctx->costs, sizeof(float) * size_t(N) * ctx->K))) return rc;
    if ((rc = ensure_device_buffer(ctx, &ctx->weights, sizeof(float) * size_t(N) * ctx->K))) return rc;
    if ((rc = ensure_device_buffer(ctx, &ctx->blk_stats, sizeof(float) * 4 * size_t(N) * n_blk))) return rc;
    if ((rc = ensure_device_buffer(ctx, &ctx->partials,
                                   sizeof(float) * 4 * size_t(N) * rs.n_chunks * rs.NB))) return rc;
    const size_t down = sizeof(float) * (size_t(N) * ctx->H * ctx->C + 3 * size_t(N));
    if ((rc = ensure_device_buffer(ctx, &ctx->down, down))) return rc;
    if ((rc = ensure_pinned_buffer(ctx, &ctx->down_host, down))) return rc;
    return MPPI_HIP_OK;
}

hipEvent_t take_event(mppi_hip_ctx* ctx) {
    if (!ctx->ev_pool.empty()) {
        hipEvent_t e = ctx->ev_pool.back();
        ctx->ev_pool.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    return e;
}

// Fold finished event pairs into the profile sums (after a stream sync).
void collect_profile(mppi_hip_ctx* ctx) {
    auto fold = [&](std::vector<std::pair<hipEvent_t, hipEvent_t>>& v, double* acc) {
        for (auto& pr : v) {
            float ms = 0.0f;
            if (hipEventElapsedTime(&ms, pr.first, pr.second) == hipSuccess) *acc += ms;
            ctx->ev_pool.push_back(pr.first);
            ctx->ev_pool.push_back(pr.second);
        }
        v.clear();
    };
    fold(ctx->pending_rollout, &ctx->prof_rollout_ms);
    fold(ctx->pending_reduce, &ctx->prof_reduce_ms);
}

struct StepBuffers {
    mppi_agent_state* agents;
    float* x;
    float* u_nom;
    const uint8_t* model;
    const uint8_t* cost;
    float* u_out;
    float* diag;
    float* hist_states = nullptr;
    float* hist_controls = nullptr;
    float* hist_costs = nullptr;
    float* terminal = nullptr;
};

// Enqueue one MPPI iteration (rollout + reduce) for N agents.
int launch_iteration(mppi_hip_ctx* ctx, uint32_t N, const StepBuffers& b, bool closed_loop) {
    const uint32_t n_blk = (ctx->K + kBlock - 1) / kBlock;
    ReduceShape rs = reduce_shape(ctx, N);
    const float* cbase = static_cast<const float*>(ctx->consts.ptr);
    const size_t ns = ctx->noise_scale.size();

    mppi_rollout_args ra{};
    ra.x0 = b.x;
    ra.u_nom = b.u_nom;
    ra.model = b.model;
    ra.cost = b.cost;
    ra.noise_scale = cbase;
    ra.u_min = cbase + ns;
    ra.u_max = cbase + ns + ctx->C;
    ra.costs = static_cast<float*>(ctx->costs.ptr);
    ra.blk_stats = static_cast<float*>(ctx->blk_stats.ptr);
    ra.agents = b.agents;
    ra.K = ctx->K;
    ra.H = ctx->H;
    ra.N = N;
    ra.n_blk = n_blk;
    ra.key0 = ctx->key0;
    ra.key1 = ctx->key1;
    ra.lambda = ctx->lambda;

    mppi_reduce_args rd{};
    rd.costs = ra.costs;
    rd.agents = b.agents;
    rd.u_nom = b.u_nom;
    rd.u_out = b.u_out;
    rd.noise_scale = ra.noise_scale;
    rd.u_min = ra.u_min;
    rd.u_max = ra.u_max;
    rd.weights = static_cast<float*>(ctx->weights.ptr);
    rd.partials = static_cast<float*>(ctx->partials.ptr);
    rd.diag = b.diag;
    rd.x_state = b.x;
    rd.hist_states = b.hist_states;
    rd.hist_controls = b.hist_controls;
    rd.hist_costs = b.hist_costs;
    rd.terminal = b.terminal;
    rd.model = b.model;
    rd.K = ctx->K;
    rd.H = ctx->H;
    rd.N = N;
    rd.NB = rs.NB;
    rd.n_chunks = rs.n_chunks;
    rd.n_groups = rs.n_groups;
    rd.spt = rs.spt;
    rd.closed_loop = closed_loop ? 1u : 0u;
    rd.key0 = ctx->key0;
    rd.key1 = ctx->key1;
    rd.lambda = ctx->lambda;

    void* rparams[] = {&ra};
    void* dparams[] = {&rd};
    hipEvent_t e0 = nullptr, e1 = nullptr, e2 = nullptr, e3 = nullptr;
    if (ctx->profiling) {
        e0 = take_event(ctx); e1 = take_event(ctx); e2 = take_event(ctx); e3 = take_event(ctx);
        if (e0) (void)hipEventRecord(e0, ctx->stream);
    }
    HIP_TRY(ctx, hipModuleLaunchKernel(ctx->mod->rollout, n_blk, N, 1, kBlock, 1, 1, 0,
                                       ctx->stream, rparams, nullptr));
    if (ctx->profiling) {
        if (e1) (void)hipEventRecord(e1, ctx->stream);
        if (e2) (void)hipEventRecord(e2, ctx->stream);
    }
    HIP_TRY(ctx, hipModuleLaunchKernel(ctx->mod->reduce, rs.n_chunks, rs.n_groups, N, kBlock, 1, 1,
                                       0, ctx->stream, dparams, nullptr));
    if (ctx->profiling) {
        if (e3) (void)hipEventRecord(e3, ctx->stream);
        if (e0 && e1) ctx->pending_rollout.push_back({e0, e1});
        if (e2 && e3) ctx->pending_reduce.push_back({e2, e3});
        ctx->prof_launches += 1;
    }
    return MPPI_HIP_OK;
}

int check_ctx(mppi_hip_ctx* ctx) {
    if (!ctx) return fail(nullptr, MPPI_HIP_E_INVALID, "context is null.");
    if (!ctx->linked)
        return fail(ctx, MPPI_HIP_E_NOT_INITIALIZED, "Backend is not initialized: no plugin linked.");
    HIP_TRY(ctx, hipSetDevice(ctx->device));
    return MPPI_HIP_OK;
}

// Resolve the parameter blob for a call: override (copied into the upload
// block), else the bound blob, else the zero blob.
const uint8_t* resolve_blob(mppi_hip_ctx* ctx, const void* ovr, size_t ovr_bytes,
                            const DeviceBuffer& bound, bool is_bound, size_t up_off) {
    if (ovr && ovr_bytes > 0) return static_cast<const uint8_t*>(ctx->up.ptr) + up_off;
    if (is_bound) return static_cast<const uint8_t*>(bound.ptr);
    const float* cbase = static_cast<const float*>(ctx->consts.ptr);
    return reinterpret_cast<const uint8_t*>(cbase + ctx->noise_scale.size() + 2 * ctx->C);
}

int check_override(mppi_hip_ctx* ctx, const void* p, size_t bytes, size_t layout, const char* what) {
    if (bytes == 0) return MPPI_HIP_OK;
    if (!p) return fail(ctx, MPPI_HIP_E_INVALID, std::string(what) + " override data is null but bytes > 0.");
    if (bytes != layout)
        return fail(ctx, MPPI_HIP_E_INVALID,
                    std::string(what) + " override bytes (" + std::to_string(bytes) +
                        ") does not match the linked layout (" + std::to_string(layout) + ").");
    return MPPI_HIP_OK;
}

// Fill the pinned upload block for N agents and copy it to the device.
int upload(mppi_hip_ctx* ctx, uint32_t N, uint64_t step, const float* x0, const float* u_nom,
           const void* mo, size_t mo_bytes, const void* co, size_t co_bytes, UploadPlan* plan) {
    *plan = plan_upload(ctx, N, mo_bytes, co_bytes);
    int rc;
    if ((rc = ensure_device_buffer(ctx, &ctx->up, plan->total))) return rc;
    if ((rc = ensure_pinned_buffer(ctx, &ctx->up_host, plan->total))) return rc;
    uint8_t* h = static_cast<uint8_t*>(ctx->up_host.ptr);
    auto* ag = reinterpret_cast<mppi_agent_state*>(h + plan->agents);
    for (uint32_t a = 0; a < N; ++a) {
        ag[a] = mppi_agent_state{};
        ag[a].step = uint32_t(step);
    }
    std::memcpy(h + plan->x0, x0, sizeof(float) * size_t(N) * ctx->D);
    const size_t hc = size_t(ctx->H) * ctx->C;
    if (u_nom) std::memcpy(h + plan->u_nom, u_nom, sizeof(float) * size_t(N) * hc);
    else std::memset(h + plan->u_nom, 0, sizeof(float) * size_t(N) * hc);
    if (mo && mo_bytes) std::memcpy(h + plan->model, mo, mo_bytes);
    if (co && co_bytes) std::memcpy(h + plan->cost, co, co_bytes);
    HIP_TRY(ctx, hipMemcpyAsync(ctx->up.ptr, h, plan->total, hipMemcpyHostToDevice, ctx->stream));
    return MPPI_HIP_OK;
}

int finish(mppi_hip_ctx* ctx) {
    HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    if (ctx->profiling) collect_profile(ctx);
    return MPPI_HIP_OK;
}

int link_common(mppi_hip_ctx* ctx, const std::string& src, const std::string& src_name,
                const std::string& plugin_dir, const char* dynamics, const char* stage,
                const char* terminal, size_t model_bytes, size_t cost_bytes, uint64_t seed) {
    std::string dyn = dynamics ? dynamics : "";
    std::string sc = stage ? stage : "";
    std::string tc = terminal ? terminal : "";
    if (dyn.empty())
        return fail(ctx, MPPI_HIP_E_PLUGIN, "Required callable name for dynamics is empty.");
    for (const std::string* n : {&dyn, &sc, &tc}) {
        if (!n->empty() && !valid_identifier(*n))
            return fail(ctx, MPPI_HIP_E_PLUGIN, "Callable name '" + *n + "' is not a C identifier.");
    }
    if (!has_identifier(src, dyn))
        return fail(ctx, MPPI_HIP_E_PLUGIN,
                    "Required callable '" + dyn + "' not found in plugin source '" + src_name + "'.");
    if (!sc.empty() && !has_identifier(src, sc))
        return fail(ctx, MPPI_HIP_E_PLUGIN,
                    "Callable '" + sc + "' not found in plugin source '" + src_name + "'.");
    if (!tc.empty() && !has_identifier(src, tc))
        return fail(ctx, MPPI_HIP_E_PLUGIN,
                    "Callable '" + tc + "' not found in plugin source '" + src_name + "'.");

    std::vector<std::string> opts = {
        "--offload-arch=" + ctx->arch,
        "-O3",
        "-std=c++17",
        "-ffp-contract=off",
        "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
        "-DMPPI_SDIM=" + std::to_string(ctx->D),
        "-DMPPI_CDIM=" + std::to_string(ctx->C),
        "-DMPPI_DYNAMICS=" + dyn,
        "-DMPPI_COV_FULL=" + std::to_string(ctx->cov_form == MPPI_HIP_COV_FULL ? 1 : 0),
    };
    if (!sc.empty()) opts.push_back("-DMPPI_STAGE_COST=" + sc);
    if (!tc.empty()) opts.push_back("-DMPPI_TERMINAL_COST=" + tc);
    if (!plugin_dir.empty()) opts.push_back("-I" + plugin_dir);

    std::string keysrc = std::string(kKernelSource) + "\x01" + kLaunchHeader + "\x01" + kCoreHeader +
                         "\x01" + kMathHeader + "\x01" + kPhiloxHeader + "\x01" + src + "\x01" +
                         mppi_hip_version();
    for (auto& o : opts) keysrc += "\x02" + o;
    char keybuf[40];
    std::snprintf(keybuf, sizeof(keybuf), "%016llx%016llx",
                  (unsigned long long)fnv1a(keysrc, 0xcbf29ce484222325ULL),
                  (unsigned long long)fnv1a(keysrc, 0x84222325cbf29ce4ULL));
    const std::string key = keybuf;

    std::shared_ptr<LoadedModule> mod;
    {
        std::lock_guard<std::mutex> lock(g_cache_mutex);
        auto it = g_modules.find({key, ctx->device});
        if (it != g_modules.end()) mod = it->second;
    }
    if (!mod) {
        std::vector<char> code;
        {
            std::lock_guard<std::mutex> lock(g_cache_mutex);
            auto it = g_code_objects.find(key);
            if (it != g_code_objects.end()) code = it->second;
        }
        const std::string dir = cache_dir();
        const std::string cpath = dir.empty() ? "" : dir + "/" + key + ".co";
        if (code.empty() && !cpath.empty()) {
            std::string bytes;
            if (read_file(cpath, &bytes)) code.assign(bytes.begin(), bytes.end());
        }
        if (code.empty()) {
            const char* hdr_src[] = {kCoreHeader, kMathHeader, kPhiloxHeader, kLaunchHeader,
                                     src.c_str(), kCoreHeader};
            const char* hdr_names[] = {"mppi_core.h", "mppi_math.h", "mppi_philox.h",
                                       "mppi_launch.h", "mppi_user_plugin.hip",
                                       "mppi_hip/mppi_core.h"};
            hiprtcProgram prog;
            if (hiprtcCreateProgram(&prog, kKernelSource, "mppi_kernels.hip", 6, hdr_src,
                                    hdr_names) != HIPRTC_SUCCESS)
                return fail(ctx, MPPI_HIP_E_COMPILE, "hiprtcCreateProgram failed.");
            std::vector<const char*> copts;
            for (auto& o : opts) copts.push_back(o.c_str());
            hiprtcResult r = hiprtcCompileProgram(prog, int(copts.size()), copts.data());
            size_t log_size = 0;
            hiprtcGetProgramLogSize(prog, &log_size);
            std::string log(log_size, '\0');
            if (log_size) hiprtcGetProgramLog(prog, &log[0]);
            if (r != HIPRTC_SUCCESS) {
                hiprtcDestroyProgram(&prog);
                return fail(ctx, MPPI_HIP_E_COMPILE,
                            "Plugin '" + src_name + "' failed to compile: " + log);
            }
            size_t code_size = 0;
            hiprtcGetCodeSize(prog, &code_size);
            code.resize(code_size);
            hiprtcGetCode(prog, code.data());
            hiprtcDestroyProgram(&prog);
            if (!cpath.empty()) {
                mkdirs(dir);
                std::string tmp = cpath + ".tmp" + std::to_string(uint64_t(std::rand()));
                std::ofstream f(tmp, std::ios::binary);
                if (f) {
                    f.write(code.data(), std::streamsize(code.size()));
                    f.close();
                    std::rename(tmp.c_str(), cpath.c_str());
                }
            }
        }
        {
            std::lock_guard<std::mutex> lock(g_cache_mutex);
            g_code_objects[key] = code;
        }
        auto m = std::make_shared<LoadedModule>();
        hipError_t e = hipModuleLoadData(&m->module, code.data());
        if (e != hipSuccess)
            return fail(ctx, MPPI_HIP_E_COMPILE,
                        std::string("hipModuleLoadData failed: ") + hipGetErrorString(e));
        if (hipModuleGetFunction(&m->rollout, m->module, "mppi_rollout") != hipSuccess ||
            hipModuleGetFunction(&m->reduce, m->module, "mppi_reduce") != hipSuccess)
            return fail(ctx, MPPI_HIP_E_COMPILE, "Required library kernels missing from code object.");
        std::lock_guard<std::mutex> lock(g_cache_mutex);
        auto it = g_modules.find({key, ctx->device});
        if (it != g_modules.end()) {
            (void)hipModuleUnload(m->module);
            mod = it->second;
        } else {
            g_modules[{key, ctx->device}] = m;
            mod = m;
        }
    }

    ctx->mod = mod;
    ctx->model_bytes = model_bytes;
    ctx->cost_bytes = cost_bytes;
    ctx->key0 = uint32_t(seed);
    ctx->key1 = uint32_t(seed >> 32);
    ctx->model_is_bound = false;
    ctx->cost_is_bound = false;
    ctx->linked = true;
    return MPPI_HIP_OK;
}

}  // namespace

extern "C" {

const char* mppi_hip_version(void) { return "mppi_hip 0.1.0"; }

const char* mppi_hip_last_error(const mppi_hip_ctx* ctx) {
    return ctx ? ctx->error.c_str() : g_thread_error.c_str();
}

int mppi_hip_create(const mppi_hip_config* cfg, mppi_hip_ctx** out) {
    if (!out) return fail(nullptr, MPPI_HIP_E_INVALID, "out_ctx is null.");
    *out = nullptr;
    if (!cfg) return fail(nullptr, MPPI_HIP_E_INVALID, "config is null.");
    if (cfg->state_dim == 0) return fail(nullptr, MPPI_HIP_E_INVALID, "state_dim must be > 0.");
    if (cfg->control_dim == 0) return fail(nullptr, MPPI_HIP_E_INVALID, "control_dim must be > 0.");
    if (cfg->horizon == 0) return fail(nullptr, MPPI_HIP_E_INVALID, "horizon must be > 0.");
    if (cfg->sample_count == 0) return fail(nullptr, MPPI_HIP_E_INVALID, "sample_count must be > 0.");
    if (cfg->state_dim > 64) return fail(nullptr, MPPI_HIP_E_INVALID, "state_dim must be <= 64.");
    if (cfg->control_dim > 16) return fail(nullptr, MPPI_HIP_E_INVALID, "control_dim must be <= 16.");
    if (uint64_t(cfg->horizon) * cfg->control_dim > kMaxHC)
        return fail(nullptr, MPPI_HIP_E_INVALID, "horizon * control_dim must be <= 4096.");
    if (!(cfg->lambda > 0.0f)) return fail(nullptr, MPPI_HIP_E_INVALID, "lambda must be > 0.");
    if (!cfg->sigma) return fail(nullptr, MPPI_HIP_E_INVALID, "sigma must not be null.");
    if (!cfg->u_min) return fail(nullptr, MPPI_HIP_E_INVALID, "u_min must not be null.");
    if (!cfg->u_max) return fail(nullptr, MPPI_HIP_E_INVALID, "u_max must not be null.");
    if (cfg->covariance_form != MPPI_HIP_COV_DIAGONAL && cfg->covariance_form != MPPI_HIP_COV_FULL)
        return fail(nullptr, MPPI_HIP_E_INVALID, "covariance_form is not DIAGONAL or FULL.");
    if (cfg->numeric_mode != MPPI_HIP_FP32)
        return fail(nullptr, MPPI_HIP_E_INVALID,
                    "numeric_mode: only FP32 is implemented by this backend.");
    for (uint32_t i = 0; i < cfg->control_dim; ++i)
        if (cfg->u_min[i] > cfg->u_max[i])
            return fail(nullptr, MPPI_HIP_E_INVALID, "u_min[" + std::to_string(i) + "] > u_max.");

    std::unique_ptr<mppi_hip_ctx> ctx(new mppi_hip_ctx());
    ctx->D = cfg->state_dim;
    ctx->C = cfg->control_dim;
    ctx->H = cfg->horizon;
    ctx->K = cfg->sample_count;
    ctx->lambda = cfg->lambda;
    ctx->cov_form = cfg->covariance_form;
    ctx->numeric_mode = cfg->numeric_mode;
    const uint32_t C = cfg->control_dim;
    if (cfg->covariance_form == MPPI_HIP_COV_FULL) {
        std::vector<float> cov(cfg->sigma, cfg->sigma + size_t(C) * C);
        for (uint32_t i = 0; i < C; ++i)
            for (uint32_t j = 0; j < i; ++j)
                if (std::fabs(cov[i * C + j] - cov[j * C + i]) >
                    1e-6f * std::max(1.0f, std::fabs(cov[i * C + j])))
                    return fail(nullptr, MPPI_HIP_E_INVALID, "sigma (full covariance) is not symmetric.");
        if (!cholesky(cov, C, &ctx->noise_scale))
            return fail(nullptr, MPPI_HIP_E_INVALID,
                        "sigma (full covariance) is not positive definite.");
    } else {
        ctx->noise_scale.assign(cfg->sigma, cfg->sigma + C);
    }
    ctx->u_min.assign(cfg->u_min, cfg->u_min + C);
    ctx->u_max.assign(cfg->u_max, cfg->u_max + C);

    int count = 0;
    hipError_t e = hipGetDeviceCount(&count);
    if (e != hipSuccess || count == 0)
        return fail(nullptr, MPPI_HIP_E_DEVICE,
                    std::string("No HIP device available: ") + hipGetErrorString(e));
    int dev = cfg->device;
    if (dev < 0) {
        if (hipGetDevice(&dev) != hipSuccess) dev = 0;
    }
    if (dev >= count) return fail(nullptr, MPPI_HIP_E_DEVICE, "device ordinal out of range.");
    ctx->device = dev;
    if ((e = hipSetDevice(dev)) != hipSuccess)
        return fail(nullptr, MPPI_HIP_E_DEVICE, std::string("hipSetDevice failed: ") + hipGetErrorString(e));
    hipDeviceProp_t prop;
    if ((e = hipGetDeviceProperties(&prop, dev)) != hipSuccess)
        return fail(nullptr, MPPI_HIP_E_DEVICE,
                    std::string("hipGetDeviceProperties failed: ") + hipGetErrorString(e));
    std::string arch = prop.gcnArchName;
    ctx->arch = arch.substr(0, arch.find(':'));
    if ((e = hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking)) != hipSuccess)
        return fail(nullptr, MPPI_HIP_E_DEVICE,
                    std::string("hipStreamCreate failed: ") + hipGetErrorString(e));

    // Constants: noise scale | u_min | u_max | 256-byte zero blob.
    const size_t ns = ctx->noise_scale.size();
    std::vector<float> host(ns + 2 * C + 64, 0.0f);
    std::copy(ctx->noise_scale.begin(), ctx->noise_scale.end(), host.begin());
    std::copy(ctx->u_min.begin(), ctx->u_min.end(), host.begin() + ns);
    std::copy(ctx->u_max.begin(), ctx->u_max.end(), host.begin() + ns + C);
    mppi_hip_ctx* raw = ctx.get();
    if (ensure_device_buffer(raw, &ctx->consts, host.size() * sizeof(float)))
        return fail(nullptr, MPPI_HIP_E_NO_MEMORY, raw->error);
    if ((e = hipMemcpy(ctx->consts.ptr, host.data(), host.size() * sizeof(float),
                       hipMemcpyHostToDevice)) != hipSuccess)
        return fail(nullptr, MPPI_HIP_E_DEVICE, std::string("hipMemcpy failed: ") + hipGetErrorString(e));
    *out = ctx.release();
    return MPPI_HIP_OK;
}

void mppi_hip_destroy(mppi_hip_ctx* ctx) {
    if (!ctx) return;
    (void)hipSetDevice(ctx->device);
    if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
    for (DeviceBuffer* b : {&ctx->consts, &ctx->model_bound, &ctx->cost_bound, &ctx->up, &ctx->down,
                            &ctx->costs, &ctx->blk_stats, &ctx->weights, &ctx->partials,
                            &ctx->hist_states, &ctx->hist_controls, &ctx->hist_costs,
                            &ctx->terminal})
        b->release();
    ctx->up_host.release();
    ctx->down_host.release();
    collect_profile(ctx);
    for (hipEvent_t e : ctx->ev_pool) (void)hipEventDestroy(e);
    if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
    delete ctx;
}

int mppi_hip_link_plugin(mppi_hip_ctx* ctx, const char* plugin_path, const char* dynamics,
                         const char* stage_cost, const char* terminal_cost, size_t model_bytes,
                         size_t cost_bytes, uint64_t seed) {
    if (!ctx) return fail(nullptr, MPPI_HIP_E_INVALID, "context is null.");
    std::string path = plugin_path ? plugin_path : "";
    if (path.empty()) return fail(ctx, MPPI_HIP_E_PLUGIN, "user_metallib_path is empty.");
    std::string src;
    if (!read_file(path, &src))
        return fail(ctx, MPPI_HIP_E_PLUGIN, "Failed to load plugin source '" + path + "'.");
    std::string dir = path.find('/') == std::string::npos ? "." : path.substr(0, path.rfind('/'));
    HIP_TRY(ctx, hipSetDevice(ctx->device));
    return link_common(ctx, src, path, dir, dynamics, stage_cost, terminal_cost, model_bytes,
                       cost_bytes, seed);
}

int mppi_hip_link_plugin_source(mppi_hip_ctx* ctx, const char* source, const char* source_name,
                                const char* dynamics, const char* stage_cost,
                                const char* terminal_cost, size_t model_bytes, size_t cost_bytes,
                                uint64_t seed) {
    if (!ctx) return fail(nullptr, MPPI_HIP_E_INVALID, "context is null.");
    if (!source || !*source) return fail(ctx, MPPI_HIP_E_PLUGIN, "plugin source is empty.");
    HIP_TRY(ctx, hipSetDevice(ctx->device));
    return link_common(ctx, source, source_name ? source_name : "<source>", "", dynamics,
                       stage_cost, terminal_cost, model_bytes, cost_bytes, seed);
}

int mppi_hip_bind_params(mppi_hip_ctx* ctx, const void* model, size_t model_bytes,
                         const void* cost, size_t cost_bytes) {
    int rc = check_ctx(ctx);
    if (rc) return rc;
    if (model_bytes != 0 && model_bytes != ctx->model_bytes)
        return fail(ctx, MPPI_HIP_E_INVALID,
                    "model_params.bytes (" + std::to_string(model_bytes) +
                        ") does not match ParamLayout::model_params_size (" +
                        std::to_string(ctx->model_bytes) + ").");
    if (cost_bytes != 0 && cost_bytes != ctx->cost_bytes)
        return fail(ctx, MPPI_HIP_E_INVALID,
                    "cost_params.bytes (" + std::to_string(cost_bytes) +
                        ") does not match ParamLayout::cost_params_size (" +
                        std::to_string(ctx->cost_bytes) + ").");
    if (model_bytes && !model) return fail(ctx, MPPI_HIP_E_INVALID, "model_params.data is null but bytes > 0.");
    if (cost_bytes && !cost) return fail(ctx, MPPI_HIP_E_INVALID, "cost_params.data is null but bytes > 0.");
    if (model_bytes) {
        if ((rc = ensure_device_buffer(ctx, &ctx->model_bound, model_bytes))) return rc;
        HIP_TRY(ctx, hipMemcpyAsync(ctx->model_bound.ptr, model, model_bytes, hipMemcpyHostToDevice,
                                    ctx->stream));
    }
    if (cost_bytes) {
        if ((rc = ensure_device_buffer(ctx, &ctx->cost_bound, cost_bytes))) return rc;
        HIP_TRY(ctx, hipMemcpyAsync(ctx->cost_bound.ptr, cost, cost_bytes, hipMemcpyHostToDevice,
                                    ctx->stream));
    }
    HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    ctx->model_is_bound = model_bytes != 0;
    ctx->cost_is_bound = cost_bytes != 0;
    return MPPI_HIP_OK;
}

int mppi_hip_step(mppi_hip_ctx* ctx, uint64_t step_index, const float* x0, const float* u_nominal,
                  const void* mo, size_t mo_bytes, const void* co, size_t co_bytes, float* u_out,
                  float* diag3) {
    int rc = check_ctx(ctx);
    if (rc) return rc;
    if (!x0) return fail(ctx, MPPI_HIP_E_INVALID, "x0 must not be null.");
    if (!u_nominal) return fail(ctx, MPPI_HIP_E_INVALID, "u_nominal must not be null.");
    if (!u_out) return fail(ctx, MPPI_HIP_E_INVALID, "u_out must not be null.");
    if ((rc = check_override(ctx, mo, mo_bytes, ctx->model_bytes, "model_params"))) return rc;
    if ((rc = check_override(ctx, co, co_bytes, ctx->cost_bytes, "cost_params"))) return rc;
    if ((rc = ensure_work_buffers(ctx, 1))) return rc;
    UploadPlan plan;
    if ((rc = upload(ctx, 1, step_index, x0, u_nominal, mo, mo_bytes, co, co_bytes, &plan))) return rc;
    uint8_t* up = static_cast<uint8_t*>(ctx->up.ptr);
    StepBuffers b;
    b.agents = reinterpret_cast<mppi_agent_state*>(up + plan.agents);
    b.x = reinterpret_cast<float*>(up + plan.x0);
    b.u_nom = reinterpret_cast<float*>(up + plan.u_nom);
    b.model = resolve_blob(ctx, mo, mo_bytes, ctx->model_bound, ctx->model_is_bound, plan.model);
    b.cost = resolve_blob(ctx, co, co_bytes, ctx->cost_bound, ctx->cost_is_bound, plan.cost);
    const size_t hc = size_t(ctx->H) * ctx->C;
    b.u_out = static_cast<float*>(ctx->down.ptr);
    b.diag = b.u_out + hc;
    if ((rc = launch_iteration(ctx, 1, b, false))) return rc;
    HIP_TRY(ctx, hipMemcpyAsync(ctx->down_host.ptr, ctx->down.ptr, sizeof(float) * (hc + 3),
                                hipMemcpyDeviceToHost, ctx->stream));
    if ((rc = finish(ctx))) return rc;
    const float* h = static_cast<const float*>(ctx->down_host.ptr);
    std::memcpy(u_out, h, sizeof(float) * hc);
    if (diag3) std::memcpy(diag3, h + hc, sizeof(float) * 3);
    ctx->last_agents = 1;
    return MPPI_HIP_OK;
}

namespace {

int run_closed_loop(mppi_hip_ctx* ctx, uint32_t N, uint32_t T, uint64_t start_step,
                    const float* x0, const float* u_nom, const void* mo, size_t mo_bytes,
                    const void* co, size_t co_bytes, float* states_out, float* controls_out,
                    float* costs_out, float* terminal_out) {
    int rc;
    if ((rc = check_override(ctx, mo, mo_bytes, ctx->model_bytes, "model_params"))) return rc;
    if ((rc = check_override(ctx, co, co_bytes, ctx->cost_bytes, "cost_params"))) return rc;
    if ((rc = ensure_work_buffers(ctx, N))) return rc;
    const size_t D = ctx->D, C = ctx->C;
    if (states_out && (rc = ensure_device_buffer(ctx, &ctx->hist_states, sizeof(float) * T * N * D))) return rc;
    if (controls_out && (rc = ensure_device_buffer(ctx, &ctx->hist_controls, sizeof(float) * T * N * C))) return rc;
    if (costs_out && (rc = ensure_device_buffer(ctx, &ctx->hist_costs, sizeof(float) * T * N))) return rc;
    if (terminal_out && (rc = ensure_device_buffer(ctx, &ctx->terminal, sizeof(float) * N * D))) return rc;
    UploadPlan plan;
    if ((rc = upload(ctx, N, start_step, x0, u_nom, mo, mo_bytes, co, co_bytes, &plan))) return rc;
    uint8_t* up = static_cast<uint8_t*>(ctx->up.ptr);
    StepBuffers b;
    b.agents = reinterpret_cast<mppi_agent_state*>(up + plan.agents);
    b.x = reinterpret_cast<float*>(up + plan.x0);
    b.u_nom = reinterpret_cast<float*>(up + plan.u_nom);
    b.model = resolve_blob(ctx, mo, mo_bytes, ctx->model_bound, ctx->model_is_bound, plan.model);
    b.cost = resolve_blob(ctx, co, co_bytes, ctx->cost_bound, ctx->cost_is_bound, plan.cost);
    b.u_out = static_cast<float*>(ctx->down.ptr);
    b.diag = b.u_out + size_t(N) * ctx->H * C;
    b.hist_states = states_out ? static_cast<float*>(ctx->hist_states.ptr) : nullptr;
    b.hist_controls = controls_out ? static_cast<float*>(ctx->hist_controls.ptr) : nullptr;
    b.hist_costs = costs_out ? static_cast<float*>(ctx->hist_costs.ptr) : nullptr;
    b.terminal = terminal_out ? static_cast<float*>(ctx->terminal.ptr) : nullptr;
    for (uint32_t s = 0; s < T; ++s)
        if ((rc = launch_iteration(ctx, N, b, true))) return rc;
    if (states_out)
        HIP_TRY(ctx, hipMemcpyAsync(states_out, ctx->hist_states.ptr, sizeof(float) * T * N * D,
                                    hipMemcpyDeviceToHost, ctx->stream));
    if (controls_out)
        HIP_TRY(ctx, hipMemcpyAsync(controls_out, ctx->hist_controls.ptr, sizeof(float) * T * N * C,
                                    hipMemcpyDeviceToHost, ctx->stream));
    if (costs_out)
        HIP_TRY(ctx, hipMemcpyAsync(costs_out, ctx->hist_costs.ptr, sizeof(float) * T * N,
                                    hipMemcpyDeviceToHost, ctx->stream));
    if (terminal_out)
        HIP_TRY(ctx, hipMemcpyAsync(terminal_out, ctx->terminal.ptr, sizeof(float) * N * D,
                                    hipMemcpyDeviceToHost, ctx->stream));
    if ((rc = finish(ctx))) return rc;
    ctx->last_agents = N;
    return MPPI_HIP_OK;
}

}  // namespace

int mppi_hip_simulate(mppi_hip_ctx* ctx, uint32_t num_steps, uint64_t start_step, const float* x0,
                      const float* u_nominal, const void* mo, size_t mo_bytes, const void* co,
                      size_t co_bytes, float* states_out, float* controls_out, float* costs_out) {
    int rc = check_ctx(ctx);
    if (rc) return rc;
    if (num_steps == 0) return fail(ctx, MPPI_HIP_E_INVALID, "num_steps must be > 0.");
    if (!x0) return fail(ctx, MPPI_HIP_E_INVALID, "x0 must not be null.");
    if (!u_nominal) return fail(ctx, MPPI_HIP_E_INVALID, "u_nominal must not be null.");
    return run_closed_loop(ctx, 1, num_steps, start_step, x0, u_nominal, mo, mo_bytes, co, co_bytes,
                           states_out, controls_out, costs_out, nullptr);
}

int mppi_hip_simulate_batch(mppi_hip_ctx* ctx, uint32_t num_agents, uint32_t num_steps,
                            uint64_t start_step, const float* x0_packed,
                            const float* u_nominal_packed, float* states_out, float* controls_out,
                            float* costs_out, float* terminal_out) {
    int rc = check_ctx(ctx);
    if (rc) return rc;
    if (num_agents == 0) return fail(ctx, MPPI_HIP_E_INVALID, "num_agents must be > 0.");
    if (num_agents > 65535) return fail(ctx, MPPI_HIP_E_INVALID, "num_agents must be <= 65535.");
    if (num_steps == 0) return fail(ctx, MPPI_HIP_E_INVALID, "num_steps must be > 0.");
    if (!x0_packed) return fail(ctx, MPPI_HIP_E_INVALID, "x0_packed must not be null.");
    return run_closed_loop(ctx, num_agents, num_steps, start_step, x0_packed, u_nominal_packed,
                           nullptr, 0, nullptr, 0, states_out, controls_out, costs_out, terminal_out);
}

int mppi_hip_read_costs(mppi_hip_ctx* ctx, uint32_t agent, float* out, size_t count) {
    int rc = check_ctx(ctx);
    if (rc) return rc;
    if (!out || count != ctx->K) return fail(ctx, MPPI_HIP_E_INVALID, "read_costs: out must hold sample_count floats.");
    if (agent >= ctx->last_agents) return fail(ctx, MPPI_HIP_E_INVALID, "read_costs: agent out of range.");
    HIP_TRY(ctx, hipMemcpyAsync(out, static_cast<float*>(ctx->costs.ptr) + size_t(agent) * ctx->K,
                                sizeof(float) * count, hipMemcpyDeviceToHost, ctx->stream));
    HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    return MPPI_HIP_OK;
}

int mppi_hip_read_weights(mppi_hip_ctx* ctx, uint32_t agent, float* out, size_t count) {
    int rc = check_ctx(ctx);
    if (rc) return rc;
    if (!out || count != ctx->K) return fail(ctx, MPPI_HIP_E_INVALID, "read_weights: out must hold sample_count floats.");
    if (agent >= ctx->last_agents) return fail(ctx, MPPI_HIP_E_INVALID, "read_weights: agent out of range.");
    HIP_TRY(ctx, hipMemcpyAsync(out, static_cast<float*>(ctx->weights.ptr) + size_t(agent) * ctx->K,
                                sizeof(float) * count, hipMemcpyDeviceToHost, ctx->stream));
    HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    return MPPI_HIP_OK;
}

int mppi_hip_set_reduce_samples_per_thread(mppi_hip_ctx* ctx, uint32_t spt) {
    if (!ctx) return fail(nullptr, MPPI_HIP_E_INVALID, "context is null.");
    if (spt > 1024) return fail(ctx, MPPI_HIP_E_INVALID, "samples per thread must be <= 1024.");
    ctx->spt_override = spt;
    return MPPI_HIP_OK;
}

int mppi_hip_profile_enable(mppi_hip_ctx* ctx, int enable) {
    if (!ctx) return fail(nullptr, MPPI_HIP_E_INVALID, "context is null.");
    ctx->profiling = enable != 0;
    return MPPI_HIP_OK;
}

int mppi_hip_profile_read(mppi_hip_ctx* ctx, double* rollout_ms, double* reduce_ms,
                          uint64_t* launches) {
    if (!ctx) return fail(nullptr, MPPI_HIP_E_INVALID, "context is null.");
    if (ctx->stream) HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    collect_profile(ctx);
    if (rollout_ms) *rollout_ms = ctx->prof_rollout_ms;
    if (reduce_ms) *reduce_ms = ctx->prof_reduce_ms;
    if (launches) *launches = ctx->prof_launches;
    ctx->prof_rollout_ms = ctx->prof_reduce_ms = 0.0;
    ctx->prof_launches = 0;
    return MPPI_HIP_OK;
}

}  // extern "C"
