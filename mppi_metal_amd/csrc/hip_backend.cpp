// hip_backend.cpp — implementation of the C ABI in include/mppi_hip.h:
// contexts, plugin linking, the single-agent step, the device-resident closed
// loop and the uniform multi-agent batch.  The uneven multi-agent step is in
// batch_backend.cpp.
//
// Replaces the reference's Objective-C++ MetalBackend (src/metal_backend.mm).
// What the reference does with metallibs, MTLLinkedFunctions and visible
// function tables, this does with one hipRTC translation unit per
// (plugin source, callable names, dimensions): the library kernels
// (mppi_kernels.hip) and the plugin are compiled together for the device's
// architecture, the code object is loaded with hipModuleLoadData and cached
// in-process and on disk by a hash of everything that went into it.
//
// Per MPPI step the backend issues one host->device copy (agent descriptors
// and state, x0, nominal controls, parameter overrides, packed in pinned
// memory), the two kernels (mppi_rollout, mppi_reduce) and one device->host
// copy (u_next and the diagnostics), then synchronises the context's stream.
#include "backend.h"

#include <hip/hip_ext.h>
#include <hip/hiprtc.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <map>
#include <mutex>
#include <sstream>
#include <sys/stat.h>
#include <unistd.h>

namespace mppi_hip_impl {

namespace {
// Sources embedded at build time (generated by mppi_metal_amd/build.py).
#include "_jit_sources.inc"
}  // namespace

thread_local std::string g_thread_error;

size_t align_up(size_t x, size_t a) { return (x + a - 1) / a * a; }

void DeviceBuffer::release() {
    if (ptr) (void)hipFree(ptr);
    ptr = nullptr;
    bytes = 0;
}

void PinnedBuffer::release() {
    if (ptr) (void)hipHostFree(ptr);
    ptr = nullptr;
    bytes = 0;
}

int fail(mppi_hip_ctx* ctx, int code, const std::string& msg) {
    if (ctx) ctx->error = msg;
    else g_thread_error = msg;
    return code;
}

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

int check_ctx(mppi_hip_ctx* ctx) {
    if (!ctx) return fail(nullptr, MPPI_HIP_E_INVALID, "context is null.");
    if (!ctx->linked)
        return fail(ctx, MPPI_HIP_E_NOT_INITIALIZED, "Backend is not initialized: no plugin linked.");
    MPPI_HIP_TRY(ctx, hipSetDevice(ctx->device));
    return MPPI_HIP_OK;
}

namespace {

uint64_t fnv1a(const std::string& s, uint64_t h) {
    for (unsigned char c : s) {
        h ^= c;
        h *= 0x100000001b3ULL;
    }
    return h;
}

std::mutex g_cache_mutex;
std::map<std::string, std::vector<char>> g_code_objects;  // key -> code object
std::map<std::pair<std::string, int>, std::shared_ptr<LoadedModule>> g_modules;  // (key, device)
// One lock per code-object key, held while that object is read from disk or
// compiled: contexts created together on several host threads compile a
// plugin once, the others wait and take the result.
std::map<std::string, std::shared_ptr<std::mutex>> g_key_locks;
std::atomic<uint64_t> g_tmp_counter{0};

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
    auto is_id = [](char c) {
        return (c >= 'a' && c <= 'z') || (c >= 'A' && c <= 'Z') || (c >= '0' && c <= '9') || c == '_';
    };
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

// Lower Cholesky factor of a symmetric positive definite row-major matrix,
// computed in double and rounded to float.
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

// hipRTC options of the library kernels for one plugin configuration.
// FP32 keeps every operation exactly rounded as written (bit-exact with the
// CPU oracle); MIXED and FP16 compile the fast hardware transcendentals
// (MPPI_FAST_MATH), fused multiply-adds and approximate division; FP16 also
// rounds the control noise to binary16 (MPPI_NOISE_F16).
std::vector<std::string> jit_options(const std::string& arch, uint32_t D, uint32_t C, bool cov_full,
                                     bool batch, bool surrogate, const std::string& dyn,
                                     const std::string& stage, const std::string& term,
                                     const std::string& include_dir, uint32_t numeric_mode) {
    const bool fast = numeric_mode != MPPI_HIP_FP32;
    std::vector<std::string> opts = {
        "--offload-arch=" + arch,
        "-O3",
        "-std=c++17",
        fast ? "-ffp-contract=fast" : "-ffp-contract=off",
        "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
        "-DMPPI_SDIM=" + std::to_string(D),
        "-DMPPI_CDIM=" + std::to_string(C),
        "-DMPPI_DYNAMICS=" + dyn,
        "-DMPPI_COV_FULL=" + std::to_string(cov_full ? 1 : 0),
        "-DMPPI_BATCH_KERNELS=" + std::to_string(batch ? 1 : 0),
        "-DMPPI_SURROGATE_KERNELS=" + std::to_string(surrogate ? 1 : 0),
    };
    if (!stage.empty()) opts.push_back("-DMPPI_STAGE_COST=" + stage);
    if (!term.empty()) opts.push_back("-DMPPI_TERMINAL_COST=" + term);
    if (!include_dir.empty()) opts.push_back("-I" + include_dir);
    if (fast) {
        opts.push_back("-fno-hip-fp32-correctly-rounded-divide-sqrt");
        opts.push_back("-DMPPI_FAST_MATH=1");
    }
    if (numeric_mode == MPPI_HIP_FP16) opts.push_back("-DMPPI_NOISE_F16=1");
    return opts;
}

// Compile the library kernels with the plugin source (no device needed).
// Returns false with the compiler log on failure.
bool jit_compile(const std::string& plugin_src, const std::vector<std::string>& opts,
                 std::vector<char>* code, std::string* log) {
    const char* hdr_src[] = {kCoreHeader, kMathHeader, kPhiloxHeader, kLaunchHeader,
                             plugin_src.c_str(), kCoreHeader, kMlpHeader};
    const char* hdr_names[] = {"mppi_core.h", "mppi_math.h", "mppi_philox.h",
                               "mppi_launch.h", "mppi_user_plugin.hip", "mppi_hip/mppi_core.h",
                               "mppi_mlp.h"};
    hiprtcProgram prog;
    if (hiprtcCreateProgram(&prog, kKernelSource, "mppi_kernels.hip", 7, hdr_src, hdr_names) !=
        HIPRTC_SUCCESS) {
        *log = "hiprtcCreateProgram failed.";
        return false;
    }
    std::vector<const char*> copts;
    for (auto& o : opts) copts.push_back(o.c_str());
    hiprtcResult r = hiprtcCompileProgram(prog, int(copts.size()), copts.data());
    size_t log_size = 0;
    hiprtcGetProgramLogSize(prog, &log_size);
    log->assign(log_size, '\0');
    if (log_size) hiprtcGetProgramLog(prog, &(*log)[0]);
    if (r != HIPRTC_SUCCESS) {
        hiprtcDestroyProgram(&prog);
        return false;
    }
    size_t code_size = 0;
    hiprtcGetCodeSize(prog, &code_size);
    code->resize(code_size);
    hiprtcGetCode(prog, code->data());
    hiprtcDestroyProgram(&prog);
    return true;
}

}  // namespace

int load_module(mppi_hip_ctx* ctx, bool batch, bool surrogate, std::shared_ptr<LoadedModule>* out) {
    const std::vector<std::string> opts =
        jit_options(ctx->arch, ctx->D, ctx->C, ctx->cov_form == MPPI_HIP_COV_FULL, batch, surrogate,
                    ctx->dyn_name, ctx->stage_name, ctx->term_name, ctx->plugin_dir, ctx->numeric_mode);

    std::string keysrc = std::string(kKernelSource) + "\x01" + kLaunchHeader + "\x01" +
                         kCoreHeader + "\x01" + kMathHeader + "\x01" + kPhiloxHeader + "\x01" +
                         kMlpHeader + "\x01" + ctx->plugin_src + "\x01" + mppi_hip_version();
    for (auto& o : opts) keysrc += "\x02" + o;
    char keybuf[40];
    std::snprintf(keybuf, sizeof(keybuf), "%016llx%016llx",
                  (unsigned long long)fnv1a(keysrc, 0xcbf29ce484222325ULL),
                  (unsigned long long)fnv1a(keysrc, 0x84222325cbf29ce4ULL));
    const std::string key = keybuf;

    {
        std::lock_guard<std::mutex> lock(g_cache_mutex);
        auto it = g_modules.find({key, ctx->device});
        if (it != g_modules.end()) {
            *out = it->second;
            return MPPI_HIP_OK;
        }
    }
    std::shared_ptr<std::mutex> key_lock;
    {
        std::lock_guard<std::mutex> lock(g_cache_mutex);
        auto& l = g_key_locks[key];
        if (!l) l = std::make_shared<std::mutex>();
        key_lock = l;
    }
    std::lock_guard<std::mutex> key_guard(*key_lock);
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
        std::string log;
        if (!jit_compile(ctx->plugin_src, opts, &code, &log))
            return fail(ctx, MPPI_HIP_E_COMPILE, "Plugin '" + ctx->plugin_name + "' failed to compile: " + log);
        if (!cpath.empty()) {
            mkdirs(dir);
            std::string tmp = cpath + ".tmp" + std::to_string(uint64_t(::getpid())) + "." +
                              std::to_string(g_tmp_counter.fetch_add(1));
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
        return fail(ctx, MPPI_HIP_E_COMPILE, std::string("hipModuleLoadData failed: ") + hipGetErrorString(e));
    bool ok = hipModuleGetFunction(&m->rollout, m->module, "mppi_rollout") == hipSuccess &&
              hipModuleGetFunction(&m->reduce, m->module, "mppi_reduce") == hipSuccess &&
              hipModuleGetFunction(&m->finalize, m->module, "mppi_finalize_agents") == hipSuccess &&
              hipModuleGetFunction(&m->rollout_fused, m->module, "mppi_rollout_fused") == hipSuccess &&
              hipModuleGetFunction(&m->fused_combine, m->module, "mppi_fused_combine") == hipSuccess &&
              hipModuleGetFunction(&m->rollout_fused_one, m->module, "mppi_rollout_fused_one") == hipSuccess &&
              hipModuleGetFunction(&m->fused_combine_one, m->module, "mppi_fused_combine_one") == hipSuccess &&
              hipModuleGetFunction(&m->weights, m->module, "mppi_weights") == hipSuccess;
    if (ok && batch) {
        const char* blocks[3] = {"64", "128", "256"};
        const char* modes[2] = {"table", "persist"};
        for (int b = 0; b < 3 && ok; ++b)
            for (int md = 0; md < 2 && ok; ++md) {
                std::string r = std::string("mppi_rollout_") + modes[md] + "_b" + blocks[b];
                std::string d = std::string("mppi_reduce_") + modes[md] + "_b" + blocks[b];
                std::string st = std::string("mppi_step_") + modes[md] + "_b" + blocks[b];
                ok = hipModuleGetFunction(&m->rollout_b[b][md], m->module, r.c_str()) == hipSuccess &&
                     hipModuleGetFunction(&m->reduce_b[b][md], m->module, d.c_str()) == hipSuccess &&
                     hipModuleGetFunction(&m->step_b[b][md], m->module, st.c_str()) == hipSuccess;
            }
    }
    if (ok && surrogate)
        ok = hipModuleGetFunction(&m->surrogate_data, m->module, "mppi_surrogate_data") == hipSuccess &&
             hipModuleGetFunction(&m->rollout_surrogate, m->module, "mppi_rollout_surrogate") == hipSuccess;
    if (!ok) {
        (void)hipModuleUnload(m->module);
        return fail(ctx, MPPI_HIP_E_COMPILE, "Required library kernels missing from code object.");
    }
    std::lock_guard<std::mutex> lock(g_cache_mutex);
    auto it = g_modules.find({key, ctx->device});
    if (it != g_modules.end()) {
        (void)hipModuleUnload(m->module);
        *out = it->second;
    } else {
        g_modules[{key, ctx->device}] = m;
        *out = m;
    }
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

void fill_common_args(mppi_hip_ctx* ctx, mppi_step_args* a) {
    const float* cbase = static_cast<const float*>(ctx->consts.ptr);
    const size_t ns = ctx->noise_scale.size();
    a->noise_scale = cbase;
    a->u_min = cbase + ns;
    a->u_max = cbase + ns + ctx->C;
    a->key0 = ctx->key0;
    a->key1 = ctx->key1;
    a->lambda = ctx->lambda;
    a->wait_limit = ctx->wait_limit;
    a->wait_bias = ctx->wait_bias;
}

int check_status(mppi_hip_ctx* ctx, const uint32_t* status, uint32_t n) {
    for (uint32_t i = 0; i < n; ++i)
        if (status[i] & MPPI_STATUS_WAIT_TIMEOUT)
            return fail(ctx, MPPI_HIP_E_DEVICE,
                        "agent " + std::to_string(i) + ": a rollout wavefront of the one-pass step gave up "
                        "waiting for its noise (the bounded wait ran out); the results are not valid.");
    return MPPI_HIP_OK;
}

namespace {

// Samples per reduce thread: about two reduce workgroups per compute unit
// over the launch (measured best on gfx950 at K = 65536, H = 64: 512
// workgroups beat 1024 and 256), at most 64 samples per thread.
uint32_t default_spt(const mppi_hip_ctx* ctx, uint32_t K, uint32_t n_groups, uint32_t N) {
    if (ctx->spt_override) return ctx->spt_override;
    const uint64_t target = 2 * uint64_t(ctx->num_cus > 0 ? ctx->num_cus : 256);
    uint64_t work = uint64_t(K) * n_groups * N;
    uint64_t spt = (work + uint64_t(kBlock) * target - 1) / (uint64_t(kBlock) * target);
    return uint32_t(std::max<uint64_t>(1, std::min<uint64_t>(64, spt)));
}

// Floats of partial sums one agent of shape d needs: the reduce's chunk
// partials or the one-pass step's slice-major tile sums, rounded up to whole
// 128-byte lines so that every agent's rows start on a line.
uint64_t partial_floats(const mppi_agent_desc& d, uint32_t C) {
    const uint64_t reduce = uint64_t(d.n_chunks) * d.NB * 4;
    const uint64_t fused = uint64_t(d.n_blk) * mppi_combine_slices(d.H, C) * mppi_combine_width(C);
    return (std::max(reduce, fused) + 31) / 32 * 32;
}

// Shape of agent 0 of N agents sharing the context's K and H (offsets 0).
mppi_agent_desc uniform_desc(const mppi_hip_ctx* ctx, uint32_t N) {
    mppi_agent_desc d{};
    d.K = ctx->K;
    d.H = ctx->H;
    d.n_blk = (ctx->K + kBlock - 1) / kBlock;
    d.NB = (ctx->H * ctx->C + 3) / 4;
    d.n_groups = (d.NB + kReduceBPG - 1) / kReduceBPG;
    d.spt = default_spt(ctx, ctx->K, d.n_groups, N);
    d.n_chunks = (ctx->K + kBlock * d.spt - 1) / (kBlock * d.spt);
    return d;
}

// Descriptors of N agents sharing the context's K and H (grid-mapped path).
std::vector<mppi_agent_desc> uniform_descs(const mppi_hip_ctx* ctx, uint32_t N) {
    const mppi_agent_desc d = uniform_desc(ctx, N);
    std::vector<mppi_agent_desc> v(N, d);
    for (uint32_t a = 0; a < N; ++a) {
        v[a].slot = a;
        v[a].cost_off = uint64_t(a) * ctx->K;
        v[a].w_off = v[a].cost_off;
        v[a].seq_off = uint64_t(a) * ctx->H * ctx->C;
        v[a].out_off = v[a].seq_off;
        v[a].blk_off = uint64_t(a) * d.n_blk;
        v[a].part_off = uint64_t(a) * partial_floats(d, ctx->C);
    }
    return v;
}

int ensure_work_buffers(mppi_hip_ctx* ctx, const std::vector<mppi_agent_desc>& d) {
    const mppi_agent_desc& last = d.back();
    const uint32_t N = uint32_t(d.size());
    int rc;
    const size_t nk = last.cost_off + last.K;
    if ((rc = ensure_device_buffer(ctx, &ctx->costs, sizeof(float) * nk))) return rc;
    if ((rc = ensure_device_buffer(ctx, &ctx->weights, sizeof(float) * nk))) return rc;
    if ((rc = ensure_device_buffer(ctx, &ctx->blk_stats, sizeof(float) * 4 * (last.blk_off + last.n_blk)))) return rc;
    if ((rc = ensure_device_buffer(ctx, &ctx->partials,
                                   sizeof(float) * (last.part_off + partial_floats(last, ctx->C))))) return rc;
    // u_out | diag (3 per agent) | status (one word per agent)
    const size_t down = sizeof(float) * (last.seq_off + size_t(last.H) * ctx->C + 4 * size_t(N));
    if ((rc = ensure_device_buffer(ctx, &ctx->down, down))) return rc;
    if ((rc = ensure_pinned_buffer(ctx, &ctx->down_host, down))) return rc;
    return MPPI_HIP_OK;
}

// Layout of the per-call upload block (identical on host and device).
struct UploadPlan {
    size_t desc = 0, agents = 0, x0 = 0, u_nom = 0, model = 0, cost = 0, total = 0;
    size_t u_nom2 = 0, device_total = 0;   // closed loop: second warm-start buffer, device only
};

UploadPlan plan_upload(const mppi_hip_ctx* ctx, uint32_t N, size_t model_bytes, size_t cost_bytes) {
    UploadPlan p;
    size_t off = 0;
    p.desc = off;
    off = align_up(off + sizeof(mppi_agent_desc) * N, 256);
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
    p.u_nom2 = off;
    off = align_up(off + sizeof(float) * size_t(N) * ctx->H * ctx->C, 256);
    p.device_total = off;
    return p;
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

// Resolve the parameter blob for a call: override (copied into the upload
// block), else the bound blob, else the zero blob.
const uint8_t* resolve_blob(mppi_hip_ctx* ctx, const void* ovr, size_t ovr_bytes,
                            const DeviceBuffer& bound, bool is_bound, size_t up_off) {
    if (ovr && ovr_bytes > 0) return static_cast<const uint8_t*>(ctx->up.ptr) + up_off;
    if (is_bound) return static_cast<const uint8_t*>(bound.ptr);
    const float* cbase = static_cast<const float*>(ctx->consts.ptr);
    return reinterpret_cast<const uint8_t*>(cbase + ctx->noise_scale.size() + 2 * ctx->C);
}

// Fill the pinned upload block for N agents and copy it to the device.
int upload(mppi_hip_ctx* ctx, const std::vector<mppi_agent_desc>& desc, uint64_t step,
           const float* x0, const float* u_nom, const void* mo, size_t mo_bytes, const void* co,
           size_t co_bytes, UploadPlan* plan) {
    const uint32_t N = uint32_t(desc.size());
    *plan = plan_upload(ctx, N, mo_bytes, co_bytes);
    int rc;
    if ((rc = ensure_device_buffer(ctx, &ctx->up, plan->device_total))) return rc;
    if ((rc = ensure_pinned_buffer(ctx, &ctx->up_host, plan->total))) return rc;
    uint8_t* h = static_cast<uint8_t*>(ctx->up_host.ptr);
    std::memcpy(h + plan->desc, desc.data(), sizeof(mppi_agent_desc) * N);
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
    MPPI_HIP_TRY(ctx, hipMemcpyAsync(ctx->up.ptr, h, plan->total, hipMemcpyHostToDevice, ctx->stream));
    return MPPI_HIP_OK;
}

// The one-pass step applies when the normals of a 256-sample workgroup fit
// in LDS (H*C <= 128), K is below the FP64 threshold, and the surrogate is
// not in use.  AUTO takes it when the rollout grid has at most one workgroup
// per compute unit: the kernel's LDS then costs no occupancy, since each
// compute unit runs one workgroup anyway.
bool use_fused(const mppi_hip_ctx* ctx, const mppi_agent_desc& d, uint32_t N, bool surrogate) {
    if (ctx->step_path == MPPI_HIP_PATH_TWO_KERNEL || surrogate) return false;
    if (size_t(d.H) * ctx->C > kFusedMaxHC || d.K >= (1u << 18) || d.n_blk > kBlock) return false;
    if (ctx->step_path == MPPI_HIP_PATH_FUSED) return true;
    return uint64_t(d.n_blk) * N <= uint64_t(ctx->num_cus);
}

// Launch f (grid in workgroups of `block` threads) on the context's stream.
// With a start and stop event (profiling), the runtime attaches them to the
// dispatch itself, so the pair times the kernel alone, as a kernel trace
// does, rather than the event packets and the launch gap around it.
hipError_t launch_kernel(hipStream_t stream, hipFunction_t f, uint32_t gx, uint32_t gy, uint32_t gz,
                         void** params, hipEvent_t start, hipEvent_t stop, uint32_t block = kBlock) {
    if (start && stop)
        return hipExtModuleLaunchKernel(f, gx * block, gy, gz, block, 1, 1, 0, stream, params, nullptr,
                                        start, stop, 0);
    return hipModuleLaunchKernel(f, gx, gy, gz, block, 1, 1, 0, stream, params, nullptr);
}

// Enqueue one MPPI iteration for N agents of shape d (grid-mapped rollout +
// reduce, or the one-pass step) on `stream` (the context's when null).
// Without `bookkeep` nothing else is touched: no profiling events are taken
// and the context's pending per-sample weights are left alone.
int launch_iteration(mppi_hip_ctx* ctx, const mppi_agent_desc& d, uint32_t N, const mppi_step_args& a,
                     hipStream_t stream = nullptr, bool bookkeep = true) {
    hipStream_t st = stream ? stream : ctx->stream;
    mppi_step_args args = a;
    void* params[] = {&args};
    hipEvent_t e0 = nullptr, e1 = nullptr, e2 = nullptr, e3 = nullptr;
    const bool prof = bookkeep && ctx->profiling;
    if (prof) {
        e0 = take_event(ctx); e1 = take_event(ctx); e2 = take_event(ctx); e3 = take_event(ctx);
    }
    alignas(16) unsigned char sur_args[1024];
    hipFunction_t sur_fn = nullptr;
    const bool sur = surrogate_rollout(ctx, args, &sur_fn, sur_args, sizeof(sur_args));
    if (bookkeep) ctx->weights_pending = false;
    if (use_fused(ctx, d, N, sur)) {
        // two threads per sample: noise producers and rollout consumers; one
        // agent described in the arguments takes the kernels that read its
        // descriptor and step counter from there
        const bool one = N == 1 && args.one;
        MPPI_HIP_TRY(ctx, launch_kernel(st, one ? ctx->mod->rollout_fused_one : ctx->mod->rollout_fused, d.n_blk,
                                        N, 1, params, e0, e1, 2 * kBlock));
        // the combine's workgroups: one per slice of timesteps
        MPPI_HIP_TRY(ctx, launch_kernel(st, one ? ctx->mod->fused_combine_one : ctx->mod->fused_combine,
                                        mppi_combine_slices(d.H, ctx->C), N, 1, params, e2, e3));
        if (bookkeep) {
            ctx->weights_pending = true;
            ctx->weights_args = args;
            ctx->weights_grid_x = (d.K + kBlock - 1) / kBlock;
            ctx->weights_grid_y = N;
        }
        if (prof) {
            if (e0 && e1) ctx->pending_rollout.push_back({e0, e1});
            if (e2 && e3) ctx->pending_reduce.push_back({e2, e3});
            ctx->prof_launches += 1;
        }
        return MPPI_HIP_OK;
    }
    if (sur) {
        void* sparams[] = {sur_args};
        MPPI_HIP_TRY(ctx, launch_kernel(st, sur_fn, (d.K + MPPI_SUR_BLOCK - 1) / MPPI_SUR_BLOCK, N, 1, sparams,
                                        e0, e1, MPPI_SUR_BLOCK));
    } else {
        MPPI_HIP_TRY(ctx, launch_kernel(st, ctx->mod->rollout, d.n_blk, N, 1, params, e0, e1));
    }
    MPPI_HIP_TRY(ctx, launch_kernel(st, ctx->mod->reduce, d.n_chunks, d.n_groups, N, params, e2, e3));
    // the update: one workgroup per agent (timed with the reduce)
    hipEvent_t e4 = prof ? take_event(ctx) : nullptr, e5 = prof ? take_event(ctx) : nullptr;
    MPPI_HIP_TRY(ctx, launch_kernel(st, ctx->mod->finalize, N, 1, 1, params, e4, e5));
    if (prof) {
        if (e0 && e1) ctx->pending_rollout.push_back({e0, e1});
        if (e2 && e3) ctx->pending_reduce.push_back({e2, e3});
        if (e4 && e5) ctx->pending_reduce.push_back({e4, e5});
        ctx->prof_launches += 1;
    }
    return MPPI_HIP_OK;
}

int finish(mppi_hip_ctx* ctx) {
    MPPI_HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    if (ctx->profiling) collect_profile(ctx);
    return MPPI_HIP_OK;
}

// Set up the step arguments of the grid-mapped path after an upload.
mppi_step_args plain_args(mppi_hip_ctx* ctx, const UploadPlan& plan, uint32_t N, const void* mo,
                          size_t mo_bytes, const void* co, size_t co_bytes) {
    uint8_t* up = static_cast<uint8_t*>(ctx->up.ptr);
    mppi_step_args a{};
    fill_common_args(ctx, &a);
    a.desc = reinterpret_cast<const mppi_agent_desc*>(up + plan.desc);
    a.agents = reinterpret_cast<mppi_agent_state*>(up + plan.agents);
    a.N = N;
    a.x = reinterpret_cast<float*>(up + plan.x0);
    a.u_nom = reinterpret_cast<float*>(up + plan.u_nom);
    a.u_out = static_cast<float*>(ctx->down.ptr);
    a.model = resolve_blob(ctx, mo, mo_bytes, ctx->model_bound, ctx->model_is_bound, plan.model);
    a.cost = resolve_blob(ctx, co, co_bytes, ctx->cost_bound, ctx->cost_is_bound, plan.cost);
    a.costs = static_cast<float*>(ctx->costs.ptr);
    a.weights = static_cast<float*>(ctx->weights.ptr);
    a.blk_stats = static_cast<float*>(ctx->blk_stats.ptr);
    a.partials = static_cast<float*>(ctx->partials.ptr);
    a.diag = a.u_out + size_t(N) * ctx->H * ctx->C;
    a.status_out = reinterpret_cast<uint32_t*>(a.diag + 3 * size_t(N));
    return a;
}

int run_closed_loop(mppi_hip_ctx* ctx, uint32_t N, uint32_t T, uint64_t start_step,
                    const float* x0, const float* u_nom, const void* mo, size_t mo_bytes,
                    const void* co, size_t co_bytes, float* states_out, float* controls_out,
                    float* costs_out, float* terminal_out) {
    int rc;
    if ((rc = check_override(ctx, mo, mo_bytes, ctx->model_bytes, "model_params"))) return rc;
    if ((rc = check_override(ctx, co, co_bytes, ctx->cost_bytes, "cost_params"))) return rc;
    std::vector<mppi_agent_desc> desc = uniform_descs(ctx, N);
    if ((rc = ensure_work_buffers(ctx, desc))) return rc;
    const size_t D = ctx->D, C = ctx->C;
    // One block for everything the loop returns, so it comes back in one
    // transfer (through pinned memory) however many outputs are asked for.
    const size_t o_ctl = align_up(sizeof(float) * T * N * D, 256);
    const size_t o_cost = o_ctl + align_up(sizeof(float) * T * N * C, 256);
    const size_t o_term = o_cost + align_up(sizeof(float) * T * N, 256);
    const size_t o_stat = o_term + align_up(sizeof(float) * N * D, 256);
    const size_t hist_bytes = o_stat + sizeof(uint32_t) * N;
    if ((rc = ensure_device_buffer(ctx, &ctx->hist, hist_bytes))) return rc;
    if ((rc = ensure_pinned_buffer(ctx, &ctx->hist_host, hist_bytes))) return rc;
    uint8_t* hd = static_cast<uint8_t*>(ctx->hist.ptr);
    const uint8_t* hh = static_cast<const uint8_t*>(ctx->hist_host.ptr);
    UploadPlan plan;
    if ((rc = upload(ctx, desc, start_step, x0, u_nom, mo, mo_bytes, co, co_bytes, &plan))) return rc;
    mppi_step_args a = plain_args(ctx, plan, N, mo, mo_bytes, co, co_bytes);
    a.closed_loop = 1;
    a.hist_states = states_out ? reinterpret_cast<float*>(hd) : nullptr;
    a.hist_controls = controls_out ? reinterpret_cast<float*>(hd + o_ctl) : nullptr;
    a.hist_costs = costs_out ? reinterpret_cast<float*>(hd + o_cost) : nullptr;
    a.terminal = terminal_out ? reinterpret_cast<float*>(hd + o_term) : nullptr;
    a.status_out = reinterpret_cast<uint32_t*>(hd + o_stat);
    // The warm start alternates between two device buffers: step s reads one
    // and its update writes the shifted sequence into the other.
    uint8_t* up = static_cast<uint8_t*>(ctx->up.ptr);
    float* warm[2] = {reinterpret_cast<float*>(up + plan.u_nom), reinterpret_cast<float*>(up + plan.u_nom2)};
    a.one = N == 1 ? 1u : 0u;
    a.desc0 = desc[0];
    for (uint32_t s = 0; s < T; ++s) {
        a.u_nom = warm[s & 1u];
        a.u_nom_next = warm[(s + 1) & 1u];
        a.step0 = uint32_t(start_step + s);   // the device counter's value at step s
        if ((rc = launch_iteration(ctx, desc[0], uint32_t(desc.size()), a))) return rc;
    }
    // The requested parts and the status words, in one transfer: from the
    // first requested part to the end of the block.
    const size_t from = states_out ? 0 : controls_out ? o_ctl : costs_out ? o_cost : terminal_out ? o_term : o_stat;
    MPPI_HIP_TRY(ctx, hipMemcpyAsync(ctx->hist_host.ptr, hd + from, hist_bytes - from, hipMemcpyDeviceToHost,
                                     ctx->stream));
    if ((rc = finish(ctx))) return rc;
    // part at offset o (o >= from) of the device block
    auto at = [&](size_t o) { return hh + (o - from); };
    if ((rc = check_status(ctx, reinterpret_cast<const uint32_t*>(at(o_stat)), N))) return rc;
    if (states_out) std::memcpy(states_out, at(0), sizeof(float) * T * N * D);
    if (controls_out) std::memcpy(controls_out, at(o_ctl), sizeof(float) * T * N * C);
    if (costs_out) std::memcpy(costs_out, at(o_cost), sizeof(float) * T * N);
    if (terminal_out) std::memcpy(terminal_out, at(o_term), sizeof(float) * N * D);
    ctx->last_agents = N;
    ctx->last_costs = static_cast<const float*>(ctx->costs.ptr);
    ctx->last_weights = static_cast<const float*>(ctx->weights.ptr);
    ctx->last_desc = desc;
    return MPPI_HIP_OK;
}

}  // namespace
}  // namespace mppi_hip_impl

using namespace mppi_hip_impl;

extern "C" {

const char* mppi_hip_version(void) { return "mppi_hip 0.2.0"; }

int mppi_hip_compile_check(const char* plugin_source, const char* include_dir, const char* dynamics_fn,
                           const char* stage_cost_fn, const char* terminal_cost_fn, uint32_t state_dim,
                           uint32_t control_dim, int covariance_full, uint32_t numeric_mode,
                           uint32_t kernel_sets, char* log, size_t log_capacity) {
    if (log && log_capacity) log[0] = '\0';
    auto put_log = [&](const std::string& s) {
        if (log && log_capacity) {
            const size_t n = std::min(s.size(), log_capacity - 1);
            std::memcpy(log, s.data(), n);
            log[n] = '\0';
        }
    };
    if (!plugin_source || !dynamics_fn || !*dynamics_fn || state_dim == 0 || control_dim == 0 ||
        numeric_mode > MPPI_HIP_MIXED) {
        put_log("plugin_source and dynamics_fn are required; dimensions must be > 0; numeric_mode "
                "must be FP32, FP16 or MIXED.");
        return MPPI_HIP_E_INVALID;
    }
    const std::vector<std::string> opts =
        jit_options("gfx950", state_dim, control_dim, covariance_full != 0, (kernel_sets & 1u) != 0,
                    (kernel_sets & 2u) != 0, dynamics_fn, stage_cost_fn ? stage_cost_fn : "",
                    terminal_cost_fn ? terminal_cost_fn : "", include_dir ? include_dir : "",
                    numeric_mode);
    std::vector<char> code;
    std::string clog;
    if (!jit_compile(plugin_source, opts, &code, &clog)) {
        put_log(clog);
        return MPPI_HIP_E_COMPILE;
    }
    return code.empty() ? MPPI_HIP_E_COMPILE : MPPI_HIP_OK;
}

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
    if (cfg->numeric_mode != MPPI_HIP_FP32 && cfg->numeric_mode != MPPI_HIP_FP16 &&
        cfg->numeric_mode != MPPI_HIP_MIXED)
        return fail(nullptr, MPPI_HIP_E_INVALID, "numeric_mode is not FP32, FP16 or MIXED.");
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
            return fail(nullptr, MPPI_HIP_E_INVALID, "sigma (full covariance) is not positive definite.");
    } else {
        ctx->noise_scale.assign(cfg->sigma, cfg->sigma + C);
    }
    ctx->u_min.assign(cfg->u_min, cfg->u_min + C);
    ctx->u_max.assign(cfg->u_max, cfg->u_max + C);

    int count = 0;
    hipError_t e = hipGetDeviceCount(&count);
    if (e != hipSuccess || count == 0)
        return fail(nullptr, MPPI_HIP_E_DEVICE, std::string("No HIP device available: ") + hipGetErrorString(e));
    int dev = cfg->device;
    if (dev < 0 && hipGetDevice(&dev) != hipSuccess) dev = 0;
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
    ctx->num_cus = prop.multiProcessorCount > 0 ? prop.multiProcessorCount : 256;
    if ((e = hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking)) != hipSuccess)
        return fail(nullptr, MPPI_HIP_E_DEVICE, std::string("hipStreamCreate failed: ") + hipGetErrorString(e));
    // Device address of the no-allocation entry point's scratch array on
    // this device (once per device, with the device current).
    noalloc_setup(dev);

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
                            &ctx->costs, &ctx->blk_stats, &ctx->weights, &ctx->partials, &ctx->hist})
        b->release();
    ctx->up_host.release();
    ctx->down_host.release();
    ctx->hist_host.release();
    collect_profile(ctx);
    for (hipEvent_t e : ctx->ev_pool) (void)hipEventDestroy(e);
    if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
    delete ctx;
}

namespace {

int link_common(mppi_hip_ctx* ctx, const std::string& src, const std::string& src_name,
                const std::string& plugin_dir, const char* dynamics, const char* stage,
                const char* terminal, size_t model_bytes, size_t cost_bytes, uint64_t seed) {
    std::string dyn = dynamics ? dynamics : "";
    std::string sc = stage ? stage : "";
    std::string tc = terminal ? terminal : "";
    if (dyn.empty()) return fail(ctx, MPPI_HIP_E_PLUGIN, "Required callable name for dynamics is empty.");
    for (const std::string* n : {&dyn, &sc, &tc})
        if (!n->empty() && !valid_identifier(*n))
            return fail(ctx, MPPI_HIP_E_PLUGIN, "Callable name '" + *n + "' is not a C identifier.");
    if (!has_identifier(src, dyn))
        return fail(ctx, MPPI_HIP_E_PLUGIN,
                    "Required callable '" + dyn + "' not found in plugin source '" + src_name + "'.");
    if (!sc.empty() && !has_identifier(src, sc))
        return fail(ctx, MPPI_HIP_E_PLUGIN, "Callable '" + sc + "' not found in plugin source '" + src_name + "'.");
    if (!tc.empty() && !has_identifier(src, tc))
        return fail(ctx, MPPI_HIP_E_PLUGIN, "Callable '" + tc + "' not found in plugin source '" + src_name + "'.");
    ctx->plugin_src = src;
    ctx->plugin_name = src_name;
    ctx->plugin_dir = plugin_dir;
    ctx->dyn_name = dyn;
    ctx->stage_name = sc;
    ctx->term_name = tc;
    ctx->linked = false;
    ctx->batch_mod.reset();
    std::shared_ptr<LoadedModule> mod;
    ctx->sur.reset();
    int rc = load_module(ctx, false, false, &mod);
    if (rc) return rc;
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
    MPPI_HIP_TRY(ctx, hipSetDevice(ctx->device));
    return link_common(ctx, src, path, dir, dynamics, stage_cost, terminal_cost, model_bytes,
                       cost_bytes, seed);
}

int mppi_hip_link_plugin_source(mppi_hip_ctx* ctx, const char* source, const char* source_name,
                                const char* dynamics, const char* stage_cost,
                                const char* terminal_cost, size_t model_bytes, size_t cost_bytes,
                                uint64_t seed) {
    if (!ctx) return fail(nullptr, MPPI_HIP_E_INVALID, "context is null.");
    if (!source || !*source) return fail(ctx, MPPI_HIP_E_PLUGIN, "plugin source is empty.");
    MPPI_HIP_TRY(ctx, hipSetDevice(ctx->device));
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
        MPPI_HIP_TRY(ctx, hipMemcpyAsync(ctx->model_bound.ptr, model, model_bytes, hipMemcpyHostToDevice,
                                         ctx->stream));
    }
    if (cost_bytes) {
        if ((rc = ensure_device_buffer(ctx, &ctx->cost_bound, cost_bytes))) return rc;
        MPPI_HIP_TRY(ctx, hipMemcpyAsync(ctx->cost_bound.ptr, cost, cost_bytes, hipMemcpyHostToDevice,
                                         ctx->stream));
    }
    MPPI_HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
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
    std::vector<mppi_agent_desc> desc = uniform_descs(ctx, 1);
    if ((rc = ensure_work_buffers(ctx, desc))) return rc;
    UploadPlan plan;
    if ((rc = upload(ctx, desc, step_index, x0, u_nominal, mo, mo_bytes, co, co_bytes, &plan))) return rc;
    mppi_step_args a = plain_args(ctx, plan, 1, mo, mo_bytes, co, co_bytes);
    a.one = 1;
    a.desc0 = desc[0];
    a.step0 = uint32_t(step_index);
    // The update, the diagnostics and the status word go straight to pinned
    // host memory (the same layout as the device block): the device writes
    // them over the bus, and they are visible to the host once the stream has
    // synchronised, so a step needs no device-to-host copy.
    const size_t hc = size_t(ctx->H) * ctx->C;
    float* hout = static_cast<float*>(ctx->down_host.ptr);
    a.u_out = hout;
    a.diag = hout + hc;
    a.status_out = reinterpret_cast<uint32_t*>(hout + hc + 3);
    if ((rc = launch_iteration(ctx, desc[0], uint32_t(desc.size()), a))) return rc;
    if ((rc = finish(ctx))) return rc;
    const float* h = static_cast<const float*>(ctx->down_host.ptr);
    if ((rc = check_status(ctx, reinterpret_cast<const uint32_t*>(h + hc + 3), 1))) return rc;
    std::memcpy(u_out, h, sizeof(float) * hc);
    if (diag3) std::memcpy(diag3, h + hc, sizeof(float) * 3);
    ctx->last_agents = 1;
    ctx->last_costs = static_cast<const float*>(ctx->costs.ptr);
    ctx->last_weights = static_cast<const float*>(ctx->weights.ptr);
    ctx->last_desc = desc;
    return MPPI_HIP_OK;
}

int mppi_hip_step_noalloc(mppi_hip_ctx* ctx, uint64_t step_index, const float* x0_dev, const float* u_nominal_dev,
                          float* u_out_dev, float* diag3_dev, void* stream) {
    int rc = check_ctx(ctx);
    if (rc) return rc;
    if (!x0_dev || !u_nominal_dev || !u_out_dev)
        return fail(ctx, MPPI_HIP_E_INVALID, "x0, u_nominal and u_out must not be null.");
    std::string why;
    uint8_t* base = static_cast<uint8_t*>(noalloc_scratch(ctx->device, &why));
    if (!base) return fail(ctx, MPPI_HIP_E_DEVICE, "mppi_hip_step_noalloc: no scratch on this device: " + why);
    // Scratch layout: descriptor | agent state | status word | costs |
    // workgroup statistics | partial sums (offsets 256-byte aligned).
    const mppi_agent_desc shape = uniform_desc(ctx, 1);
    const size_t off_agent = 256, off_status = 512, off_costs = 768;
    const size_t off_stats = off_costs + align_up(sizeof(float) * shape.K, 256);
    const size_t off_part = off_stats + align_up(sizeof(float) * 4 * shape.n_blk, 256);
    const size_t total = off_part + align_up(sizeof(float) * partial_floats(shape, ctx->C), 256);
    if (total > kNoallocScratchBytes)
        return fail(ctx, MPPI_HIP_E_INVALID,
                    "mppi_hip_step_noalloc: sample_count and horizon need " + std::to_string(total) +
                        " bytes of scratch; the limit is " + std::to_string(kNoallocScratchBytes) +
                        " bytes (8 MiB) per device.");
    mppi_step_args a{};
    fill_common_args(ctx, &a);
    a.desc = reinterpret_cast<const mppi_agent_desc*>(base);
    a.agents = reinterpret_cast<mppi_agent_state*>(base + off_agent);
    a.status_out = reinterpret_cast<uint32_t*>(base + off_status);
    a.N = 1;
    a.x = const_cast<float*>(x0_dev);              // read only: open loop
    a.u_nom = const_cast<float*>(u_nominal_dev);   // read only: open loop
    a.u_out = u_out_dev;
    a.diag = diag3_dev;
    a.model = resolve_blob(ctx, nullptr, 0, ctx->model_bound, ctx->model_is_bound, 0);
    a.cost = resolve_blob(ctx, nullptr, 0, ctx->cost_bound, ctx->cost_is_bound, 0);
    a.costs = reinterpret_cast<float*>(base + off_costs);
    a.weights = nullptr;
    a.blk_stats = reinterpret_cast<float*>(base + off_stats);
    a.partials = reinterpret_cast<float*>(base + off_part);
    a.one = 1;
    a.desc0 = shape;
    a.step0 = uint32_t(step_index);
    hipStream_t st = stream ? static_cast<hipStream_t>(stream) : ctx->stream;
    MPPI_HIP_TRY(ctx, noalloc_init(st, shape, uint32_t(step_index), const_cast<mppi_agent_desc*>(a.desc), a.agents,
                                   a.status_out));
    if ((rc = launch_iteration(ctx, shape, 1, a, st, false))) return rc;
    // read_costs / read_weights refer to the last call that used the
    // context's own buffers; after this call they report no iteration.
    ctx->last_agents = 0;
    ctx->weights_pending = false;
    if (stream) return MPPI_HIP_OK;
    uint32_t status = 0;
    MPPI_HIP_TRY(ctx, hipMemcpyAsync(&status, a.status_out, sizeof(uint32_t), hipMemcpyDeviceToHost, ctx->stream));
    MPPI_HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    return check_status(ctx, &status, 1);
}

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

namespace {
int read_per_sample(mppi_hip_ctx* ctx, const float* base, uint32_t agent, float* out, size_t count,
                    const char* what) {
    int rc = check_ctx(ctx);
    if (rc) return rc;
    if (!base || agent >= ctx->last_agents || agent >= ctx->last_desc.size())
        return fail(ctx, MPPI_HIP_E_INVALID, std::string(what) + ": no iteration for this agent.");
    const mppi_agent_desc& d = ctx->last_desc[agent];
    if (!out || count != d.K)
        return fail(ctx, MPPI_HIP_E_INVALID, std::string(what) + ": out must hold the agent's sample count.");
    const bool weights = std::strcmp(what, "read_weights") == 0;
    if (weights && ctx->weights_pending) {
        // A one-pass step: form the weights of its iteration now.
        mppi_step_args wa = ctx->weights_args;
        void* params[] = {&wa};
        MPPI_HIP_TRY(ctx, hipModuleLaunchKernel(ctx->mod->weights, ctx->weights_grid_x, ctx->weights_grid_y, 1,
                                                kBlock, 1, 1, 0, ctx->stream, params, nullptr));
        ctx->weights_pending = false;
    }
    const uint64_t off = weights ? d.w_off : d.cost_off;
    MPPI_HIP_TRY(ctx, hipMemcpyAsync(out, base + off, sizeof(float) * count,
                                     hipMemcpyDeviceToHost, ctx->stream));
    MPPI_HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    return MPPI_HIP_OK;
}
}  // namespace

int mppi_hip_read_costs(mppi_hip_ctx* ctx, uint32_t agent, float* out, size_t count) {
    return read_per_sample(ctx, ctx ? ctx->last_costs : nullptr, agent, out, count, "read_costs");
}

int mppi_hip_read_weights(mppi_hip_ctx* ctx, uint32_t agent, float* out, size_t count) {
    return read_per_sample(ctx, ctx ? ctx->last_weights : nullptr, agent, out, count, "read_weights");
}

int mppi_hip_set_step_path(mppi_hip_ctx* ctx, uint32_t path) {
    if (!ctx) return fail(nullptr, MPPI_HIP_E_INVALID, "ctx is null.");
    if (path > MPPI_HIP_PATH_FUSED) return fail(ctx, MPPI_HIP_E_INVALID, "step path must be AUTO, TWO_KERNEL or FUSED.");
    ctx->step_path = path;
    return MPPI_HIP_OK;
}

int mppi_hip_set_reduce_samples_per_thread(mppi_hip_ctx* ctx, uint32_t spt) {
    if (!ctx) return fail(nullptr, MPPI_HIP_E_INVALID, "context is null.");
    if (spt > 1024) return fail(ctx, MPPI_HIP_E_INVALID, "samples per thread must be <= 1024.");
    ctx->spt_override = spt;
    return MPPI_HIP_OK;
}

int mppi_hip_set_wait_limit(mppi_hip_ctx* ctx, uint32_t polls, int force_timeout) {
    if (!ctx) return fail(nullptr, MPPI_HIP_E_INVALID, "context is null.");
    ctx->wait_limit = polls;
    ctx->wait_bias = force_timeout ? 0x40000000u : 0u;
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
    if (ctx->stream) MPPI_HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    collect_profile(ctx);
    if (rollout_ms) *rollout_ms = ctx->prof_rollout_ms;
    if (reduce_ms) *reduce_ms = ctx->prof_reduce_ms;
    if (launches) *launches = ctx->prof_launches;
    ctx->prof_rollout_ms = ctx->prof_reduce_ms = 0.0;
    ctx->prof_launches = 0;
    return MPPI_HIP_OK;
}

}  // extern "C"
