// surrogate.hip — the learned-dynamics surrogate: its plugin-independent
// training and evaluation kernels (compiled ahead of time for gfx950 into
// libmppi_hip.so) and the mppi_hip_surrogate_* entry points of the C ABI.
// The network and its matrix-core layout are described in mppi_mlp.h; the
// plugin-dependent kernels (training data from the plugin's dynamics, the
// rollout with the network as dynamics) are in mppi_kernels.hip.
//
// One optimiser step over a batch of n samples (n a multiple of 64) is four
// launches:
//   mppi_mlp_train_fwd_bwd  per 16-sample block: forward pass (FP16 MFMA),
//                           MSE error against the normalised targets,
//                           backward pass through the layers (BF16 MFMA with
//                           the transposed weight fragments), writing the
//                           inputs, activations and pre-activation gradients
//                           as bf16, feature-major ([feature][sample]), plus
//                           per-wave loss partials.
//   mppi_mlp_weight_grad    weight and bias gradients as MFMA GEMMs that
//                           contract over samples (A = dZ^T, B = activations;
//                           biases with B = ones), split over sample chunks;
//                           FP32 partials in parameter order.
//   mppi_mlp_adam           sums the chunk partials in a fixed order and
//                           applies Adam to the FP32 master parameters.
//   mppi_mlp_pack           rebuilds the A fragments (FP16 forward, BF16
//                           backward) from the masters.
#include <hip/hip_runtime.h>

#include "backend.h"
#include "mppi_mlp.h"

#include <cmath>
#include <cstring>
#include <random>
#include <vector>

namespace {

constexpr int HID = MPPI_MLP_HID;

struct train_args {
    const short* frags;
    const float* params;      // FP32 masters
    mppi_mlp_dims dims;
    const float* x;           // [n][in] normalised inputs
    const float* t;           // [n][out] raw targets (state deltas)
    const float* out_mu;      // [out]
    const float* out_inv_sd;  // [out]
    short* xt;                // [16][n]
    short* h1t;               // [128][n]
    short* h2t;               // [128][n]
    short* dz1t;              // [128][n]
    short* dz2t;              // [128][n]
    short* dyt;               // [16][n]
    float* loss_part;         // [n / 16]
    uint32_t n;
    float grad_scale;         // 2 / (n_total * out)
};

struct grad_args {
    const short* xt;
    const short* h1t;
    const short* h2t;
    const short* dz1t;
    const short* dz2t;
    const short* dyt;
    float* partial;           // [ksplit][n_params]
    mppi_mlp_dims dims;
    uint32_t n;
    uint32_t kchunk;          // samples per split (multiple of 16)
};

struct adam_args {
    float* params;
    float* m;
    float* v;
    const float* partial;
    mppi_mlp_dims dims;
    uint32_t ksplit;
    float lr_t;               // lr * sqrt(1 - beta2^t) / (1 - beta1^t)
    float beta1, beta2, eps;
    float* grad_out;          // summed gradient (optional, for checks)
    uint32_t apply;           // 0: only sum into grad_out
    const float* loss_part;   // per-block squared errors of the step
    uint32_t n_parts;
    float loss_scale;         // 1 / (n * out)
    float* loss_out;          // batch MSE of the step (optional)
};

struct pack_args {
    const float* params;
    mppi_mlp_dims dims;
    short* frags;
};

struct forward_args {
    const short* frags;
    const float* params;
    mppi_mlp_dims dims;
    const float* x;           // [n][in] normalised
    float* y;                 // [n][out] normalised outputs (with bias)
    uint32_t n;
};

}  // namespace

extern "C" __global__ void __launch_bounds__(256) mppi_mlp_train_fwd_bwd(train_args a) {
    const uint32_t lane = threadIdx.x & 63, g = lane >> 4, c = lane & 15;
    const uint32_t s0 = (blockIdx.x * 4 + (threadIdx.x >> 6)) * 16;
    if (s0 >= a.n) return;
    const uint32_t n = a.n, s = s0 + c;
    const mppi_mlp_dims& d = a.dims;
    const mppi_v4f zero = {0.0f, 0.0f, 0.0f, 0.0f};

    mppi_v4s xb;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        const uint32_t k = 4 * g + j;
        const float v = k < d.in ? a.x[size_t(s) * d.in + k] : 0.0f;
        xb[j] = mppi_f2h(v);
        a.xt[size_t(k) * n + s] = mppi_f2bf(v);
    }
    mppi_v4s h1b[8], h2b[8], y1b[8], y2b[8];
    mppi_v4f y = mppi_mlp_block(a.frags, a.params + d.b1, a.params + d.b2, xb, lane, h1b, h2b, y1b, y2b);
#pragma unroll
    for (int mb = 0; mb < 8; ++mb)
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const size_t row = 16 * mb + 4 * g + j;
            a.h1t[row * n + s] = mppi_f2bf(mppi_h2f(h1b[mb][j]));
            a.h2t[row * n + s] = mppi_f2bf(mppi_h2f(h2b[mb][j]));
        }
    // MSE against normalised targets.
    float loss = 0.0f;
    mppi_v4s dyb;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        const uint32_t r = 4 * g + j;
        float dy = 0.0f;
        if (r < d.out) {
            const float yv = y[j] + a.params[d.b3 + r];
            const float tn = (a.t[size_t(s) * d.out + r] - a.out_mu[r]) * a.out_inv_sd[r];
            const float e = yv - tn;
            loss += e * e;
            dy = a.grad_scale * e;
        }
        dyb[j] = mppi_f2bf(dy);
        a.dyt[size_t(r) * n + s] = dyb[j];
    }
    // dH2 = W3^T dY, dZ2 = dH2 T'(z2), T' from the stored y (mppi_act_deriv)
    mppi_v4s dz2b[8];
#pragma unroll
    for (int mb = 0; mb < 8; ++mb) {
        mppi_v4f acc = mppi_mfma_bf16(mppi_frag(a.frags, MPPI_MLP_FRAG_W3T + mb, lane), dyb, zero);
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const float yv = mppi_h2f(y2b[mb][j]);
            dz2b[mb][j] = mppi_f2bf(acc[j] * mppi_act_deriv(yv));
            a.dz2t[size_t(16 * mb + 4 * g + j) * n + s] = dz2b[mb][j];
        }
    }
    // dH1 = W2^T dZ2, dZ1 = dH1 T'(z1)
#pragma unroll
    for (int mb = 0; mb < 8; ++mb) {
        mppi_v4f acc = zero;
#pragma unroll
        for (int kb = 0; kb < 8; ++kb)
            acc = mppi_mfma_bf16(mppi_frag(a.frags, MPPI_MLP_FRAG_W2T + mb * 8 + kb, lane), dz2b[kb], acc);
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const float yv = mppi_h2f(y1b[mb][j]);
            a.dz1t[size_t(16 * mb + 4 * g + j) * n + s] = mppi_f2bf(acc[j] * mppi_act_deriv(yv));
        }
    }
    // loss partial of this 16-sample block
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) loss += __shfl_xor(loss, o, 64);
    if (lane == 0) a.loss_part[s0 / 16] = loss;
}

// Tiles: 0..7 dW1 (out block), 8..71 dW2 (ob * 8 + ib), 72..79 dW3 (in
// block), 80..87 db1, 88..95 db2, 96 db3.  Grid (97, ksplit), one wave each.
extern "C" __global__ void __launch_bounds__(64) mppi_mlp_weight_grad(grad_args a) {
    const uint32_t lane = threadIdx.x, g = lane >> 4, c = lane & 15;
    const uint32_t tile = blockIdx.x, ks = blockIdx.y;
    const mppi_mlp_dims& d = a.dims;
    const short* at;
    const short* bt = nullptr;   // null: B = ones (bias gradient)
    uint32_t ob = 0, ib = 0, kind;
    if (tile < 8) { kind = 0; ob = tile; at = a.dz1t; bt = a.xt; }
    else if (tile < 72) { kind = 1; ob = (tile - 8) / 8; ib = (tile - 8) % 8; at = a.dz2t; bt = a.h1t; }
    else if (tile < 80) { kind = 2; ib = tile - 72; at = a.dyt; bt = a.h2t; }
    else if (tile < 88) { kind = 3; ob = tile - 80; at = a.dz1t; }
    else if (tile < 96) { kind = 4; ob = tile - 88; at = a.dz2t; }
    else { kind = 5; at = a.dyt; }
    const uint32_t k_begin = ks * a.kchunk;
    const uint32_t k_end = k_begin + a.kchunk < a.n ? k_begin + a.kchunk : a.n;
    mppi_v4f acc = {0.0f, 0.0f, 0.0f, 0.0f};
    mppi_v4s ones;
#pragma unroll
    for (int j = 0; j < 4; ++j) ones[j] = short(0x3f80);
    const short* arow = at + size_t(16 * ob + c) * a.n;
    const short* brow = bt ? bt + size_t(16 * ib + c) * a.n : nullptr;
    for (uint32_t k0 = k_begin; k0 < k_end; k0 += 16) {
        const mppi_v4s av = *reinterpret_cast<const mppi_v4s*>(arow + k0 + 4 * g);
        const mppi_v4s bv = brow ? *reinterpret_cast<const mppi_v4s*>(brow + k0 + 4 * g) : ones;
        acc = mppi_mfma_bf16(av, bv, acc);
    }
    float* out = a.partial + size_t(ks) * d.n_params;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        const uint32_t r = 4 * g + j;   // row of the output tile
        switch (kind) {
            case 0:
                if (c < d.in) out[d.w1 + (16 * ob + r) * d.in + c] = acc[j];
                break;
            case 1:
                out[d.w2 + (16 * ob + r) * HID + 16 * ib + c] = acc[j];
                break;
            case 2:
                if (r < d.out) out[d.w3 + r * HID + 16 * ib + c] = acc[j];
                break;
            case 3:
                if (c == 0) out[d.b1 + 16 * ob + r] = acc[j];
                break;
            case 4:
                if (c == 0) out[d.b2 + 16 * ob + r] = acc[j];
                break;
            default:
                if (c == 0 && r < d.out) out[d.b3 + r] = acc[j];
                break;
        }
    }
}

extern "C" __global__ void __launch_bounds__(256) mppi_mlp_adam(adam_args a) {
    if (blockIdx.x == 0 && a.loss_out) {
        __shared__ float s_l[4];
        float l = 0.0f;
        for (uint32_t i = threadIdx.x; i < a.n_parts; i += 256) l += a.loss_part[i];
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) l += __shfl_xor(l, o, 64);
        if ((threadIdx.x & 63) == 0) s_l[threadIdx.x >> 6] = l;
        __syncthreads();
        if (threadIdx.x == 0) *a.loss_out = ((s_l[0] + s_l[1]) + (s_l[2] + s_l[3])) * a.loss_scale;
    }
    const uint32_t p = blockIdx.x * 256 + threadIdx.x;
    if (p >= a.dims.n_params) return;
    float gsum = 0.0f;
    for (uint32_t k = 0; k < a.ksplit; ++k) gsum += a.partial[size_t(k) * a.dims.n_params + p];
    if (a.grad_out) a.grad_out[p] = gsum;
    if (!a.apply) return;
    const float m = a.beta1 * a.m[p] + (1.0f - a.beta1) * gsum;
    const float v = a.beta2 * a.v[p] + (1.0f - a.beta2) * gsum * gsum;
    a.m[p] = m;
    a.v[p] = v;
    a.params[p] -= a.lr_t * m / (__builtin_sqrtf(v) + a.eps);
}

// Column order of K-block kb of the 32x32 fragments: the order in which a
// 32x32 result tile's registers hold its rows (mppi_mlp.h).
__device__ inline uint32_t mlp_y_unit(uint32_t kb, uint32_t p) {
    return 16 * kb + 8 * ((p % 8) / 4) + 4 * (p / 8) + p % 4;
}

// One thread per fragment element (152 x16 fragments of 256, then the 32x32
// set: 8 layer-1 fragments of 256, 32 + 8 fragments of 512).
extern "C" __global__ void __launch_bounds__(256) mppi_mlp_pack(pack_args a) {
    const uint32_t e = blockIdx.x * 256 + threadIdx.x;
    if (e >= MPPI_MLP_TOTAL_ELEMS) return;
    const mppi_mlp_dims& d = a.dims;
    const float* P = a.params;
    if (e >= MPPI_MLP_Y_BASE) {
        const uint32_t q = e - MPPI_MLP_Y_BASE;
        float v = 0.0f;
        if (q < MPPI_MLP_Y_W2) {                      // W1Y, kb1 * 4 + mb
            const uint32_t f = q / 256, lane = (q / 4) % 64, j = q % 4;
            const uint32_t i = lane % 32, k = 8 * (f / 4) + 4 * (lane / 32) + j;
            if (k < d.in) v = P[d.w1 + (32 * (f % 4) + i) * d.in + k];
        } else {
            const uint32_t r = q < MPPI_MLP_Y_W3 ? q - MPPI_MLP_Y_W2 : q - MPPI_MLP_Y_W3;
            const uint32_t f = r / 512, lane = (r / 8) % 64, j = r % 8;
            const uint32_t i = lane % 32, p = 8 * (lane / 32) + j;
            if (q < MPPI_MLP_Y_W3) {
                v = P[d.w2 + (32 * (f / 8) + i) * HID + mlp_y_unit(f % 8, p)];                     // W2Y, mb * 8 + kb
            } else if (d.out <= 8) {                                                               // W3Z, kb
                const uint32_t m = lane % 16, g = lane / 16;
                if (m / 8 == g % 2 && m % 8 < d.out) v = P[d.w3 + (m % 8) * HID + mlp_y_unit(f, 8 * (g / 2) + j)];
            } else if (i < d.out) {
                v = P[d.w3 + i * HID + mlp_y_unit(f, p)];                                          // W3Y, kb
            }
        }
        a.frags[e] = mppi_f2h(v);
        return;
    }
    const uint32_t f = e / MPPI_MLP_FRAG_ELEMS, lane = (e / 4) % 64, j = e % 4;
    const uint32_t i = lane % 16, k = 4 * (lane / 16) + j;   // A[i][k] of the fragment
    float v = 0.0f;
    if (f < MPPI_MLP_FRAG_W2) {                       // W1, mb = f
        if (k < d.in) v = P[d.w1 + (16 * f + i) * d.in + k];
    } else if (f < MPPI_MLP_FRAG_W3) {                // W2
        const uint32_t q = f - MPPI_MLP_FRAG_W2, mb = q / 8, kb = q % 8;
        v = P[d.w2 + (16 * mb + i) * HID + 16 * kb + k];
    } else if (f < MPPI_MLP_FRAG_W2T) {               // W3, kb
        const uint32_t kb = f - MPPI_MLP_FRAG_W3;
        if (i < d.out) v = P[d.w3 + i * HID + 16 * kb + k];
    } else if (f < MPPI_MLP_FRAG_W3T) {               // W2^T
        const uint32_t q = f - MPPI_MLP_FRAG_W2T, mb = q / 8, kb = q % 8;
        v = P[d.w2 + (16 * kb + k) * HID + 16 * mb + i];
    } else {                                          // W3^T, mb
        const uint32_t mb = f - MPPI_MLP_FRAG_W3T;
        if (k < d.out) v = P[d.w3 + k * HID + 16 * mb + i];
    }
    a.frags[e] = f < MPPI_MLP_FRAG_W2T ? mppi_f2h(v) : mppi_f2bf(v);
}

// Inference over n samples (n a multiple of 16): y = MLP(x) + b3.
extern "C" __global__ void __launch_bounds__(256) mppi_mlp_forward(forward_args a) {
    const uint32_t lane = threadIdx.x & 63, g = lane >> 4, c = lane & 15;
    const uint32_t s0 = (blockIdx.x * 4 + (threadIdx.x >> 6)) * 16;
    if (s0 >= a.n) return;
    const uint32_t s = s0 + c;
    const mppi_mlp_dims& d = a.dims;
    mppi_v4s xb;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        const uint32_t k = 4 * g + j;
        xb[j] = mppi_f2h(k < d.in ? a.x[size_t(s) * d.in + k] : 0.0f);
    }
    mppi_v4f y = mppi_mlp_block(a.frags, a.params + d.b1, a.params + d.b2, xb, lane, nullptr, nullptr, nullptr,
                                nullptr);
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        const uint32_t r = 4 * g + j;
        if (r < d.out) a.y[size_t(s) * d.out + r] = y[j] + a.params[d.b3 + r];
    }
}

// The hidden activation T (mppi_mlp.h) of n pre-activations, as binary16
// bits (one thread per pair; n even).
extern "C" __global__ void __launch_bounds__(256) mppi_mlp_activation(const float* z, uint32_t* out, uint32_t pairs) {
    const uint32_t p = blockIdx.x * 256 + threadIdx.x;
    if (p < pairs) out[p] = mppi_act2(z[2 * p], z[2 * p + 1]);
}

// ===========================================================================
// Host side
// ===========================================================================
namespace mppi_hip_impl {

struct SurrogateState {
    std::shared_ptr<LoadedModule> mod;   // JIT module with the surrogate kernels
    mppi_mlp_dims dims{};
    mppi_mlp_norm norm{};
    float box_lo[16] = {}, box_hi[16] = {};   // sampling box of the training inputs
    bool ready = false, enabled = false;
    uint32_t adam_t = 0;
    DeviceBuffer params, m, v, frags, partial, grad_sum, loss_part, loss_hist, out_norm;
    DeviceBuffer xt, h1t, h2t, dz1t, dz2t, dyt;
    DeviceBuffer data_x, data_t, io_x, io_y;
    ~SurrogateState() {
        for (DeviceBuffer* b : {&params, &m, &v, &frags, &partial, &grad_sum, &loss_part, &loss_hist,
                                &out_norm, &xt, &h1t, &h2t, &dz1t, &dz2t, &dyt, &data_x,
                                &data_t, &io_x, &io_y})
            b->release();
    }
};

namespace {

constexpr uint32_t kSplit = 8;   // sample chunks of the weight-gradient GEMMs
// Batches are whole 16-sample blocks in each of the kSplit chunks.
constexpr uint32_t kBatchQuantum = 16 * kSplit;

mppi_mlp_dims dims_for(uint32_t in, uint32_t out) {
    mppi_mlp_dims d{};
    d.in = in;
    d.out = out;
    d.w1 = 0;
    d.b1 = d.w1 + HID * in;
    d.w2 = d.b1 + HID;
    d.b2 = d.w2 + HID * HID;
    d.w3 = d.b2 + HID;
    d.b3 = d.w3 + out * HID;
    d.n_params = d.b3 + out;
    return d;
}

int ensure_state(mppi_hip_ctx* ctx) {
    int rc = check_ctx(ctx);
    if (rc) return rc;
    if (ctx->D + ctx->C > 16)
        return fail(ctx, MPPI_HIP_E_INVALID, "surrogate: state_dim + control_dim must be <= 16.");
    if (!ctx->sur) ctx->sur = std::make_shared<SurrogateState>();
    SurrogateState& s = *ctx->sur;
    if (!s.mod && (rc = load_module(ctx, false, true, &s.mod))) return rc;
    s.dims = dims_for(ctx->D + ctx->C, ctx->D);
    const size_t P = s.dims.n_params;
    if ((rc = ensure_device_buffer(ctx, &s.params, sizeof(float) * P))) return rc;
    if ((rc = ensure_device_buffer(ctx, &s.m, sizeof(float) * P))) return rc;
    if ((rc = ensure_device_buffer(ctx, &s.v, sizeof(float) * P))) return rc;
    if ((rc = ensure_device_buffer(ctx, &s.grad_sum, sizeof(float) * P))) return rc;
    if ((rc = ensure_device_buffer(ctx, &s.partial, sizeof(float) * P * kSplit))) return rc;
    if ((rc = ensure_device_buffer(ctx, &s.frags, sizeof(short) * MPPI_MLP_TOTAL_ELEMS))) return rc;
    if ((rc = ensure_device_buffer(ctx, &s.out_norm, sizeof(float) * 32))) return rc;
    return MPPI_HIP_OK;
}

int ensure_activations(mppi_hip_ctx* ctx, uint32_t n) {
    SurrogateState& s = *ctx->sur;
    int rc;
    for (DeviceBuffer* b : {&s.h1t, &s.h2t, &s.dz1t, &s.dz2t})
        if ((rc = ensure_device_buffer(ctx, b, sizeof(short) * HID * n))) return rc;
    for (DeviceBuffer* b : {&s.xt, &s.dyt})
        if ((rc = ensure_device_buffer(ctx, b, sizeof(short) * 16 * n))) return rc;
    if ((rc = ensure_device_buffer(ctx, &s.loss_part, sizeof(float) * (n / 16)))) return rc;
    return MPPI_HIP_OK;
}

// 1 / period in FP32 (0 for an input that is not periodic).
float inv_period(float period) {
    return period > 0.0f ? 1.0f / period : 0.0f;
}

// Host form of mppi_mlp_input (same operations, FP32).
float normalise_input(const mppi_mlp_norm& n, uint32_t k, float v) {
    float d = v - n.in_center[k];
    if (n.in_period[k] > 0.0f) d -= n.in_period[k] * std::nearbyint(d * inv_period(n.in_period[k]));
    return d * n.in_inv_half[k];
}

// Upload the output normalisation (mean, 1/sd) used by the training kernel.
int upload_out_norm(mppi_hip_ctx* ctx) {
    SurrogateState& s = *ctx->sur;
    float host[32] = {0};
    for (uint32_t i = 0; i < s.dims.out; ++i) {
        host[i] = s.norm.out_mu[i];
        host[16 + i] = 1.0f / s.norm.out_sd[i];
    }
    MPPI_HIP_TRY(ctx, hipMemcpyAsync(s.out_norm.ptr, host, sizeof(host), hipMemcpyHostToDevice, ctx->stream));
    return MPPI_HIP_OK;
}

int pack(mppi_hip_ctx* ctx) {
    SurrogateState& s = *ctx->sur;
    pack_args pa{static_cast<const float*>(s.params.ptr), s.dims, static_cast<short*>(s.frags.ptr)};
    hipLaunchKernelGGL(mppi_mlp_pack, dim3(MPPI_MLP_TOTAL_ELEMS / 256), dim3(256), 0, ctx->stream, pa);
    MPPI_HIP_TRY(ctx, hipGetLastError());
    return MPPI_HIP_OK;
}

// One forward/backward pass over n samples at x / t, then either an Adam
// step (apply) or only the summed gradient in grad_sum.
int train_step(mppi_hip_ctx* ctx, const float* x, const float* t, uint32_t n, bool apply, float lr,
               float* loss_out) {
    SurrogateState& s = *ctx->sur;
    const float* onorm = static_cast<const float*>(s.out_norm.ptr);
    train_args ta{};
    ta.frags = static_cast<const short*>(s.frags.ptr);
    ta.params = static_cast<const float*>(s.params.ptr);
    ta.dims = s.dims;
    ta.x = x;
    ta.t = t;
    ta.out_mu = onorm;
    ta.out_inv_sd = onorm + 16;
    ta.xt = static_cast<short*>(s.xt.ptr);
    ta.h1t = static_cast<short*>(s.h1t.ptr);
    ta.h2t = static_cast<short*>(s.h2t.ptr);
    ta.dz1t = static_cast<short*>(s.dz1t.ptr);
    ta.dz2t = static_cast<short*>(s.dz2t.ptr);
    ta.dyt = static_cast<short*>(s.dyt.ptr);
    ta.loss_part = static_cast<float*>(s.loss_part.ptr);
    ta.n = n;
    ta.grad_scale = 2.0f / (float(n) * float(s.dims.out));
    hipLaunchKernelGGL(mppi_mlp_train_fwd_bwd, dim3(n / 64), dim3(256), 0, ctx->stream, ta);
    MPPI_HIP_TRY(ctx, hipGetLastError());
    grad_args ga{};
    ga.xt = ta.xt;
    ga.h1t = ta.h1t;
    ga.h2t = ta.h2t;
    ga.dz1t = ta.dz1t;
    ga.dz2t = ta.dz2t;
    ga.dyt = ta.dyt;
    ga.partial = static_cast<float*>(s.partial.ptr);
    ga.dims = s.dims;
    ga.n = n;
    ga.kchunk = n / kSplit;
    hipLaunchKernelGGL(mppi_mlp_weight_grad, dim3(97, kSplit), dim3(64), 0, ctx->stream, ga);
    MPPI_HIP_TRY(ctx, hipGetLastError());
    adam_args aa{};
    aa.params = static_cast<float*>(s.params.ptr);
    aa.m = static_cast<float*>(s.m.ptr);
    aa.v = static_cast<float*>(s.v.ptr);
    aa.partial = ga.partial;
    aa.dims = s.dims;
    aa.ksplit = kSplit;
    aa.beta1 = 0.9f;
    aa.beta2 = 0.999f;
    aa.eps = 1e-8f;
    aa.apply = apply ? 1u : 0u;
    aa.grad_out = apply ? nullptr : static_cast<float*>(s.grad_sum.ptr);
    if (apply) {
        s.adam_t += 1;
        const double b1t = std::pow(0.9, double(s.adam_t)), b2t = std::pow(0.999, double(s.adam_t));
        aa.lr_t = float(lr * std::sqrt(1.0 - b2t) / (1.0 - b1t));
    }
    aa.loss_part = ta.loss_part;
    aa.n_parts = n / 16;
    aa.loss_scale = 1.0f / (float(n) * float(s.dims.out));
    aa.loss_out = loss_out;
    hipLaunchKernelGGL(mppi_mlp_adam, dim3((s.dims.n_params + 255) / 256), dim3(256), 0, ctx->stream, aa);
    MPPI_HIP_TRY(ctx, hipGetLastError());
    if (apply) return pack(ctx);
    return MPPI_HIP_OK;
}

// Generate M transitions with the plugin dynamics into data_x / data_t.
int generate(mppi_hip_ctx* ctx, uint32_t M, uint32_t stream, uint64_t seed, const float* lo,
             const float* hi, DeviceBuffer* xb, DeviceBuffer* tb) {
    SurrogateState& s = *ctx->sur;
    const uint32_t in = s.dims.in, out = s.dims.out;
    int rc;
    if ((rc = ensure_device_buffer(ctx, xb, sizeof(float) * size_t(M) * in))) return rc;
    if ((rc = ensure_device_buffer(ctx, tb, sizeof(float) * size_t(M) * out))) return rc;
    mppi_surrogate_data_args da{};
    da.x = static_cast<float*>(xb->ptr);
    da.t = static_cast<float*>(tb->ptr);
    da.model = ctx->model_is_bound ? static_cast<const uint8_t*>(ctx->model_bound.ptr)
                                   : reinterpret_cast<const uint8_t*>(static_cast<const float*>(ctx->consts.ptr) +
                                                                      ctx->noise_scale.size() + 2 * ctx->C);
    da.M = M;
    da.key0 = uint32_t(seed);
    da.key1 = uint32_t(seed >> 32);
    da.stream = stream;
    for (uint32_t k = 0; k < in; ++k) {
        da.lo[k] = lo[k];
        da.hi[k] = hi[k];
        da.center[k] = s.norm.in_center[k];
        da.inv_half[k] = s.norm.in_inv_half[k];
        da.period[k] = s.norm.in_period[k];
        da.inv_period[k] = inv_period(s.norm.in_period[k]);
    }
    void* params[] = {&da};
    MPPI_HIP_TRY(ctx, hipModuleLaunchKernel(s.mod->surrogate_data, (M + 255) / 256, 1, 1, 256, 1, 1, 0,
                                            ctx->stream, params, nullptr));
    return MPPI_HIP_OK;
}

void sampling_box(mppi_hip_ctx* ctx, const float* state_lo, const float* state_hi, float* lo, float* hi) {
    for (uint32_t k = 0; k < ctx->D; ++k) {
        lo[k] = state_lo[k];
        hi[k] = state_hi[k];
    }
    for (uint32_t k = 0; k < ctx->C; ++k) {
        lo[ctx->D + k] = ctx->u_min[k];
        hi[ctx->D + k] = ctx->u_max[k];
    }
}

}  // namespace

bool surrogate_rollout(mppi_hip_ctx* ctx, const mppi_step_args& a, hipFunction_t* fn, void* storage,
                       size_t bytes) {
    if (!ctx->sur || !ctx->sur->enabled || !ctx->sur->ready || bytes < sizeof(mppi_surrogate_rollout_args))
        return false;
    SurrogateState& s = *ctx->sur;
    mppi_surrogate_rollout_args ra{};
    ra.s = a;
    ra.frags = static_cast<const short*>(s.frags.ptr);
    ra.params = static_cast<const float*>(s.params.ptr);
    ra.b1 = s.dims.b1;
    ra.b2 = s.dims.b2;
    ra.b3 = s.dims.b3;
    for (int k = 0; k < 16; ++k) {
        ra.in_center[k] = s.norm.in_center[k];
        ra.in_inv_half[k] = s.norm.in_inv_half[k];
        ra.in_period[k] = s.norm.in_period[k];
        ra.in_inv_period[k] = inv_period(s.norm.in_period[k]);
        ra.out_mu[k] = s.norm.out_mu[k];
        ra.out_sd[k] = s.norm.out_sd[k];
    }
    std::memcpy(storage, &ra, sizeof(ra));
    *fn = s.mod->rollout_surrogate;
    return true;
}

}  // namespace mppi_hip_impl

using namespace mppi_hip_impl;

extern "C" {

int mppi_hip_surrogate_train(mppi_hip_ctx* ctx, const mppi_hip_surrogate_config* cfg, float* loss_history) {
    if (!cfg) return fail(ctx, MPPI_HIP_E_INVALID, "surrogate config is null.");
    int rc = ensure_state(ctx);
    if (rc) return rc;
    if (!cfg->state_lo || !cfg->state_hi) return fail(ctx, MPPI_HIP_E_INVALID, "surrogate: state_lo / state_hi are null.");
    if (cfg->batch == 0 || cfg->batch % kBatchQuantum)
        return fail(ctx, MPPI_HIP_E_INVALID, "surrogate: batch must be a positive multiple of 128.");
    if (cfg->steps == 0) return fail(ctx, MPPI_HIP_E_INVALID, "surrogate: steps must be > 0.");
    if (!(cfg->learning_rate > 0.0f)) return fail(ctx, MPPI_HIP_E_INVALID, "surrogate: learning_rate must be > 0.");
    SurrogateState& s = *ctx->sur;
    const uint32_t in = s.dims.in, out = s.dims.out, B = cfg->batch;
    const uint32_t n_batches = std::max<uint32_t>(1, (cfg->num_transitions + B - 1) / B);
    const uint32_t M = n_batches * B;
    float lo[16], hi[16];
    sampling_box(ctx, cfg->state_lo, cfg->state_hi, lo, hi);
    s.norm = mppi_mlp_norm{};
    for (uint32_t k = 0; k < in; ++k) {
        if (!(hi[k] > lo[k])) return fail(ctx, MPPI_HIP_E_INVALID, "surrogate: empty sampling range.");
        const uint32_t mode = (k < ctx->D && cfg->state_input) ? cfg->state_input[k] : MPPI_HIP_SURROGATE_INPUT_BOX;
        if (mode > MPPI_HIP_SURROGATE_INPUT_IGNORED)
            return fail(ctx, MPPI_HIP_E_INVALID, "surrogate: unknown state input mode.");
        s.norm.in_center[k] = 0.5f * (lo[k] + hi[k]);
        s.norm.in_inv_half[k] = mode == MPPI_HIP_SURROGATE_INPUT_IGNORED ? 0.0f : 2.0f / (hi[k] - lo[k]);
        s.norm.in_period[k] = mode == MPPI_HIP_SURROGATE_INPUT_PERIODIC ? hi[k] - lo[k] : 0.0f;
        s.box_lo[k] = lo[k];
        s.box_hi[k] = hi[k];
    }
    if ((rc = generate(ctx, M, 0u, cfg->seed, lo, hi, &s.data_x, &s.data_t))) return rc;
    // Output normalisation from the first transitions.
    const uint32_t ns = std::min<uint32_t>(M, 65536);
    std::vector<float> t(size_t(ns) * out);
    MPPI_HIP_TRY(ctx, hipMemcpyAsync(t.data(), s.data_t.ptr, sizeof(float) * t.size(), hipMemcpyDeviceToHost, ctx->stream));
    MPPI_HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    for (uint32_t k = 0; k < out; ++k) {
        double m = 0.0, v = 0.0;
        for (uint32_t i = 0; i < ns; ++i) m += t[size_t(i) * out + k];
        m /= ns;
        for (uint32_t i = 0; i < ns; ++i) {
            const double e = t[size_t(i) * out + k] - m;
            v += e * e;
        }
        s.norm.out_mu[k] = float(m);
        s.norm.out_sd[k] = float(std::max(std::sqrt(v / ns), 1e-6));
    }
    // Initialisation: weights uniform +-sqrt(6 / (fan_in + fan_out)), those
    // of the hidden layers divided by 5.5 (about the slope of T at the centre
    // of its linear region, 2 A = 5.54); hidden biases 1/2 (that centre),
    // output biases 0.
    std::vector<float> P(s.dims.n_params, 0.0f);
    std::mt19937_64 gen(cfg->seed ^ 0x5eed5eedULL);
    auto init = [&](uint32_t off, uint32_t rows, uint32_t cols, float scale) {
        const float a = scale * std::sqrt(6.0f / float(rows + cols));
        std::uniform_real_distribution<float> U(-a, a);
        for (uint32_t i = 0; i < rows * cols; ++i) P[off + i] = U(gen);
    };
    init(s.dims.w1, HID, in, 1.0f / 5.5f);
    init(s.dims.w2, HID, HID, 1.0f / 5.5f);
    init(s.dims.w3, out, HID, 1.0f);
    for (uint32_t i = 0; i < HID; ++i) P[s.dims.b1 + i] = P[s.dims.b2 + i] = 0.5f;
    MPPI_HIP_TRY(ctx, hipMemcpyAsync(s.params.ptr, P.data(), sizeof(float) * P.size(), hipMemcpyHostToDevice, ctx->stream));
    MPPI_HIP_TRY(ctx, hipMemsetAsync(s.m.ptr, 0, sizeof(float) * P.size(), ctx->stream));
    MPPI_HIP_TRY(ctx, hipMemsetAsync(s.v.ptr, 0, sizeof(float) * P.size(), ctx->stream));
    s.adam_t = 0;
    if ((rc = upload_out_norm(ctx))) return rc;
    if ((rc = pack(ctx))) return rc;
    if ((rc = ensure_activations(ctx, B))) return rc;
    if ((rc = ensure_device_buffer(ctx, &s.loss_hist, sizeof(float) * cfg->steps))) return rc;
    const double pi = 3.14159265358979323846;
    for (uint32_t step = 0; step < cfg->steps; ++step) {
        const uint32_t b = step % n_batches;
        const float lr = float(cfg->learning_rate * (0.1 + 0.9 * 0.5 * (1.0 + std::cos(pi * step / cfg->steps))));
        rc = train_step(ctx, static_cast<const float*>(s.data_x.ptr) + size_t(b) * B * in,
                        static_cast<const float*>(s.data_t.ptr) + size_t(b) * B * out, B, true, lr,
                        static_cast<float*>(s.loss_hist.ptr) + step);
        if (rc) return rc;
    }
    if (loss_history)
        MPPI_HIP_TRY(ctx, hipMemcpyAsync(loss_history, s.loss_hist.ptr, sizeof(float) * cfg->steps,
                                         hipMemcpyDeviceToHost, ctx->stream));
    MPPI_HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    s.ready = true;
    return MPPI_HIP_OK;
}

namespace {
// Normalised network outputs (without denormalisation) for n inputs in io_x.
int forward_io(mppi_hip_ctx* ctx, uint32_t n) {
    SurrogateState& s = *ctx->sur;
    forward_args fa{static_cast<const short*>(s.frags.ptr), static_cast<const float*>(s.params.ptr), s.dims,
                    static_cast<const float*>(s.io_x.ptr), static_cast<float*>(s.io_y.ptr), n};
    hipLaunchKernelGGL(mppi_mlp_forward, dim3((n / 16 + 3) / 4), dim3(256), 0, ctx->stream, fa);
    MPPI_HIP_TRY(ctx, hipGetLastError());
    return MPPI_HIP_OK;
}
}  // namespace

int mppi_hip_surrogate_evaluate(mppi_hip_ctx* ctx, uint32_t num_samples, uint64_t seed, float* rmse,
                                float* relative_rmse) {
    int rc = check_ctx(ctx);
    if (rc) return rc;
    if (!ctx->sur || !ctx->sur->ready) return fail(ctx, MPPI_HIP_E_NOT_INITIALIZED, "surrogate is not trained.");
    SurrogateState& s = *ctx->sur;
    const uint32_t n = std::max<uint32_t>(16, (num_samples + 15) / 16 * 16);
    const uint32_t out = s.dims.out;
    DeviceBuffer tt;
    // stream 1: never used by training
    if ((rc = generate(ctx, n, 1u, seed, s.box_lo, s.box_hi, &s.io_x, &tt))) return rc;
    if ((rc = ensure_device_buffer(ctx, &s.io_y, sizeof(float) * size_t(n) * out))) { tt.release(); return rc; }
    if ((rc = forward_io(ctx, n))) { tt.release(); return rc; }
    std::vector<float> y(size_t(n) * out), t(size_t(n) * out);
    hipError_t e1 = hipMemcpyAsync(y.data(), s.io_y.ptr, sizeof(float) * y.size(), hipMemcpyDeviceToHost, ctx->stream);
    hipError_t e2 = hipMemcpyAsync(t.data(), tt.ptr, sizeof(float) * t.size(), hipMemcpyDeviceToHost, ctx->stream);
    hipError_t e3 = hipStreamSynchronize(ctx->stream);
    tt.release();
    if (e1 != hipSuccess || e2 != hipSuccess || e3 != hipSuccess)
        return fail(ctx, MPPI_HIP_E_DEVICE, "surrogate evaluate: copy failed.");
    for (uint32_t k = 0; k < out; ++k) {
        double se = 0.0, m = 0.0, v = 0.0;
        for (uint32_t i = 0; i < n; ++i) m += t[size_t(i) * out + k];
        m /= n;
        for (uint32_t i = 0; i < n; ++i) {
            const double pred = double(y[size_t(i) * out + k]) * s.norm.out_sd[k] + s.norm.out_mu[k];
            const double err = pred - t[size_t(i) * out + k];
            const double dv = t[size_t(i) * out + k] - m;
            se += err * err;
            v += dv * dv;
        }
        const double r = std::sqrt(se / n), sd = std::sqrt(v / n);
        if (rmse) rmse[k] = float(r);
        if (relative_rmse) relative_rmse[k] = float(sd > 0 ? r / sd : 0.0);
    }
    return MPPI_HIP_OK;
}

int mppi_hip_surrogate_predict(mppi_hip_ctx* ctx, uint32_t n, const float* states, const float* controls,
                               float* next_states) {
    int rc = check_ctx(ctx);
    if (rc) return rc;
    if (!ctx->sur || !ctx->sur->ready) return fail(ctx, MPPI_HIP_E_NOT_INITIALIZED, "surrogate is not trained.");
    if (!states || !controls || !next_states || n == 0) return fail(ctx, MPPI_HIP_E_INVALID, "surrogate predict: bad arguments.");
    SurrogateState& s = *ctx->sur;
    const uint32_t in = s.dims.in, out = s.dims.out, D = ctx->D, C = ctx->C;
    const uint32_t np = (n + 15) / 16 * 16;
    std::vector<float> x(size_t(np) * in, 0.0f);
    for (uint32_t i = 0; i < n; ++i) {
        for (uint32_t k = 0; k < D; ++k) x[size_t(i) * in + k] = normalise_input(s.norm, k, states[size_t(i) * D + k]);
        for (uint32_t k = 0; k < C; ++k)
            x[size_t(i) * in + D + k] = normalise_input(s.norm, D + k, controls[size_t(i) * C + k]);
    }
    if ((rc = ensure_device_buffer(ctx, &s.io_x, sizeof(float) * x.size()))) return rc;
    if ((rc = ensure_device_buffer(ctx, &s.io_y, sizeof(float) * size_t(np) * out))) return rc;
    MPPI_HIP_TRY(ctx, hipMemcpyAsync(s.io_x.ptr, x.data(), sizeof(float) * x.size(), hipMemcpyHostToDevice, ctx->stream));
    if ((rc = forward_io(ctx, np))) return rc;
    std::vector<float> y(size_t(np) * out);
    MPPI_HIP_TRY(ctx, hipMemcpyAsync(y.data(), s.io_y.ptr, sizeof(float) * y.size(), hipMemcpyDeviceToHost, ctx->stream));
    MPPI_HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    for (uint32_t i = 0; i < n; ++i)
        for (uint32_t k = 0; k < D; ++k)
            next_states[size_t(i) * D + k] =
                states[size_t(i) * D + k] + (y[size_t(i) * out + k] * s.norm.out_sd[k] + s.norm.out_mu[k]);
    return MPPI_HIP_OK;
}

int mppi_hip_surrogate_use(mppi_hip_ctx* ctx, int enable) {
    int rc = check_ctx(ctx);
    if (rc) return rc;
    if (enable && (!ctx->sur || !ctx->sur->ready))
        return fail(ctx, MPPI_HIP_E_NOT_INITIALIZED, "surrogate is not trained or imported.");
    if (ctx->sur) ctx->sur->enabled = enable != 0;
    return MPPI_HIP_OK;
}

int mppi_hip_surrogate_export(mppi_hip_ctx* ctx, float* params, size_t capacity, size_t* n_params,
                              float* norm) {
    int rc = check_ctx(ctx);
    if (rc) return rc;
    if (!ctx->sur || !ctx->sur->ready) return fail(ctx, MPPI_HIP_E_NOT_INITIALIZED, "surrogate is not trained.");
    SurrogateState& s = *ctx->sur;
    if (n_params) *n_params = s.dims.n_params;
    if (params) {
        if (capacity < s.dims.n_params) return fail(ctx, MPPI_HIP_E_INVALID, "surrogate export: capacity too small.");
        MPPI_HIP_TRY(ctx, hipMemcpyAsync(params, s.params.ptr, sizeof(float) * s.dims.n_params,
                                         hipMemcpyDeviceToHost, ctx->stream));
        MPPI_HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    }
    if (norm) std::memcpy(norm, &s.norm, sizeof(s.norm));
    return MPPI_HIP_OK;
}

int mppi_hip_surrogate_import(mppi_hip_ctx* ctx, const float* params, size_t n_params, const float* norm) {
    int rc = ensure_state(ctx);
    if (rc) return rc;
    SurrogateState& s = *ctx->sur;
    if (!params || !norm || n_params != s.dims.n_params)
        return fail(ctx, MPPI_HIP_E_INVALID, "surrogate import: expected " + std::to_string(s.dims.n_params) +
                                                 " parameters and a normalisation block.");
    std::memcpy(&s.norm, norm, sizeof(s.norm));
    for (uint32_t k = 0; k < s.dims.in; ++k) {
        const float half = s.norm.in_period[k] > 0.0f ? 0.5f * s.norm.in_period[k]
                           : s.norm.in_inv_half[k] > 0.0f ? 1.0f / s.norm.in_inv_half[k] : 0.0f;
        s.box_lo[k] = s.norm.in_center[k] - half;
        s.box_hi[k] = s.norm.in_center[k] + half;
    }
    for (uint32_t k = 0; k < s.dims.out; ++k)
        if (!(s.norm.out_sd[k] > 0.0f)) return fail(ctx, MPPI_HIP_E_INVALID, "surrogate import: output sd must be > 0.");
    MPPI_HIP_TRY(ctx, hipMemcpyAsync(s.params.ptr, params, sizeof(float) * n_params, hipMemcpyHostToDevice, ctx->stream));
    if ((rc = upload_out_norm(ctx))) return rc;
    if ((rc = pack(ctx))) return rc;
    MPPI_HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    s.ready = true;
    return MPPI_HIP_OK;
}

int mppi_hip_surrogate_gradients(mppi_hip_ctx* ctx, uint32_t n, const float* x, const float* targets,
                                 float* grads, float* loss) {
    int rc = check_ctx(ctx);
    if (rc) return rc;
    if (!ctx->sur || !ctx->sur->ready) return fail(ctx, MPPI_HIP_E_NOT_INITIALIZED, "surrogate is not trained or imported.");
    if (n == 0 || n % kBatchQuantum || !x || !targets)
        return fail(ctx, MPPI_HIP_E_INVALID, "surrogate gradients: n must be a positive multiple of 128.");
    SurrogateState& s = *ctx->sur;
    const uint32_t in = s.dims.in, out = s.dims.out;
    if ((rc = ensure_activations(ctx, n))) return rc;
    DeviceBuffer dx, dt, dl;
    auto cleanup = [&]() { dx.release(); dt.release(); dl.release(); };
    if ((rc = ensure_device_buffer(ctx, &dx, sizeof(float) * size_t(n) * in)) ||
        (rc = ensure_device_buffer(ctx, &dt, sizeof(float) * size_t(n) * out)) ||
        (rc = ensure_device_buffer(ctx, &dl, sizeof(float)))) { cleanup(); return rc; }
    hipError_t e = hipMemcpyAsync(dx.ptr, x, sizeof(float) * size_t(n) * in, hipMemcpyHostToDevice, ctx->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(dt.ptr, targets, sizeof(float) * size_t(n) * out, hipMemcpyHostToDevice, ctx->stream);
    if (e != hipSuccess) { cleanup(); return fail(ctx, MPPI_HIP_E_DEVICE, "surrogate gradients: upload failed."); }
    rc = train_step(ctx, static_cast<const float*>(dx.ptr), static_cast<const float*>(dt.ptr), n, false, 0.0f,
                    static_cast<float*>(dl.ptr));
    if (!rc) {
        if (grads) e = hipMemcpyAsync(grads, s.grad_sum.ptr, sizeof(float) * s.dims.n_params, hipMemcpyDeviceToHost, ctx->stream);
        if (e == hipSuccess && loss) e = hipMemcpyAsync(loss, dl.ptr, sizeof(float), hipMemcpyDeviceToHost, ctx->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
        if (e != hipSuccess) rc = fail(ctx, MPPI_HIP_E_DEVICE, "surrogate gradients: readback failed.");
    }
    (void)hipStreamSynchronize(ctx->stream);
    cleanup();
    return rc;
}

int mppi_hip_surrogate_activation(mppi_hip_ctx* ctx, uint32_t n, const float* z, float* out) {
    int rc = check_ctx(ctx);
    if (rc) return rc;
    if (n == 0 || !z || !out) return fail(ctx, MPPI_HIP_E_INVALID, "surrogate activation: bad arguments.");
    const uint32_t pairs = (n + 1) / 2;
    std::vector<float> zin(size_t(pairs) * 2, 0.0f);
    std::memcpy(zin.data(), z, sizeof(float) * n);
    std::vector<uint32_t> bits(pairs);
    DeviceBuffer dz, dt;
    auto cleanup = [&]() { dz.release(); dt.release(); };
    if ((rc = ensure_device_buffer(ctx, &dz, sizeof(float) * zin.size())) ||
        (rc = ensure_device_buffer(ctx, &dt, sizeof(uint32_t) * pairs))) { cleanup(); return rc; }
    hipError_t e = hipMemcpyAsync(dz.ptr, zin.data(), sizeof(float) * zin.size(), hipMemcpyHostToDevice, ctx->stream);
    if (e == hipSuccess) {
        hipLaunchKernelGGL(mppi_mlp_activation, dim3((pairs + 255) / 256), dim3(256), 0, ctx->stream,
                           static_cast<const float*>(dz.ptr), static_cast<uint32_t*>(dt.ptr), pairs);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipMemcpyAsync(bits.data(), dt.ptr, sizeof(uint32_t) * pairs, hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    (void)hipStreamSynchronize(ctx->stream);
    cleanup();
    if (e != hipSuccess) return fail(ctx, MPPI_HIP_E_DEVICE, "surrogate activation: device error.");
    for (uint32_t i = 0; i < n; ++i) {
        const uint16_t h = uint16_t(bits[i / 2] >> (16 * (i % 2)));
        out[i] = float(__builtin_bit_cast(_Float16, h));
    }
    return MPPI_HIP_OK;
}

}  // extern "C"
