// calibrate.hip — measured issue ceilings of one MI355X for the roofline
// figures: the rate at which the SIMDs retire plain (unpacked) FP32 VALU
// instructions and 32x32x16 FP16 matrix instructions (the shape of the
// surrogate rollout's products), and how far a wavefront overlaps the two
// when its matrix instructions are spread one per sixteen vector
// instructions.  Each probe kernel runs long unrolled streams of independent
// instructions from eight wavefronts per SIMD, so the figure is the
// hardware's issue rate for that kind, not a latency.  The published FP32
// vector peak counts packed FP32 instructions (two FMAs per lane per
// instruction), which this library never emits; the plain-instruction
// ceiling is the one its VALU-bound kernels are measured against.
#include "backend.h"

#include <hip/hip_runtime.h>

namespace {

constexpr int kProbeThreads = 256;
constexpr int kValuChains = 16;      // independent FMA chains per lane
constexpr int kMfmaChains = 4;       // independent 32x32 accumulators per wavefront

typedef _Float16 probe_v8h __attribute__((ext_vector_type(8)));
typedef float probe_v16f __attribute__((ext_vector_type(16)));

// Per iteration: NM matrix instructions (v_mfma_f32_32x32x16_f16) over
// kMfmaChains accumulators, each followed by NV / NM FP32 FMAs over
// kValuChains chains (NM = 0: NV FMAs).
template <int NM, int NV>
__global__ void __launch_bounds__(kProbeThreads) mppi_probe(float* out, int iters, float a, float b) {
    probe_v16f acc[kMfmaChains];
#pragma unroll
    for (int i = 0; i < kMfmaChains; ++i)
#pragma unroll
        for (int j = 0; j < 16; ++j) acc[i][j] = 0.0f;
    probe_v8h A, B;
#pragma unroll
    for (int j = 0; j < 8; ++j) {
        A[j] = _Float16(a * float(threadIdx.x & 31) + float(j));
        B[j] = _Float16(a - 0.01f * float(j));
    }
    float x[kValuChains];
#pragma unroll
    for (int i = 0; i < kValuChains; ++i) x[i] = float(threadIdx.x) * 1e-3f + float(i);
    constexpr int kPer = NM > 0 ? NV / NM : NV;
    for (int it = 0; it < iters; ++it) {
#pragma unroll
        for (int m = 0; m < (NM > 0 ? NM : 1); ++m) {
            if constexpr (NM > 0)
                acc[m % kMfmaChains] = __builtin_amdgcn_mfma_f32_32x32x16_f16(A, B, acc[m % kMfmaChains], 0, 0, 0);
#pragma unroll
            for (int v = 0; v < kPer; ++v) x[(m * kPer + v) % kValuChains] = __builtin_fmaf(x[(m * kPer + v) % kValuChains], a, b);
            if constexpr (NM > 0 && NV > 0) {
                __builtin_amdgcn_sched_group_barrier(0x008, 1, 0);     // one matrix instruction,
                __builtin_amdgcn_sched_group_barrier(0x002, kPer, 0);  // then its vector instructions
            }
        }
    }
    float s = 0.0f;
#pragma unroll
    for (int i = 0; i < kMfmaChains; ++i)
#pragma unroll
        for (int j = 0; j < 16; ++j) s += acc[i][j];
#pragma unroll
    for (int i = 0; i < kValuChains; ++i) s += x[i];
    out[size_t(blockIdx.x) * kProbeThreads + threadIdx.x] = s;
}

}  // namespace

extern "C" int mppi_hip_calibrate(int device, double* valu_lane_ops_per_s, double* mfma_f16_flops_per_s,
                                  double* mixed_over_sum) {
    if (!valu_lane_ops_per_s || !mfma_f16_flops_per_s) return MPPI_HIP_E_INVALID;
    int n_dev = 0;
    if (hipGetDeviceCount(&n_dev) != hipSuccess || device < 0 || device >= n_dev) return MPPI_HIP_E_DEVICE;
    if (hipSetDevice(device) != hipSuccess) return MPPI_HIP_E_DEVICE;
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) != hipSuccess) return MPPI_HIP_E_DEVICE;
    // one wavefront per SIMD per workgroup, eight workgroups per compute unit
    const int blocks = prop.multiProcessorCount * 8;
    float* out = nullptr;
    if (hipMalloc(&out, sizeof(float) * size_t(blocks) * kProbeThreads) != hipSuccess) return MPPI_HIP_E_NO_MEMORY;
    hipStream_t st = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    int rc = MPPI_HIP_OK;
    if (hipStreamCreate(&st) != hipSuccess || hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess)
        rc = MPPI_HIP_E_DEVICE;
    auto timed = [&](auto kernel, int iters) -> double {
        hipLaunchKernelGGL(kernel, dim3(blocks), dim3(kProbeThreads), 0, st, out, iters, 0.9999f, 1e-4f);   // warm-up
        if (hipEventRecord(e0, st) != hipSuccess) return -1.0;
        hipLaunchKernelGGL(kernel, dim3(blocks), dim3(kProbeThreads), 0, st, out, iters, 0.9999f, 1e-4f);
        if (hipEventRecord(e1, st) != hipSuccess || hipEventSynchronize(e1) != hipSuccess) return -1.0;
        float ms = 0.0f;
        if (hipEventElapsedTime(&ms, e0, e1) != hipSuccess) return -1.0;
        return double(ms) * 1e-3;
    };
    if (rc == MPPI_HIP_OK) {
        const int it = 4000;
        const double tv = timed(mppi_probe<0, 128>, it);     // VALU only
        const double tm = timed(mppi_probe<16, 0>, it);      // matrix only
        const double tm4 = timed(mppi_probe<4, 0>, it);      // the mixed probe's parts alone
        const double tv64 = timed(mppi_probe<0, 64>, it);
        const double tmv = timed(mppi_probe<4, 64>, it);     // one matrix instruction per sixteen FMAs
        if (tv <= 0.0 || tm <= 0.0 || tm4 <= 0.0 || tv64 <= 0.0 || tmv <= 0.0 || hipGetLastError() != hipSuccess) {
            rc = MPPI_HIP_E_DEVICE;
        } else {
            const double lanes = double(blocks) * kProbeThreads;
            *valu_lane_ops_per_s = lanes * double(it) * 128.0 / tv;
            const double waves = double(blocks) * (kProbeThreads / 64);
            *mfma_f16_flops_per_s = waves * double(it) * 16.0 * (2.0 * 32 * 32 * 16) / tm;
            if (mixed_over_sum) *mixed_over_sum = tmv / (tm4 + tv64);
        }
    }
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (st) (void)hipStreamDestroy(st);
    (void)hipFree(out);
    return rc;
}
