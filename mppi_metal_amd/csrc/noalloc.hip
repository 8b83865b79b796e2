// noalloc.hip — device side of mppi_hip_step_noalloc (include/mppi_hip.h):
// the file-scope scratch array that holds a call's work buffers, its
// per-device address, and the kernel that initialises a call's descriptor
// and agent state in it.
//
// `scratch` is one array per device (each GPU the library runs on holds its
// own copy until the process ends).  Host code never uses the array's name
// as an address: its device address on each device is taken once, with that
// device current, by hipGetSymbolAddress when the first context is created
// on the device, and every later use (kernel arguments) goes through that
// pointer.  Every call on one device shares that device's array and finds
// the previous call's contents in it: the init kernel rewrites the
// descriptor, the agent state (step counter, arrival counters, status) and
// the status word before the step's kernels run, and every other region is
// written before it is read within the step.  Calls that overlap on one
// device race for the array.
#include "backend.h"

#include <hip/hip_runtime.h>

#include <map>
#include <mutex>

static __device__ __align__(256) unsigned char scratch[mppi_hip_impl::kNoallocScratchBytes];

namespace {

std::mutex g_scratch_mutex;
std::map<int, void*> g_scratch_ptr;       // device -> device address of `scratch`
std::map<int, std::string> g_scratch_err; // device -> why the address is missing

// One thread writes the call's descriptor, a fresh agent state at step
// `step` and a cleared status word.
__global__ void mppi_noalloc_init(mppi_agent_desc d, uint32_t step, mppi_agent_desc* desc_out,
                                  mppi_agent_state* agent_out, uint32_t* status_out) {
    if (threadIdx.x != 0) return;
    *desc_out = d;
    mppi_agent_state s{};
    s.step = step;
    *agent_out = s;
    *status_out = 0u;
}

}  // namespace

namespace mppi_hip_impl {

void noalloc_setup(int device) {
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if (g_scratch_ptr.count(device)) return;
    void* ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr) {
        g_scratch_err[device] = std::string("hipGetSymbolAddress(scratch) failed: ") + hipGetErrorString(e);
        return;
    }
    g_scratch_ptr[device] = ptr;
}

void* noalloc_scratch(int device, std::string* why) {
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    auto it = g_scratch_ptr.find(device);
    if (it != g_scratch_ptr.end()) return it->second;
    auto er = g_scratch_err.find(device);
    *why = er != g_scratch_err.end() ? er->second : "the scratch address of this device was never taken";
    return nullptr;
}

hipError_t noalloc_init(hipStream_t stream, const mppi_agent_desc& d, uint32_t step, mppi_agent_desc* desc_out,
                        mppi_agent_state* agent_out, uint32_t* status_out) {
    hipLaunchKernelGGL(mppi_noalloc_init, dim3(1), dim3(64), 0, stream, d, step, desc_out, agent_out, status_out);
    return hipGetLastError();
}

}  // namespace mppi_hip_impl
