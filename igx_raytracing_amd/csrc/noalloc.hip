// Fixed per-device scratch of the allocation-free closest-hit entry point
// (rt_trace_closest_noalloc).  The traversal's spill stack and its dynamic
// work counter, which rt_trace_closest keeps in per-stream buffers it
// allocates on first use, live here in a file-scope device array instead.
//
// Host code never passes `scratch` itself to a kernel: in host code the name
// is a host-side shadow, not a device address.  rt_ctx_create resolves the
// array's address on each device once (hipGetSymbolAddress, with that device
// current) and every launch uses that per-device address.
#include <hip/hip_runtime.h>

#include <cstdint>
#include <mutex>
#include <string>

#include "noalloc.hpp"

namespace rtx {
namespace {

static __device__ __align__(256) unsigned char scratch[kNoallocScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_mu;
void *g_ptr[kMaxDevices] = {};

// Zeroes the work counter on the launch stream (a kernel, not a memset, so
// a caller may capture the call in a graph and replay it).
__global__ void k_zero_words(uint32_t *p, uint32_t n) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0u;
}

}  // namespace

int noalloc_scratch_init(int device, std::string &err) {
    if (device < 0 || device >= kMaxDevices) {
        err = "device index beyond the scratch table";
        return -1;
    }
    std::lock_guard<std::mutex> g(g_mu);
    if (g_ptr[device]) return 0;
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
        err = std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e);
        return -1;
    }
    // Once per device: the kernels keep the counter at zero between
    // launches, and each call clears it again before its launch.
    hipLaunchKernelGGL(k_zero_words, dim3(1), dim3(64), 0, nullptr, static_cast<uint32_t *>(p),
                       static_cast<uint32_t>(kNoallocHeaderBytes / sizeof(uint32_t)));
    hipError_t e2 = hipGetLastError();
    if (e2 == hipSuccess) e2 = hipDeviceSynchronize();
    if (e2 != hipSuccess) {
        err = std::string("scratch clear: ") + hipGetErrorString(e2);
        return -1;
    }
    g_ptr[device] = p;
    return 0;
}

void *noalloc_scratch(int device) {
    if (device < 0 || device >= kMaxDevices) return nullptr;
    std::lock_guard<std::mutex> g(g_mu);
    return g_ptr[device];
}

int noalloc_clear_work(uint32_t *work, hipStream_t stream) {
    hipLaunchKernelGGL(k_zero_words, dim3(1), dim3(64), 0, stream, work, 2u);
    return static_cast<int>(hipGetLastError());
}

}  // namespace rtx
