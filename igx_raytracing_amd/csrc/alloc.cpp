// Allocation wrappers with failure injection (see alloc.hpp).
#include "alloc.hpp"

#include <cstdlib>
#include <mutex>
#include <string>

namespace rtx {
namespace {

std::mutex g_mu;
std::string g_spec;
unsigned long long g_count = 0, g_fail_at = 0;

// Whether this allocation is the one RT_FAIL_DEVICE_ALLOC names.
bool injected_failure() {
    const char *env = std::getenv("RT_FAIL_DEVICE_ALLOC");
    std::lock_guard<std::mutex> g(g_mu);
    const std::string spec = env ? env : "";
    if (spec != g_spec) {
        g_spec = spec;
        g_count = 0;
        g_fail_at = spec.empty() ? 0 : std::strtoull(spec.c_str(), nullptr, 10);
    }
    return g_fail_at != 0 && ++g_count == g_fail_at;
}

}  // namespace

hipError_t dev_malloc_raw(void **p, size_t bytes) {
    if (injected_failure()) {
        *p = nullptr;
        return hipErrorOutOfMemory;
    }
    return hipMalloc(p, bytes);
}

hipError_t host_malloc_pinned_raw(void **p, size_t bytes) {
    if (injected_failure()) {
        *p = nullptr;
        return hipErrorOutOfMemory;
    }
    return hipHostMalloc(p, bytes, hipHostMallocDefault);
}

}  // namespace rtx
