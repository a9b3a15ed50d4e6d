// Device and pinned host allocations of the library, in one place so that
// their failure paths can be exercised: with RT_FAIL_DEVICE_ALLOC=k in the
// environment, the k-th such allocation after the variable was last set (or
// changed) fails with hipErrorOutOfMemory without allocating.  Unset, or 0,
// every call goes straight to hipMalloc / hipHostMalloc.
#pragma once
#include <hip/hip_runtime.h>

#include <cstddef>

namespace rtx {

hipError_t dev_malloc_raw(void **p, size_t bytes);
hipError_t host_malloc_pinned_raw(void **p, size_t bytes);

template <typename T>
inline hipError_t dev_malloc(T **p, size_t bytes) {
    return dev_malloc_raw(reinterpret_cast<void **>(p), bytes);
}
template <typename T>
inline hipError_t host_malloc_pinned(T **p, size_t bytes) {
    return host_malloc_pinned_raw(reinterpret_cast<void **>(p), bytes);
}

}  // namespace rtx
