// The allocation-free entry point's per-device scratch (noalloc.hip).
#pragma once
#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <string>

#include "rt/rt_c.h"

namespace rtx {

// Layout: [0, kNoallocHeaderBytes) the dynamic schedule's work counter (two
// words, zero between launches), then the spill stack (uint2 entries).
constexpr size_t kNoallocScratchBytes = RT_NOALLOC_SCRATCH_BYTES;
constexpr size_t kNoallocHeaderBytes = 256;
constexpr size_t kNoallocSpillEntries = (kNoallocScratchBytes - kNoallocHeaderBytes) / 8;

// Resolves the device address of the scratch on `device`, which must be the
// current device, and zeroes its header once.  0 on success.
int noalloc_scratch_init(int device, std::string &err);
// The address noalloc_scratch_init resolved, or null.
void *noalloc_scratch(int device);
// Zeroes the two work-counter words on `stream` with a kernel.  Returns a
// hipError_t value.
int noalloc_clear_work(uint32_t *work, hipStream_t stream);

}  // namespace rtx
