// Layout and helpers of the fixed device scratch used by hcg_track_batch_device_noalloc
// (defined in noalloc.hip).
#pragma once
#include <hip/hip_runtime.h>

#include <cstddef>

#include "hcgpu.h"

namespace hcg {

struct NoallocScratch {
    int32_t next_path;               // work-queue head
    int32_t pad_;
    unsigned long long counters[4];  // steps, Newton iterations, counted FP64 flops, learned accepted
    hcg_options opts;                // the launch's options
};

constexpr size_t kScratchBytes = 1024;
static_assert(sizeof(NoallocScratch) <= kScratchBytes, "scratch layout exceeds its fixed size");

// Device address of the current device's scratch (hipGetSymbolAddress, once per device);
// the caller has made `device` current.  Returns 0, or -1 with *err set.
int noalloc_scratch(int device, NoallocScratch** out, const char** err);
// Clear the queue head and counters and store the options, on `stream`, with a kernel.
int noalloc_init(hipStream_t stream, NoallocScratch* s, const hcg_options* o);

}  // namespace hcg
