// Fixed device scratch of hcg_track_batch_device_noalloc: the launch state the tracker
// kernel needs beyond the caller's inputs and outputs (work-queue head, the four launch
// counters and the device copy of the options) lives in one file-scope device array, so a
// launch through that entry point allocates nothing.
//
// Each GPU the library runs on holds its own copy of `scratch` (kScratchBytes) until the
// process ends.  Host code never passes the array's name to a kernel: it takes the device
// address from hipGetSymbolAddress once per device (noalloc_scratch) and uses that.  Every
// call on one GPU shares the array, so calls that overlap on one GPU race; each call clears
// the queue head and counters and writes the options with a small kernel (no memset node, no
// host-to-device copy of pageable memory), so the array's earlier contents never matter.
#include <hip/hip_runtime.h>

#include <mutex>

#include "hcg_noalloc.h"

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_mu;
void* g_ptr[kMaxDevices] = {};
}  // namespace

static __device__ __align__(256) unsigned char scratch[hcg::kScratchBytes];

__global__ void hcg_scratch_init_kernel(hcg::NoallocScratch* s, hcg_options o) {
    if (threadIdx.x == 0) {
        s->next_path = 0;
        for (int k = 0; k < 4; k++) s->counters[k] = 0ULL;
        s->opts = o;
    }
}

namespace hcg {

int noalloc_scratch(int device, NoallocScratch** out, const char** err) {
    if (device < 0 || device >= kMaxDevices) {
        *err = "device index beyond the scratch table";
        return -1;
    }
    std::lock_guard<std::mutex> lk(g_mu);
    if (!g_ptr[device]) {
        void* p = nullptr;
        hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));  // the current device's copy
        if (e != hipSuccess || !p) {
            *err = hipGetErrorString(e);
            return -1;
        }
        g_ptr[device] = p;
    }
    *out = (NoallocScratch*)g_ptr[device];
    return 0;
}

int noalloc_init(hipStream_t stream, NoallocScratch* s, const hcg_options* o) {
    hipLaunchKernelGGL(hcg_scratch_init_kernel, dim3(1), dim3(64), 0, stream, s, *o);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

}  // namespace hcg
