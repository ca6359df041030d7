// Work distribution of the persistent tracker kernels over the lanes of a wave.
//
// A lane whose path has ended does not start its next path at once: starting a path
// (tracker and predictor initialization, a condition estimate) costs a few tracker steps,
// and while one lane does it the other 63 lanes of the wave wait.  Idle lanes therefore
// wait until at least `refill` of the wave's live lanes are idle (or all of them are),
// then take their next paths from the work queue together, with one atomic per wave.
#pragma once
#include "hcgpu_types.h"

#define HCG_DEFAULT_REFILL 8

__device__ __forceinline__ int refill_threshold(const hcg_options* o) {
    int r = o->refill_idle;
    return r > 0 ? r : HCG_DEFAULT_REFILL;
}

// Called by every live lane of the wave, converged: the mask of the lanes that take a new
// path in this loop iteration (all idle lanes, or none).
__device__ __forceinline__ unsigned long long wave_refill_mask(bool idle, int refill) {
    const unsigned long long im = __ballot(idle ? 1 : 0), lm = __ballot(1);
    return (im != 0ull && (__popcll(im) >= refill || im == lm)) ? im : 0ull;
}

// Called by exactly the lanes of mask m (non-empty, this lane in m): consecutive queue
// indices, in lane order, from one atomicAdd on the queue head.
__device__ __forceinline__ int wave_fetch(unsigned long long m, int32_t* next) {
    const int lane = (int)__lane_id();
    const int leader = __ffsll((unsigned long long)m) - 1;
    int base = 0;
    if (lane == leader) base = atomicAdd(next, __popcll(m));
    base = __shfl(base, leader);
    return base + __popcll(m & ((1ull << lane) - 1ull));
}
