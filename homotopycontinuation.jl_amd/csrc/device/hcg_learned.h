// Learned step-size proposer hooks of the tracker (stubs: plain rule only).
#pragma once

__device__ __forceinline__ void learned_record(Lane& L, const NewtRes& res) { L.last_theta = res.theta; }
__device__ __forceinline__ double learned_adjust(Lane& L, const NewtRes& res, double ds) { return ds; }
