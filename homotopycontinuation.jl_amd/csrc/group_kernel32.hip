// Lane-group path tracker kernel for systems of 17 to 24 variables (precompiled for gfx950).
//
// One path per group of 32 lanes (half a wave): lane r holds entry r of every vector and row r
// of the Jacobian and its LU, as in group_kernel.hip, whose tracker, interpreter and dispatcher
// this file builds with GL = 32.  One workgroup = 4 waves = 8 groups.  Broadcasts within a group
// cross the two 16-lane rows of its half wave, so they go through the LDS crossbar (ds_bpermute)
// instead of DPP row broadcasts.  The slot programs keep 16 instructions per slot; lanes 16..31
// take no slot instructions.
#define GL 32
#include <hip/hip_runtime.h>

#include "device/hcg_group_run.h"
#include "hcg_group_launch.h"

namespace {

constexpr int kBlock = 256;

template <int NN>
__global__ void __launch_bounds__(kBlock, 2) hcg_group_kernel32(hcg_multi_kargs m, const hcg_gsys* sysd, int lds_per_path) {
    extern __shared__ __align__(16) char smem[];
    hcg_group_run<NN>(m, sysd, lds_per_path, smem, (int)threadIdx.x);
}

template <int NN> hipError_t allow_lds32(size_t bytes) {
    return hipFuncSetAttribute(reinterpret_cast<const void*>(&hcg_group_kernel32<NN>),
                               hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes);
}

}  // namespace

namespace hcg {

hipError_t group_launch32(int nn, int grid, int lds_per_path, bool mlp_lds, hipStream_t stream,
                          const hcg_multi_kargs& m, const hcg_gsys* d_sys) {
    const size_t shmem = (mlp_lds ? (size_t)HCG_GROUP_MLP_LDS : 0) + (size_t)(kBlock / GL) * lds_per_path;
    if (nn != 24) return hipErrorInvalidValue;
    hipError_t e = allow_lds32<24>(shmem);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(hcg_group_kernel32<24>, dim3(grid), dim3(kBlock), shmem, stream, m, d_sys, lds_per_path);
    return hipGetLastError();
}

hipError_t group_kernel32_blocks_per_cu(int nn, size_t shmem, int* blocks) {
    if (nn != 24) return hipErrorInvalidValue;
    hipError_t e = allow_lds32<24>(shmem);
    if (e != hipSuccess) return e;
    return hipOccupancyMaxActiveBlocksPerMultiprocessor(blocks, hcg_group_kernel32<24>, kBlock, shmem);
}

hipError_t group_kernel32_attributes(int nn, hipFuncAttributes* attr) {
    if (nn != 24) return hipErrorInvalidValue;
    return hipFuncGetAttributes(attr, reinterpret_cast<const void*>(&hcg_group_kernel32<24>));
}

}  // namespace hcg
