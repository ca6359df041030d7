// Lane-group path tracker kernel (precompiled for gfx950; no per-system compilation).
//
// One workgroup = 4 waves = 16 groups of 16 lanes; each group tracks one path at a time and
// takes the next path from the launch's work queue when its path ends.  The system's tapes
// arrive as slot programs (hcg_gsys, built by group.cpp from group_sched.cpp's schedules);
// each group owns `lds_per_path` bytes of LDS for their registers.  The tracker itself is
// hcg_group_body.h; this file holds the evaluation dispatcher (the one place each slot
// interpreter is instantiated), the path loop and the host-side launcher.
#include <hip/hip_runtime.h>

#include "device/hcg_group_run.h"
#include "hcg_group_launch.h"

namespace {

constexpr int kBlock = 256;
constexpr int kMlpLdsBytes = HCG_GROUP_MLP_LDS;

template <int NN>
__global__ void __launch_bounds__(kBlock, 2) hcg_group_kernel(hcg_multi_kargs m, const hcg_gsys* sysd, int lds_per_path) {
    extern __shared__ __align__(16) char smem[];
    hcg_group_run<NN>(m, sysd, lds_per_path, smem, (int)threadIdx.x);
}

}  // namespace

namespace hcg {

template <int NN> hipError_t allow_lds(size_t bytes) {
    return hipFuncSetAttribute(reinterpret_cast<const void*>(&hcg_group_kernel<NN>),
                               hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes);
}

hipError_t group_launch(int nn, int grid, int lds_per_path, bool mlp_lds, hipStream_t stream,
                        const hcg_multi_kargs& m, const hcg_gsys* d_sys) {
    if (nn > 16) return group_launch32(nn, grid, lds_per_path, mlp_lds, stream, m, d_sys);
    // LDS: the step network's BF16 weights (learned step only), then each group's registers
    const size_t shmem = (mlp_lds ? (size_t)kMlpLdsBytes : 0) + (size_t)(kBlock / HCG_GROUP_LANES) * lds_per_path;
    hipError_t e = nn == 4 ? allow_lds<4>(shmem) : nn == 8 ? allow_lds<8>(shmem) : nn == 12 ? allow_lds<12>(shmem)
                                                                              : allow_lds<16>(shmem);
    if (e != hipSuccess) return e;
    switch (nn) {
    case 4: hipLaunchKernelGGL(hcg_group_kernel<4>, dim3(grid), dim3(kBlock), shmem, stream, m, d_sys, lds_per_path); break;
    case 8: hipLaunchKernelGGL(hcg_group_kernel<8>, dim3(grid), dim3(kBlock), shmem, stream, m, d_sys, lds_per_path); break;
    case 12: hipLaunchKernelGGL(hcg_group_kernel<12>, dim3(grid), dim3(kBlock), shmem, stream, m, d_sys, lds_per_path); break;
    case 16: hipLaunchKernelGGL(hcg_group_kernel<16>, dim3(grid), dim3(kBlock), shmem, stream, m, d_sys, lds_per_path); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

// Workgroups of the kernel instance that fit on one CU with `shmem` bytes of dynamic LDS each
// (registers and LDS both counted); each workgroup is one wave per SIMD.
hipError_t group_kernel_blocks_per_cu(int nn, size_t shmem, int* blocks) {
    if (nn > 16) return group_kernel32_blocks_per_cu(nn, shmem, blocks);
    hipError_t e = nn == 4 ? allow_lds<4>(shmem) : nn == 8 ? allow_lds<8>(shmem) : nn == 12 ? allow_lds<12>(shmem)
                                                                              : allow_lds<16>(shmem);
    if (e != hipSuccess) return e;
    switch (nn) {
    case 4: return hipOccupancyMaxActiveBlocksPerMultiprocessor(blocks, hcg_group_kernel<4>, kBlock, shmem);
    case 8: return hipOccupancyMaxActiveBlocksPerMultiprocessor(blocks, hcg_group_kernel<8>, kBlock, shmem);
    case 12: return hipOccupancyMaxActiveBlocksPerMultiprocessor(blocks, hcg_group_kernel<12>, kBlock, shmem);
    case 16: return hipOccupancyMaxActiveBlocksPerMultiprocessor(blocks, hcg_group_kernel<16>, kBlock, shmem);
    default: return hipErrorInvalidValue;
    }
}

hipError_t group_kernel_attributes(int nn, hipFuncAttributes* attr) {
    if (nn > 16) return group_kernel32_attributes(nn, attr);
    switch (nn) {
    case 4: return hipFuncGetAttributes(attr, reinterpret_cast<const void*>(&hcg_group_kernel<4>));
    case 8: return hipFuncGetAttributes(attr, reinterpret_cast<const void*>(&hcg_group_kernel<8>));
    case 12: return hipFuncGetAttributes(attr, reinterpret_cast<const void*>(&hcg_group_kernel<12>));
    case 16: return hipFuncGetAttributes(attr, reinterpret_cast<const void*>(&hcg_group_kernel<16>));
    default: return hipErrorInvalidValue;
    }
}

}  // namespace hcg
