// Launcher of the precompiled lane-group tracker kernels (group_kernel.hip, group_kernel32.hip).
#pragma once
#include <hip/hip_runtime.h>

#include "hcg_group.h"

namespace hcg {

// mlp_lds: stage the step network's weights in LDS (when the learned step is on); the kernel
// stages them exactly when the launch's options ask for the learned step and weights are given.
hipError_t group_launch(int nn, int grid, int lds_per_path, bool mlp_lds, hipStream_t stream,
                        const hcg_multi_kargs& m, const hcg_gsys* d_sys);
hipError_t group_kernel_attributes(int nn, hipFuncAttributes* attr);
hipError_t group_kernel_blocks_per_cu(int nn, size_t shmem, int* blocks);
// the 32-lane instance (nn = 24; group_kernel32.hip), called by the three above
hipError_t group_launch32(int nn, int grid, int lds_per_path, bool mlp_lds, hipStream_t stream,
                          const hcg_multi_kargs& m, const hcg_gsys* d_sys);
hipError_t group_kernel32_attributes(int nn, hipFuncAttributes* attr);
hipError_t group_kernel32_blocks_per_cu(int nn, size_t shmem, int* blocks);

}  // namespace hcg
