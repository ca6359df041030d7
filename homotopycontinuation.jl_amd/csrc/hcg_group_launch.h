// Launcher of the precompiled lane-group tracker kernel (group_kernel.hip).
#pragma once
#include <hip/hip_runtime.h>

#include "hcg_group.h"

namespace hcg {

hipError_t group_launch(int nn, int grid, int lds_per_path, hipStream_t stream, const hcg_multi_kargs& m,
                        const hcg_gsys* d_sys);
hipError_t group_kernel_attributes(int nn, hipFuncAttributes* attr);
hipError_t group_kernel_blocks_per_cu(int nn, size_t shmem, int* blocks);

}  // namespace hcg
