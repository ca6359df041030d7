// Learned step-size proposer: host interface of the hand-written HIP training kernels.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>

#include "hcgpu.h"

namespace hcg {

// Device workspace reused across calls (grow-only).
struct MlpWork {
    float* params = nullptr;   // HCG_MLP_NPARAMS
    float* grad = nullptr;     // HCG_MLP_NPARAMS
    float* adam_m = nullptr;
    float* adam_v = nullptr;
    float* x = nullptr;        // cap * 8
    float* y = nullptr;        // cap * 2
    float* w = nullptr;        // cap * 2
    float* pred = nullptr;     // cap * 2
    float* loss = nullptr;     // 1
    int32_t* perm = nullptr;   // cap
    long cap = 0;
};

void mlp_work_free(MlpWork* w);
int mlp_forward_backward(MlpWork* w, hipStream_t s, const float* params, int32_t n, const float* x, const float* y,
                         const float* wt, float* out_pred, float* loss, float* grad, std::string* err);
int mlp_train(MlpWork* w, hipStream_t s, float* params, int32_t n, const float* x, const float* y, const float* wt,
              int32_t epochs, int32_t batch, float lr, uint64_t seed, float* final_loss, std::string* err);

}  // namespace hcg
