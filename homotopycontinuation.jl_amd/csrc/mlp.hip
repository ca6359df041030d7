// Training kernels of the learned step-size proposer (8 -> 64 -> 64 GELU -> 2).
//
// Forward uses BF16-rounded copies of the FP32 master weights (round to nearest even)
// with FP32 accumulation; gradients flow to the master weights unchanged (straight-
// through).  Loss per record: w0 * (o0 - y0)^2 / 2 + w1 * BCEWithLogits(o1, y1),
// averaged over the batch.  One wave per block handles 64 records: activations are
// staged in LDS, thread t owns record t in the row-wise passes and neuron t in the
// column-wise (weight-gradient) reductions; block partial gradients are added to
// the global gradient with atomics.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <numeric>
#include <random>
#include <vector>

#include "hcg_mlp.h"

namespace {

constexpr int IN = HCG_MLP_IN, H = HCG_MLP_HIDDEN, OUT = HCG_MLP_OUT;
constexpr int OFF_W1 = 0, OFF_B1 = OFF_W1 + H * IN, OFF_W2 = OFF_B1 + H, OFF_B2 = OFF_W2 + H * H,
              OFF_W3 = OFF_B2 + H, OFF_B3 = OFF_W3 + OUT * H, NP = OFF_B3 + OUT;
static_assert(NP == HCG_MLP_NPARAMS, "parameter layout");
constexpr int S = 64;  // records per block

__device__ __forceinline__ float bf16r(float x) {
    unsigned u = __float_as_uint(x);
    if ((u & 0x7f800000u) == 0x7f800000u) return x;  // inf / nan unchanged
    unsigned r = u + 0x7fffu + ((u >> 16) & 1u);
    return __uint_as_float(r & 0xffff0000u);
}
__device__ __forceinline__ float gelu(float a) { return 0.5f * a * (1.0f + erff(a * 0.70710678118654752f)); }
__device__ __forceinline__ float gelu_d(float a) {
    return 0.5f * (1.0f + erff(a * 0.70710678118654752f)) + a * 0.39894228040143268f * expf(-0.5f * a * a);
}

// grad += d loss / d params for records idx[0..cnt) (or 0..cnt when idx == nullptr);
// pred (optional) receives the outputs; loss_acc accumulates the summed loss.
__global__ void __launch_bounds__(S) mlp_fwd_bwd(const float* __restrict__ P, const float* __restrict__ X,
                                                  const float* __restrict__ Y, const float* __restrict__ Wt,
                                                  const int32_t* __restrict__ idx, int cnt, float inv_n,
                                                  float* __restrict__ grad, float* __restrict__ pred,
                                                  float* __restrict__ loss_acc) {
    __shared__ float w[NP];
    __shared__ float a1[S][H + 1];
    __shared__ float a2[S][H + 1];
    __shared__ float xs[S][IN];
    __shared__ float dO[S][OUT];
    const int t = threadIdx.x;
    for (int i = t; i < NP; i += S) w[i] = bf16r(P[i]);
    const int r = blockIdx.x * S + t;
    const bool valid = r < cnt;
    const int rec = valid ? (idx ? idx[r] : r) : 0;
    for (int i = 0; i < IN; i++) xs[t][i] = valid ? X[(size_t)rec * IN + i] : 0.0f;
    __syncthreads();
    // forward, thread t = record t
    for (int j = 0; j < H; j++) {
        float s = w[OFF_B1 + j];
        for (int i = 0; i < IN; i++) s = fmaf(w[OFF_W1 + j * IN + i], xs[t][i], s);
        a1[t][j] = s;
    }
    for (int j = 0; j < H; j++) {
        float s = w[OFF_B2 + j];
        for (int k = 0; k < H; k++) s = fmaf(w[OFF_W2 + j * H + k], gelu(a1[t][k]), s);
        a2[t][j] = s;
    }
    float o[OUT];
    for (int j = 0; j < OUT; j++) {
        float s = w[OFF_B3 + j];
        for (int k = 0; k < H; k++) s = fmaf(w[OFF_W3 + j * H + k], gelu(a2[t][k]), s);
        o[j] = s;
    }
    float l = 0.0f;
    if (valid) {
        float y0 = Y[(size_t)rec * OUT], y1 = Y[(size_t)rec * OUT + 1];
        float w0 = Wt[(size_t)rec * OUT], w1 = Wt[(size_t)rec * OUT + 1];
        float e = o[0] - y0;
        float sg = 1.0f / (1.0f + expf(-o[1]));
        // BCE with logits, stable form: max(o,0) - o*y + log(1 + exp(-|o|))
        float bce = fmaxf(o[1], 0.0f) - o[1] * y1 + log1pf(expf(-fabsf(o[1])));
        l = w0 * 0.5f * e * e + w1 * bce;
        dO[t][0] = w0 * e * inv_n;
        dO[t][1] = w1 * (sg - y1) * inv_n;
        if (pred) { pred[(size_t)r * OUT] = o[0]; pred[(size_t)r * OUT + 1] = o[1]; }
    } else {
        dO[t][0] = dO[t][1] = 0.0f;
    }
    if (loss_acc) {
        for (int off = 32; off > 0; off >>= 1) l += __shfl_down(l, off, 64);
        if (t == 0) atomicAdd(loss_acc, l);
    }
    __syncthreads();
    // layer 3 gradients: thread t = hidden unit t
    {
        float g0 = 0.0f, g1 = 0.0f;
        for (int s = 0; s < S; s++) {
            float h = gelu(a2[s][t]);
            g0 = fmaf(dO[s][0], h, g0);
            g1 = fmaf(dO[s][1], h, g1);
        }
        atomicAdd(&grad[OFF_W3 + t], g0);
        atomicAdd(&grad[OFF_W3 + H + t], g1);
        if (t < OUT) {
            float gb = 0.0f;
            for (int s = 0; s < S; s++) gb += dO[s][t];
            atomicAdd(&grad[OFF_B3 + t], gb);
        }
    }
    // dA2 (row-wise, overwrite a2 in place after reading it)
    __syncthreads();
    {
        float d[H];
        for (int k = 0; k < H; k++) {
            float s = dO[t][0] * w[OFF_W3 + k] + dO[t][1] * w[OFF_W3 + H + k];
            d[k] = s * gelu_d(a2[t][k]);
        }
        for (int k = 0; k < H; k++) a2[t][k] = d[k];
    }
    __syncthreads();
    // layer 2 gradients: thread t = output unit t of layer 2
    {
        float acc[H];
        for (int k = 0; k < H; k++) acc[k] = 0.0f;
        float gb = 0.0f;
        for (int s = 0; s < S; s++) {
            float da = a2[s][t];
            gb += da;
            for (int k = 0; k < H; k++) acc[k] = fmaf(da, gelu(a1[s][k]), acc[k]);
        }
        for (int k = 0; k < H; k++) atomicAdd(&grad[OFF_W2 + t * H + k], acc[k]);
        atomicAdd(&grad[OFF_B2 + t], gb);
    }
    // dA1 (row-wise)
    {
        float d[H];
        for (int k = 0; k < H; k++) {
            float s = 0.0f;
            for (int j = 0; j < H; j++) s = fmaf(a2[t][j], w[OFF_W2 + j * H + k], s);
            d[k] = s * gelu_d(a1[t][k]);
        }
        __syncthreads();
        for (int k = 0; k < H; k++) a1[t][k] = d[k];
    }
    __syncthreads();
    // layer 1 gradients: thread t = hidden unit t
    {
        float acc[IN];
        for (int i = 0; i < IN; i++) acc[i] = 0.0f;
        float gb = 0.0f;
        for (int s = 0; s < S; s++) {
            float da = a1[s][t];
            gb += da;
            for (int i = 0; i < IN; i++) acc[i] = fmaf(da, xs[s][i], acc[i]);
        }
        for (int i = 0; i < IN; i++) atomicAdd(&grad[OFF_W1 + t * IN + i], acc[i]);
        atomicAdd(&grad[OFF_B1 + t], gb);
    }
}

__global__ void fill_zero(float* p, int n) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0.0f;
}

__global__ void adam_step(float* __restrict__ P, const float* __restrict__ G, float* __restrict__ M, float* __restrict__ V,
                          int n, float lr, float b1, float b2, float eps, float c1, float c2) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float g = G[i];
    float m = b1 * M[i] + (1.0f - b1) * g;
    float v = b2 * V[i] + (1.0f - b2) * g * g;
    M[i] = m;
    V[i] = v;
    P[i] -= lr * (m / c1) / (sqrtf(v / c2) + eps);
}

#define CK(x)                                                            \
    do {                                                                 \
        hipError_t e_ = (x);                                             \
        if (e_ != hipSuccess) { *err = std::string(#x ": ") + hipGetErrorString(e_); return HCG_ERR_HIP; } \
    } while (0)

int ensure(hcg::MlpWork* w, long n, std::string* err) {
    if (!w->params) {
        CK(hipMalloc(&w->params, sizeof(float) * NP));
        CK(hipMalloc(&w->grad, sizeof(float) * NP));
        CK(hipMalloc(&w->adam_m, sizeof(float) * NP));
        CK(hipMalloc(&w->adam_v, sizeof(float) * NP));
        CK(hipMalloc(&w->loss, sizeof(float)));
    }
    if (n > w->cap) {
        hipFree(w->x); hipFree(w->y); hipFree(w->w); hipFree(w->pred); hipFree(w->perm);
        w->x = w->y = w->w = w->pred = nullptr; w->perm = nullptr;
        CK(hipMalloc(&w->x, sizeof(float) * IN * n));
        CK(hipMalloc(&w->y, sizeof(float) * OUT * n));
        CK(hipMalloc(&w->w, sizeof(float) * OUT * n));
        CK(hipMalloc(&w->pred, sizeof(float) * OUT * n));
        CK(hipMalloc(&w->perm, sizeof(int32_t) * n));
        w->cap = n;
    }
    return HCG_OK;
}

}  // namespace

namespace hcg {

void mlp_work_free(MlpWork* w) {
    void* ps[] = {w->params, w->grad, w->adam_m, w->adam_v, w->x, w->y, w->w, w->pred, w->loss, w->perm};
    for (void* p : ps) if (p) hipFree(p);
    *w = MlpWork();
}

int mlp_forward_backward(MlpWork* w, hipStream_t st, const float* params, int32_t n, const float* x, const float* y,
                         const float* wt, float* out_pred, float* loss, float* grad, std::string* err) {
    int rc = ensure(w, std::max<int32_t>(n, 1), err);
    if (rc) return rc;
    CK(hipMemcpyAsync(w->params, params, sizeof(float) * NP, hipMemcpyHostToDevice, st));
    if (n > 0) {
        CK(hipMemcpyAsync(w->x, x, sizeof(float) * IN * n, hipMemcpyHostToDevice, st));
        CK(hipMemcpyAsync(w->y, y, sizeof(float) * OUT * n, hipMemcpyHostToDevice, st));
        CK(hipMemcpyAsync(w->w, wt, sizeof(float) * OUT * n, hipMemcpyHostToDevice, st));
    }
    fill_zero<<<(NP + 255) / 256, 256, 0, st>>>(w->grad, NP);
    fill_zero<<<1, 64, 0, st>>>(w->loss, 1);
    if (n > 0)
        mlp_fwd_bwd<<<(n + S - 1) / S, S, 0, st>>>(w->params, w->x, w->y, w->w, nullptr, n, 1.0f / n, w->grad, w->pred, w->loss);
    CK(hipGetLastError());
    float l = 0.0f;
    CK(hipMemcpyAsync(&l, w->loss, sizeof(float), hipMemcpyDeviceToHost, st));
    if (grad) CK(hipMemcpyAsync(grad, w->grad, sizeof(float) * NP, hipMemcpyDeviceToHost, st));
    if (out_pred && n > 0) CK(hipMemcpyAsync(out_pred, w->pred, sizeof(float) * OUT * n, hipMemcpyDeviceToHost, st));
    CK(hipStreamSynchronize(st));
    if (loss) *loss = n > 0 ? l / n : 0.0f;
    return HCG_OK;
}

int mlp_train(MlpWork* w, hipStream_t st, float* params, int32_t n, const float* x, const float* y, const float* wt,
              int32_t epochs, int32_t batch, float lr, uint64_t seed, float* final_loss, std::string* err) {
    int rc = ensure(w, n, err);
    if (rc) return rc;
    CK(hipMemcpyAsync(w->params, params, sizeof(float) * NP, hipMemcpyHostToDevice, st));
    CK(hipMemcpyAsync(w->x, x, sizeof(float) * IN * n, hipMemcpyHostToDevice, st));
    CK(hipMemcpyAsync(w->y, y, sizeof(float) * OUT * n, hipMemcpyHostToDevice, st));
    CK(hipMemcpyAsync(w->w, wt, sizeof(float) * OUT * n, hipMemcpyHostToDevice, st));
    fill_zero<<<(NP + 255) / 256, 256, 0, st>>>(w->adam_m, NP);
    fill_zero<<<(NP + 255) / 256, 256, 0, st>>>(w->adam_v, NP);
    std::vector<int32_t> perm(n);
    std::iota(perm.begin(), perm.end(), 0);
    std::mt19937_64 rng(seed);
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    long step = 0;
    float last = 0.0f;
    for (int ep = 0; ep < epochs; ep++) {
        std::shuffle(perm.begin(), perm.end(), rng);
        CK(hipMemcpyAsync(w->perm, perm.data(), sizeof(int32_t) * n, hipMemcpyHostToDevice, st));
        fill_zero<<<1, 64, 0, st>>>(w->loss, 1);
        for (int off = 0; off < n; off += batch) {
            int cnt = std::min(batch, n - off);
            fill_zero<<<(NP + 255) / 256, 256, 0, st>>>(w->grad, NP);
            mlp_fwd_bwd<<<(cnt + S - 1) / S, S, 0, st>>>(w->params, w->x, w->y, w->w, w->perm + off, cnt, 1.0f / cnt,
                                                         w->grad, nullptr, w->loss);
            step++;
            float c1 = 1.0f - std::pow(b1, (float)step), c2 = 1.0f - std::pow(b2, (float)step);
            adam_step<<<(NP + 255) / 256, 256, 0, st>>>(w->params, w->grad, w->adam_m, w->adam_v, NP, lr, b1, b2, eps, c1, c2);
        }
        CK(hipGetLastError());
        CK(hipMemcpyAsync(&last, w->loss, sizeof(float), hipMemcpyDeviceToHost, st));
        CK(hipStreamSynchronize(st));
    }
    CK(hipMemcpyAsync(params, w->params, sizeof(float) * NP, hipMemcpyDeviceToHost, st));
    CK(hipStreamSynchronize(st));
    if (final_loss) *final_loss = last / n;
    return HCG_OK;
}

}  // namespace hcg
