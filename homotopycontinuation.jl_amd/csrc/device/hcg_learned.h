// Learned step-size proposer, fused into the tracker.
//
// After every tracker step the eight per-step features below are formed (the same
// function feeds the step log the network is trained on).  When the learned proposer is
// on, an MLP 8 -> 64 -> 64 (GELU) -> 2 runs per lane in FP32 from BF16 weights staged in
// LDS.  Output 0 is the log of the step relative to the plain rule's proposal; the
// proposal is only ever lengthened, by at most options.learned_max_mult, and stays under
// the plain rule's growth limit.  Output 1 is the at-infinity logit (logged, not acted
// on).  A learned step must pass the same Newton acceptance test as a plain one; after a
// rejected step the plain rule alone proposes until a step is accepted again, so the
// plain rule remains both the fallback and the bound.
#pragma once

#define HCG_FEATS 8
#define HCG_LN10 2.302585092994046

__device__ __forceinline__ float feat_clean(double v) {
    if (isnan(v)) return 0.0f;
    return (float)jclamp(v, -80.0, 80.0);
}

// Features of the step just taken (ds0 = its size), with the tracker state after the
// accept / reject update: log Δs, Newton contraction θ, log(e Δs^4) with e the
// predictor's local error constant, t, log10 cond(H_x, ẋ), log ω, log μ, log τ.
__device__ void step_features(Lane& L, const NewtRes& res, double ds0) {
    L.feat[0] = feat_clean(hc_log(ds0));
    L.feat[1] = feat_clean(res.theta);
    L.feat[2] = feat_clean(hc_log(L.local_error * ipow(ds0, 4)));
    L.feat[3] = feat_clean(seg_t(L).re);
    L.feat[4] = feat_clean(hc_log(L.cond_H_xd) / HCG_LN10);
    L.feat[5] = feat_clean(hc_log(L.omega));
    L.feat[6] = feat_clean(hc_log(L.mu));
    L.feat[7] = feat_clean(hc_log(L.tau));
}

// parameter layout: W1[64][8] b1[64] W2[64][64] b2[64] W3[2][64] b3[2]
#define MLP_W1 0
#define MLP_B1 (MLP_W1 + HCG_MLP_HIDDEN * HCG_MLP_IN)
#define MLP_W2 (MLP_B1 + HCG_MLP_HIDDEN)
#define MLP_B2 (MLP_W2 + HCG_MLP_HIDDEN * HCG_MLP_HIDDEN)
#define MLP_W3 (MLP_B2 + HCG_MLP_HIDDEN)
#define MLP_B3 (MLP_W3 + HCG_MLP_OUT * HCG_MLP_HIDDEN)

__device__ __forceinline__ float bf16_to_f32(unsigned short h) { return __uint_as_float(((unsigned)h) << 16); }
__device__ __forceinline__ unsigned short f32_to_bf16(float x) {  // round to nearest even
    unsigned u = __float_as_uint(x);
    if ((u & 0x7f800000u) == 0x7f800000u) return (unsigned short)(u >> 16);
    return (unsigned short)((u + 0x7fffu + ((u >> 16) & 1u)) >> 16);
}
__device__ __forceinline__ float mlp_gelu(float a) { return 0.5f * a * (1.0f + erff(a * 0.70710678118654752f)); }

// One inference; w: BF16 weights in LDS (broadcast reads), x: 8 features.
__device__ __noinline__ void mlp_forward(const unsigned short* __restrict__ w, const float* x, float* out) {
    float g1[HCG_MLP_HIDDEN];
#pragma unroll
    for (int j = 0; j < HCG_MLP_HIDDEN; j++) {
        float s = bf16_to_f32(w[MLP_B1 + j]);
#pragma unroll
        for (int i = 0; i < HCG_MLP_IN; i++) s = fmaf(bf16_to_f32(w[MLP_W1 + j * HCG_MLP_IN + i]), x[i], s);
        g1[j] = mlp_gelu(s);
    }
    float o0 = bf16_to_f32(w[MLP_B3]), o1 = bf16_to_f32(w[MLP_B3 + 1]);
    for (int j = 0; j < HCG_MLP_HIDDEN; j++) {
        float s = bf16_to_f32(w[MLP_B2 + j]);
#pragma unroll
        for (int k = 0; k < HCG_MLP_HIDDEN; k++) s = fmaf(bf16_to_f32(w[MLP_W2 + j * HCG_MLP_HIDDEN + k]), g1[k], s);
        float h = mlp_gelu(s);
        o0 = fmaf(bf16_to_f32(w[MLP_W3 + j]), h, o0);
        o1 = fmaf(bf16_to_f32(w[MLP_W3 + HCG_MLP_HIDDEN + j]), h, o1);
    }
    out[0] = o0;
    out[1] = o1;
}

// Learned proposal for the step after an accepted one; `ds` is the plain rule's
// proposal after all of its limits.
__device__ double learned_adjust(Lane& L, const NewtRes& res, double ds) {
    L.ds_plain = ds;
    L.last_learned = 0;
    if (!L.mlp_w || L.extended_prec || L.singular_endgame || L.last_failed > 0) return ds;
    if (fabs(seg_t(L).re) <= L.o->learned_min_t) return ds;
    float out[2];
    mlp_forward(L.mlp_w, L.feat, out);
    L.flops += 2.0 * (HCG_MLP_IN * HCG_MLP_HIDDEN + HCG_MLP_HIDDEN * HCG_MLP_HIDDEN + HCG_MLP_OUT * HCG_MLP_HIDDEN);
    L.inf_logit = out[1];
    double lm = jclamp((double)out[0], 0.0, hc_log(L.o->learned_max_mult));
    if (!(lm > 0.0)) return ds;
    double dsl = jmin(jmin(ds * hc_exp(lm), 10 * L.ds_prev), L.o->max_step_size);
    if (!(dsl > ds)) return ds;
    L.last_learned = 1;
    L.learned_proposed++;
    return dsl;
}
