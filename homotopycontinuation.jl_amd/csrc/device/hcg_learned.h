// Learned step-size proposer, fused into the tracker.
//
// After every tracker step the eight per-step features below are formed (the same
// function feeds the step log the network is trained on).  When the learned proposer is
// on, an MLP 8 -> 64 -> 64 (GELU) -> 2 runs per lane in FP32 from BF16 weights staged in
// LDS.  Output 0 is the log of the step relative to the plain rule's proposal; the
// proposal is only ever lengthened, by at most options.learned_max_mult, and stays under
// the plain rule's growth limit.  Output 1 is the at-infinity logit: when options
// learned_inf_t > 0, a path whose logit exceeds learned_inf_logit starts its endgame once
// t <= learned_inf_t (the endgame's own valuation checks still decide every path).  A learned step must pass the same Newton acceptance test as a plain one; after a
// rejected step the plain rule alone proposes until a step is accepted again, so the
// plain rule remains both the fallback and the bound.
// (Included once per system by hcg_tracker_body.h.)

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

// Learned proposal for the step after an accepted one; `ds` is the plain rule's
// proposal after all of its limits.
__device__ double learned_adjust(Lane& L, const NewtRes& res, double ds) {
    L.ds_plain = ds;
    L.last_learned = 0;
    if (!L.mlp_w || L.extended_prec || L.singular_endgame || L.last_failed > 0) return ds;
    if (fabs(seg_t(L).re) <= L.o->learned_min_t) return ds;
    float out[2];
    mlp_forward(L.mlp_w, L.feat, out);
    L.inf_logit = out[1];
    {   // at-infinity flag: start the endgame (its own checks decide) at a larger t
        const hcg_options* o = L.o;
        double tn = fabs(seg_t(L).re);
        if (o->learned_inf_t > 0.0 && out[1] > o->learned_inf_logit && tn <= o->learned_inf_t && tn > L.eg_start)
            L.eg_start = tn;
    }
    double lm = jclamp((double)out[0], 0.0, hc_log(L.o->learned_max_mult));
    if (!(lm > 0.0)) return ds;
    double dsl = jmin(jmin(ds * hc_exp(lm), 10 * L.ds_prev), L.o->max_step_size);
    if (!(dsl > ds)) return ds;
    L.last_learned = 1;
    L.learned_proposed++;
    return dsl;
}
