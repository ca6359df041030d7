// The lane-group tracker's per-group driver: evaluation dispatcher (the one place each slot
// interpreter is instantiated), path loop over the launch's work queue, PathResult records.
// Included by the kernel (group_kernel.hip).
#pragma once
#include "hcg_group_body.h"
#include "hcgpu.h"

// LDS of a workgroup: the learned proposer's BF16 weights first, then one region per group
#define HCG_GROUP_MLP_LDS (((HCG_MLP_NPARAMS * 2 + 15) / 16) * 16)

template <int NN>
__device__ __forceinline__ void g_do_eval(GT<NN>& T) {
    const hcg_gsys* S = T.sys;
    const int r = T.r, n = T.n;
    GEval E;
    E.lds = T.lds;
    E.t = T.ev_t;
    E.p = T.p;
    E.q = T.q;
    const cplx t = T.ev_t;
    const cplx g = mkc(S->gamma_re, S->gamma_im);
    const cplx ts = c_mul(g, t), tt = mkc(1.0 - t.re, -t.im);
    const int rr = r < n ? r : 0;
    switch (T.req) {
    case GREQ_EJ: {
        g_bind(E, &S->ej);
        g_run_c(E, r, n, T.ev_x);
        if (S->kind == HCG_PARAMETER) {
            T.ev_u = g_read_c(E, g_gl_i(E.P->u_ref[0], rr));
            G_FOR(NN, J, T.A[J] = J < n ? g_read_c(E, g_gl_i(E.P->U_ref[0], rr * n + (J < n ? J : 0))) : mkc(0.0, 0.0););
        } else {  // u = γ t G + (1 - t) F entry by entry (src/homotopies/straight_line_homotopy.jl:95-124)
            T.ev_u = c_muladd(ts, g_read_c(E, g_gl_i(E.P->u_ref[0], rr)), c_mul(tt, g_read_c(E, g_gl_i(E.P->u_ref[1], rr))));
            G_FOR(NN, J, {
                if (J < n) {
                    const cplx gj = g_read_c(E, g_gl_i(E.P->U_ref[0], rr * n + J)), fj = g_read_c(E, g_gl_i(E.P->U_ref[1], rr * n + J));
                    T.A[J] = c_muladd(ts, gj, c_mul(tt, fj));
                } else {
                    T.A[J] = mkc(0.0, 0.0);
                }
            });
        }
        break;
    }
    case GREQ_E:
    case GREQ_T1:
        g_bind(E, &S->e);
        if (T.req == GREQ_T1 && S->kind == HCG_PARAMETER) {  // taylor!(Val(1)) through the series
            g_run_s<1>(E, r, n, T.tx0, T.tx0, T.tx0);
            T.ev_u = g_read_s<1>(E, g_gl_i(E.P->u_ref[0], rr), 1);
            break;
        }
        g_run_c(E, r, n, T.req == GREQ_T1 ? T.tx0 : T.ev_x);
        if (S->kind == HCG_PARAMETER) {
            T.ev_u = g_read_c(E, g_gl_i(E.P->u_ref[0], rr));
        } else if (T.req == GREQ_T1) {  // taylor_1!: γ G(x) - F(x)
            T.ev_u = c_sub(c_mul(g, g_read_c(E, g_gl_i(E.P->u_ref[0], rr))), g_read_c(E, g_gl_i(E.P->u_ref[1], rr)));
        } else {
            T.ev_u = c_muladd(ts, g_read_c(E, g_gl_i(E.P->u_ref[0], rr)), c_mul(tt, g_read_c(E, g_gl_i(E.P->u_ref[1], rr))));
        }
        break;
    case GREQ_DD:
        g_bind(E, &S->e);
        g_run_dd(E, r, n, T.ev_x);
        if (S->kind == HCG_PARAMETER) {
            T.ev_u = cdd_round(g_read_dd(E, g_gl_i(E.P->u_ref[0], rr)));
        } else {
            T.ev_u = cdd_round(cdd_add(cdd_mul_c(ts, g_read_dd(E, g_gl_i(E.P->u_ref[0], rr))),
                                       cdd_mul_c(tt, g_read_dd(E, g_gl_i(E.P->u_ref[1], rr)))));
        }
        break;
    case GREQ_T2:
    case GREQ_T3: {
        g_bind(E, &S->e);
        const int K = T.req == GREQ_T2 ? 2 : 3;
        cplx gK1, gK, fK1, fK;
        if (K == 2) {
            g_run_s<2>(E, r, n, T.tx0, T.tx1, T.tx1);
            fK = g_read_s<2>(E, g_gl_i(E.P->u_ref[E.P->ntapes - 1], rr), 2);
            fK1 = g_read_s<2>(E, g_gl_i(E.P->u_ref[E.P->ntapes - 1], rr), 1);
            gK = S->kind == HCG_PARAMETER ? fK : g_read_s<2>(E, g_gl_i(E.P->u_ref[0], rr), 2);
            gK1 = S->kind == HCG_PARAMETER ? fK1 : g_read_s<2>(E, g_gl_i(E.P->u_ref[0], rr), 1);
        } else {
            g_run_s<3>(E, r, n, T.tx0, T.tx1, T.tx2);
            fK = g_read_s<3>(E, g_gl_i(E.P->u_ref[E.P->ntapes - 1], rr), 3);
            fK1 = g_read_s<3>(E, g_gl_i(E.P->u_ref[E.P->ntapes - 1], rr), 2);
            gK = S->kind == HCG_PARAMETER ? fK : g_read_s<3>(E, g_gl_i(E.P->u_ref[0], rr), 3);
            gK1 = S->kind == HCG_PARAMETER ? fK1 : g_read_s<3>(E, g_gl_i(E.P->u_ref[0], rr), 2);
        }
        if (S->kind == HCG_PARAMETER) {
            T.ev_u = fK;
        } else {  // γ (G_{K-1} + t G_K) + ((1 - t) F_K - F_{K-1})
            const cplx st = c_mul(g, c_add(gK1, c_mul(t, gK)));
            const cplx tg = c_sub(c_mul(tt, fK), fK1);
            T.ev_u = c_add(st, tg);
        }
        break;
    }
    case GREQ_COND:  // cond(WS, row_scaling, col_scaling), src/linear_algebra.jl:745-774
        T.ev_cond = T.ws_cond(true, T.row_scaling, true, T.col_scaling);
        T.count_eval(T.req);
        T.req = GREQ_NONE;
        return;
    default:
        break;
    }
    if (T.req == GREQ_EJ) T.factor_after_eval(T.ev_factor);
    if (r >= n) T.ev_u = mkc(0.0, 0.0);
    T.count_eval(T.req);
    T.req = GREQ_NONE;
}

// The request type a wave serves in this pass of its path loop.  The four groups of a wave
// run a slot program together only when their requests agree; groups with another pending
// request wait for a later pass.  Corrector requests come first and the predictor's Taylor
// requests last, so a group that has finished its step waits at the next step's first Taylor
// request until the other groups of its wave arrive there, and from then on the groups step
// together.  Every pass serves at least one group (the type is one that some group is
// waiting with), and a waiting group waits at most until each other group of its wave reaches
// its next Taylor request or ends its path.  Only the order of evaluations across groups
// changes, never a path's own sequence, so results do not change.
__device__ __forceinline__ int g_wave_pick(int req) {
    if (__ballot(req == GREQ_COND)) return GREQ_COND;
    if (__ballot(req == GREQ_DD)) return GREQ_DD;
    if (__ballot(req == GREQ_E)) return GREQ_E;
    if (__ballot(req == GREQ_EJ)) return GREQ_EJ;
    if (__ballot(req == GREQ_T3)) return GREQ_T3;
    if (__ballot(req == GREQ_T2)) return GREQ_T2;
    return GREQ_T1;
}

// PathResult(endgame_tracker, ...), src/endgame_tracker.jl:847-888 (residual from the last
// evaluation, which path_run made at the reported point)
template <int NN>
__device__ __forceinline__ void g_write_result(GT<NN>& T, const hcg_kargs& a, int k) {
    const bool ok = T.eg_code == GEG_SUCCESS;
    const double t = ok ? 0.0 : T.seg_t().re;
    const cplx s = ok ? T.solution : T.x;
    const double residual = T.norm_inf_v(T.res, false);
    if (T.r == 0) {
        hcg_path_result R;
        R.code = T.eg_code;
        R.singular = T.eg_singular;
        R.winding_number = T.eg_winding;
        R.extended_precision = T.extended_prec;
        R.extended_precision_used = T.used_ext;
        R.accepted_steps = T.accepted;
        R.rejected_steps = T.rejected;
        R.steps_eg = T.steps_eg;
        R.valuation_valid = !(t > T.eg_start);
        R.newton_iters = T.newton_iters;
        R.learned_proposed = T.learned_proposed;
        R.learned_accepted = T.learned_accepted;
        R.t = t;
        R.accuracy = T.eg_accuracy;
        R.residual = residual;
        R.cond = T.eg_cond;
        R.omega = T.omega;
        R.mu = T.mu;
        R.last_t = T.last_t;
        a.results[k] = R;
    }
    if (T.r < T.n) {
        const size_t o = (size_t)k * T.n + T.r;
        a.solutions[2 * o] = s.re;
        a.solutions[2 * o + 1] = s.im;
        a.last_points[2 * o] = T.last_point.re;
        a.last_points[2 * o + 1] = T.last_point.im;
        a.valuations[o] = T.val_x;
    }
}

template <int NN>
__device__ __forceinline__ void g_write_invalid(int r, int n, const hcg_kargs& a, int k) {
    if (r == 0) {
        hcg_path_result R;
        R.code = GEG_UNKNOWN;
        R.singular = R.winding_number = R.extended_precision = R.extended_precision_used = 0;
        R.accepted_steps = R.rejected_steps = R.steps_eg = R.valuation_valid = R.newton_iters = 0;
        R.learned_proposed = R.learned_accepted = 0;
        R.t = a.t1; R.accuracy = R.residual = R.cond = HCG_NAN; R.omega = R.mu = R.last_t = HCG_NAN;
        a.results[k] = R;
    }
    if (r < n) {
        const size_t o = (size_t)k * n + r;
        a.solutions[2 * o] = a.solutions[2 * o + 1] = HCG_NAN;
        a.last_points[2 * o] = a.last_points[2 * o + 1] = HCG_NAN;
        a.valuations[o] = HCG_NAN;
    }
}

// The work of one lane of a workgroup: lanes GL g .. GL g + GL - 1 (g = tid / GL) form a group that
// tracks paths from the launch's queue until it is empty; smem: the workgroup's LDS (the
// network's BF16 weights, then lds_per_path bytes per group).
template <int NN>
__device__ __forceinline__ void hcg_group_run(const hcg_multi_kargs& m, const hcg_gsys* sysd, int lds_per_path,
                                              char* smem, int tid) {
    const hcg_options* opts = m.sys[0].opts;
    const unsigned short* mw = stage_mlp(m.sys[0].mlp, opts, (unsigned short*)smem);
    GT<NN> T;
    T.r = g_lane();
    // the step network's weights take the front of LDS only when the learned step is on
    T.lds = smem + (mw ? HCG_GROUP_MLP_LDS : 0) + (tid / GL) * lds_per_path;
    T.mlp_w = mw;
    T.o = opts;
    T.cond_H_xd = 1.0;
    T.req = GREQ_NONE;
    T.last_point = mkc(0.0, 0.0);
    T.last_t = HCG_NAN;
    T.log_records = m.sys[0].log_records;
    T.log_count = m.sys[0].log_count;
    T.log_capacity = m.sys[0].log_capacity;
    double fl = 0.0;
    long long steps = 0, newton = 0, lacc = 0;
    int k = -1, s = 0;
    long evals = 0;
    while (true) {
        if (k < 0) {
            int kk = 0;
            if (T.r == 0) kk = atomicAdd(m.next_path, 1);
            k = gs_i(kk, 0);
            if (k >= m.total) break;
            s = 0;
            while (s + 1 < m.nsys && k >= m.offset[s + 1]) s++;
            const hcg_kargs& a = m.sys[s];
            const int kl = k - m.offset[s];
            T.sys = &sysd[s];
            T.n = T.sys->n;
            T.kind = T.sys->kind;
            T.path = kl;
            T.flops = 0.0;
            T.x = T.r < T.n ? mkc(a.starts[2 * ((size_t)kl * T.n + T.r)], a.starts[2 * ((size_t)kl * T.n + T.r) + 1])
                            : mkc(0.0, 0.0);
            T.p = T.q = nullptr;
            if (T.kind == HCG_PARAMETER) {
                const int gq = a.group_of_path ? a.group_of_path[kl] : 0;
                if (gq < 0 || gq >= a.ngroups) {  // device-resident input: reject out-of-range groups
                    g_write_invalid<NN>(T.r, T.n, a, kl);
                    k = -1;
                    continue;
                }
                T.p = a.start_params;
                T.q = a.target_params + 2 * T.sys->npar * gq;
            }
            T.co_path = T.co_egstep = T.co_tstep = T.co_newton = T.co_initn = T.co_ext = T.co_pred = T.co_tinit =
                T.co_refine = T.co_upprec = T.co_tstop = T.co_condat = T.co_singpost = T.co_result = 0;
            T.co_addsmp = T.co_sw = T.co_ainf = 0;
            evals = 0;
        }
        // a request the wave did not serve in the last pass is still pending
        const bool live = T.req != GREQ_NONE || T.path_run(m.sys[s].t1);
        if (live) {
            const int pick = g_wave_pick(T.req);
            if (T.req != GREQ_NONE && T.req != pick) continue;
            g_do_eval<NN>(T);
            // a hard bound on the evaluations of one path (far above what the step limits
            // allow), so that every group reaches the end of its loop whatever happens
            if (++evals <= 64L * (opts->max_steps + opts->max_endgame_steps) + 4096) continue;
            T.eg_code = GEG_UNKNOWN;
            T.ev_x = T.x;
            T.res = mkc(HCG_NAN, HCG_NAN);
        }
        g_write_result<NN>(T, m.sys[s], k - m.offset[s]);
        fl += T.flops;
        steps += T.accepted + T.rejected;
        newton += T.newton_iters;
        lacc += T.learned_accepted;
        k = -1;
    }
    if (T.r == 0) {
        unsigned long long* c = m.sys[0].counters;
        atomicAdd(&c[0], (unsigned long long)steps);
        atomicAdd(&c[1], (unsigned long long)newton);
        atomicAdd(&c[2], (unsigned long long)fl);
        atomicAdd(&c[3], (unsigned long long)lacc);
    }
}

