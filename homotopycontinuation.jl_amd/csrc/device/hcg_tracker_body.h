// Tracker body: everything of one system's tracker (Lane state, linear algebra,
// predictor, corrector, step control, endgame, result records, per-lane driver).
// No include guard on purpose: the multi-system program includes it once per
// system, inside a namespace, with that system's HCG_* macros defined.
#ifndef HCG_N
#error "HCG_N must be defined by the generated program"
#endif
#define NN HCG_N

// TrackerCode, src/tracker.jl:166-176
enum { TC_TRACKING = 0, TC_SUCCESS, TC_MAX_STEPS, TC_ACCURACY_LIMIT, TC_ILL_CONDITIONED, TC_INVALID_START,
       TC_INVALID_START_SINGULAR, TC_STEP_TOO_SMALL, TC_UNKNOWN };
// EndgameTrackerCode, src/endgame_tracker.jl:100-117
enum { EG_TRACKING = 0, EG_SUCCESS, EG_AT_INFINITY, EG_AT_ZERO, EG_ACCURACY_LIMIT, EG_INVALID_START,
       EG_INVALID_START_SINGULAR, EG_ILL_CONDITIONED, EG_MAX_STEPS, EG_MAX_EXT_STEPS, EG_MAX_WINDING,
       EG_STEP_TOO_SMALL, EG_UNKNOWN };
enum { NEWT_CONVERGED = 0, NEWT_TERMINATED, NEWT_MAX_ITERS, NEWT_SINGULARITY };

#define IDX(i, j) ((j) * NN + (i))

// FP64 flop counts of the dense linear algebra (complex: mul 6, muladd 8, div 11)
#define FL_LU (11.0 * (NN * (NN - 1) / 2) + 8.0 * ((NN - 1) * NN * (2 * NN - 1) / 6))
#define FL_SOLVE (8.0 * NN * (NN - 1) + 11.0 * NN)
#define FL_MIXED (80.0 * NN * NN + 26.0 * NN + FL_SOLVE)
#define FL_FIXED (8.0 * NN * NN + 4.0 * NN + FL_SOLVE)

struct Lane {
    const hcg_options* o;
    // ParameterHomotopy data
#if HCG_KIND == 1
    const double* p;  // start parameters (device memory, shared by all paths)
    const double* q;  // this path's target parameters
#endif
    // TrackerState
    cplx x[NN], xh[NN], xb[NN];
    cplx seg_start, seg_target; double seg_absD, seg_s, seg_s1; int seg_forward;
    double ds_prev, accuracy, omega, omega_prev, mu, tau;
    int extended_prec, used_ext, refined_ext, keep_ext;
    double w[NN];
    int code;
    int accepted, rejected, last_failed, ext_acc, ext_rej;
    // MatrixWorkspace
    cplx A[NN * NN], LU[NN * NN];
    int ipiv[NN];
    double rs[NN];
    int factorized, scaled;
    cplx dx[NN], r[NN], tmp[NN], rdx[NN];  // rdx: the MatrixWorkspace's own correction buffer (δx)
    // Predictor
    int hermite, order, winding;
    double trust_region, local_error;
    cplx tx0[NN], tx1[NN], tx2[NN], tx3[NN], ptx0[NN], ptx1[NN];
    cplx pred_t, pred_prev_t, pred_s, pred_prev_s;
    double txn[4];
    int newton_iters;
    // EndgameTrackerState
    int eg_code, singular_endgame;
    double val_x[NN], val_tx[NN], dval_x[NN], dval_tx[NN], vx2[NN], vx1[NN], ldx2[NN], ldx1[NN];
    double logt2, logt1;
    int eg_winding;
    cplx solution[NN];
    double eg_accuracy, eg_cond;
    int eg_singular;
    int steps_eg, ext_steps_eg_start;
    int jtzf0, jtzf1;
    cplx last_point[NN]; double last_t;
    double row_scaling[NN], col_scaling[NN];
    double ai_starts[NN], ai_tols[NN], ai_abs[NN], ai_conds[NN];
    cplx smp0[3][NN], smp1[3][NN];
    int smp_order[3];
    double sample_times[3], sample_conds[3];
    int singular_steps;
    cplx prediction[NN], prev_prediction[NN];
    // flattened singular_endgame_step!
    int sing_phase;
    double sing_t;
    // learned proposer bookkeeping
    int learned_proposed, learned_accepted, last_learned;
    double cond_H_xd, ds_plain;
    double eg_start;  // endgame_start, or an earlier start for paths flagged as heading to infinity
    float feat[8], inf_logit;
    const unsigned short* mlp_w;  // BF16 weights in LDS, or null (plain rule only)
    double flops;  // counted FP64 flops of this path
    int guard;     // loop iterations of the current path (each performs >= 1 tracker step)
    // step log
    float* log_records;
    int* log_count;
    int log_capacity, path;
};

// ------------------------------- homotopy -------------------------------

__device__ __noinline__ void H_eval(Lane& L, const cplx* x, cplx t, cplx* u) {
    L.flops += HCG_FL_EVAL;
#if HCG_KIND == 0
    cplx u2[NN];
    sysG_eval(x, u);
    sysF_eval(x, u2);
    cplx ts = c_mul(mkc(HCG_GAMMA_RE, HCG_GAMMA_IM), t), tt = mkc(1.0 - t.re, -t.im);
    for (int i = 0; i < NN; i++) u[i] = c_muladd(ts, u[i], c_mul(tt, u2[i]));
#else
    sysF_eval(x, L.p, L.q, t, u);
#endif
}

__device__ __noinline__ void H_eval_dd(Lane& L, const cplx* x, cplx t, cplx* u) {
    L.flops += HCG_FL_DD;
    cdd xd[NN], u1[NN];
    for (int i = 0; i < NN; i++) xd[i] = cdd_from(x[i]);
#if HCG_KIND == 0
    cdd u2[NN];
    sysG_eval_dd(xd, u1);
    sysF_eval_dd(xd, u2);
    cplx ts = c_mul(mkc(HCG_GAMMA_RE, HCG_GAMMA_IM), t), tt = mkc(1.0 - t.re, -t.im);
    for (int i = 0; i < NN; i++) u[i] = cdd_round(cdd_add(cdd_mul_c(ts, u1[i]), cdd_mul_c(tt, u2[i])));
#else
    sysF_eval_dd(xd, L.p, L.q, t, u1);
    for (int i = 0; i < NN; i++) u[i] = cdd_round(u1[i]);
#endif
}

// evaluate_and_jacobian!; U column-major
__device__ __noinline__ void H_eval_jac(Lane& L, const cplx* x, cplx t, cplx* u, cplx* U) {
    L.flops += HCG_FL_EJ;
#if HCG_KIND == 0
    // u = ts G + tt F entry by entry, in one generated function (no second Jacobian in memory)
    sysH_eval_jac(x, c_mul(mkc(HCG_GAMMA_RE, HCG_GAMMA_IM), t), mkc(1.0 - t.re, -t.im), u, U);
#else
    sysF_eval_jac(x, L.p, L.q, t, u, U);
#endif
}

// taylor!(u, Val(K), H, tx, t) with tx = (tx0, ..., tx_{K-1})
__device__ __noinline__ void H_taylor(Lane& L, int K, cplx t, cplx* u) {
    L.flops += K == 1 ? HCG_FL_T1 : (K == 2 ? HCG_FL_T2 : HCG_FL_T3);
#if HCG_KIND == 0
    // u = γ (G_{K-1} + t G_K) + ((1 - t) F_K - F_{K-1}) (K = 1: γ G - F), one generated
    // function per order, tail-called
    const cplx g = mkc(HCG_GAMMA_RE, HCG_GAMMA_IM), omt = mkc(1.0 - t.re, -t.im);
    if (K == 1) sysH_taylor1(L.tx0, g, t, omt, u);
    else if (K == 2) sysH_taylor2(L.tx0, L.tx1, g, t, omt, u);
    else sysH_taylor3(L.tx0, L.tx1, L.tx2, g, t, omt, u);
#else
    // tail calls (no local buffers): no register save and restore around the generated body
    if (K == 1) sysF_taylor1(L.tx0, L.p, L.q, t, u);
    else if (K == 2) sysF_taylor2(L.tx0, L.tx1, L.p, L.q, t, u);
    else sysF_taylor3(L.tx0, L.tx1, L.tx2, L.p, L.q, t, u);
#endif
}

// ----------------------------- linear algebra -----------------------------
// The matrix workspace holds the Jacobian A and its LU factors.  The factorization reads A
// (row-scaled by rs when `scaled` is set) and writes LU: no separate copy of A into LU.
__device__ __forceinline__ void ws_updated(Lane& L) { L.factorized = 0; L.scaled = 0; }

#if NN <= 12
// Loops over the compile-time size are unrolled for small systems, so that vectors and
// matrix columns are held in registers with compile-time indices.
#define HCG_UNROLL _Pragma("unroll")
// exchange v[k] and v[p] for a run-time p >= k without dynamic register indexing
__device__ __forceinline__ void reg_swap(cplx (&v)[NN], int k, int p) {
    cplx t = v[k], nk = t;
    HCG_UNROLL
    for (int m = k + 1; m < NN; m++) {
        bool s = p == m;
        cplx vm = v[m];
        nk = mkc(s ? vm.re : nk.re, s ? vm.im : nk.im);  // component selects keep v in registers
        v[m] = mkc(s ? t.re : vm.re, s ? t.im : vm.im);
    }
    v[k] = nk;
}
// lu!, src/linear_algebra.jl:130-184, in left-looking (column by column) order.  Every entry
// receives the same updates with the same operands in the same order as in the reference's
// right-looking loop: a column's row interchanges are applied before its updates, and the
// multipliers of earlier columns are interchanged when a pivot is chosen.  So the factors
// and pivots are identical.  The working column and the multipliers are register arrays with
// compile-time indices (interchanges by selects); A is read once and LU written once.
__device__ __noinline__ void lu_factor(Lane& L) {
    const cplx* A = L.A;
    cplx* F = L.LU;
    L.flops += FL_LU;
    const bool sc = L.scaled;
    double r[NN];
    HCG_UNROLL
    for (int i = 0; i < NN; i++) r[i] = sc ? L.rs[i] : 1.0;
    int piv[NN];
    cplx M[NN][NN];  // M[k][i]: multiplier of row i in column k (i > k)
    HCG_UNROLL
    for (int j = 0; j < NN; j++) {
        cplx c[NN];
        HCG_UNROLL
        for (int i = 0; i < NN; i++) { cplx a = A[IDX(i, j)]; c[i] = sc ? c_scale(r[i], a) : a; }
        HCG_UNROLL
        for (int k = 0; k < j; k++) reg_swap(c, k, piv[k]);
        HCG_UNROLL
        for (int k = 0; k < j; k++) {
            cplx ukj = c[k];
            HCG_UNROLL
            for (int i = k + 1; i < NN; i++) c[i] = c_submul(M[k][i], ukj, c[i]);
        }
        int kp = j;
        double amax = c_abs2(c[j]);
        HCG_UNROLL
        for (int i = j + 1; i < NN; i++) {
            double absi = c_abs2(c[i]);
            if (absi > amax) { kp = i; amax = absi; }
        }
        piv[j] = kp;
        L.ipiv[j] = kp;
        reg_swap(c, j, kp);
        HCG_UNROLL
        for (int k = 0; k < j; k++) reg_swap(M[k], j, kp);
        if (amax != 0.0) {
            cplx d = c[j];
            HCG_UNROLL
            for (int i = j + 1; i < NN; i++) c[i] = c_div(c[i], d);
        }
        HCG_UNROLL
        for (int i = 0; i <= j; i++) F[IDX(i, j)] = c[i];  // U part: final
        HCG_UNROLL
        for (int i = j + 1; i < NN; i++) M[j][i] = c[i];
    }
    HCG_UNROLL
    for (int k = 0; k < NN; k++) {
        HCG_UNROLL
        for (int i = k + 1; i < NN; i++) F[IDX(i, k)] = M[k][i];
    }
    L.factorized = 1;
}
// lu_ldiv!, src/linear_algebra.jl:268-316: x = LU \ (P (rs .* b)) (rs == nullptr: no scaling);
// x may alias b.  The vector stays in registers.
__device__ __noinline__ void lu_solve_rs(Lane& L, cplx* x, const cplx* b, const double* rs) {
    const cplx* F = L.LU;
    L.flops += FL_SOLVE;
    cplx v[NN];
    HCG_UNROLL
    for (int i = 0; i < NN; i++) v[i] = rs ? c_scale(rs[i], b[i]) : b[i];
    HCG_UNROLL
    for (int i = 0; i < NN; i++) reg_swap(v, i, L.ipiv[i]);
    HCG_UNROLL
    for (int j = 0; j < NN; j++) {
        cplx xj = v[j];
        HCG_UNROLL
        for (int i = j + 1; i < NN; i++) v[i] = c_submul(F[IDX(i, j)], xj, v[i]);
    }
    HCG_UNROLL
    for (int j = NN - 1; j >= 0; j--) {
        cplx xj = c_div(v[j], F[IDX(j, j)]);
        v[j] = xj;
        HCG_UNROLL
        for (int i = 0; i < j; i++) v[i] = c_submul(F[IDX(i, j)], xj, v[i]);
    }
    HCG_UNROLL
    for (int i = 0; i < NN; i++) x[i] = v[i];
}
#else
#define HCG_UNROLL
// lu!, src/linear_algebra.jl:130-184 (right-looking, in place on LU = diag(rs) A)
__device__ __noinline__ void lu_factor(Lane& L) {
    cplx* F = L.LU;
    L.flops += FL_LU;
    if (L.scaled) for (int k = 0; k < NN * NN; k++) F[k] = c_scale(L.rs[k % NN], L.A[k]);
    else for (int k = 0; k < NN * NN; k++) F[k] = L.A[k];
    for (int k = 0; k < NN; k++) {
        int kp = k;
        double amax = c_abs2(F[IDX(k, k)]);
        for (int i = k + 1; i < NN; i++) {
            double absi = c_abs2(F[IDX(i, k)]);
            if (absi > amax) { kp = i; amax = absi; }
        }
        L.ipiv[k] = kp;
        if (amax != 0.0) {
            if (k != kp)
                for (int i = 0; i < NN; i++) { cplx t = F[IDX(k, i)]; F[IDX(k, i)] = F[IDX(kp, i)]; F[IDX(kp, i)] = t; }
            cplx Fkk = F[IDX(k, k)];
            for (int i = k + 1; i < NN; i++) F[IDX(i, k)] = c_div(F[IDX(i, k)], Fkk);
        }
        for (int j = k + 1; j < NN; j++) {
            cplx Fkj = F[IDX(k, j)];
            for (int i = k + 1; i < NN; i++) F[IDX(i, j)] = c_submul(F[IDX(i, k)], Fkj, F[IDX(i, j)]);
        }
    }
    L.factorized = 1;
}
// lu_ldiv!, src/linear_algebra.jl:268-316: x = LU \ (P (rs .* b)) (rs == nullptr: no scaling)
__device__ __noinline__ void lu_solve_rs(Lane& L, cplx* x, const cplx* b, const double* rs) {
    const cplx* F = L.LU;
    L.flops += FL_SOLVE;
    if (rs) for (int i = 0; i < NN; i++) x[i] = c_scale(rs[i], b[i]);
    else if (x != b) for (int i = 0; i < NN; i++) x[i] = b[i];
    for (int i = 0; i < NN; i++) { int p = L.ipiv[i]; if (i != p) { cplx t = x[i]; x[i] = x[p]; x[p] = t; } }
    for (int j = 0; j < NN; j++) {
        cplx xj = x[j];
        for (int i = j + 1; i < NN; i++) x[i] = c_submul(F[IDX(i, j)], xj, x[i]);
    }
    for (int j = NN - 1; j >= 0; j--) {
        cplx xj = c_div(x[j], F[IDX(j, j)]);
        x[j] = xj;
        for (int i = 0; i < j; i++) x[i] = c_submul(F[IDX(i, j)], xj, x[i]);
    }
}
#endif
__device__ __forceinline__ void lu_solve(Lane& L, cplx* x) { lu_solve_rs(L, x, x, nullptr); }
// lu_ldiv_adj!, src/linear_algebra.jl:318-354
__device__ __noinline__ void lu_solve_adj(Lane& L, cplx* x) {
    const cplx* A = L.LU;
    L.flops += FL_SOLVE;
    for (int j = 0; j < NN; j++) {
        cplx z = x[j];
        for (int i = 0; i < j; i++) z = c_submul(c_conj(A[IDX(i, j)]), x[i], z);
        x[j] = c_div(z, c_conj(A[IDX(j, j)]));
    }
    for (int j = NN - 1; j >= 0; j--) {
        cplx z = x[j];
        for (int i = NN - 1; i > j; i--) z = c_submul(c_conj(A[IDX(i, j)]), x[i], z);
        x[j] = z;
    }
    for (int i = NN - 1; i >= 0; i--) { int p = L.ipiv[i]; if (i != p) { cplx t = x[i]; x[i] = x[p]; x[p] = t; } }
}
// LA.ldiv!(x, WS, b), src/linear_algebra.jl:389-408
__device__ __noinline__ void ws_ldiv(Lane& L, cplx* x, const cplx* b) {
    if (NN == 1) { x[0] = c_div(b[0], L.A[0]); return; }
    if (!L.factorized) lu_factor(L);
    lu_solve_rs(L, x, b, L.scaled ? L.rs : nullptr);
}
// skeel_row_scaling!, src/linear_algebra.jl:432-459
__device__ __noinline__ void skeel_row_scaling(double* d, const cplx* A, const double* c, double threshold) {
    double r[NN];  // row sums in registers, written to d once
    HCG_UNROLL
    for (int i = 0; i < NN; i++) r[i] = 0.0;
    HCG_UNROLL
    for (int j = 0; j < NN; j++) {
        double cj = c[j];
        HCG_UNROLL
        for (int i = 0; i < NN; i++) r[i] += c_fabs(A[IDX(i, j)]) * cj;
    }
    double m = r[0];
    HCG_UNROLL
    for (int i = 1; i < NN; i++) m = jmax(m, r[i]);
    double s = threshold + m;
    HCG_UNROLL
    for (int i = 0; i < NN; i++) {
        int e = 0;
        if (isfinite(r[i])) frexp(r[i], &e);
        d[i] = (double)e < s ? 1.0 : ldexp(1.0, -e);
    }
}
// weighted / plain inf-norms, src/norm.jl:150-234 (w == nullptr: InfNorm)
__device__ __noinline__ double norm_inf(const cplx* x, const double* w) {
    double dmax = w ? c_abs2(c_divr(x[0], w[0])) : c_abs2(x[0]);
    for (int i = 1; i < NN; i++) dmax = max_fast(dmax, w ? c_abs2(c_divr(x[i], w[i])) : c_abs2(x[i]));
    double d = sqrt(dmax);
    if (isinf(d)) {
        dmax = w ? c_hypot(c_divr(x[0], w[0])) : c_hypot(x[0]);
        for (int i = 1; i < NN; i++) dmax = jmax(dmax, w ? c_hypot(c_divr(x[i], w[i])) : c_hypot(x[i]));
        return dmax;
    }
    return d;
}
__device__ __noinline__ double dist_inf(const cplx* x, const cplx* y, const double* w) {
    double dmax = w ? c_abs2(c_divr(c_sub(x[0], y[0]), w[0])) : c_abs2(c_sub(x[0], y[0]));
    for (int i = 1; i < NN; i++) dmax = max_fast(dmax, w ? c_abs2(c_divr(c_sub(x[i], y[i]), w[i])) : c_abs2(c_sub(x[i], y[i])));
    double d = sqrt(dmax);
    if (isinf(d)) {
        if (w) {
            dmax = c_hypot(c_divr(c_sub(x[0], y[0]), w[0]));
            for (int i = 1; i < NN; i++) dmax = jmax(dmax, c_hypot(c_divr(c_sub(x[i], y[i]), w[i])));
        } else {  // the reference's unweighted fallback mixes abs and abs2 (src/norm.jl:160-163)
            dmax = c_hypot(c_sub(x[0], y[0]));
            for (int i = 1; i < NN; i++) dmax = jmax(dmax, c_abs2(c_sub(x[i], y[i])));
        }
        return dmax;
    }
    return d;
}
// mixed_precision_iterative_refinement!, src/linear_algebra.jl:528-544 (residual in double-double)
__device__ __noinline__ double mixed_refine(Lane& L, cplx* x, const cplx* b, const double* w) {
    L.flops += FL_MIXED - FL_SOLVE;
    cdd r[NN];
    cdd z = mkcdd(mkdd(0.0, 0.0), mkdd(0.0, 0.0));
    for (int i = 0; i < NN; i++) r[i] = z;
    for (int j = 0; j < NN; j++) {
        cdd xj = cdd_from(x[j]);
        for (int i = 0; i < NN; i++) r[i] = cdd_add(r[i], cdd_mul_c(L.A[IDX(i, j)], xj));
    }
    for (int i = 0; i < NN; i++) L.tmp[i] = cdd_round(cdd_sub_c(r[i], b[i]));
    ws_ldiv(L, L.rdx, L.tmp);
    for (int i = 0; i < NN; i++) x[i] = c_sub(x[i], L.rdx[i]);
    return norm_inf(L.rdx, w) / norm_inf(x, w);
}
// fixed_precision_iterative_refinement!, src/linear_algebra.jl:553-567
__device__ __noinline__ double fixed_refine(Lane& L, cplx* x, const cplx* b, const double* w) {
    L.flops += FL_FIXED - FL_SOLVE;
    cplx acc[NN];  // A x accumulated in registers
    HCG_UNROLL
    for (int i = 0; i < NN; i++) acc[i] = mkc(0.0, 0.0);
    HCG_UNROLL
    for (int j = 0; j < NN; j++) {
        cplx xj = x[j];
        HCG_UNROLL
        for (int i = 0; i < NN; i++) acc[i] = c_muladd(L.A[IDX(i, j)], xj, acc[i]);
    }
    HCG_UNROLL
    for (int i = 0; i < NN; i++) L.tmp[i] = c_sub(acc[i], b[i]);
    ws_ldiv(L, L.rdx, L.tmp);
    for (int i = 0; i < NN; i++) x[i] = c_sub(x[i], L.rdx[i]);
    return norm_inf(L.rdx, w) / norm_inf(x, w);
}
// iterative_refinement!, src/linear_algebra.jl:864-885
__device__ __noinline__ double iterative_refinement(Lane& L, cplx* x, const cplx* b, const double* w, double tol, int max_iters) {
    double d = mixed_refine(L, x, b, w);
    for (int i = 2; i <= max_iters; i++) {
        double d2 = mixed_refine(L, x, b, w);
        if (d2 < tol) return d2;
        else if (d2 > 0.5 * d) return d2;
        d = d2;
    }
    return d;
}
// Jacobian ldiv! with WeightedNorm (src/linear_algebra.jl:846-862)
__device__ __noinline__ void jac_ldiv_w(Lane& L, cplx* x, const cplx* b) {
    if (!L.factorized) {
        skeel_row_scaling(L.rs, L.A, L.w, -30.0);  // lu_factor applies it to the rows of A
        L.scaled = 1;
    }
    ws_ldiv(L, x, b);
}
// inverse_inf_norm_est, src/linear_algebra.jl:585-682
__device__ __noinline__ double inverse_inf_norm_est(Lane& L, const double* dl, const double* dr) {
    if (!L.factorized) lu_factor(L);
    const double* rsc = L.scaled ? L.rs : nullptr;
    cplx y[NN];
    double x[NN];
    for (int i = 0; i < NN; i++) { x[i] = 1.0 / NN; if (dr) x[i] /= dr[i]; }
    for (int i = 0; i < NN; i++) y[i] = mkc(x[i], 0.0);
    lu_solve_adj(L, y);
    for (int i = 0; i < NN; i++) { if (dl) y[i] = c_divr(y[i], dl[i]); if (rsc) y[i] = c_scale(rsc[i], y[i]); }
    double g = 0.0;
    for (int i = 0; i < NN; i++) g += c_fabs(y[i]);
    for (int i = 0; i < NN; i++) {
        y[i] = c_divr(y[i], c_fabs(y[i]));
        if (dl) y[i] = c_divr(y[i], dl[i]);
        if (rsc) y[i] = c_divr(y[i], rsc[i]);
    }
    lu_solve(L, y);
    for (int i = 0; i < NN; i++) x[i] = dr ? y[i].re / dr[i] : y[i].re;
    int k = 2;
    while (true) {
        int j = 0;
        double mx = fabs(x[0]);
        for (int i = 1; i < NN; i++) if (fabs(x[i]) > mx) { j = i; mx = fabs(x[i]); }
        for (int i = 0; i < NN; i++) x[i] = 0.0;
        x[j] = 1.0;
        if (dr) for (int i = 0; i < NN; i++) x[i] /= dr[i];
        for (int i = 0; i < NN; i++) y[i] = mkc(x[i], 0.0);
        lu_solve_adj(L, y);
        for (int i = 0; i < NN; i++) { if (dl) y[i] = c_divr(y[i], dl[i]); if (rsc) y[i] = c_scale(rsc[i], y[i]); }
        double gb = g;
        g = 0.0;
        for (int i = 0; i < NN; i++) g += c_fabs(y[i]);
        if (g <= gb) { g = gb; break; }
        for (int i = 0; i < NN; i++) {
            y[i] = c_divr(y[i], c_fabs(y[i]));
            if (dl) y[i] = c_divr(y[i], dl[i]);
            if (rsc) y[i] = c_divr(y[i], rsc[i]);
        }
        lu_solve(L, y);
        for (int i = 0; i < NN; i++) x[i] = dr ? y[i].re / dr[i] : y[i].re;
        k += 1;
        double xn = fabs(x[0]);
        for (int i = 1; i < NN; i++) xn = jmax(xn, fabs(x[i]));
        if (x[j] == xn || k > 2) break;
    }
    return nanmin(g, HCG_INF);
}
// inf_norm, src/linear_algebra.jl:684-707
__device__ __noinline__ double inf_norm_ws(const Lane& L, const double* dl, const double* dr) {
    double nrm = -HCG_INF;
    for (int i = 0; i < NN; i++) {
        double ni = 0.0;
        for (int j = 0; j < NN; j++) ni += dr ? c_fabs(L.A[IDX(i, j)]) * dr[j] : c_fabs(L.A[IDX(i, j)]);
        if (dl) ni *= dl[i];
        nrm = max_fast(nrm, ni);
    }
    return nrm;
}
// LA.cond(WS, d_l, d_r), src/linear_algebra.jl:745-774
__device__ __noinline__ double ws_cond(Lane& L, const double* dl, const double* dr) {
    if (NN == 1) {
        double a = c_hypot(L.A[0]);
        if (!dl && !dr) return 1.0 / a;
        if (!dl) return 1.0 / (a * dr[0]);
        if (!dr) return 1.0 / (dl[0] * a);
        return 1.0 / (dl[0] * a * dr[0]);
    }
    return inverse_inf_norm_est(L, dl, dr) * inf_norm_ws(L, dl, dr);
}
// LinearAlgebra.rank(J, rtol): one-sided Jacobi singular values of a copy of J (in LU)
__device__ __noinline__ int rank_rtol(Lane& L, double rtol) {
    cplx* A = L.LU;
    for (int k = 0; k < NN * NN; k++) A[k] = L.A[k];
    for (int sweep = 0; sweep < 60; sweep++) {
        double off = 0.0;
        for (int p = 0; p < NN; p++)
            for (int q = p + 1; q < NN; q++) {
                double app = 0, aqq = 0; cplx apq = mkc(0, 0);
                for (int i = 0; i < NN; i++) {
                    app += c_abs2(A[IDX(i, p)]); aqq += c_abs2(A[IDX(i, q)]);
                    apq = c_add(apq, c_mul(c_conj(A[IDX(i, p)]), A[IDX(i, q)]));
                }
                double g = sqrt(c_abs2(apq));
                if (g == 0 || g <= 1e-300) continue;
                double rel = g / sqrt(app * aqq);
                if (rel > off) off = rel;
                if (rel < 1e-15) continue;
                double zeta = (aqq - app) / (2 * g);
                double tt = (zeta >= 0 ? 1.0 : -1.0) / (fabs(zeta) + sqrt(1 + zeta * zeta));
                double cs = 1 / sqrt(1 + tt * tt), sn = cs * tt;
                cplx ph = mkc(apq.re / g, apq.im / g);
                for (int i = 0; i < NN; i++) {
                    cplx ap = A[IDX(i, p)], aq = A[IDX(i, q)];
                    cplx aqr = c_mul(aq, c_conj(ph));
                    A[IDX(i, p)] = c_sub(c_scale(cs, ap), c_scale(sn, aqr));
                    A[IDX(i, q)] = c_mul(c_add(c_scale(sn, ap), c_scale(cs, aqr)), ph);
                }
            }
        if (off < 1e-15) break;
    }
    double smax = 0, s[NN];
    for (int p = 0; p < NN; p++) { double v = 0; for (int i = 0; i < NN; i++) v += c_abs2(A[IDX(i, p)]); s[p] = sqrt(v); if (s[p] > smax) smax = s[p]; }
    int r = 0;
    for (int p = 0; p < NN; p++) if (s[p] > rtol * smax) r++;
    L.factorized = 0;
    return r;
}

// ----------------------------- segment stepper -----------------------------
__device__ __forceinline__ void seg_init(Lane& L, cplx start, cplx target) {
    L.seg_start = start; L.seg_target = target;
    cplx d = c_sub(target, start);
    L.seg_absD = c_hypot(d);
    L.seg_forward = c_hypot(start) < c_hypot(target);
    L.seg_s = L.seg_s1 = L.seg_forward ? 0.0 : L.seg_absD;
}
__device__ __forceinline__ bool seg_is_done(const Lane& L) { return L.seg_forward ? L.seg_s == L.seg_absD : L.seg_s == 0.0; }
__device__ __forceinline__ void seg_propose(Lane& L, double ds) {
    if (L.seg_forward) L.seg_s1 = jmin(L.seg_s + ds, L.seg_absD);
    else L.seg_s1 = jmax(L.seg_s - ds, 0.0);
}
__device__ __forceinline__ double seg_dist(const Lane& L) { return L.seg_forward ? L.seg_absD - L.seg_s : L.seg_s; }
__device__ __forceinline__ double seg_ds(const Lane& L) { return L.seg_forward ? L.seg_s1 - L.seg_s : L.seg_s - L.seg_s1; }
__device__ __forceinline__ cplx seg_th(const Lane& L, double s) {
    cplx st = L.seg_start, tg = L.seg_target; double D = L.seg_absD;
    if (L.seg_forward) {
        if (s == 0.0) return st;
        if (s == D) return tg;
        return c_add(st, c_scale(s / D, c_sub(tg, st)));
    }
    if (s == D) return st;
    if (s == 0.0) return tg;
    return c_add(tg, c_scale(s / D, c_sub(st, tg)));
}
__device__ __forceinline__ cplx seg_t(const Lane& L) { return seg_th(L, L.seg_s); }
__device__ __forceinline__ cplx seg_t1(const Lane& L) { return seg_th(L, L.seg_s1); }
__device__ __forceinline__ cplx seg_dt(const Lane& L) {
    if (L.seg_forward) return c_scale((L.seg_s1 - L.seg_s) / L.seg_absD, c_sub(L.seg_target, L.seg_start));
    return c_scale((L.seg_s - L.seg_s1) / L.seg_absD, c_sub(L.seg_target, L.seg_start));
}

// ----------------------------- weighted norm -----------------------------
__device__ __noinline__ void wnorm_init(Lane& L) {
    const hcg_options* o = L.o;
    double pn = norm_inf(L.x, nullptr);
    for (int i = 0; i < NN; i++) {
        double wi = c_fabs(L.x[i]);
        if (wi < o->scale_min * pn) wi = o->scale_min * pn;
        else if (wi > o->scale_max * pn) wi = o->scale_max * pn;
        L.w[i] = jmax(wi, o->scale_abs_min);
    }
}
__device__ __noinline__ void wnorm_update(Lane& L, const cplx* x) {
    const hcg_options* o = L.o;
    double nx = norm_inf(x, L.w);
    for (int i = 0; i < NN; i++) {
        double wi = (c_fabs(x[i]) + L.w[i]) / 2;
        if (wi < o->scale_min * nx) wi = o->scale_min * nx;
        else if (wi > o->scale_max * nx) wi = o->scale_max * nx;
        if (isfinite(wi)) L.w[i] = jmax(wi, o->scale_abs_min);
    }
}

// ----------------------------- Newton corrector -----------------------------
struct NewtRes { int code; double accuracy; int iters; double omega, theta, mu_low, norm_dx0; };

__device__ __forceinline__ NewtRes mkres(int code, double acc, int iters, double omega, double theta, double mu_low, double ndx0) {
    NewtRes R; R.code = code; R.accuracy = acc; R.iters = iters; R.omega = omega; R.theta = theta; R.mu_low = mu_low; R.norm_dx0 = ndx0;
    return R;
}

// newton!, src/newton_corrector.jl:80-205 (x̄ = L.xb; every iterate aliases it)
__device__ __noinline__ NewtRes newton(Lane& L, const cplx* x0, cplx t, double mu, double omega, int ext, int accurate_mu, int first_correction) {
    double a = L.o->a, ha = h_of(a);
    cplx* xb = L.xb;
    if (xb != x0) for (int k = 0; k < NN; k++) xb[k] = x0[k];
    double mu_low = HCG_NAN, theta = HCG_NAN, ndx = HCG_NAN, ndx_prev = HCG_NAN, ndx0 = HCG_NAN;
    double abar = a;
    for (int i = 0; i <= 10; i++) {
        L.newton_iters++;
        H_eval_jac(L, xb, t, L.r, L.A);
        if (ext) H_eval_dd(L, xb, t, L.r);
        ws_updated(L);
        jac_ldiv_w(L, L.dx, L.r);
        if (ext) iterative_refinement(L, L.dx, L.r, L.w, abar * abar, 3);
        ndx = norm_inf(L.dx, L.w);
        if (isnan(ndx)) return mkres(NEWT_SINGULARITY, norm_inf(L.dx, L.w), i + 1, omega, theta, mu_low, ndx0);
        for (int k = 0; k < NN; k++) xb[k] = c_sub(xb[k], L.dx[k]);
        if (i == 0) ndx0 = ndx;
        if (i == 1) omega = 2 * ndx / (ndx_prev * ndx_prev);
        if (i >= 1) theta = ndx / ndx_prev;
        if ((i >= 1 && theta > abar) || (i == 0 && !first_correction && 0.125 * ndx0 * omega > ha)) {
            return mkres(NEWT_TERMINATED, norm_inf(L.dx, L.w), i + 1, omega, theta, mu_low, ndx0);
        } else if (omega * (ndx * ndx) < 2 * mu * sqrt(1 - 2 * ha)) {
            H_eval_jac(L, xb, t, L.r, L.A);
            ws_updated(L);
            if (ext) {
                ws_ldiv(L, L.dx, L.r);
                mu_low = norm_inf(L.dx, L.w);
                H_eval_dd(L, xb, t, L.r);
            }
            ws_ldiv(L, L.dx, L.r);
            if (ext) iterative_refinement(L, L.dx, L.r, L.w, abar * abar, 3);
            for (int k = 0; k < NN; k++) xb[k] = c_sub(xb[k], L.dx[k]);
            double ndx1 = norm_inf(L.dx, L.w);
            if (isnan(ndx1)) return mkres(NEWT_SINGULARITY, norm_inf(L.dx, L.w), i + 1, omega, theta, mu_low, ndx0);
            if (ndx1 > sqrt(ndx)) {
                theta = ndx1 / ndx;
                return mkres(NEWT_TERMINATED, ndx1, i + 2, omega, theta, mu_low, ndx0);
            }
            if (ndx1 > 2 * mu && ext) {
                H_eval_dd(L, xb, t, L.r);
                ws_ldiv(L, L.dx, L.r);
                ndx = ndx1;
                ndx1 = norm_inf(L.dx, L.w);
                mu = ndx1;
            } else if (ndx1 > 2 * mu || accurate_mu) {
                H_eval(L, xb, t, L.r);
                ws_ldiv(L, L.dx, L.r);
                mu = norm_inf(L.dx, L.w);
            } else {
                mu = ndx1;
            }
            if (i == 0) {
                double wb = 2 * ndx / (ndx1 * ndx1);
                if (wb < omega) omega = wb; else omega *= 0.25;
            }
            return mkres(NEWT_CONVERGED, mu, i + 2, omega, theta, mu_low, ndx0);
        }
        ndx_prev = ndx;
        if (i >= 1) abar *= abar;
    }
    return mkres(NEWT_MAX_ITERS, mu, 11, omega, theta, mu_low, ndx0);
}

// init_newton!, src/newton_corrector.jl:207-286
__device__ __noinline__ int init_newton(Lane& L, cplx t, int ext, double* omega_out, double* mu_out) {
    double a = L.o->a;
    cplx* xb = L.xb;
    const cplx* x0 = L.x;
    H_eval_jac(L, x0, t, L.r, L.A);
    if (ext) H_eval_dd(L, x0, t, L.r);
    ws_updated(L);
    jac_ldiv_w(L, L.dx, L.r);
    double v = norm_inf(L.dx, L.w) + DEPS;
    int valid = 0;
    double omega = HCG_NAN, mu = HCG_NAN;
    double eps = sqrt(v);
    for (int k = 1; k <= 3; k++) {
        for (int i = 0; i < NN; i++) xb[i] = mkc(x0[i].re + eps * L.w[i], x0[i].im);
        H_eval_jac(L, xb, t, L.r, L.A);
        if (ext) H_eval_dd(L, xb, t, L.r);
        ws_updated(L);
        jac_ldiv_w(L, L.dx, L.r);
        for (int i = 0; i < NN; i++) xb[i] = c_sub(xb[i], L.dx[i]);
        double ndx0 = norm_inf(L.dx, L.w);
        if (ext) H_eval_dd(L, xb, t, L.r);
        else H_eval(L, xb, t, L.r);
        jac_ldiv_w(L, L.dx, L.r);
        for (int i = 0; i < NN; i++) xb[i] = c_sub(xb[i], L.dx[i]);
        double ndx1 = norm_inf(L.dx, L.w) + DEPS;
        if (ndx1 < a * ndx0) {
            omega = 2 * ndx1 / (ndx0 * ndx0);
            mu = ndx1;
            double a7 = ipow(a, 7);
            if (omega * mu > a7) {
                NewtRes rr = newton(L, xb, t, a7 / omega, omega, ext, 1, 0);
                if (rr.code == NEWT_CONVERGED) { valid = 1; omega = rr.omega; mu = rr.accuracy; }
                else valid = 0;
            } else {
                valid = 1;
                break;
            }
        } else {
            eps *= sqrt(eps);
        }
    }
    *omega_out = omega; *mu_out = mu;
    return valid;
}

// extended_prec_refinement_step!, src/newton_corrector.jl:55-78; writes x̄ = x - Δx into xout
__device__ __noinline__ double ext_refine_step(Lane& L, cplx* xout, const cplx* x, cplx t, int simple) {
    cplx xin[NN];
    for (int i = 0; i < NN; i++) xin[i] = x[i];
    H_eval_jac(L, xin, t, L.r, L.A);
    H_eval_dd(L, xin, t, L.r);
    ws_updated(L);
    jac_ldiv_w(L, L.dx, L.r);
    iterative_refinement(L, L.dx, L.r, L.w, 1e-8, 3);
    for (int i = 0; i < NN; i++) xout[i] = c_sub(xin[i], L.dx[i]);
    if (simple) {
        H_eval_dd(L, xout, t, L.r);
        jac_ldiv_w(L, L.dx, L.r);
    }
    return norm_inf(L.dx, L.w);
}

// ------------------------------- predictor -------------------------------
__device__ cplx t_to_s_plane(cplx t, int m) {
    double r = c_fabs(t);
    if (t.im == 0.0 && t.re > 0.0) return mkc(nthroot(r, m), 0.0);
    double th = atan2(t.im, t.re);
    th = fmod(th, 2 * HCG_PI); if (th < 0) th += 2 * HCG_PI;
    double rr = nthroot(r, m);
    return mkc(rr * cos(th / m), rr * sin(th / m));
}
__device__ __forceinline__ cplx cpowi(cplx z, int n) { cplx r = mkc(1.0, 0.0); for (int i = 0; i < n; i++) r = c_mul(r, z); return r; }

// update!(predictor, ...), src/predictor.jl:158-284; xhat may be nullptr
__device__ __noinline__ void pred_update(Lane& L, cplx t, const cplx* xhat) {
    cplx* u = L.r;
    cplx xt[NN];
    int m = L.winding;
    for (int i = 0; i < NN; i++) { L.ptx0[i] = L.tx0[i]; L.ptx1[i] = L.tx1[i]; }
    L.pred_prev_t = L.pred_t; L.pred_t = t;
    if (m > 1) { L.pred_prev_s = L.pred_s; L.pred_s = t_to_s_plane(t, m); }
    if (!xhat) L.local_error = HCG_NAN;
    else L.local_error = dist_inf(xhat, L.x, L.w) / ipow(c_fabs(c_sub(t, L.pred_prev_t)), L.order);
    for (int i = 0; i < NN; i++) L.tx0[i] = L.x[i];
    L.txn[0] = norm_inf(L.x, L.w);

    H_taylor(L, 1, t, u);
    for (int i = 0; i < NN; i++) u[i] = c_neg(u[i]);
    cplx ub[NN];
    for (int i = 0; i < NN; i++) ub[i] = u[i];
    ws_ldiv(L, xt, ub);
    double delta = fixed_refine(L, xt, ub, L.w);
    L.cond_H_xd = delta / DEPS;
    if (delta > 1e-10) iterative_refinement(L, xt, ub, nullptr, 1e-10, 5);
    L.txn[1] = norm_inf(xt, L.w);
    for (int i = 0; i < NN; i++) L.tx1[i] = xt[i];
    if (m > 1) {
        L.hermite = 1;
        L.order = 4;
        L.trust_region = L.txn[0] / L.txn[1];
        if (isnan(L.local_error)) { double q = L.txn[1] / L.txn[0]; L.local_error = q * q * q; }
        return;
    }
    H_taylor(L, 2, t, u);
    for (int i = 0; i < NN; i++) ub[i] = c_neg(u[i]);
    ws_ldiv(L, xt, ub);
    if (delta > 1e-10) iterative_refinement(L, xt, ub, L.w, 1e-10, 4);
    L.txn[2] = norm_inf(xt, L.w);
    for (int i = 0; i < NN; i++) L.tx2[i] = xt[i];

    H_taylor(L, 3, t, u);
    for (int i = 0; i < NN; i++) ub[i] = c_neg(u[i]);
    ws_ldiv(L, xt, ub);
    if (delta > 1e-4) iterative_refinement(L, xt, ub, L.w, 1e-4, 3);
    L.txn[3] = norm_inf(xt, L.w);
    for (int i = 0; i < NN; i++) L.tx3[i] = xt[i];

    double tau = HCG_INF;
    for (int i = 0; i < NN; i++) {
        double c1 = c_fabs(L.tx1[i]), c2 = c_fabs(L.tx2[i]), c3 = c_fabs(L.tx3[i]);
        double lam = jmax(1e-6, c1);
        c1 /= lam; c2 /= lam * lam; c3 /= lam * lam * lam;
        double tol = 1e-14 * jmax(jmax(c1, c2), c3);
        if (!((c1 <= tol && c2 <= tol && c3 <= tol) || c2 <= tol)) {
            double ti = (c2 / c3) / lam;
            if (ti < tau) tau = ti;
        }
    }
    if (!isfinite(tau)) tau = L.txn[2] / L.txn[3];
    if (!isfinite(tau)) {
        double mx = L.txn[0];
        for (int k = 1; k < 4; k++) mx = jmax(mx, L.txn[k]);
        tau = L.txn[0] / mx;
    }
    L.hermite = 0;
    L.order = 4;
    L.trust_region = tau;
    if (isnan(L.local_error)) { double it = 1.0 / tau; double it2 = it * it; L.local_error = it2 * it2; }
}

__device__ __noinline__ void hermite_real(cplx* out, const cplx* a0, const cplx* a1, double t0, const cplx* b0, const cplx* b1, double t1, double t) {
    double s = (t - t0) / (t1 - t0);
    double oms = 1 - s;
    double h00 = (1 + 2 * s) * (oms * oms), h10 = (t - t0) * (oms * oms);
    double h01 = (s * s) * (3 - 2 * s), h11 = (t - t0) * s * (s - 1);
    for (int i = 0; i < NN; i++)
        out[i] = c_add(c_add(c_add(c_scale(h00, a0[i]), c_scale(h10, a1[i])), c_scale(h01, b0[i])), c_scale(h11, b1[i]));
}
__device__ __noinline__ void hermite_cplx(cplx* out, const cplx* a0, const cplx* a1, cplx t0, const cplx* b0, const cplx* b1, cplx t1, cplx t) {
    cplx s = c_div(c_sub(t, t0), c_sub(t1, t0));
    cplx oms = c_sub(mkc(1, 0), s);
    cplx oms2 = c_mul(oms, oms);
    cplx h00 = c_mul(c_add(mkc(1, 0), c_scale(2, s)), oms2), h10 = c_mul(c_sub(t, t0), oms2);
    cplx h01 = c_mul(c_mul(s, s), c_sub(mkc(3, 0), c_scale(2, s))), h11 = c_mul(c_mul(c_sub(t, t0), s), c_sub(s, mkc(1, 0)));
    for (int i = 0; i < NN; i++)
        out[i] = c_add(c_add(c_add(c_mul(h00, a0[i]), c_mul(h10, a1[i])), c_mul(h01, b0[i])), c_mul(h11, b1[i]));
}

// predict!, src/predictor.jl:286-329
__device__ __noinline__ void predict(Lane& L, cplx t, cplx dt) {
    if (!L.hermite) {
        double lam = L.trust_region, lam2 = lam * lam, lam3 = lam * lam * lam;
        for (int i = 0; i < NN; i++) {
            cplx x0 = L.tx0[i], x1 = L.tx1[i], x2 = L.tx2[i], x3 = L.tx3[i];
            double c = c_fabs(x0), c1 = c_fabs(x1), c2 = c_fabs(x2), c3 = c_fabs(x3);
            double a1 = c1 * lam, a2 = c2 * lam2, a3 = c3 * lam3;
            double tau = 1e-12 * sqrt(c * c + a1 * a1 + a2 * a2 + a3 * a3);
            if (c3 * lam3 <= tau || c2 * lam2 <= tau) {
                L.xh[i] = c_add(x0, c_mul(dt, c_add(x1, c_mul(dt, x2))));
            } else {
                cplx di = c_sub(mkc(1, 0), c_div(c_mul(dt, x3), x2));
                L.xh[i] = c_add(x0, c_mul(dt, c_add(x1, c_div(c_mul(dt, x2), di))));
            }
        }
    } else {
        int m = L.winding;
        cplx ps = t_to_s_plane(L.pred_prev_t, m), s = t_to_s_plane(t, m), s1 = t_to_s_plane(c_add(t, dt), m);
        cplx psm, sm;
        if (m == 2) { psm = c_scale(2, ps); sm = c_scale(2, s); }
        else { psm = c_scale((double)m, cpowi(ps, m - 1)); sm = c_scale((double)m, cpowi(s, m - 1)); }
        cplx a1[NN], b1[NN];
        for (int i = 0; i < NN; i++) { a1[i] = c_mul(psm, L.ptx1[i]); b1[i] = c_mul(sm, L.tx1[i]); }
        if (ps.im == 0.0 && s.im == 0.0 && s1.im == 0.0) hermite_real(L.xh, L.ptx0, a1, ps.re, L.tx0, b1, s.re, s1.re);
        else hermite_cplx(L.xh, L.ptx0, a1, ps, L.tx0, b1, s, s1);
    }
}

// ------------------------------- tracker -------------------------------
__device__ __forceinline__ int ext_steps(const Lane& L) { return L.ext_acc + L.ext_rej; }

// initial_step_size, src/tracker.jl:520-539
__device__ double initial_step_size(const Lane& L) {
    const hcg_options* o = L.o;
    double a = o->beta_a * o->a;
    double e = L.local_error;
    if (isinf(e)) e = 1e5;
    double ds1 = nthroot((sqrt(1 + 2 * h_of(a)) - 1) / (L.omega * e), L.order) / o->beta_omega_p;
    double ds2 = o->beta_tau * L.trust_region;
    return jmin(jmin(nanmin(ds1, ds2), o->max_step_size), o->max_initial_step_size);
}

#include "hcg_learned.h"

// update_stepsize!, src/tracker.jl:541-588 (the plain rule; the learned proposer may
// only shrink or enlarge a converged step within the limits it checks, see hcg_learned.h)
__device__ __noinline__ void update_stepsize(Lane& L, const NewtRes& res) {
    const hcg_options* o = L.o;
    double a = o->beta_a * o->a;
    double omega = jclamp(L.omega + 2 * (L.omega - L.omega_prev), L.omega, 8 * L.omega);
    double ds;
    if (res.code == NEWT_CONVERGED) {
        double ds1 = nthroot((sqrt(1 + 2 * h_of(a)) - 1) / (omega * L.local_error), L.order) / o->beta_omega_p;
        double ds2 = o->beta_tau * L.tau;
        if (seg_dist(L) < ds2) ds2 = o->strict_beta_tau * L.tau;
        ds = jmin(nanmin(ds1, ds2), o->max_step_size);
        ds = jmin(ds, 10 * L.ds_prev);
        if (L.last_failed > 0) ds = jmin(ds, L.ds_prev);
        ds = learned_adjust(L, res, ds);
    } else {
        int j = res.iters - 2;
        double Th = (j >= 0) ? nthroot(res.theta, 1 << j) : 1.0;
        double hT = h_of(Th), ha = h_of(0.5 * a);
        if (isnan(Th) || res.code == NEWT_SINGULARITY || isnan(res.accuracy) || res.iters == 1 || hT < ha)
            ds = 0.25 * seg_ds(L);
        else
            ds = nthroot((sqrt(1 + 2 * h_of(0.5 * a)) - 1) / (sqrt(1 + 2 * h_of(Th)) - 1), L.order) * seg_ds(L);
    }
    seg_propose(L, ds);
}

// check_terminated!, src/tracker.jl:591-619
__device__ void check_terminated(Lane& L) {
    const hcg_options* o = L.o;
    double tol_acc = (L.extended_prec || !o->extended_precision) ? ipow(o->a, (1 << o->min_newton_iters) - 1) * h_of(o->a) : HCG_INF;
    cplx t1 = seg_t1(L), t = seg_t(L);
    if (seg_is_done(L)) L.code = TC_SUCCESS;
    else if (L.accepted + L.rejected >= o->max_steps) L.code = TC_MAX_STEPS;
    else if (L.omega * L.mu > tol_acc) L.code = TC_ACCURACY_LIMIT;
    else if (seg_ds(L) < o->min_step_size) L.code = TC_STEP_TOO_SMALL;
    else if (c_fabs(c_sub(t1, t)) <= 2 * jeps(c_fabs(t))) L.code = TC_STEP_TOO_SMALL;
    else if (o->min_rel_step_size > 0 && !(t1.re == L.seg_target.re && t1.im == L.seg_target.im) &&
             c_fabs(c_sub(t1, t)) < c_fabs(t) * o->min_rel_step_size)
        L.code = TC_STEP_TOO_SMALL;
}

// init!(tracker, x₁, t₁, t₀), src/tracker.jl:639-754 (L.x holds x₁)
__device__ __noinline__ int tracker_init(Lane& L, cplx t1, cplx t0) {
    const hcg_options* o = L.o;
    seg_init(L, t1, t0);
    L.ds_prev = 0.0;
    L.accuracy = DEPS;
    L.omega = 1.0;
    L.keep_ext = 0;
    wnorm_init(L);
    L.code = TC_TRACKING;
    L.accepted = L.rejected = L.ext_acc = L.ext_rej = 0;
    L.last_failed = 0;
    cplx t = seg_t(L);
    int extp = 0;
    double omega = HCG_NAN, mu = HCG_NAN;
    int valid = init_newton(L, t, extp, &omega, &mu);
    if (!valid) {
        extp = 1;
        valid = init_newton(L, t, 1, &omega, &mu);
    }
    L.used_ext = L.extended_prec = extp;
    if (!isnan(omega)) L.omega = omega;
    if (valid) {
        L.accuracy = mu;
        L.mu = jmax(mu, DEPS);
    } else {
        H_eval_jac(L, L.x, t, L.r, L.A);
        bool anynan = false;
        for (int i = 0; i < NN * NN; i++) if (c_isnan(L.A[i])) anynan = true;
        int corank = anynan ? 0 : NN - rank_rtol(L, 1e-14);
        L.code = corank > 0 ? TC_INVALID_START_SINGULAR : TC_INVALID_START;
        return 0;
    }
    L.tau = HCG_INF;
    H_eval_jac(L, L.x, t, L.r, L.A);
    ws_updated(L);
    // init!(predictor), src/predictor.jl:124-133
    L.winding = 1;
    L.pred_t = L.pred_prev_t = mkc(HCG_NAN, 0.0);
    L.pred_s = L.pred_prev_s = mkc(HCG_NAN, 0.0);
    L.trust_region = L.local_error = HCG_NAN;
    pred_update(L, t, nullptr);
    L.tau = L.trust_region;
    double ds = jmax(initial_step_size(L), o->min_step_size);
    seg_propose(L, ds);
    L.omega_prev = L.omega;
    return L.code == TC_TRACKING;
}
// init!(tracker, t₀), src/tracker.jl:756-766
__device__ void tracker_init_t0(Lane& L, cplx t0) {
    L.code = TC_TRACKING;
    seg_init(L, seg_t(L), t0);
    seg_propose(L, initial_step_size(L));
    L.ds_prev = 0.0;
}
// update_precision! / use_extended_precision!, src/tracker.jl:768-813
__device__ __noinline__ void update_precision(Lane& L, double mu_low) {
    const hcg_options* o = L.o;
    if (!o->extended_precision) return;
    double a = o->a;
    if (L.extended_prec && !L.keep_ext && !isnan(mu_low) && mu_low > L.mu) {
        if (mu_low * L.omega < ipow(a, 7) * h_of(a)) { L.extended_prec = 0; L.mu = mu_low; }
    } else if (L.mu * L.omega > ipow(a, 5) * h_of(a)) {
        if (L.extended_prec) return;
        L.extended_prec = 1;
        L.used_ext = 1;
        cplx t = seg_t(L);
        double mu = L.mu;
        for (int i = 0; i < 2; i++) mu = ext_refine_step(L, L.x, L.x, t, 0);
        L.mu = jmax(mu, DEPS);
    }
}
// refine_current_solution!, src/tracker.jl:815-844
__device__ __noinline__ double refine_current_solution(Lane& L, double min_tol, int nsteps) {
    cplx t = seg_t(L);
    double mu = L.accuracy;
    double mub = ext_refine_step(L, L.xb, L.x, t, 0);
    if (mub < mu) { for (int i = 0; i < NN; i++) L.x[i] = L.xb[i]; mu = mub; }
    int k = 1;
    while (mu > min_tol && k <= nsteps) {
        mub = ext_refine_step(L, L.xb, L.x, t, 1);
        if (mub < mu) { for (int i = 0; i < NN; i++) L.x[i] = L.xb[i]; mu = mub; }
        k++;
    }
    return mu;
}

__device__ __noinline__ void log_step(Lane& L, const NewtRes& res, int acc) {
    int idx = atomicAdd(L.log_count, 1);
    if (idx >= L.log_capacity) return;
    float* r = L.log_records + (size_t)idx * HCG_STEP_RECORD_FLOATS;
    r[0] = (float)L.path;
    r[1] = (float)(L.accepted + L.rejected - 1);
    for (int i = 0; i < HCG_FEATS; i++) r[2 + i] = L.feat[i];
    r[10] = (float)acc;
    r[11] = (float)res.iters;
    r[12] = (float)hc_log(seg_ds(L));
    r[13] = (float)(L.extended_prec + 2 * L.singular_endgame + 4 * L.last_learned);
    r[14] = (float)(hc_log(L.accuracy) / HCG_LN10);
    r[15] = (float)(isnan(L.ds_plain) ? hc_log(seg_ds(L)) : hc_log(L.ds_plain));
}

// step!(tracker), src/tracker.jl:851-926; returns !last_step_failed
__device__ __noinline__ int tracker_step(Lane& L) {
    const hcg_options* o = L.o;
    cplx t = seg_t(L), dt = seg_dt(L), t1 = seg_t1(L);
    const double ds0 = seg_ds(L);
    predict(L, t, dt);
    wnorm_update(L, L.xh);
    NewtRes res = newton(L, L.xh, t1, L.mu, L.omega, L.extended_prec, 0, L.accepted == 0);
    if (res.code == NEWT_CONVERGED && L.last_learned) L.learned_accepted++;
    if (res.code == NEWT_CONVERGED) {
        for (int i = 0; i < NN; i++) L.x[i] = L.xb[i];
        L.ds_prev = seg_ds(L);
        L.seg_s = L.seg_s1;
        L.accuracy = res.accuracy;
        L.mu = jmax(res.accuracy, DEPS);
        L.omega_prev = L.omega;
        L.omega = jmax(jmax(res.omega, 0.5 * L.omega), 0.1);
        update_precision(L, res.mu_low);
        if (seg_is_done(L) && o->extended_precision && L.accuracy > 1e-14) {
            L.accuracy = refine_current_solution(L, 1e-14, 3);
            L.refined_ext = 1;
        }
        pred_update(L, seg_t(L), L.xh);
        L.tau = L.trust_region;
        L.accepted++;
        L.ext_acc += L.extended_prec;
        L.last_failed = 0;
    } else {
        L.rejected++;
        L.ext_rej += L.extended_prec;
        L.last_failed++;
    }
    if (L.log_records || L.mlp_w) step_features(L, res, ds0);
    L.last_learned = 0;
    L.ds_plain = HCG_NAN;
    update_stepsize(L, res);
    if (L.log_records) log_step(L, res, res.code == NEWT_CONVERGED);
    check_terminated(L);
    return !(L.last_failed > 0);
}

// ------------------------------- valuation -------------------------------
__device__ __forceinline__ double nu1(cplx x, cplx xd, double t) {
    double xx = x.re * x.re + x.im * x.im;
    double mu = x.re * xd.re + x.im * xd.im;
    return t * (mu / xx);
}
__device__ __forceinline__ void nu_nu1(cplx x, cplx xd, cplx x2, double t, double* v, double* v1) {
    double xx = x.re * x.re + x.im * x.im;
    double mu = x.re * xd.re + x.im * xd.im;
    double l = mu / xx;
    double mu1 = x.re * x2.re + xd.re * xd.re + x.im * x2.im + xd.im * xd.im;
    double l1 = mu1 / xx - 2 * (l * l);
    *v = t * l; *v1 = t * l1 + l;
}
__device__ __forceinline__ double finite_diff(double v, double s, double v2, double s2, double v1, double s1) {
    double D1 = s - s1, D2 = s - s2, D12 = s1 - s2;
    return (D2 * v1) / (D12 * D1) - ((D12 + D2) * v2) / (D12 * D2) - (D12 * v) / (D1 * D2);
}
// update!(val, pred, t), src/valuation.jl:82-124
__device__ __noinline__ void val_update(Lane& L, double t) {
    double logt = hc_log(t);
    double logt2 = L.logt2, logt1 = L.logt1;
    bool diff = L.winding > 1 && !isnan(logt2);
    for (int i = 0; i < NN; i++) {
        cplx xi = L.tx0[i], xdi = L.tx1[i], x2i = L.tx2[i], x3i = L.tx3[i];
        double ldxi = hc_log(c_fabs(xdi));
        if (diff) {
            double vi = nu1(xi, xdi, t);
            double dvi = finite_diff(vi, logt, L.vx2[i], logt2, L.vx1[i], logt1);
            L.val_x[i] = vi; L.dval_x[i] = dvi;
            double vdx = finite_diff(ldxi, logt, L.ldx2[i], logt2, L.ldx1[i], logt1);
            double dvdx = finite_diff(vdx, logt, 0.0, logt2, 0.0, logt1);
            L.val_tx[i] = vdx + 1; L.dval_tx[i] = dvdx;
        } else {
            double vi, v1;
            nu_nu1(xi, xdi, c_scale(2, x2i), t, &vi, &v1);
            L.val_x[i] = vi; L.dval_x[i] = t * v1;
            L.vx1[i] = L.vx2[i]; L.vx2[i] = vi;
            double vdx, v1b;
            nu_nu1(xdi, c_scale(2, x2i), c_scale(6, x3i), t, &vdx, &v1b);
            L.val_tx[i] = vdx + 1; L.dval_tx[i] = t * v1b;
        }
        L.ldx1[i] = L.ldx2[i]; L.ldx2[i] = ldxi;
    }
    L.logt1 = logt2; L.logt2 = logt;
}
__device__ __forceinline__ double val_eps_inf(const Lane& L, int i) {
    double vx = L.val_x[i];
    return jmax(jmax(fabs(1.0 - L.val_tx[i] / vx), fabs(L.dval_x[i] / vx)), fabs(L.dval_tx[i] / L.val_tx[i]));
}
// is_finite(val), src/valuation.jl:175-205
__device__ bool val_is_finite(const Lane& L, double finite_tol, bool zero_is_finite, int max_winding) {
    double delta = 1.0 / max_winding;
    for (int i = 0; i < NN; i++) {
        double vx = L.val_x[i];
        if (fabs(vx) < finite_tol) {
            if (!(fabs(L.dval_x[i]) < finite_tol) || L.val_tx[i] < 0.5 * delta) return false;
        } else if (zero_is_finite && vx > (delta - finite_tol)) {
            if (!(val_eps_inf(L, i) < finite_tol)) return false;
        } else return false;
    }
    return true;
}
// estimate_winding_number, src/valuation.jl:207-228
__device__ int estimate_winding_number(const Lane& L, int max_w, double* err_out) {
    int m = 1; double min_err = HCG_INF;
    for (int k = 1; k <= max_w; k++) {
        double err = 0.0;
        for (int i = 0; i < NN; i++) {
            double mv = k * L.val_tx[i];
            double ei = fabs(rint(mv) - mv);
            err = ei > err ? ei : err;
        }
        if (err < min_err) { m = k; min_err = err; }
    }
    *err_out = min_err;
    return m;
}

// -------------------------------- endgame --------------------------------
__device__ __noinline__ double tracker_cond_at(Lane& L, const cplx* x, cplx t, const double* dl, const double* dr) {
    H_eval_jac(L, x, t, L.r, L.A);
    ws_updated(L);
    return ws_cond(L, dl, dr);
}
__device__ __forceinline__ void eg_row_scaling(Lane& L) { skeel_row_scaling(L.row_scaling, L.A, L.col_scaling, L.o->scaling_threshold); }

__device__ __forceinline__ int tc_to_eg(int c) {
    switch (c) {
    case TC_SUCCESS: return EG_SUCCESS;
    case TC_MAX_STEPS: return EG_MAX_STEPS;
    case TC_ACCURACY_LIMIT: return EG_ACCURACY_LIMIT;
    case TC_ILL_CONDITIONED: return EG_ILL_CONDITIONED;
    case TC_INVALID_START: return EG_INVALID_START;
    case TC_INVALID_START_SINGULAR: return EG_INVALID_START_SINGULAR;
    case TC_STEP_TOO_SMALL: return EG_STEP_TOO_SMALL;
    case TC_UNKNOWN: return EG_UNKNOWN;
    default: return EG_TRACKING;
    }
}

// init!(endgame_tracker, x, t₁), src/endgame_tracker.jl:260-294
__device__ __noinline__ void eg_init(Lane& L, double t1) {
    L.newton_iters = 0;
    L.refined_ext = 0;
    L.learned_proposed = L.learned_accepted = 0;
    L.last_learned = 0;
    L.inf_logit = 0.0f;
    L.ds_plain = HCG_NAN;
    L.eg_start = L.o->endgame_start;
    for (int i = 0; i < 8; i++) L.feat[i] = 0.0f;
    for (int i = 0; i < NN; i++) { L.tx0[i] = L.tx1[i] = L.tx2[i] = L.tx3[i] = mkc(0.0, 0.0); }
    tracker_init(L, mkc(t1, 0.0), mkc(0.0, 0.0));
    L.eg_code = tc_to_eg(L.code);
    L.singular_endgame = 0;
    L.sing_phase = 0;
    L.jtzf0 = L.jtzf1 = 0;
    for (int i = 0; i < NN; i++) {
        L.val_x[i] = L.val_tx[i] = L.dval_x[i] = L.dval_tx[i] = 0.0;
        L.vx2[i] = L.vx1[i] = L.ldx2[i] = L.ldx1[i] = 0.0;
        L.solution[i] = mkc(HCG_NAN, HCG_NAN);
        L.row_scaling[i] = 1.0; L.col_scaling[i] = 1.0; L.ai_starts[i] = HCG_NAN;
    }
    L.logt2 = L.logt1 = HCG_NAN;
    L.eg_winding = 0;
    L.eg_accuracy = HCG_NAN; L.eg_cond = HCG_NAN; L.eg_singular = 0;
    L.steps_eg = 0;
    L.ext_steps_eg_start = 0x7fffffff;
    L.singular_steps = 0;
    L.last_t = HCG_NAN;
    for (int k = 0; k < 3; k++) L.smp_order[k] = k;
}

// tracking_stopped!, src/endgame_tracker.jl:695-721
__device__ __noinline__ void tracking_stopped(Lane& L) {
    const hcg_options* o = L.o;
    L.eg_accuracy = L.accuracy;
    if (L.eg_code == EG_SUCCESS && L.eg_accuracy > 1e-14) refine_current_solution(L, 1e-14, o->refine_steps);
    for (int i = 0; i < NN; i++) L.solution[i] = L.x[i];
    L.eg_winding = 0;
    if (L.eg_code == EG_SUCCESS) {
        for (int i = 0; i < NN; i++) L.col_scaling[i] = L.w[i];
        eg_row_scaling(L);
        L.eg_cond = tracker_cond_at(L, L.solution, mkc(0, 0), L.row_scaling, L.col_scaling);
        L.eg_singular = L.eg_cond > o->sing_cond || L.eg_accuracy > o->sing_accuracy;
    }
}
// add_sample!, src/endgame_tracker.jl:630-662 (samples addressed through smp_order)
__device__ __noinline__ void add_sample(Lane& L, double t) {
    int m = L.eg_winding;
    double s = nthroot(t, m);
    double mu = m * ipow(s, m - 1);
    double kappa = ws_cond(L, L.row_scaling, L.col_scaling);
    int slot;
    if (L.singular_steps <= 2) slot = L.singular_steps;
    else {
        int o0 = L.smp_order[0];
        L.smp_order[0] = L.smp_order[1]; L.sample_times[0] = L.sample_times[1]; L.sample_conds[0] = L.sample_conds[1];
        L.smp_order[1] = L.smp_order[2]; L.sample_times[1] = L.sample_times[2]; L.sample_conds[1] = L.sample_conds[2];
        L.smp_order[2] = o0;
        slot = 2;
    }
    int b = L.smp_order[slot];
    for (int i = 0; i < NN; i++) { L.smp0[b][i] = L.tx0[i]; L.smp1[b][i] = c_scale(mu, L.tx1[i]); }
    L.sample_times[slot] = s;
    L.sample_conds[slot] = kappa;
}
// predict_endpoint!, src/endgame_tracker.jl:664-693
__device__ __noinline__ double predict_endpoint(Lane& L) {
    if (L.singular_steps < 2) return HCG_INF;
    int b0 = L.smp_order[0], b1 = L.smp_order[1], b2 = L.smp_order[2];
    if (L.singular_steps == 2)
        hermite_real(L.prediction, L.smp0[b0], L.smp1[b0], L.sample_times[0], L.smp0[b1], L.smp1[b1], L.sample_times[1], 0.0);
    for (int i = 0; i < NN; i++) L.prev_prediction[i] = L.prediction[i];
    hermite_real(L.prediction, L.smp0[b1], L.smp1[b1], L.sample_times[1], L.smp0[b2], L.smp1[b2], L.sample_times[2], 0.0);
    double p = L.sample_times[2] / L.sample_times[1];
    double p2 = p * p;
    double err = dist_inf(L.prediction, L.prev_prediction, nullptr) / fabs(p2 * p2 - 1);
    double ns = norm_inf(L.prediction, nullptr);
    if (ns > 1e-8) err /= ns;
    return err;
}
// check_finite!, src/endgame_tracker.jl:402-422
__device__ bool check_finite(Lane& L) {
    const hcg_options* o = L.o;
    if (!val_is_finite(L, o->val_finite_tol, !o->zero_is_at_infinity, o->max_winding_number)) return false;
    double merr;
    int m = estimate_winding_number(L, o->max_winding_number, &merr);
    if (merr < o->val_finite_tol) {
        if (m == 1 && !L.jtzf0) return false;
        L.eg_winding = m;
        return true;
    }
    return false;
}
// check_at_infinity!, src/endgame_tracker.jl:424-499
__device__ __noinline__ bool check_at_infinity(Lane& L) {
    const hcg_options* o = L.o;
    if (!o->at_infinity_check) return false;
    bool zero_is_finite = !o->zero_is_at_infinity;
    for (int i = 0; i < NN; i++) {
        double vx = L.val_x[i];
        double e = val_eps_inf(L, i);
        if (isnan(e)) L.ai_tols[i] = HCG_INF;
        else if (vx + e < -o->val_finite_tol) L.ai_tols[i] = e;
        else if (!zero_is_finite && vx - e > -o->val_finite_tol) L.ai_tols[i] = e;
        else L.ai_tols[i] = HCG_INF;
    }
    double kappa = HCG_NAN;
    double t = seg_t(L).re;
    for (int i = 0; i < NN; i++) {
        if (L.ai_tols[i] < o->val_at_infinity_tol) {
            if (isnan(L.ai_starts[i])) {
                bool allnan = true;
                for (int k = 0; k < NN; k++) if (!isnan(L.ai_starts[k])) allnan = false;
                if (allnan) {
                    for (int k = 0; k < NN; k++) L.col_scaling[k] = L.w[k];
                    eg_row_scaling(L);
                }
                kappa = ws_cond(L, L.row_scaling, L.col_scaling);
                L.ai_conds[i] = kappa;
                L.ai_abs[i] = c_fabs(L.x[i]);
                L.ai_starts[i] = t;
            } else {
                if (isnan(kappa)) kappa = ws_cond(L, L.row_scaling, L.col_scaling);
                double vi = L.val_x[i];
                bool at_zero = L.val_x[i] > 0;
                double cond_growth = kappa / L.ai_conds[i];
                double coord_growth = at_zero ? L.ai_abs[i] / c_fabs(L.x[i]) : c_fabs(L.x[i]) / L.ai_abs[i];
                if (coord_growth > jclamp(hc_pow(0.25, 4 * vi), 20, o->min_coord_growth) &&
                    (cond_growth > o->min_cond_growth || kappa > jmax(1e8, o->min_cond))) {
                    L.eg_code = at_zero ? EG_AT_ZERO : EG_AT_INFINITY;
                    return true;
                }
            }
        } else if (!isnan(L.ai_starts[i])) {
            L.ai_starts[i] = HCG_NAN;
        }
    }
    return false;
}
// switch_to_singular!, src/endgame_tracker.jl:501-524
__device__ __noinline__ void switch_to_singular(Lane& L) {
    L.singular_endgame = 1;
    double t = seg_t(L).re;
    bool allone = true;
    for (int i = 0; i < NN; i++) if (L.row_scaling[i] != 1.0) allone = false;
    if (allone) {
        for (int i = 0; i < NN; i++) L.col_scaling[i] = L.w[i];
        eg_row_scaling(L);
    }
    add_sample(L, t);
    L.singular_steps = 0;
    L.winding = L.eg_winding;
    L.ai_conds[0] = L.sample_conds[0];
    L.keep_ext = 1;
}
__device__ void switch_to_regular(Lane& L) {
    L.singular_endgame = 0;
    L.winding = 1;
    tracker_init_t0(L, mkc(0, 0));
}

// rest of singular_endgame_step! after its inner tracking loop, src/endgame_tracker.jl:555-628
__device__ __noinline__ void singular_post(Lane& L, bool max_steps) {
    const hcg_options* o = L.o;
    double lam = 0.25;
    L.singular_steps += 1;
    if (!max_steps) {
        if (L.code != TC_SUCCESS) {
            L.eg_code = tc_to_eg(L.code);
            tracking_stopped(L);
            return;
        }
        val_update(L, lam * L.sing_t);
        double merr;
        int mh = estimate_winding_number(L, o->max_winding_number, &merr);
        if (mh != L.eg_winding || merr > 0.1) { switch_to_regular(L); return; }
        add_sample(L, lam * L.sing_t);
        if (!(L.singular_steps >= 2)) return;
        double acc = predict_endpoint(L);
        if (L.singular_steps == 2) {
            L.eg_accuracy = acc;
            for (int i = 0; i < NN; i++) L.solution[i] = L.prediction[i];
            return;
        } else if (acc < L.eg_accuracy && L.eg_accuracy > 1e-12) {
            L.eg_accuracy = acc;
            for (int i = 0; i < NN; i++) L.solution[i] = L.prediction[i];
            return;
        }
    }
    int m = L.eg_winding;
    double kappa = L.sample_conds[2];
    double zero_cond = 1.0 / (m + 1);
    for (int i = 0; i < NN; i++)
        L.solution[i] = (L.val_x[i] < zero_cond) ? L.prediction[i] : mkc(copysign(0.0, L.prediction[i].re), copysign(0.0, L.prediction[i].im));
    double kappa0 = tracker_cond_at(L, L.solution, mkc(0, 0), L.row_scaling, L.col_scaling);
    double J0 = inf_norm_ws(L, L.row_scaling, nullptr);
    if (L.eg_accuracy < o->singular_min_accuracy &&
        (((m > 1 && kappa > o->min_cond && nanmax(kappa0, 1.0 / J0) > kappa) || (m == 1 && kappa0 > 1e12)) ||
         max_steps || (NN == 1 && 1.0 / J0 < o->min_cond))) {
        L.eg_cond = jmax(kappa0, 1.0 / J0);
        L.eg_singular = 1;
        L.eg_code = EG_SUCCESS;
    } else if (!max_steps) {
        switch_to_regular(L);
    }
}

// step!(endgame_tracker), src/endgame_tracker.jl:329-399, with singular_endgame_step!'s
// inner `while is_tracking: step!` loop flattened to one tracker step per call
__device__ __noinline__ void eg_step(Lane& L) {
    const hcg_options* o = L.o;
    if (!(L.singular_endgame && L.sing_phase == 1)) {
        if (L.steps_eg >= o->max_endgame_steps) { L.eg_code = EG_MAX_STEPS; return; }
        if (ext_steps(L) - L.ext_steps_eg_start > o->max_endgame_extended_steps) {
            bool allfinite = true;
            for (int i = 0; i < NN; i++) if (c_isnan(L.solution[i])) allfinite = false;
            if (allfinite && L.eg_winding != 0 && L.eg_accuracy < o->singular_min_accuracy) {
                L.eg_cond = tracker_cond_at(L, L.solution, mkc(0, 0), L.row_scaling, L.col_scaling);
                L.eg_singular = 1;
                L.eg_code = EG_SUCCESS;
                return;
            }
            L.eg_code = EG_MAX_EXT_STEPS;
            return;
        }
        for (int i = 0; i < NN; i++) L.last_point[i] = L.x[i];
        L.last_t = seg_t(L).re;
    }
    if (L.singular_endgame) {
        if (L.sing_phase == 0) {
            L.sing_t = seg_t(L).re;
            tracker_init_t0(L, mkc(0.25 * L.sing_t, 0.0));
            L.sing_phase = 1;
        }
        // one iteration of `while is_tracking(tracker.state.code)`
        tracker_step(L);
        bool max_steps = false;
        if ((L.steps_eg += 1) >= o->max_endgame_steps) { L.eg_code = EG_MAX_STEPS; max_steps = true; }
        else if (ext_steps(L) - L.ext_steps_eg_start > o->max_endgame_extended_steps) { L.eg_code = EG_MAX_EXT_STEPS; max_steps = true; }
        if (!max_steps && L.code == TC_TRACKING) return;
        L.sing_phase = 0;
        singular_post(L, max_steps);
        return;
    }
    cplx tp = seg_t1(L);
    int jump = tp.re == 0.0 && tp.im == 0.0;
    int ok = tracker_step(L);
    L.eg_code = tc_to_eg(L.code);
    if (L.eg_code != EG_TRACKING) { tracking_stopped(L); return; }
    L.jtzf0 = L.jtzf1; L.jtzf1 = jump;
    double t = seg_t(L).re;
    if (!(t <= L.eg_start)) return;
    if (L.steps_eg == 0) L.ext_steps_eg_start = ext_steps(L);
    L.steps_eg += 1;
    if (!ok) return;
    val_update(L, t);
    if (check_finite(L)) { switch_to_singular(L); return; }
    check_at_infinity(L);
}

// PathResult(endgame_tracker, ...), src/endgame_tracker.jl:847-888
__device__ __noinline__ void write_result(Lane& L, const hcg_kargs& a, int k) {
    const hcg_options* o = L.o;
    double t;
    const cplx* s;
    if (L.eg_code == EG_SUCCESS) { t = 0.0; s = L.solution; H_eval(L, L.solution, mkc(0, 0), L.r); }
    else { t = seg_t(L).re; s = L.x; H_eval(L, L.x, seg_t(L), L.r); }
    hcg_path_result R;
    R.code = L.eg_code;
    R.singular = L.eg_singular;
    R.winding_number = L.eg_winding;
    R.extended_precision = L.extended_prec;
    R.extended_precision_used = L.used_ext;
    R.accepted_steps = L.accepted;
    R.rejected_steps = L.rejected;
    R.steps_eg = L.steps_eg;
    R.valuation_valid = !(t > L.eg_start);
    R.newton_iters = L.newton_iters;
    R.learned_proposed = L.learned_proposed;
    R.learned_accepted = L.learned_accepted;
    R.t = t;
    R.accuracy = L.eg_accuracy;
    R.residual = norm_inf(L.r, nullptr);
    R.cond = L.eg_cond;
    R.omega = L.omega;
    R.mu = L.mu;
    R.last_t = L.last_t;
    a.results[k] = R;
    for (int i = 0; i < NN; i++) {
        size_t o = (size_t)k * NN + i;
        a.solutions[2 * o] = s[i].re; a.solutions[2 * o + 1] = s[i].im;
        a.last_points[2 * o] = L.last_point[i].re; a.last_points[2 * o + 1] = L.last_point[i].im;
        a.valuations[o] = L.val_x[i];
    }
}

// record of a path whose inputs were rejected on the device (out-of-range parameter group)
__device__ __noinline__ void write_result_invalid(Lane& L, const hcg_kargs& a, int k) {
    hcg_path_result R;
    R.code = EG_UNKNOWN;
    R.singular = R.winding_number = R.extended_precision = R.extended_precision_used = 0;
    R.accepted_steps = R.rejected_steps = R.steps_eg = R.valuation_valid = R.newton_iters = 0;
    R.learned_proposed = R.learned_accepted = 0;
    R.t = a.t1; R.accuracy = R.residual = R.cond = HCG_NAN; R.omega = R.mu = R.last_t = HCG_NAN;
    a.results[k] = R;
    for (int i = 0; i < NN; i++) {
        a.solutions[2 * ((size_t)k * NN + i)] = HCG_NAN; a.solutions[2 * ((size_t)k * NN + i) + 1] = HCG_NAN;
        a.last_points[2 * ((size_t)k * NN + i)] = HCG_NAN; a.last_points[2 * ((size_t)k * NN + i) + 1] = HCG_NAN;
        a.valuations[(size_t)k * NN + i] = HCG_NAN;
    }
}


// ------------------------ per-lane path driver ------------------------
// Shared by the single-system kernel below and by the multi-system kernel, which
// includes this body once per system (each in its own namespace).
__device__ __forceinline__ void lane_setup(Lane& L, const hcg_kargs& a, const unsigned short* mlp_w) {
    L.o = a.opts;
    L.mlp_w = mlp_w;
    L.log_records = a.log_records;
    L.log_count = a.log_count;
    L.log_capacity = a.log_capacity;
    L.cond_H_xd = 1.0;
}

// Start path k; returns 0 when the path cannot start (its result is already written).
__device__ __noinline__ int lane_begin(Lane& L, const hcg_kargs& a, int k) {
    for (int i = 0; i < NN; i++) L.x[i] = mkc(a.starts[2 * ((size_t)k * NN + i)], a.starts[2 * ((size_t)k * NN + i) + 1]);
    L.path = k;
    L.flops = 0.0;
#if HCG_KIND == 1
    {
        int g = a.group_of_path ? a.group_of_path[k] : 0;
        if (g < 0 || g >= a.ngroups) {  // device-resident input: reject out-of-range groups
            L.eg_code = EG_UNKNOWN;
            L.accepted = L.rejected = L.steps_eg = L.newton_iters = 0;
            write_result_invalid(L, a, k);
            return 0;
        }
        L.p = a.start_params;
        L.q = a.target_params + 2 * HCG_NPAR * g;
    }
#endif
    eg_init(L, a.t1);
    L.guard = 0;
    return 1;
}

__device__ __forceinline__ int lane_ended(const Lane& L) { return L.eg_code != EG_TRACKING; }

// One loop iteration of a running path (one tracker step), with the per-path guard.
__device__ __forceinline__ void lane_iter(Lane& L) {
    eg_step(L);
    if (++L.guard > 4 * (L.o->max_steps + L.o->max_endgame_steps) + 64) L.eg_code = EG_UNKNOWN;
}

__device__ __forceinline__ void lane_finish(Lane& L, const hcg_kargs& a, int k, LaneTotals& T) {
    write_result(L, a, k);
    T.steps += L.accepted + L.rejected;
    T.newton += L.newton_iters;
    T.lacc += L.learned_accepted;
    T.flops += L.flops;
}

#undef NN
#undef IDX
#undef FL_LU
#undef FL_SOLVE
#undef FL_MIXED
#undef FL_FIXED
#undef HCG_UNROLL
