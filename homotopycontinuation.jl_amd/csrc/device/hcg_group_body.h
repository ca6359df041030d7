// Lane-group tracker: one path per group of 16 lanes.  Lane r holds entry r of every vector
// and row r of the Jacobian and of its LU factors, in registers; scalars are replicated over
// the group.  The algorithm is the per-lane tracker's (and the reference's) step for step:
//   tracker            src/tracker.jl:517-996     predictor  src/predictor.jl:124-371
//   Newton corrector   src/newton_corrector.jl:55-286
//   linear algebra     src/linear_algebra.jl:88-885   weighted norm  src/norm.jl:101-234
//   valuation          src/valuation.jl:38-228    endgame  src/endgame_tracker.jl:260-911
// with every reduction taken in the reference's order (hcg_group_prims.h), so results equal the
// oracle's bit for bit.
//
// Control flow.  Evaluations of H (FP64, double-double), of H and its Jacobian, and of Taylor
// coefficients run the system's slot programs (hcg_group_interp.h).  So that each of those
// interpreters appears once in the kernel and the path state never crosses a call, the
// tracker is written as nested coroutines: a function that needs an evaluation stores its
// request, records where it stopped and returns; the kernel's loop runs the request and
// resumes the path.  All values that live across a request are members of GT.
#pragma once
#include <type_traits>

#include "hcg_arith.h"
#include "hcg_group_interp.h"
#include "hcg_group_prims.h"
#include "hcg_group_types.h"
#include "hcg_kargs.h"
#include "hcg_mlp_infer.h"

// TrackerCode, src/tracker.jl:166-176
enum { GTC_TRACKING = 0, GTC_SUCCESS, GTC_MAX_STEPS, GTC_ACCURACY_LIMIT, GTC_ILL_CONDITIONED, GTC_INVALID_START,
       GTC_INVALID_START_SINGULAR, GTC_STEP_TOO_SMALL, GTC_UNKNOWN };
// EndgameTrackerCode, src/endgame_tracker.jl:100-117
enum { GEG_TRACKING = 0, GEG_SUCCESS, GEG_AT_INFINITY, GEG_AT_ZERO, GEG_ACCURACY_LIMIT, GEG_INVALID_START,
       GEG_INVALID_START_SINGULAR, GEG_ILL_CONDITIONED, GEG_MAX_STEPS, GEG_MAX_EXT_STEPS, GEG_MAX_WINDING,
       GEG_STEP_TOO_SMALL, GEG_UNKNOWN };
enum { GNEWT_CONVERGED = 0, GNEWT_TERMINATED, GNEWT_MAX_ITERS, GNEWT_SINGULARITY };
// evaluation requests
enum { GREQ_NONE = 0, GREQ_EJ, GREQ_E, GREQ_DD, GREQ_T1, GREQ_T2, GREQ_T3, GREQ_COND };
// what follows an evaluation of H and its Jacobian: the factorization the reference would do at
// the Jacobian's first use (row-scaled by jac_ldiv_w, or plain by ldiv!), done right away so
// that the factorization appears once in the kernel
enum { GFACT_NONE = 0, GFACT_SCALED, GFACT_PLAIN };

// coroutine helpers: a function returns 1 while it waits for an evaluation, 0 when done
#define CO_BEGIN(v) switch (v) { case 0:
#define CO_END(v) } v = 0; return 0;
#define CO_EVAL(v, kind) do { req = (kind); v = __LINE__; return 1; case __LINE__:; } while (0)
#define CO_AWAIT(v, call) do { v = __LINE__; /* fallthrough */ case __LINE__: if (call) return 1; } while (0)

struct GNewtRes { int code; double accuracy; int iters; double omega, theta, mu_low, norm_dx0; };

template <int NN>
struct GT {
    // ---------------- identity ----------------
    int r, n, kind;
    const hcg_options* o;
    const hcg_gsys* sys;
    const double* p;  // parameter homotopy: start parameters
    const double* q;  // this path's target parameters
    char* lds;        // this group's LDS registers
    const unsigned short* mlp_w;
    // ---------------- evaluation request ----------------
    int req;
    cplx ev_x, ev_t;  // point (lane entry) and t of the request
    int ev_factor;    // GFACT_* after GREQ_EJ
    cplx ev_u;        // result entry
    double ev_cond;   // result of GREQ_COND: cond(WS, row_scaling, col_scaling)
    cplx A[NN];       // Jacobian row r (result of GREQ_EJ, kept as the workspace matrix)
    // ---------------- TrackerState ----------------
    cplx x, xh, xb;
    cplx seg_start, seg_target; double seg_absD, seg_s, seg_s1; int seg_forward;
    double ds_prev, accuracy, omega, omega_prev, mu, tau;
    int extended_prec, used_ext, refined_ext, keep_ext;
    double w;
    int code;
    int accepted, rejected, last_failed, ext_acc, ext_rej;
    // ---------------- MatrixWorkspace ----------------
    cplx LU[NN];      // row of the LU factors held by this lane (physical position `pos`)
    int pos;          // physical row position of this lane's row after pivoting
    int ipiv[NN];     // pivot sequence (same in every lane of the group)
    int src_of;       // lane holding position r (for returning solutions in variable order)
    double rs;
    int factorized, scaled;
    cplx dx, res, tmp, rdx;
    // ---------------- Predictor ----------------
    int hermite, order, winding;
    double trust_region, local_error;
    cplx tx0, tx1, tx2, tx3, ptx0, ptx1;
    cplx pred_t, pred_prev_t, pred_s, pred_prev_s;
    double txn[4];
    int newton_iters;
    // ---------------- EndgameTrackerState ----------------
    int eg_code, singular_endgame;
    double val_x, val_tx, dval_x, dval_tx, vx2, vx1, ldx2, ldx1;
    double logt2, logt1;
    int eg_winding;
    cplx solution;
    double eg_accuracy, eg_cond;
    int eg_singular;
    int steps_eg, ext_steps_eg_start;
    int jtzf0, jtzf1;
    cplx last_point; double last_t;
    double row_scaling, col_scaling;
    double ai_starts, ai_tols, ai_abs, ai_conds;
    cplx smp0_0, smp0_1, smp0_2, smp1_0, smp1_1, smp1_2;
    int smp_order0, smp_order1, smp_order2;
    double sample_times[3], sample_conds[3];
    int singular_steps;
    cplx prediction, prev_prediction;
    int sing_phase;
    double sing_t;
    // ---------------- learned proposer ----------------
    int learned_proposed, learned_accepted, last_learned;
    double cond_H_xd, ds_plain;
    double eg_start;
    float feat[8], inf_logit;
    double flops;
    int guard;
    int path;
    // step log (training records; null when not logging)
    float* log_records;
    int* log_count;
    int log_capacity;
    // ---------------- coroutine states and their locals ----------------
    int co_path, co_egstep, co_tstep, co_newton, co_initn, co_ext, co_pred, co_tinit, co_refine, co_upprec,
        co_tstop, co_condat, co_singpost, co_result;
    // newton
    GNewtRes nres;
    const cplx* nw_x0;  // unused marker (x0 is copied into xb before the call)
    double nw_t_re, nw_t_im, nw_mu, nw_omega, nw_mu_low, nw_theta, nw_ndx, nw_ndx_prev, nw_ndx0, nw_abar, nw_ndx1;
    int nw_ext, nw_accurate_mu, nw_first, nw_i;
    // init_newton
    int in_ext, in_valid, in_k;
    double in_omega, in_mu, in_eps, in_v, in_ndx0, in_ndx1;
    cplx in_t;
    // ext_refine_step
    int er_simple, er_target_xb;  // writes xb (1) or x (0)
    cplx er_t, er_xin;
    double er_out;
    // predictor update
    cplx pu_t;
    int pu_have_xhat;
    double pu_delta;
    cplx pu_xt, pu_ub;
    // tracker_init
    int ti_extp, ti_valid;
    double ti_omega, ti_mu;
    cplx ti_t;
    // refine_current_solution / update_precision
    double rf_min_tol, rf_mu, rf_mub;
    int rf_nsteps, rf_k;
    cplx rf_t;
    int up_i;
    double up_mu;
    // tracker_step
    double ts_ds0;
    cplx ts_t, ts_dt, ts_t1;
    GNewtRes ts_res;
    int ts_ok;
    // cond at a point
    cplx ca_x, ca_t;
    int ca_use_rs;
    double ca_out;
    // singular_post
    int sp_max_steps;
    // add_sample, switch_to_singular, check_at_infinity
    int co_addsmp, co_sw, co_ainf;
    double as_t, sw_t, ai_kappa;
    int ai_hit, ai_first;
    // eg_step
    int es_jump;

    // ======================= vector helpers =======================
    __device__ __forceinline__ bool act() const { return r < n; }
    // weighted / plain inf-norms, src/norm.jl:150-234 (sequential max over the entries)
    __device__ __forceinline__ double norm_inf_v(cplx v, bool weighted) const {
        const double term = weighted ? c_abs2(c_divr(v, w)) : c_abs2(v);
        double dmax = g_fold<NN>(term, n, [](double a, double b) { return max_fast(a, b); });
        double d = sqrt(dmax);
        if (isinf(d)) {
            const double h = weighted ? c_hypot(c_divr(v, w)) : c_hypot(v);
            return g_fold<NN>(h, n, [](double a, double b) { return jmax(a, b); });
        }
        return d;
    }
    __device__ __forceinline__ double dist_inf_v(cplx a, cplx b, bool weighted) const {
        const cplx dv = c_sub(a, b);
        const double term = weighted ? c_abs2(c_divr(dv, w)) : c_abs2(dv);
        double dmax = g_fold<NN>(term, n, [](double u, double v) { return max_fast(u, v); });
        double d = sqrt(dmax);
        if (isinf(d)) {
            if (weighted) return g_fold<NN>(c_hypot(c_divr(dv, w)), n, [](double u, double v) { return jmax(u, v); });
            // the reference's unweighted fallback mixes abs and abs2 (src/norm.jl:160-163)
            const double h0 = gb_d<0>(c_hypot(dv));
            double acc = h0;
            const double t2 = c_abs2(dv);
            G_FOR(NN, J, if (J >= 1 && J < n) acc = jmax(acc, gb_d<J>(t2)););
            return acc;
        }
        return d;
    }
    // entry j (run time) of a row held in registers
    __device__ __forceinline__ cplx row_get(const cplx (&R)[NN], int j) const {
        cplx v = R[0];
        G_FOR(NN, J, if (J >= 1 && j == J) v = R[J];);
        return v;
    }

    // ======================= linear algebra =======================
    __device__ __forceinline__ void ws_updated() { factorized = 0; scaled = 0; }

    // lu!, src/linear_algebra.jl:130-184 on diag(rs) A (rs when `scaled`), right-looking with
    // partial pivoting on abs2.  Rows stay in their lanes: `pos` is the row's physical position
    // and the interchange at column k swaps the positions of two lanes; every entry gets the
    // reference's updates in the reference's order.
    __device__ __forceinline__ void lu_factor() {
        flops += 11.0 * (n * (n - 1) / 2) + 8.0 * ((n - 1) * n * (2 * n - 1) / 6);
        const double sc = scaled ? rs : 1.0;
        G_FOR(NN, J, LU[J] = scaled ? c_scale(sc, A[J]) : A[J];);
        pos = r;
        G_FOR(NN, K, {
            if (K < n) {
                // pivot search over positions K..n-1 in position order (first maximum of abs2)
                const double a = c_abs2(LU[K]);
                const bool cand = pos >= K && act();
                // value at position K
                const int lk = g_first(pos == K, n);
                const double amax0 = gs_d(a, lk);
                // the sequential scan: positions K+1.. in order, strictly greater replaces
                double amax = amax0;
                int kp = K;
                for (int pp = K + 1; pp < n; pp++) {
                    const int lp = g_first(pos == pp, n);
                    const double ap = gs_d(a, lp);
                    if (ap > amax) { kp = pp; amax = ap; }
                }
                (void)cand;
                ipiv[K] = kp;
                if (amax != 0.0) {
                    if (kp != K) {  // interchange the positions of rows K and kp
                        if (pos == K) pos = kp;
                        else if (pos == kp) pos = K;
                    }
                    // pivot row: the lane now at position K
                    const int lpiv = g_first(pos == K, n);
                    const cplx fkk = gs_c(LU[K], lpiv);
                    if (pos > K && act()) LU[K] = c_div(LU[K], fkk);
                }
                const int lpiv2 = g_first(pos == K, n);
                G_FOR(NN, J, {
                    if (J > K && J < n) {
                        const cplx fkj = gs_c(LU[J], lpiv2);
                        if (pos > K && act()) LU[J] = c_submul(LU[K], fkj, LU[J]);
                    }
                });
            }
        });
        // lane holding position r (the inverse of pos): solutions are returned in variable order
        src_of = r;
        G_FOR(NN, J, if (J < n && gb_i<J>(pos) == r) src_of = J;);
        factorized = 1;
    }
    // lu_ldiv!, src/linear_algebra.jl:268-316: x = LU \ (P (rsc .* b)).  Lane r enters with b_r
    // and leaves with x_r.  The interchanged vector P b has at position p the entry of the lane
    // whose row sits at position p, so each lane's own entry is already in place.
    __device__ __forceinline__ cplx lu_solve_rs(cplx b, bool use_rs, double rsc) {
        flops += 8.0 * n * (n - 1) + 11.0 * n;
        cplx v = use_rs ? c_scale(rsc, b) : b;
        // forward substitution with the unit lower factor, positions 0..n-1
        G_FOR(NN, J, {
            if (J < n) {
                const int lj = g_first(pos == J, n);
                const cplx xj = gs_c(v, lj);
                if (pos > J && act()) v = c_submul(LU[J], xj, v);
            }
        });
        // back substitution, positions n-1..0
        G_FOR(NN, JJ, {
            constexpr int J = NN - 1 - JJ;
            if (J < n) {
                const int lj = g_first(pos == J, n);
                if (pos == J) v = c_div(v, LU[J]);
                const cplx xj = gs_c(v, lj);
                if (pos < J && act()) v = c_submul(LU[J], xj, v);
            }
        });
        // position r -> lane r
        return gs_c(v, src_of);
    }
    // lu_ldiv_adj!, src/linear_algebra.jl:318-354: solve (P^T L U)^H ... as the reference:
    // forward with U^H then back with L^H, then the interchanges in reverse.  Works on
    // positions; lane r enters with entry r and leaves with entry r.
    __device__ __forceinline__ cplx lu_solve_adj(cplx b) {
        flops += 8.0 * n * (n - 1) + 11.0 * n;
        // x at position p lives in the lane whose row sits at p: move entry r to lane src_of
        // ... equivalently each lane takes the entry whose index is its position
        cplx v = gs_c(b, pos < n ? pos : 0);
        // for j = 0..n-1: z = x[j] - sum_{i<j} conj(U[i,j]) x[i]; x[j] = z / conj(U[j,j])
        G_FOR(NN, J, {
            if (J < n) {
                // lane at position J accumulates over positions i < J in order
                cplx z = v;
                G_FOR(NN, I, {
                    if (I < J) {
                        const int li = g_first(pos == I, n);
                        const cplx xi = gs_c(v, li);
                        const cplx uij = gs_c(LU[J], li);  // U[i, J] from the lane at position i
                        if (pos == J) z = c_submul(c_conj(uij), xi, z);
                    }
                });
                if (pos == J) v = c_div(z, c_conj(LU[J]));
            }
        });
        // for j = n-1..0: z = x[j] - sum_{i=n-1..j+1} conj(L[i,j]) x[i]; x[j] = z
        G_FOR(NN, JJ, {
            constexpr int J = NN - 1 - JJ;
            if (J < n) {
                cplx z = v;
                G_FOR(NN, II, {
                    constexpr int I = NN - 1 - II;
                    if (I > J && I < n) {
                        const int li = g_first(pos == I, n);
                        const cplx xi = gs_c(v, li);
                        const cplx lij = gs_c(LU[J], li);  // L[i, J] from the lane at position i
                        if (pos == J) z = c_submul(c_conj(lij), xi, z);
                    }
                });
                if (pos == J) v = z;
            }
        });
        // back to index order, then the interchanges in reverse: x[i] <-> x[ipiv[i]]
        cplx e = gs_c(v, src_of);
        G_FOR(NN, II, {
            constexpr int I = NN - 1 - II;
            if (I < n) {
                const int pI = ipiv[I];
                if (pI != I) {
                    const cplx ei = gs_c(e, I), ep = gs_c(e, pI);
                    if (r == I) e = ep;
                    else if (r == pI) e = ei;
                }
            }
        });
        return e;
    }
    // LA.ldiv!(x, WS, b), src/linear_algebra.jl:389-408.  The workspace was factorized right
    // after the Jacobian's evaluation (GFACT_*), which is when the reference factorizes it at
    // its first use; an unfactorized workspace here would be a bug, and gives NaN.
    __device__ __forceinline__ cplx ws_ldiv(cplx b) {
        if (n == 1) { const cplx a0 = gb_c<0>(A[0]); return c_div(b, a0); }
        if (!factorized) return mkc(HCG_NAN, HCG_NAN);
        return lu_solve_rs(b, scaled != 0, rs);
    }
    // the factorization that follows a Jacobian evaluation (see GFACT_*)
    __device__ __forceinline__ void factor_after_eval(int mode) {
        ws_updated();
        if (mode == GFACT_SCALED) {  // jac_ldiv_w: Skeel row scaling with the current weights
            rs = skeel_row_scaling(w, -30.0);
            scaled = 1;
        }
        if (mode != GFACT_NONE && n > 1) lu_factor();
    }
    // skeel_row_scaling!, src/linear_algebra.jl:432-459; c: column weights (entry per lane)
    __device__ __forceinline__ double skeel_row_scaling(double c, double threshold) const {
        double rsum = 0.0;
        G_FOR(NN, J, if (J < n) rsum += c_fabs(A[J]) * gb_d<J>(c););
        const double m = g_fold<NN>(rsum, n, [](double a, double b) { return jmax(a, b); });
        const double s = threshold + m;
        int e = 0;
        if (isfinite(rsum)) frexp(rsum, &e);
        return (double)e < s ? 1.0 : ldexp(1.0, -e);
    }
    // mixed_precision_iterative_refinement!, src/linear_algebra.jl:528-544 (residual in double-double)
    __device__ __forceinline__ double mixed_refine(cplx& xv, cplx b, bool weighted) {
        flops += (80.0 * n * n + 26.0 * n);
        cdd acc = mkcdd(mkdd(0.0, 0.0), mkdd(0.0, 0.0));
        G_FOR(NN, J, if (J < n) { const cdd xj = cdd_from(gb_c<J>(xv)); acc = cdd_add(acc, cdd_mul_c(A[J], xj)); });
        tmp = cdd_round(cdd_sub_c(acc, b));
        rdx = ws_ldiv(tmp);
        xv = c_sub(xv, rdx);
        return norm_inf_v(rdx, weighted) / norm_inf_v(xv, weighted);
    }
    // fixed_precision_iterative_refinement!, src/linear_algebra.jl:553-567
    __device__ __forceinline__ double fixed_refine(cplx& xv, cplx b, bool weighted) {
        flops += (8.0 * n * n + 4.0 * n);
        cplx acc = mkc(0.0, 0.0);
        G_FOR(NN, J, if (J < n) acc = c_muladd(A[J], gb_c<J>(xv), acc););
        tmp = c_sub(acc, b);
        rdx = ws_ldiv(tmp);
        xv = c_sub(xv, rdx);
        return norm_inf_v(rdx, weighted) / norm_inf_v(xv, weighted);
    }
    // iterative_refinement!, src/linear_algebra.jl:864-885
    __device__ __forceinline__ double iterative_refinement(cplx& xv, cplx b, bool weighted, double tol, int max_iters) {
        double d = mixed_refine(xv, b, weighted);
        for (int i = 2; i <= max_iters; i++) {
            double d2 = mixed_refine(xv, b, weighted);
            if (d2 < tol) return d2;
            else if (d2 > 0.5 * d) return d2;
            d = d2;
        }
        return d;
    }
    // Jacobian ldiv! with WeightedNorm (src/linear_algebra.jl:846-862); the scaling and the
    // factorization were done after the evaluation (GFACT_SCALED)
    __device__ __forceinline__ cplx jac_ldiv_w(cplx b) { return ws_ldiv(b); }
    // inverse_inf_norm_est, src/linear_algebra.jl:585-682; dl, dr: entries (or 1 when absent)
    __device__ __forceinline__ double inverse_inf_norm_est(bool has_dl, double dl, bool has_dr, double dr) {
        if (!factorized) return HCG_NAN;
        const bool has_rs = scaled != 0;
        double xr = 1.0 / n;
        if (has_dr) xr /= dr;
        cplx y = mkc(xr, 0.0);
        y = lu_solve_adj(y);
        if (has_dl) y = c_divr(y, dl);
        if (has_rs) y = c_scale(rs, y);
        double g = g_sum<NN>(c_fabs(y), n, 0.0);
        y = c_divr(y, c_fabs(y));
        if (has_dl) y = c_divr(y, dl);
        if (has_rs) y = c_divr(y, rs);
        y = lu_solve_rs(y, false, 1.0);
        xr = has_dr ? y.re / dr : y.re;
        int k = 2;
        while (true) {
            // j = first index of the largest |x_j| (strictly greater replaces)
            int j = 0;
            double mx = gb_d<0>(fabs(xr));
            G_FOR(NN, I, if (I >= 1 && I < n) { const double v = gb_d<I>(fabs(xr)); if (v > mx) { j = I; mx = v; } });
            xr = (r == j) ? 1.0 : 0.0;
            if (has_dr) xr /= dr;
            y = mkc(xr, 0.0);
            y = lu_solve_adj(y);
            if (has_dl) y = c_divr(y, dl);
            if (has_rs) y = c_scale(rs, y);
            const double gb = g;
            g = g_sum<NN>(c_fabs(y), n, 0.0);
            if (g <= gb) { g = gb; break; }
            y = c_divr(y, c_fabs(y));
            if (has_dl) y = c_divr(y, dl);
            if (has_rs) y = c_divr(y, rs);
            y = lu_solve_rs(y, false, 1.0);
            xr = has_dr ? y.re / dr : y.re;
            k += 1;
            const double xn = g_fold<NN>(fabs(xr), n, [](double a, double b) { return jmax(a, b); });
            const double xj = gs_d(xr, j);
            if (xj == xn || k > 2) break;
        }
        return nanmin(g, HCG_INF);
    }
    // inf_norm, src/linear_algebra.jl:684-707
    __device__ __forceinline__ double inf_norm_ws(bool has_dl, double dl, bool has_dr, double dr) const {
        double ni = 0.0;
        G_FOR(NN, J, if (J < n) ni += has_dr ? c_fabs(A[J]) * gb_d<J>(dr) : c_fabs(A[J]););
        if (has_dl) ni *= dl;
        double nrm = -HCG_INF;
        nrm = max_fast(nrm, gb_d<0>(ni));
        G_FOR(NN, I, if (I >= 1 && I < n) nrm = max_fast(nrm, gb_d<I>(ni)););
        return nrm;
    }
    // LA.cond(WS, d_l, d_r), src/linear_algebra.jl:745-774
    __device__ __forceinline__ double ws_cond(bool has_dl, double dl, bool has_dr, double dr) {
        if (n == 1) {
            const double a = c_hypot(gb_c<0>(A[0]));
            const double dl0 = gb_d<0>(dl), dr0 = gb_d<0>(dr);
            if (!has_dl && !has_dr) return 1.0 / a;
            if (!has_dl) return 1.0 / (a * dr0);
            if (!has_dr) return 1.0 / (dl0 * a);
            return 1.0 / (dl0 * a * dr0);
        }
        return inverse_inf_norm_est(has_dl, dl, has_dr, dr) * inf_norm_ws(has_dl, dl, has_dr, dr);
    }
    // LinearAlgebra.rank(J, rtol): one-sided Jacobi singular values of a copy of A (the LU
    // registers); column sums taken over the rows in order
    __device__ __forceinline__ int rank_rtol(double rtol) {
        G_FOR(NN, J, LU[J] = A[J];);
        for (int sweep = 0; sweep < 60; sweep++) {
            double off = 0.0;
            for (int pp = 0; pp < n; pp++)
                for (int qq = pp + 1; qq < n; qq++) {
                    const cplx ap = row_get(LU, pp), aq = row_get(LU, qq);
                    double app = g_sum<NN>(c_abs2(ap), n, 0.0), aqq = g_sum<NN>(c_abs2(aq), n, 0.0);
                    const cplx prod = c_mul(c_conj(ap), aq);
                    double apq_re = 0.0, apq_im = 0.0;
                    {   // apq = sum_i conj(A_ip) A_iq, as c_add in order
                        cplx acc = mkc(0.0, 0.0);
                        G_FOR(NN, I, if (I < n) acc = c_add(acc, gb_c<I>(prod)););
                        apq_re = acc.re; apq_im = acc.im;
                    }
                    const cplx apq = mkc(apq_re, apq_im);
                    double g = sqrt(c_abs2(apq));
                    if (g == 0 || g <= 1e-300) continue;
                    double rel = g / sqrt(app * aqq);
                    if (rel > off) off = rel;
                    if (rel < 1e-15) continue;
                    double zeta = (aqq - app) / (2 * g);
                    double tt = (zeta >= 0 ? 1.0 : -1.0) / (fabs(zeta) + sqrt(1 + zeta * zeta));
                    double cs = 1 / sqrt(1 + tt * tt), sn = cs * tt;
                    cplx ph = mkc(apq.re / g, apq.im / g);
                    const cplx aqr = c_mul(aq, c_conj(ph));
                    const cplx np_ = c_sub(c_scale(cs, ap), c_scale(sn, aqr));
                    const cplx nq_ = c_mul(c_add(c_scale(sn, ap), c_scale(cs, aqr)), ph);
                    G_FOR(NN, J, { if (J == pp) LU[J] = np_; if (J == qq) LU[J] = nq_; });
                }
            if (off < 1e-15) break;
        }
        double smax = 0.0, s[NN];
        G_FOR(NN, PP, {
            if (PP < n) {
                s[PP] = sqrt(g_sum<NN>(c_abs2(LU[PP]), n, 0.0));
                if (s[PP] > smax) smax = s[PP];
            }
        });
        int rk = 0;
        G_FOR(NN, PP, if (PP < n && s[PP] > rtol * smax) rk++;);
        factorized = 0;
        return rk;
    }

    // ======================= segment stepper (src/utils.jl:300-401) =======================
    __device__ __forceinline__ void seg_init(cplx start, cplx target) {
        seg_start = start; seg_target = target;
        cplx d = c_sub(target, start);
        seg_absD = c_hypot(d);
        seg_forward = c_hypot(start) < c_hypot(target);
        seg_s = seg_s1 = seg_forward ? 0.0 : seg_absD;
    }
    __device__ __forceinline__ bool seg_is_done() const { return seg_forward ? seg_s == seg_absD : seg_s == 0.0; }
    __device__ __forceinline__ void seg_propose(double ds) {
        if (seg_forward) seg_s1 = jmin(seg_s + ds, seg_absD);
        else seg_s1 = jmax(seg_s - ds, 0.0);
    }
    __device__ __forceinline__ double seg_dist() const { return seg_forward ? seg_absD - seg_s : seg_s; }
    __device__ __forceinline__ double seg_ds() const { return seg_forward ? seg_s1 - seg_s : seg_s - seg_s1; }
    __device__ __forceinline__ cplx seg_th(double s) const {
        cplx st = seg_start, tg = seg_target; double D = seg_absD;
        if (seg_forward) {
            if (s == 0.0) return st;
            if (s == D) return tg;
            return c_add(st, c_scale(s / D, c_sub(tg, st)));
        }
        if (s == D) return st;
        if (s == 0.0) return tg;
        return c_add(tg, c_scale(s / D, c_sub(st, tg)));
    }
    __device__ __forceinline__ cplx seg_t() const { return seg_th(seg_s); }
    __device__ __forceinline__ cplx seg_t1() const { return seg_th(seg_s1); }
    __device__ __forceinline__ cplx seg_dt() const {
        if (seg_forward) return c_scale((seg_s1 - seg_s) / seg_absD, c_sub(seg_target, seg_start));
        return c_scale((seg_s - seg_s1) / seg_absD, c_sub(seg_target, seg_start));
    }

    // ======================= weighted norm =======================
    __device__ __forceinline__ void wnorm_init() {
        const double pn = norm_inf_v(x, false);
        double wi = c_fabs(x);
        if (wi < o->scale_min * pn) wi = o->scale_min * pn;
        else if (wi > o->scale_max * pn) wi = o->scale_max * pn;
        w = act() ? jmax(wi, o->scale_abs_min) : 1.0;
    }
    __device__ __forceinline__ void wnorm_update(cplx xv) {
        const double nx = norm_inf_v(xv, true);
        double wi = (c_fabs(xv) + w) / 2;
        if (wi < o->scale_min * nx) wi = o->scale_min * nx;
        else if (wi > o->scale_max * nx) wi = o->scale_max * nx;
        if (isfinite(wi) && act()) w = jmax(wi, o->scale_abs_min);
    }

    __device__ __forceinline__ void count_eval(int req_kind) {
        switch (req_kind) {
        case GREQ_EJ: flops += sys->fl_ej; break;
        case GREQ_E: flops += sys->fl_eval; break;
        case GREQ_DD: flops += sys->fl_dd; break;
        case GREQ_T1: flops += sys->fl_t1; break;
        case GREQ_T2: flops += sys->fl_t2; break;
        case GREQ_T3: flops += sys->fl_t3; break;
        default: break;  // GREQ_COND: counted by the linear algebra itself
        }
    }

    // ======================= Newton corrector =======================
    static __device__ __forceinline__ GNewtRes mkres(int c, double acc, int it, double om, double th, double ml, double n0) {
        GNewtRes R; R.code = c; R.accuracy = acc; R.iters = it; R.omega = om; R.theta = th; R.mu_low = ml; R.norm_dx0 = n0;
        return R;
    }
    // newton!, src/newton_corrector.jl:80-205; the start point is in xb on entry; every iterate
    // aliases xb.  Arguments: nw_t_*, nw_mu, nw_omega, nw_ext, nw_accurate_mu, nw_first.
    __device__ int newton() {
        CO_BEGIN(co_newton)
        nw_mu_low = HCG_NAN; nw_theta = HCG_NAN; nw_ndx = HCG_NAN; nw_ndx_prev = HCG_NAN; nw_ndx0 = HCG_NAN;
        nw_abar = o->a;
        for (nw_i = 0; nw_i <= 10; nw_i++) {
            newton_iters++;
            ev_x = xb; ev_t = mkc(nw_t_re, nw_t_im); ev_factor = GFACT_SCALED;
            CO_EVAL(co_newton, GREQ_EJ);
            res = ev_u;
            if (nw_ext) {
                ev_x = xb;
                CO_EVAL(co_newton, GREQ_DD);
                res = ev_u;
            }
            dx = jac_ldiv_w(res);
            if (nw_ext) iterative_refinement(dx, res, true, nw_abar * nw_abar, 3);
            nw_ndx = norm_inf_v(dx, true);
            if (isnan(nw_ndx)) { nres = mkres(GNEWT_SINGULARITY, norm_inf_v(dx, true), nw_i + 1, nw_omega, nw_theta, nw_mu_low, nw_ndx0); break; }
            xb = c_sub(xb, dx);
            if (nw_i == 0) nw_ndx0 = nw_ndx;
            if (nw_i == 1) nw_omega = 2 * nw_ndx / (nw_ndx_prev * nw_ndx_prev);
            if (nw_i >= 1) nw_theta = nw_ndx / nw_ndx_prev;
            if ((nw_i >= 1 && nw_theta > nw_abar) ||
                (nw_i == 0 && !nw_first && 0.125 * nw_ndx0 * nw_omega > h_of(o->a))) {
                nres = mkres(GNEWT_TERMINATED, norm_inf_v(dx, true), nw_i + 1, nw_omega, nw_theta, nw_mu_low, nw_ndx0);
                break;
            } else if (nw_omega * (nw_ndx * nw_ndx) < 2 * nw_mu * sqrt(1 - 2 * h_of(o->a))) {
                ev_x = xb; ev_t = mkc(nw_t_re, nw_t_im); ev_factor = GFACT_PLAIN;
                CO_EVAL(co_newton, GREQ_EJ);
                res = ev_u;
                if (nw_ext) {
                    dx = ws_ldiv(res);
                    nw_mu_low = norm_inf_v(dx, true);
                    ev_x = xb;
                    CO_EVAL(co_newton, GREQ_DD);
                    res = ev_u;
                }
                dx = ws_ldiv(res);
                if (nw_ext) iterative_refinement(dx, res, true, nw_abar * nw_abar, 3);
                xb = c_sub(xb, dx);
                nw_ndx1 = norm_inf_v(dx, true);
                if (isnan(nw_ndx1)) { nres = mkres(GNEWT_SINGULARITY, norm_inf_v(dx, true), nw_i + 1, nw_omega, nw_theta, nw_mu_low, nw_ndx0); break; }
                if (nw_ndx1 > sqrt(nw_ndx)) {
                    nw_theta = nw_ndx1 / nw_ndx;
                    nres = mkres(GNEWT_TERMINATED, nw_ndx1, nw_i + 2, nw_omega, nw_theta, nw_mu_low, nw_ndx0);
                    break;
                }
                if (nw_ndx1 > 2 * nw_mu && nw_ext) {
                    ev_x = xb;
                    CO_EVAL(co_newton, GREQ_DD);
                    res = ev_u;
                    dx = ws_ldiv(res);
                    nw_ndx = nw_ndx1;
                    nw_ndx1 = norm_inf_v(dx, true);
                    nw_mu = nw_ndx1;
                } else if (nw_ndx1 > 2 * nw_mu || nw_accurate_mu) {
                    ev_x = xb;
                    CO_EVAL(co_newton, GREQ_E);
                    res = ev_u;
                    dx = ws_ldiv(res);
                    nw_mu = norm_inf_v(dx, true);
                } else {
                    nw_mu = nw_ndx1;
                }
                if (nw_i == 0) {
                    double wb = 2 * nw_ndx / (nw_ndx1 * nw_ndx1);
                    if (wb < nw_omega) nw_omega = wb; else nw_omega *= 0.25;
                }
                nres = mkres(GNEWT_CONVERGED, nw_mu, nw_i + 2, nw_omega, nw_theta, nw_mu_low, nw_ndx0);
                break;
            }
            nw_ndx_prev = nw_ndx;
            if (nw_i >= 1) nw_abar *= nw_abar;
        }
        if (nw_i > 10) nres = mkres(GNEWT_MAX_ITERS, nw_mu, 11, nw_omega, nw_theta, nw_mu_low, nw_ndx0);
        CO_END(co_newton)
    }

    // init_newton!, src/newton_corrector.jl:207-286; point x, in_t, in_ext; results in_valid,
    // in_omega, in_mu
    __device__ int init_newton() {
        CO_BEGIN(co_initn)
        ev_x = x; ev_t = in_t; ev_factor = GFACT_SCALED;
        CO_EVAL(co_initn, GREQ_EJ);
        res = ev_u;
        if (in_ext) { ev_x = x; CO_EVAL(co_initn, GREQ_DD); res = ev_u; }
        dx = jac_ldiv_w(res);
        in_v = norm_inf_v(dx, true) + DEPS;
        in_valid = 0;
        in_omega = HCG_NAN; in_mu = HCG_NAN;
        in_eps = sqrt(in_v);
        for (in_k = 1; in_k <= 3; in_k++) {
            xb = mkc(x.re + in_eps * w, x.im);
            ev_x = xb; ev_t = in_t; ev_factor = GFACT_SCALED;
            CO_EVAL(co_initn, GREQ_EJ);
            res = ev_u;
            if (in_ext) { ev_x = xb; CO_EVAL(co_initn, GREQ_DD); res = ev_u; }
            dx = jac_ldiv_w(res);
            xb = c_sub(xb, dx);
            in_ndx0 = norm_inf_v(dx, true);
            ev_x = xb;
            if (in_ext) { CO_EVAL(co_initn, GREQ_DD); }
            else { CO_EVAL(co_initn, GREQ_E); }
            res = ev_u;
            dx = jac_ldiv_w(res);
            xb = c_sub(xb, dx);
            in_ndx1 = norm_inf_v(dx, true) + DEPS;
            if (in_ndx1 < o->a * in_ndx0) {
                in_omega = 2 * in_ndx1 / (in_ndx0 * in_ndx0);
                in_mu = in_ndx1;
                if (in_omega * in_mu > ipow(o->a, 7)) {
                    nw_t_re = in_t.re; nw_t_im = in_t.im;
                    nw_mu = ipow(o->a, 7) / in_omega; nw_omega = in_omega;
                    nw_ext = in_ext; nw_accurate_mu = 1; nw_first = 0;
                    CO_AWAIT(co_initn, newton());
                    if (nres.code == GNEWT_CONVERGED) { in_valid = 1; in_omega = nres.omega; in_mu = nres.accuracy; }
                    else in_valid = 0;
                } else {
                    in_valid = 1;
                    break;
                }
            } else {
                in_eps *= sqrt(in_eps);
            }
        }
        CO_END(co_initn)
    }

    // extended_prec_refinement_step!, src/newton_corrector.jl:55-78 from er_xin at er_t;
    // writes x - Δx into xb (er_target_xb) or x; result er_out
    __device__ int ext_refine_step() {
        CO_BEGIN(co_ext)
        ev_x = er_xin; ev_t = er_t; ev_factor = GFACT_SCALED;
        CO_EVAL(co_ext, GREQ_EJ);
        ev_x = er_xin;
        CO_EVAL(co_ext, GREQ_DD);
        res = ev_u;
        dx = jac_ldiv_w(res);
        iterative_refinement(dx, res, true, 1e-8, 3);
        if (er_target_xb) xb = c_sub(er_xin, dx); else x = c_sub(er_xin, dx);
        if (er_simple) {
            ev_x = er_target_xb ? xb : x;
            CO_EVAL(co_ext, GREQ_DD);
            res = ev_u;
            dx = jac_ldiv_w(res);
        }
        er_out = norm_inf_v(dx, true);
        CO_END(co_ext)
    }

    // ======================= predictor =======================
    static __device__ __forceinline__ cplx t_to_s_plane(cplx t, int m) {
        double rr0 = c_fabs(t);
        if (t.im == 0.0 && t.re > 0.0) return mkc(nthroot_inl(rr0, m), 0.0);
        double th = atan2(t.im, t.re);
        th = fmod(th, 2 * HCG_PI); if (th < 0) th += 2 * HCG_PI;
        double rr = nthroot_inl(rr0, m);
        return mkc(rr * cos(th / m), rr * sin(th / m));
    }
    static __device__ __forceinline__ cplx cpowi(cplx z, int k) { cplx v = mkc(1.0, 0.0); for (int i = 0; i < k; i++) v = c_mul(v, z); return v; }

    // update!(predictor, ...), src/predictor.jl:158-284 at pu_t; pu_have_xhat: xh holds x̂
    __device__ int pred_update() {
        CO_BEGIN(co_pred)
        ptx0 = tx0; ptx1 = tx1;
        pred_prev_t = pred_t; pred_t = pu_t;
        if (winding > 1) { pred_prev_s = pred_s; pred_s = t_to_s_plane(pu_t, winding); }
        if (!pu_have_xhat) local_error = HCG_NAN;
        else local_error = dist_inf_v(xh, x, true) / ipow(c_fabs(c_sub(pu_t, pred_prev_t)), order);
        tx0 = x;
        txn[0] = norm_inf_v(x, true);
        ev_t = pu_t;
        CO_EVAL(co_pred, GREQ_T1);
        pu_ub = c_neg(ev_u);
        pu_xt = ws_ldiv(pu_ub);
        pu_delta = fixed_refine(pu_xt, pu_ub, true);
        cond_H_xd = pu_delta / DEPS;
        if (pu_delta > 1e-10) iterative_refinement(pu_xt, pu_ub, false, 1e-10, 5);
        txn[1] = norm_inf_v(pu_xt, true);
        tx1 = pu_xt;
        if (winding > 1) {
            hermite = 1;
            order = 4;
            trust_region = txn[0] / txn[1];
            if (isnan(local_error)) { double qq = txn[1] / txn[0]; local_error = qq * qq * qq; }
        } else {
            ev_t = pu_t;
            CO_EVAL(co_pred, GREQ_T2);
            pu_ub = c_neg(ev_u);
            pu_xt = ws_ldiv(pu_ub);
            if (pu_delta > 1e-10) iterative_refinement(pu_xt, pu_ub, true, 1e-10, 4);
            txn[2] = norm_inf_v(pu_xt, true);
            tx2 = pu_xt;
            ev_t = pu_t;
            CO_EVAL(co_pred, GREQ_T3);
            pu_ub = c_neg(ev_u);
            pu_xt = ws_ldiv(pu_ub);
            if (pu_delta > 1e-4) iterative_refinement(pu_xt, pu_ub, true, 1e-4, 3);
            txn[3] = norm_inf_v(pu_xt, true);
            tx3 = pu_xt;
            {
                double ti = HCG_INF;
                double c1 = c_fabs(tx1), c2 = c_fabs(tx2), c3 = c_fabs(tx3);
                double lam = jmax(1e-6, c1);
                c1 /= lam; c2 /= lam * lam; c3 /= lam * lam * lam;
                double tol = 1e-14 * jmax(jmax(c1, c2), c3);
                if (!((c1 <= tol && c2 <= tol && c3 <= tol) || c2 <= tol)) ti = (c2 / c3) / lam;
                // tau = the smallest ti over the entries (a NaN ti never replaces)
                double tauv = HCG_INF;
                G_FOR(NN, J, if (J < n) { const double v = gb_d<J>(ti); if (v < tauv) tauv = v; });
                if (!isfinite(tauv)) tauv = txn[2] / txn[3];
                if (!isfinite(tauv)) {
                    double mx = txn[0];
                    for (int k = 1; k < 4; k++) mx = jmax(mx, txn[k]);
                    tauv = txn[0] / mx;
                }
                hermite = 0;
                order = 4;
                trust_region = tauv;
                if (isnan(local_error)) { double it = 1.0 / tauv; double it2 = it * it; local_error = it2 * it2; }
            }
        }
        CO_END(co_pred)
    }

    static __device__ __forceinline__ cplx hermite_real(cplx a0, cplx a1, double t0, cplx b0, cplx b1, double t1, double t) {
        double s = (t - t0) / (t1 - t0);
        double oms = 1 - s;
        double h00 = (1 + 2 * s) * (oms * oms), h10 = (t - t0) * (oms * oms);
        double h01 = (s * s) * (3 - 2 * s), h11 = (t - t0) * s * (s - 1);
        return c_add(c_add(c_add(c_scale(h00, a0), c_scale(h10, a1)), c_scale(h01, b0)), c_scale(h11, b1));
    }
    static __device__ __forceinline__ cplx hermite_cplx(cplx a0, cplx a1, cplx t0, cplx b0, cplx b1, cplx t1, cplx t) {
        cplx s = c_div(c_sub(t, t0), c_sub(t1, t0));
        cplx oms = c_sub(mkc(1, 0), s);
        cplx oms2 = c_mul(oms, oms);
        cplx h00 = c_mul(c_add(mkc(1, 0), c_scale(2, s)), oms2), h10 = c_mul(c_sub(t, t0), oms2);
        cplx h01 = c_mul(c_mul(s, s), c_sub(mkc(3, 0), c_scale(2, s))), h11 = c_mul(c_mul(c_sub(t, t0), s), c_sub(s, mkc(1, 0)));
        return c_add(c_add(c_add(c_mul(h00, a0), c_mul(h10, a1)), c_mul(h01, b0)), c_mul(h11, b1));
    }
    // predict!, src/predictor.jl:286-329
    __device__ __forceinline__ void predict(cplx t, cplx dt) {
        if (!hermite) {
            double lam = trust_region, lam2 = lam * lam, lam3 = lam * lam * lam;
            cplx x0 = tx0, x1 = tx1, x2 = tx2, x3 = tx3;
            double c = c_fabs(x0), c1 = c_fabs(x1), c2 = c_fabs(x2), c3 = c_fabs(x3);
            double a1 = c1 * lam, a2 = c2 * lam2, a3 = c3 * lam3;
            double tauv = 1e-12 * sqrt(c * c + a1 * a1 + a2 * a2 + a3 * a3);
            if (c3 * lam3 <= tauv || c2 * lam2 <= tauv) {
                xh = c_add(x0, c_mul(dt, c_add(x1, c_mul(dt, x2))));
            } else {
                cplx di = c_sub(mkc(1, 0), c_div(c_mul(dt, x3), x2));
                xh = c_add(x0, c_mul(dt, c_add(x1, c_div(c_mul(dt, x2), di))));
            }
        } else {
            int m = winding;
            cplx ps = t_to_s_plane(pred_prev_t, m), s = t_to_s_plane(t, m), s1 = t_to_s_plane(c_add(t, dt), m);
            cplx psm, sm;
            if (m == 2) { psm = c_scale(2, ps); sm = c_scale(2, s); }
            else { psm = c_scale((double)m, cpowi(ps, m - 1)); sm = c_scale((double)m, cpowi(s, m - 1)); }
            cplx a1 = c_mul(psm, ptx1), b1 = c_mul(sm, tx1);
            if (ps.im == 0.0 && s.im == 0.0 && s1.im == 0.0) xh = hermite_real(ptx0, a1, ps.re, tx0, b1, s.re, s1.re);
            else xh = hermite_cplx(ptx0, a1, ps, tx0, b1, s, s1);
        }
    }

    // ======================= tracker =======================
    __device__ __forceinline__ int ext_steps() const { return ext_acc + ext_rej; }
    // initial_step_size, src/tracker.jl:520-539
    __device__ __forceinline__ double initial_step_size() const {
        double a = o->beta_a * o->a;
        double e = local_error;
        if (isinf(e)) e = 1e5;
        double ds1 = nthroot_inl((sqrt(1 + 2 * h_of(a)) - 1) / (omega * e), order) / o->beta_omega_p;
        double ds2 = o->beta_tau * trust_region;
        return jmin(jmin(nanmin(ds1, ds2), o->max_step_size), o->max_initial_step_size);
    }

    // step features and the learned proposal (see hcg_learned.h for the rules)
    __device__ __forceinline__ void step_features(const GNewtRes& rr, double ds0) {
        feat[0] = feat_clean(hc_log_inl(ds0));
        feat[1] = feat_clean(rr.theta);
        feat[2] = feat_clean(hc_log_inl(local_error * ipow(ds0, 4)));
        feat[3] = feat_clean(seg_t().re);
        feat[4] = feat_clean(hc_log_inl(cond_H_xd) / HCG_LN10);
        feat[5] = feat_clean(hc_log_inl(omega));
        feat[6] = feat_clean(hc_log_inl(mu));
        feat[7] = feat_clean(hc_log_inl(tau));
    }
    // the network of hcg_mlp_infer.h evaluated by the 16 lanes together (each lane 4 hidden
    // units), with every sum in the order of mlp_forward, so the output is identical
    __device__ __forceinline__ void mlp_forward_group(float* out) {
        // 128 floats of this group's register file (free between evaluations); both the
        // hidden units and the staged weights are read through local (LDS) pointers
        typedef __attribute__((address_space(3))) float lds_f32;
        typedef __attribute__((address_space(3))) const unsigned short lds_u16;
        lds_f32* hbuf = (lds_f32*)lds;
        lds_u16* wv = (lds_u16*)mlp_w;
        // lanes 0..15 hold 4 hidden units each (in a 32-lane group lanes 16..31 have none)
        const bool has_units = r < HCG_MLP_HIDDEN / 4;
#pragma unroll
        for (int u = 0; u < 4; u++) {
            if (!has_units) break;
            const int j = 4 * r + u;
            float s = bf16_to_f32(wv[MLP_B1 + j]);
#pragma unroll
            for (int i = 0; i < HCG_MLP_IN; i++) s = fmaf(bf16_to_f32(wv[MLP_W1 + j * HCG_MLP_IN + i]), feat[i], s);
            hbuf[j] = mlp_gelu(s);
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
#pragma unroll
        for (int u = 0; u < 4; u++) {
            if (!has_units) break;
            const int j = 4 * r + u;
            float s = bf16_to_f32(wv[MLP_B2 + j]);
            for (int k = 0; k < HCG_MLP_HIDDEN; k++) s = fmaf(bf16_to_f32(wv[MLP_W2 + j * HCG_MLP_HIDDEN + k]), hbuf[k], s);
            hbuf[HCG_MLP_HIDDEN + j] = mlp_gelu(s);
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        float o0 = bf16_to_f32(wv[MLP_B3]), o1 = bf16_to_f32(wv[MLP_B3 + 1]);
        for (int j = 0; j < HCG_MLP_HIDDEN; j++) {
            const float h = hbuf[HCG_MLP_HIDDEN + j];
            o0 = fmaf(bf16_to_f32(wv[MLP_W3 + j]), h, o0);
            o1 = fmaf(bf16_to_f32(wv[MLP_W3 + HCG_MLP_HIDDEN + j]), h, o1);
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        out[0] = o0;
        out[1] = o1;
    }
    __device__ __forceinline__ double learned_adjust(double ds) {
        ds_plain = ds;
        last_learned = 0;
        if (!mlp_w || extended_prec || singular_endgame || last_failed > 0) return ds;
        if (fabs(seg_t().re) <= o->learned_min_t) return ds;
        float out[2];
        mlp_forward_group(out);
        inf_logit = out[1];
        {
            double tn = fabs(seg_t().re);
            if (o->learned_inf_t > 0.0 && out[1] > o->learned_inf_logit && tn <= o->learned_inf_t && tn > eg_start)
                eg_start = tn;
        }
        double lm = jclamp((double)out[0], 0.0, hc_log_inl(o->learned_max_mult));
        if (!(lm > 0.0)) return ds;
        double dsl = jmin(jmin(ds * hc_exp_inl(lm), 10 * ds_prev), o->max_step_size);
        if (!(dsl > ds)) return ds;
        last_learned = 1;
        learned_proposed++;
        return dsl;
    }

    // one step record (layout HCG_LOG_* of hcgpu_types.h), written by lane 0 of the group
    __device__ __forceinline__ void log_step(const GNewtRes& rr, int acc) {
        const double lds_next = hc_log_inl(seg_ds());
        const double lacc = hc_log_inl(accuracy) / HCG_LN10;
        const double lplain = isnan(ds_plain) ? lds_next : hc_log_inl(ds_plain);
        if (r != 0) return;
        int idx = atomicAdd(log_count, 1);
        if (idx >= log_capacity) return;
        float* q = log_records + (size_t)idx * HCG_STEP_RECORD_FLOATS;
        q[HCG_LOG_PATH] = (float)path;
        q[HCG_LOG_STEP] = (float)(accepted + rejected - 1);
        for (int i = 0; i < HCG_FEATS; i++) q[HCG_LOG_FEATURES + i] = feat[i];
        q[HCG_LOG_ACCEPTED] = (float)acc;
        q[HCG_LOG_NEWTON_ITERS] = (float)rr.iters;
        q[HCG_LOG_NEXT_LOG_DS] = (float)lds_next;
        q[HCG_LOG_FLAGS] = (float)(extended_prec + 2 * singular_endgame + 4 * last_learned);
        q[HCG_LOG_LOG10_ACCURACY] = (float)lacc;
        q[HCG_LOG_PLAIN_LOG_DS] = (float)lplain;
    }

    // update_stepsize!, src/tracker.jl:541-588
    __device__ __forceinline__ void update_stepsize(const GNewtRes& rr) {
        double a = o->beta_a * o->a;
        double om = jclamp(omega + 2 * (omega - omega_prev), omega, 8 * omega);
        double ds;
        if (rr.code == GNEWT_CONVERGED) {
            double ds1 = nthroot_inl((sqrt(1 + 2 * h_of(a)) - 1) / (om * local_error), order) / o->beta_omega_p;
            double ds2 = o->beta_tau * tau;
            if (seg_dist() < ds2) ds2 = o->strict_beta_tau * tau;
            ds = jmin(nanmin(ds1, ds2), o->max_step_size);
            ds = jmin(ds, 10 * ds_prev);
            if (last_failed > 0) ds = jmin(ds, ds_prev);
            ds = learned_adjust(ds);
        } else {
            int j = rr.iters - 2;
            double Th = (j >= 0) ? nthroot_inl(rr.theta, 1 << j) : 1.0;
            double hT = h_of(Th), ha = h_of(0.5 * a);
            if (isnan(Th) || rr.code == GNEWT_SINGULARITY || isnan(rr.accuracy) || rr.iters == 1 || hT < ha)
                ds = 0.25 * seg_ds();
            else
                ds = nthroot_inl((sqrt(1 + 2 * h_of(0.5 * a)) - 1) / (sqrt(1 + 2 * h_of(Th)) - 1), order) * seg_ds();
        }
        seg_propose(ds);
    }
    // check_terminated!, src/tracker.jl:591-619
    __device__ __forceinline__ void check_terminated() {
        double tol_acc = (extended_prec || !o->extended_precision) ? ipow(o->a, (1 << o->min_newton_iters) - 1) * h_of(o->a) : HCG_INF;
        cplx t1 = seg_t1(), t = seg_t();
        if (seg_is_done()) code = GTC_SUCCESS;
        else if (accepted + rejected >= o->max_steps) code = GTC_MAX_STEPS;
        else if (omega * mu > tol_acc) code = GTC_ACCURACY_LIMIT;
        else if (seg_ds() < o->min_step_size) code = GTC_STEP_TOO_SMALL;
        else if (c_fabs(c_sub(t1, t)) <= 2 * jeps(c_fabs(t))) code = GTC_STEP_TOO_SMALL;
        else if (o->min_rel_step_size > 0 && !(t1.re == seg_target.re && t1.im == seg_target.im) &&
                 c_fabs(c_sub(t1, t)) < c_fabs(t) * o->min_rel_step_size)
            code = GTC_STEP_TOO_SMALL;
    }

    // init!(tracker, x₁, t₁, t₀), src/tracker.jl:639-754 (x holds x₁); t0 = 0
    __device__ int tracker_init(double t1) {
        CO_BEGIN(co_tinit)
        seg_init(mkc(t1, 0.0), mkc(0.0, 0.0));
        ds_prev = 0.0;
        accuracy = DEPS;
        omega = 1.0;
        keep_ext = 0;
        wnorm_init();
        code = GTC_TRACKING;
        accepted = rejected = ext_acc = ext_rej = 0;
        last_failed = 0;
        ti_t = seg_t();
        for (ti_extp = 0; ti_extp < 2; ti_extp++) {
            in_t = ti_t; in_ext = ti_extp;
            CO_AWAIT(co_tinit, init_newton());
            if (in_valid) break;
        }
        if (ti_extp > 1) ti_extp = 1;
        ti_valid = in_valid; ti_omega = in_omega; ti_mu = in_mu;
        used_ext = extended_prec = ti_extp;
        if (!isnan(ti_omega)) omega = ti_omega;
        if (ti_valid) {
            accuracy = ti_mu;
            mu = jmax(ti_mu, DEPS);
        } else {
            ev_x = x; ev_t = ti_t; ev_factor = GFACT_NONE;
            CO_EVAL(co_tinit, GREQ_EJ);
            res = ev_u;
            {
                bool nanrow = false;
                G_FOR(NN, J, if (J < n && c_isnan(A[J])) nanrow = true;);
                const bool anynan = g_any(nanrow, n);
                const int corank = anynan ? 0 : n - rank_rtol(1e-14);
                code = corank > 0 ? GTC_INVALID_START_SINGULAR : GTC_INVALID_START;
            }
            co_tinit = 0;
            return 0;
        }
        tau = HCG_INF;
        ev_x = x; ev_t = ti_t; ev_factor = GFACT_PLAIN;  // first used by the predictor's ldiv!
        CO_EVAL(co_tinit, GREQ_EJ);
        res = ev_u;
        // init!(predictor), src/predictor.jl:124-133
        winding = 1;
        pred_t = pred_prev_t = mkc(HCG_NAN, 0.0);
        pred_s = pred_prev_s = mkc(HCG_NAN, 0.0);
        trust_region = local_error = HCG_NAN;
        pu_t = ti_t; pu_have_xhat = 0;
        CO_AWAIT(co_tinit, pred_update());
        tau = trust_region;
        {
            double ds = jmax(initial_step_size(), o->min_step_size);
            seg_propose(ds);
        }
        omega_prev = omega;
        CO_END(co_tinit)
    }
    // init!(tracker, t₀), src/tracker.jl:756-766
    __device__ __forceinline__ void tracker_init_t0(cplx t0) {
        code = GTC_TRACKING;
        seg_init(seg_t(), t0);
        seg_propose(initial_step_size());
        ds_prev = 0.0;
    }
    // update_precision! / use_extended_precision!, src/tracker.jl:768-813 (argument up_mu = mu_low)
    __device__ int update_precision() {
        CO_BEGIN(co_upprec)
        if (o->extended_precision) {
            if (extended_prec && !keep_ext && !isnan(up_mu) && up_mu > mu) {
                if (up_mu * omega < ipow(o->a, 7) * h_of(o->a)) { extended_prec = 0; mu = up_mu; }
            } else if (mu * omega > ipow(o->a, 5) * h_of(o->a)) {
                if (!extended_prec) {
                    extended_prec = 1;
                    used_ext = 1;
                    rf_t = seg_t();
                    up_mu = mu;
                    for (up_i = 0; up_i < 2; up_i++) {
                        er_xin = x; er_t = rf_t; er_simple = 0; er_target_xb = 0;
                        CO_AWAIT(co_upprec, ext_refine_step());
                        up_mu = er_out;
                    }
                    mu = jmax(up_mu, DEPS);
                }
            }
        }
        CO_END(co_upprec)
    }
    // refine_current_solution!, src/tracker.jl:815-844 (arguments rf_min_tol, rf_nsteps; result rf_mu)
    __device__ int refine_current_solution() {
        CO_BEGIN(co_refine)
        rf_t = seg_t();
        rf_mu = accuracy;
        er_xin = x; er_t = rf_t; er_simple = 0; er_target_xb = 1;
        CO_AWAIT(co_refine, ext_refine_step());
        rf_mub = er_out;
        if (rf_mub < rf_mu) { x = xb; rf_mu = rf_mub; }
        rf_k = 1;
        while (rf_mu > rf_min_tol && rf_k <= rf_nsteps) {
            er_xin = x; er_t = rf_t; er_simple = 1; er_target_xb = 1;
            CO_AWAIT(co_refine, ext_refine_step());
            rf_mub = er_out;
            if (rf_mub < rf_mu) { x = xb; rf_mu = rf_mub; }
            rf_k++;
        }
        CO_END(co_refine)
    }

    // step!(tracker), src/tracker.jl:851-926; result ts_ok = !last_step_failed
    __device__ int tracker_step() {
        CO_BEGIN(co_tstep)
        ts_t = seg_t(); ts_dt = seg_dt(); ts_t1 = seg_t1();
        ts_ds0 = seg_ds();
        predict(ts_t, ts_dt);
        wnorm_update(xh);
        xb = xh;
        nw_t_re = ts_t1.re; nw_t_im = ts_t1.im; nw_mu = mu; nw_omega = omega;
        nw_ext = extended_prec; nw_accurate_mu = 0; nw_first = accepted == 0;
        CO_AWAIT(co_tstep, newton());
        ts_res = nres;
        if (ts_res.code == GNEWT_CONVERGED && last_learned) learned_accepted++;
        if (ts_res.code == GNEWT_CONVERGED) {
            x = xb;
            ds_prev = seg_ds();
            seg_s = seg_s1;
            accuracy = ts_res.accuracy;
            mu = jmax(ts_res.accuracy, DEPS);
            omega_prev = omega;
            omega = jmax(jmax(ts_res.omega, 0.5 * omega), 0.1);
            up_mu = ts_res.mu_low;
            CO_AWAIT(co_tstep, update_precision());
            if (seg_is_done() && o->extended_precision && accuracy > 1e-14) {
                rf_min_tol = 1e-14; rf_nsteps = 3;
                CO_AWAIT(co_tstep, refine_current_solution());
                accuracy = rf_mu;
                refined_ext = 1;
            }
            pu_t = seg_t(); pu_have_xhat = 1;
            CO_AWAIT(co_tstep, pred_update());
            tau = trust_region;
            accepted++;
            ext_acc += extended_prec;
            last_failed = 0;
        } else {
            rejected++;
            ext_rej += extended_prec;
            last_failed++;
        }
        if (log_records || mlp_w) step_features(ts_res, ts_ds0);
        last_learned = 0;
        ds_plain = HCG_NAN;
        update_stepsize(ts_res);
        if (log_records) log_step(ts_res, ts_res.code == GNEWT_CONVERGED);
        check_terminated();
        ts_ok = !(last_failed > 0);
        CO_END(co_tstep)
    }

    // ======================= valuation (src/valuation.jl) =======================
    static __device__ __forceinline__ double nu1(cplx xv, cplx xd, double t) {
        double xx = xv.re * xv.re + xv.im * xv.im;
        double m = xv.re * xd.re + xv.im * xd.im;
        return t * (m / xx);
    }
    static __device__ __forceinline__ void nu_nu1(cplx xv, cplx xd, cplx x2, double t, double* v, double* v1) {
        double xx = xv.re * xv.re + xv.im * xv.im;
        double m = xv.re * xd.re + xv.im * xd.im;
        double l = m / xx;
        double m1 = xv.re * x2.re + xd.re * xd.re + xv.im * x2.im + xd.im * xd.im;
        double l1 = m1 / xx - 2 * (l * l);
        *v = t * l; *v1 = t * l1 + l;
    }
    static __device__ __forceinline__ double finite_diff(double v, double s, double v2, double s2, double v1, double s1) {
        double D1 = s - s1, D2 = s - s2, D12 = s1 - s2;
        return (D2 * v1) / (D12 * D1) - ((D12 + D2) * v2) / (D12 * D2) - (D12 * v) / (D1 * D2);
    }
    // update!(val, pred, t), src/valuation.jl:82-124
    __device__ __forceinline__ void val_update(double t) {
        double logt = hc_log_inl(t);
        bool diff = winding > 1 && !isnan(logt2);
        double ldxi = hc_log_inl(c_fabs(tx1));
        if (diff) {
            double vi = nu1(tx0, tx1, t);
            double dvi = finite_diff(vi, logt, vx2, logt2, vx1, logt1);
            val_x = vi; dval_x = dvi;
            double vdx = finite_diff(ldxi, logt, ldx2, logt2, ldx1, logt1);
            double dvdx = finite_diff(vdx, logt, 0.0, logt2, 0.0, logt1);
            val_tx = vdx + 1; dval_tx = dvdx;
        } else {
            double vi, v1;
            nu_nu1(tx0, tx1, c_scale(2, tx2), t, &vi, &v1);
            val_x = vi; dval_x = t * v1;
            vx1 = vx2; vx2 = vi;
            double vdx, v1b;
            nu_nu1(tx1, c_scale(2, tx2), c_scale(6, tx3), t, &vdx, &v1b);
            val_tx = vdx + 1; dval_tx = t * v1b;
        }
        ldx1 = ldx2; ldx2 = ldxi;
        logt1 = logt2; logt2 = logt;
    }
    __device__ __forceinline__ double val_eps_inf() const {
        double vx = val_x;
        return jmax(jmax(fabs(1.0 - val_tx / vx), fabs(dval_x / vx)), fabs(dval_tx / val_tx));
    }
    // is_finite(val), src/valuation.jl:175-205 (every entry must pass)
    __device__ __forceinline__ bool val_is_finite(double finite_tol, bool zero_is_finite, int max_winding) const {
        double delta = 1.0 / max_winding;
        bool ok;
        double vx = val_x;
        if (fabs(vx) < finite_tol) ok = (fabs(dval_x) < finite_tol) && !(val_tx < 0.5 * delta);
        else if (zero_is_finite && vx > (delta - finite_tol)) ok = val_eps_inf() < finite_tol;
        else ok = false;
        return g_all(ok, n);
    }
    // estimate_winding_number, src/valuation.jl:207-228
    __device__ __forceinline__ int estimate_winding_number(int max_w, double* err_out) const {
        int m = 1; double min_err = HCG_INF;
        for (int k = 1; k <= max_w; k++) {
            double mv = k * val_tx;
            double ei = fabs(rint(mv) - mv);
            double err = 0.0;
            G_FOR(NN, J, if (J < n) { const double e = gb_d<J>(ei); err = e > err ? e : err; });
            if (err < min_err) { m = k; min_err = err; }
        }
        *err_out = min_err;
        return m;
    }

    // ======================= endgame =======================
    // cond at (ca_x, ca_t) with row_scaling (ca_use_rs) and col_scaling: result ca_out
    __device__ int cond_at() {
        CO_BEGIN(co_condat)
        ev_x = ca_x; ev_t = ca_t; ev_factor = GFACT_PLAIN;
        CO_EVAL(co_condat, GREQ_EJ);
        res = ev_u;
        CO_EVAL(co_condat, GREQ_COND);
        ca_out = ev_cond;
        CO_END(co_condat)
    }
    __device__ __forceinline__ void eg_row_scaling() { row_scaling = skeel_row_scaling(col_scaling, o->scaling_threshold); }
    static __device__ __forceinline__ int tc_to_eg(int c) {
        switch (c) {
        case GTC_SUCCESS: return GEG_SUCCESS;
        case GTC_MAX_STEPS: return GEG_MAX_STEPS;
        case GTC_ACCURACY_LIMIT: return GEG_ACCURACY_LIMIT;
        case GTC_ILL_CONDITIONED: return GEG_ILL_CONDITIONED;
        case GTC_INVALID_START: return GEG_INVALID_START;
        case GTC_INVALID_START_SINGULAR: return GEG_INVALID_START_SINGULAR;
        case GTC_STEP_TOO_SMALL: return GEG_STEP_TOO_SMALL;
        case GTC_UNKNOWN: return GEG_UNKNOWN;
        default: return GEG_TRACKING;
        }
    }
    // tracking_stopped!, src/endgame_tracker.jl:695-721
    __device__ int tracking_stopped() {
        CO_BEGIN(co_tstop)
        eg_accuracy = accuracy;
        if (eg_code == GEG_SUCCESS && eg_accuracy > 1e-14) {
            rf_min_tol = 1e-14; rf_nsteps = o->refine_steps;
            CO_AWAIT(co_tstop, refine_current_solution());
        }
        solution = x;
        eg_winding = 0;
        if (eg_code == GEG_SUCCESS) {
            col_scaling = w;
            eg_row_scaling();
            ca_x = solution; ca_t = mkc(0, 0);
            CO_AWAIT(co_tstop, cond_at());
            eg_cond = ca_out;
            eg_singular = eg_cond > o->sing_cond || eg_accuracy > o->sing_accuracy;
        }
        CO_END(co_tstop)
    }
    // add_sample!, src/endgame_tracker.jl:630-662 at t = as_t (samples addressed through smp_order)
    __device__ int add_sample() {
        CO_BEGIN(co_addsmp)
        CO_EVAL(co_addsmp, GREQ_COND);
        {
            const double kappa = ev_cond;
            int m = eg_winding;
            double sv = nthroot_inl(as_t, m);
            double mu_s = m * ipow(sv, m - 1);
            int slot;
            if (singular_steps <= 2) slot = singular_steps;
            else {
                int o0 = smp_order0;
                smp_order0 = smp_order1; sample_times[0] = sample_times[1]; sample_conds[0] = sample_conds[1];
                smp_order1 = smp_order2; sample_times[1] = sample_times[2]; sample_conds[1] = sample_conds[2];
                smp_order2 = o0;
                slot = 2;
            }
            const int b = slot == 0 ? smp_order0 : (slot == 1 ? smp_order1 : smp_order2);
            const cplx s1v = c_scale(mu_s, tx1);
            if (b == 0) { smp0_0 = tx0; smp1_0 = s1v; }
            else if (b == 1) { smp0_1 = tx0; smp1_1 = s1v; }
            else { smp0_2 = tx0; smp1_2 = s1v; }
            if (slot == 0) { sample_times[0] = sv; sample_conds[0] = kappa; }
            else if (slot == 1) { sample_times[1] = sv; sample_conds[1] = kappa; }
            else { sample_times[2] = sv; sample_conds[2] = kappa; }
        }
        CO_END(co_addsmp)
    }
    __device__ __forceinline__ cplx smp0(int b) const { return b == 0 ? smp0_0 : (b == 1 ? smp0_1 : smp0_2); }
    __device__ __forceinline__ cplx smp1(int b) const { return b == 0 ? smp1_0 : (b == 1 ? smp1_1 : smp1_2); }
    // predict_endpoint!, src/endgame_tracker.jl:664-693
    __device__ __forceinline__ double predict_endpoint() {
        if (singular_steps < 2) return HCG_INF;
        const int b0 = smp_order0, b1 = smp_order1, b2 = smp_order2;
        if (singular_steps == 2)
            prediction = hermite_real(smp0(b0), smp1(b0), sample_times[0], smp0(b1), smp1(b1), sample_times[1], 0.0);
        prev_prediction = prediction;
        prediction = hermite_real(smp0(b1), smp1(b1), sample_times[1], smp0(b2), smp1(b2), sample_times[2], 0.0);
        double pr = sample_times[2] / sample_times[1];
        double p2 = pr * pr;
        double err = dist_inf_v(prediction, prev_prediction, false) / fabs(p2 * p2 - 1);
        double ns = norm_inf_v(prediction, false);
        if (ns > 1e-8) err /= ns;
        return err;
    }
    // check_finite!, src/endgame_tracker.jl:402-422
    __device__ __forceinline__ bool check_finite() {
        if (!val_is_finite(o->val_finite_tol, !o->zero_is_at_infinity, o->max_winding_number)) return false;
        double merr;
        int m = estimate_winding_number(o->max_winding_number, &merr);
        if (merr < o->val_finite_tol) {
            if (m == 1 && !jtzf0) return false;
            eg_winding = m;
            return true;
        }
        return false;
    }
    // check_at_infinity!, src/endgame_tracker.jl:424-499 (a sequential loop over the entries;
    // result in ai_hit).  The loop needs cond(WS) at most once in value: the row scaling can
    // change only at the first entry that qualifies (when no entry has a start yet), before the
    // first condition number is taken, so it is computed once, up front.
    __device__ int check_at_infinity() {
        CO_BEGIN(co_ainf)
        ai_hit = 0;
        if (o->at_infinity_check) {
            {
                const bool zero_is_finite = !o->zero_is_at_infinity;
                double vx = val_x;
                double e = val_eps_inf();
                if (isnan(e)) ai_tols = HCG_INF;
                else if (vx + e < -o->val_finite_tol) ai_tols = e;
                else if (!zero_is_finite && vx - e > -o->val_finite_tol) ai_tols = e;
                else ai_tols = HCG_INF;
            }
            ai_first = g_first(ai_tols < o->val_at_infinity_tol, n);
            ai_kappa = HCG_NAN;
            if (ai_first < n) {
                if (isnan(gs_d(ai_starts, ai_first)) && g_all(isnan(ai_starts), n)) {
                    col_scaling = w;
                    eg_row_scaling();
                }
                CO_EVAL(co_ainf, GREQ_COND);
                ai_kappa = ev_cond;
            }
            {
                double kappa = HCG_NAN;
                double t = seg_t().re;
                bool hit = false;
                for (int i = 0; i < n && !hit; i++) {
                    const double tol_i = gs_d(ai_tols, i);
                    const double start_i = gs_d(ai_starts, i);
                    if (tol_i < o->val_at_infinity_tol) {
                        if (isnan(start_i)) {
                            kappa = ai_kappa;
                            const double absx = c_fabs(x);
                            if (r == i) { ai_conds = kappa; ai_abs = absx; ai_starts = t; }
                        } else {
                            if (isnan(kappa)) kappa = ai_kappa;
                            const double vi = gs_d(val_x, i);
                            const bool at_zero = vi > 0;
                            const double cond_growth = kappa / gs_d(ai_conds, i);
                            const double absx_i = gs_d(c_fabs(x), i), aiabs_i = gs_d(ai_abs, i);
                            const double coord_growth = at_zero ? aiabs_i / absx_i : absx_i / aiabs_i;
                            if (coord_growth > jclamp(hc_pow_inl(0.25, 4 * vi), 20, o->min_coord_growth) &&
                                (cond_growth > o->min_cond_growth || kappa > jmax(1e8, o->min_cond))) {
                                eg_code = at_zero ? GEG_AT_ZERO : GEG_AT_INFINITY;
                                hit = true;
                            }
                        }
                    } else if (!isnan(start_i)) {
                        if (r == i) ai_starts = HCG_NAN;
                    }
                }
                ai_hit = hit;
            }
        }
        CO_END(co_ainf)
    }
    // switch_to_singular!, src/endgame_tracker.jl:501-524
    __device__ int switch_to_singular() {
        CO_BEGIN(co_sw)
        singular_endgame = 1;
        sw_t = seg_t().re;
        if (g_all(row_scaling == 1.0, n)) {
            col_scaling = w;
            eg_row_scaling();
        }
        as_t = sw_t;
        CO_AWAIT(co_sw, add_sample());
        singular_steps = 0;
        winding = eg_winding;
        if (r == 0) ai_conds = sample_conds[0];
        keep_ext = 1;
        CO_END(co_sw)
    }
    __device__ __forceinline__ void switch_to_regular() {
        singular_endgame = 0;
        winding = 1;
        tracker_init_t0(mkc(0, 0));
    }
    // rest of singular_endgame_step! after its inner tracking loop, src/endgame_tracker.jl:555-628
    __device__ int singular_post() {
        CO_BEGIN(co_singpost)
        singular_steps += 1;
        if (!sp_max_steps) {
            if (code != GTC_SUCCESS) {
                eg_code = tc_to_eg(code);
                CO_AWAIT(co_singpost, tracking_stopped());
                co_singpost = 0;
                return 0;
            }
            val_update(0.25 * sing_t);
            {
                double merr;
                int mh = estimate_winding_number(o->max_winding_number, &merr);
                if (mh != eg_winding || merr > 0.1) { switch_to_regular(); co_singpost = 0; return 0; }
            }
            as_t = 0.25 * sing_t;
            CO_AWAIT(co_singpost, add_sample());
            if (!(singular_steps >= 2)) { co_singpost = 0; return 0; }
            {
                double acc = predict_endpoint();
                if (singular_steps == 2) {
                    eg_accuracy = acc;
                    solution = prediction;
                    co_singpost = 0;
                    return 0;
                } else if (acc < eg_accuracy && eg_accuracy > 1e-12) {
                    eg_accuracy = acc;
                    solution = prediction;
                    co_singpost = 0;
                    return 0;
                }
            }
        }
        {
            int m = eg_winding;
            double zero_cond = 1.0 / (m + 1);
            solution = (val_x < zero_cond) ? prediction : mkc(copysign(0.0, prediction.re), copysign(0.0, prediction.im));
        }
        ca_x = solution; ca_t = mkc(0, 0);
        CO_AWAIT(co_singpost, cond_at());
        {
            int m = eg_winding;
            double kappa = sample_conds[2];
            double kappa0 = ca_out;
            double J0 = inf_norm_ws(true, row_scaling, false, 1.0);
            if (eg_accuracy < o->singular_min_accuracy &&
                (((m > 1 && kappa > o->min_cond && nanmax(kappa0, 1.0 / J0) > kappa) || (m == 1 && kappa0 > 1e12)) ||
                 sp_max_steps || (n == 1 && 1.0 / J0 < o->min_cond))) {
                eg_cond = jmax(kappa0, 1.0 / J0);
                eg_singular = 1;
                eg_code = GEG_SUCCESS;
            } else if (!sp_max_steps) {
                switch_to_regular();
            }
        }
        CO_END(co_singpost)
    }

    // step!(endgame_tracker), src/endgame_tracker.jl:329-399, singular_endgame_step!'s inner
    // loop flattened to one tracker step per call
    __device__ int eg_step() {
        CO_BEGIN(co_egstep)
        if (!(singular_endgame && sing_phase == 1)) {
            if (steps_eg >= o->max_endgame_steps) { eg_code = GEG_MAX_STEPS; co_egstep = 0; return 0; }
            if (ext_steps() - ext_steps_eg_start > o->max_endgame_extended_steps) {
                es_jump = g_all(!c_isnan(solution), n);  // every entry of the solution finite
                if (es_jump && eg_winding != 0 && eg_accuracy < o->singular_min_accuracy) {
                    ca_x = solution; ca_t = mkc(0, 0);
                    CO_AWAIT(co_egstep, cond_at());
                    eg_cond = ca_out;
                    eg_singular = 1;
                    eg_code = GEG_SUCCESS;
                    co_egstep = 0;
                    return 0;
                }
                eg_code = GEG_MAX_EXT_STEPS;
                co_egstep = 0;
                return 0;
            }
            last_point = x;
            last_t = seg_t().re;
        }
        if (singular_endgame) {
            if (sing_phase == 0) {
                sing_t = seg_t().re;
                tracker_init_t0(mkc(0.25 * sing_t, 0.0));
                sing_phase = 1;
            }
            CO_AWAIT(co_egstep, tracker_step());
            sp_max_steps = 0;
            if ((steps_eg += 1) >= o->max_endgame_steps) { eg_code = GEG_MAX_STEPS; sp_max_steps = 1; }
            else if (ext_steps() - ext_steps_eg_start > o->max_endgame_extended_steps) { eg_code = GEG_MAX_EXT_STEPS; sp_max_steps = 1; }
            if (!sp_max_steps && code == GTC_TRACKING) { co_egstep = 0; return 0; }
            sing_phase = 0;
            CO_AWAIT(co_egstep, singular_post());
            co_egstep = 0;
            return 0;
        }
        {
            cplx tp = seg_t1();
            es_jump = tp.re == 0.0 && tp.im == 0.0;
        }
        CO_AWAIT(co_egstep, tracker_step());
        eg_code = tc_to_eg(code);
        if (eg_code != GEG_TRACKING) {
            CO_AWAIT(co_egstep, tracking_stopped());
            co_egstep = 0;
            return 0;
        }
        jtzf0 = jtzf1; jtzf1 = es_jump;
        {
            double t = seg_t().re;
            if (!(t <= eg_start)) { co_egstep = 0; return 0; }
            if (steps_eg == 0) ext_steps_eg_start = ext_steps();
            steps_eg += 1;
            if (!ts_ok) { co_egstep = 0; return 0; }
            val_update(t);
        }
        if (check_finite()) {
            CO_AWAIT(co_egstep, switch_to_singular());
            co_egstep = 0;
            return 0;
        }
        CO_AWAIT(co_egstep, check_at_infinity());
        CO_END(co_egstep)
    }

    // init!(endgame_tracker, x, t₁), src/endgame_tracker.jl:260-294 (after tracker_init)
    __device__ __forceinline__ void eg_init_rest() {
        eg_code = tc_to_eg(code);
        singular_endgame = 0;
        sing_phase = 0;
        jtzf0 = jtzf1 = 0;
        val_x = val_tx = dval_x = dval_tx = 0.0;
        vx2 = vx1 = ldx2 = ldx1 = 0.0;
        solution = mkc(HCG_NAN, HCG_NAN);
        row_scaling = 1.0; col_scaling = 1.0; ai_starts = HCG_NAN;
        logt2 = logt1 = HCG_NAN;
        eg_winding = 0;
        eg_accuracy = HCG_NAN; eg_cond = HCG_NAN; eg_singular = 0;
        steps_eg = 0;
        ext_steps_eg_start = 0x7fffffff;
        singular_steps = 0;
        last_t = HCG_NAN;
        smp_order0 = 0; smp_order1 = 1; smp_order2 = 2;
    }
    __device__ __forceinline__ void eg_init_start() {
        newton_iters = 0;
        refined_ext = 0;
        learned_proposed = learned_accepted = 0;
        last_learned = 0;
        inf_logit = 0.0f;
        ds_plain = HCG_NAN;
        eg_start = o->endgame_start;
        for (int i = 0; i < 8; i++) feat[i] = 0.0f;
        tx0 = tx1 = tx2 = tx3 = mkc(0.0, 0.0);
    }

    // The whole path: init, steps until the endgame tracker stops, the PathResult's residual.
    __device__ int path_run(double t1) {
        CO_BEGIN(co_path)
        eg_init_start();
        CO_AWAIT(co_path, tracker_init(t1));
        eg_init_rest();
        guard = 0;
        while (eg_code == GEG_TRACKING) {
            CO_AWAIT(co_path, eg_step());
            if (++guard > 4 * (o->max_steps + o->max_endgame_steps) + 64) eg_code = GEG_UNKNOWN;
        }
        // PathResult residual: H at the reported point
        if (eg_code == GEG_SUCCESS) { ev_x = solution; ev_t = mkc(0, 0); }
        else { ev_x = x; ev_t = seg_t(); }
        CO_EVAL(co_path, GREQ_E);
        res = ev_u;
        CO_END(co_path)
    }
};
