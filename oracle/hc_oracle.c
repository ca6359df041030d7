/*
 * ORACLE — test infrastructure only.  Only tests/, __graft_entry__.smoke() and
 * bench.py's cpu_baseline leg may load it, and only as the checker / CPU baseline.
 * The product (libhcgpu.so) never links or calls it.
 *
 * A plain-C restatement of HomotopyContinuation.jl 2.20's path tracker for square
 * polynomial systems, written from the reference's Julia source:
 *
 *   straight-line program eval / Jacobian / Taylor   src/model_kit/operations.jl:183-248,
 *                                                    src/model_kit/taylor.jl:272-920,
 *                                                    src/model_kit/instruction_interpreter.jl:444-491
 *   homotopies                                       src/homotopies/straight_line_homotopy.jl:81-154,
 *                                                    src/homotopies/parameter_homotopy.jl:66-102
 *   linear algebra (LU, Skeel, refinement, cond)     src/linear_algebra.jl:88-885
 *   weighted norm                                    src/norm.jl:101-234
 *   segment stepper                                  src/utils.jl:300-401
 *   predictor (Padé(2,1), cubic Hermite)             src/predictor.jl:124-371
 *   Newton corrector                                 src/newton_corrector.jl:55-286
 *   tracker                                          src/tracker.jl:517-996
 *   valuation                                        src/valuation.jl:38-228
 *   endgame tracker + PathResult                     src/endgame_tracker.jl:260-911
 *   threaded batch driver                            src/solve.jl:628-709
 *   certification (interval Krawczyk)                src/certification.jl:1029-1394,
 *                                                    src/interval_arithmetic.jl
 *
 * Pinning (see tests/): DD arithmetic against mpmath; SLP eval/Jacobian/Taylor
 * against closed-form polynomial evaluation and exact series composition; LU/solve
 * against numpy; cond within 10x of the true norm; norms, nthroot and the stepper
 * against the reference tests' exact values; whole tracker against the reference's
 * known-answer tests (test/tracker_test.jl, test/endgame_tracker_test.jl,
 * test/endgame_test.jl) and mpmath-refined endpoints.
 * Parity unpinned: the rounding of fused SLP Taylor rules (only their mathematical
 * value is pinned) and the complex division formula (naive, as @fastmath).
 * Transcendental SLP ops are not restated (polynomial systems only).
 */
#include "hc_oracle_core.h"
#include <stdlib.h>
#include <string.h>
#include <stdio.h>
#include <pthread.h>
#include <stdatomic.h>

/* ---------------- ABI structs (same layout as include/hcgpu.h) ---------------- */
typedef struct orc_tape {
    const int32_t* instr;
    const double* constants;
    const int32_t* u_assign;
    const int32_t* U_assign;
    int32_t ninstr, nconst;
    int32_t params_first, nparams;
    int32_t vars_first, nvars;
    int32_t t_index, nu;
    int32_t nU, output_dim;
    int32_t tape_space, pad_;
} orc_tape;
typedef struct orc_system { const orc_tape* eval_jac; const orc_tape* eval; const double* fixed_params; } orc_system;
typedef struct orc_homotopy_desc {
    int32_t kind, n;
    orc_system start, target;
    double gamma_re, gamma_im;
} orc_homotopy_desc;
typedef struct orc_options {
    int32_t max_steps, extended_precision;
    double max_step_size, max_initial_step_size, min_step_size, min_rel_step_size, terminate_cond;
    double a, beta_a, beta_omega_p, beta_tau, strict_beta_tau;
    int32_t min_newton_iters;
    int32_t max_endgame_steps;
    double endgame_start;
    int32_t max_endgame_extended_steps, max_winding_number;
    double min_cond, min_cond_growth, min_coord_growth;
    int32_t zero_is_at_infinity, at_infinity_check, only_nonsingular, refine_steps;
    double singular_min_accuracy, val_finite_tol, val_at_infinity_tol;
    double sing_cond, sing_accuracy, scaling_threshold;
    double scale_min, scale_abs_min, scale_max;
    int32_t learned_step, block_size, blocks_per_cu, log_steps;
    double learned_max_mult, learned_min_t, learned_inf_t, learned_inf_logit;
    int32_t refill_idle, reserved_i;  /* GPU lane scheduling only */
} orc_options;
typedef struct orc_batch {
    int32_t npaths, ngroups;
    const double* starts;
    const double* start_params;
    const double* target_params;
    const int32_t* group_of_path;
    double t1, t0;
} orc_batch;
typedef struct orc_path_result {
    int32_t code, singular, winding_number, extended_precision, extended_precision_used;
    int32_t accepted_steps, rejected_steps, steps_eg, valuation_valid, newton_iters;
    int32_t learned_proposed, learned_accepted;
    double t, accuracy, residual, cond, omega, mu, last_t;
} orc_path_result;

enum { OP_STOP = 0, OP_CB, OP_ACOS, OP_ASIN, OP_COS, OP_COSH, OP_EXP, OP_INV, OP_INV_NOT_ZERO, OP_INVSQR,
       OP_NEG, OP_SIN, OP_SINH, OP_SQR, OP_SQRT, OP_TAN, OP_TANH, OP_IDENTITY, OP_ADD, OP_DIV, OP_MUL, OP_SUB,
       OP_POW_INT, OP_POW, OP_ADD3, OP_MUL3, OP_MULADD, OP_MULSUB, OP_SUBMUL, OP_ADD4, OP_MUL4, OP_MULMULADD,
       OP_MULMULSUB };

/* TrackerCode (src/tracker.jl:166-176) */
enum { TC_TRACKING = 0, TC_SUCCESS, TC_MAX_STEPS, TC_ACCURACY_LIMIT, TC_ILL_CONDITIONED, TC_INVALID_START,
       TC_INVALID_START_SINGULAR, TC_STEP_TOO_SMALL, TC_UNKNOWN };
/* EndgameTrackerCode (src/endgame_tracker.jl:100-117) */
enum { EG_TRACKING = 0, EG_SUCCESS, EG_AT_INFINITY, EG_AT_ZERO, EG_ACCURACY_LIMIT, EG_INVALID_START,
       EG_INVALID_START_SINGULAR, EG_ILL_CONDITIONED, EG_MAX_STEPS, EG_MAX_EXT_STEPS, EG_MAX_WINDING,
       EG_STEP_TOO_SMALL, EG_UNKNOWN, EG_POST_CHECK_FAILED, EG_EXCESS, EG_POLYHEDRAL_FAILED };
/* NewtonCorrectorCodes (src/newton_corrector.jl:14-19) */
enum { NEWT_CONVERGED = 0, NEWT_TERMINATED, NEWT_MAX_ITERS, NEWT_SINGULARITY };

static int tc_to_eg(int c) {
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

static __thread char orc_err[256];
int orc_debug = 0;
const char* orc_last_error(void) { return orc_err; }

/* ============================ straight-line programs ============================ */
static int op_arity(int op) {
    switch (op) {
    case OP_STOP: return 0;
    case OP_CB: case OP_INV: case OP_INV_NOT_ZERO: case OP_INVSQR: case OP_NEG: case OP_SQR: case OP_IDENTITY:
        return 1;
    case OP_ADD: case OP_DIV: case OP_MUL: case OP_SUB: case OP_POW_INT: return 2;
    case OP_ADD3: case OP_MUL3: case OP_MULADD: case OP_MULSUB: case OP_SUBMUL: return 3;
    case OP_ADD4: case OP_MUL4: case OP_MULMULADD: case OP_MULMULSUB: return 4;
    default: return -1; /* transcendental ops: not restated */
    }
}

int orc_tape_check(const orc_tape* tp) {
    if (!tp || !tp->instr) { snprintf(orc_err, sizeof orc_err, "null tape"); return -1; }
    if (tp->t_index != 0) { snprintf(orc_err, sizeof orc_err, "homotopy tapes (t_index) unsupported"); return -1; }
    for (int i = 0; i < tp->ninstr; i++) {
        const int32_t* I = tp->instr + 6 * i;
        int op = I[4];
        int ar = op_arity(op);
        if (ar < 0) { snprintf(orc_err, sizeof orc_err, "unsupported op %d", op); return -1; }
        if (op == OP_STOP) continue;
        for (int k = 0; k < ar; k++) {
            if (op == OP_POW_INT && k == 1) continue;
            if (I[k] < 1 || I[k] > tp->tape_space) { snprintf(orc_err, sizeof orc_err, "bad tape index"); return -1; }
        }
        if (I[5] < 1 || I[5] > tp->tape_space) { snprintf(orc_err, sizeof orc_err, "bad output index"); return -1; }
    }
    return 0;
}

/* Base.power_by_squaring, p >= 1 */
static cplx c_pbs(cplx x, long p) {
    if (p == 1) return x;
    if (p == 0) return C(1, 0);
    if (p == 2) return cmul(x, x);
    int t = __builtin_ctzl((unsigned long)p) + 1;
    p >>= t;
    while (--t > 0) x = cmul(x, x);
    cplx y = x;
    while (p > 0) {
        t = __builtin_ctzl((unsigned long)p) + 1;
        p >>= t;
        while (--t >= 0) x = cmul(x, x);
        y = cmul(y, x);
    }
    return y;
}
static cdd cdd_pbs(cdd x, long p) {
    if (p == 1) return x;
    if (p == 0) { cdd one = {{1, 0}, {0, 0}}; return one; }
    if (p == 2) return cdd_mul(x, x);
    int t = __builtin_ctzl((unsigned long)p) + 1;
    p >>= t;
    while (--t > 0) x = cdd_mul(x, x);
    cdd y = x;
    while (p > 0) {
        t = __builtin_ctzl((unsigned long)p) + 1;
        p >>= t;
        while (--t >= 0) x = cdd_mul(x, x);
        y = cdd_mul(y, x);
    }
    return y;
}
/* op_pow_int, src/model_kit/operations.jl:230-231 */
static cplx c_pow_int(cplx x, long p) { return p > 0 ? c_pbs(x, p) : cinv(c_pbs(x, -p)); }
static cdd cdd_pow_int(cdd x, long p) { return p > 0 ? cdd_pbs(x, p) : cdd_inv(cdd_pbs(x, -p)); }
/* op_sqr / op_cb on Complex, src/model_kit/operations.jl:192-202 */
static cplx c_sqr(cplx z) { return C((z.re + z.im) * (z.re - z.im), (z.re + z.re) * z.im); }
static cplx c_cb(cplx z) {
    double a = (z.re + z.im) * (z.re - z.im), b = (z.re + z.re) * z.im;
    return C(fma(a, z.re, -(b * z.im)), fma(a, z.im, b * z.re));
}

/* One ComplexF64 execution of the tape (compiled_execute_impl semantics). */
static void tape_run_f64(const orc_tape* tp, cplx* T) {
    for (int i = 0; i < tp->ninstr; i++) {
        const int32_t* I = tp->instr + 6 * i;
        int op = I[4];
        if (op == OP_STOP) break;
        cplx a = T[I[0] - 1], b, c, d, r;
        switch (op) {
        case OP_CB: r = c_cb(a); break;
        case OP_INV: r = cinv(a); break;
        case OP_INV_NOT_ZERO: r = (a.re == 0 && a.im == 0) ? a : cinv(a); break;
        case OP_INVSQR: r = c_sqr(cinv(a)); break;
        case OP_NEG: r = cneg(a); break;
        case OP_SQR: r = c_sqr(a); break;
        case OP_IDENTITY: r = a; break;
        case OP_ADD: b = T[I[1] - 1]; r = cadd(a, b); break;
        case OP_SUB: b = T[I[1] - 1]; r = csub(a, b); break;
        case OP_MUL: b = T[I[1] - 1]; r = cmul(a, b); break;
        case OP_DIV: b = T[I[1] - 1]; r = cdiv(a, b); break;
        case OP_POW_INT: r = c_pow_int(a, I[1]); break;
        case OP_ADD3: b = T[I[1] - 1]; c = T[I[2] - 1]; r = cadd(cadd(a, b), c); break;
        case OP_MUL3: b = T[I[1] - 1]; c = T[I[2] - 1]; r = cmul(cmul(a, b), c); break;
        case OP_MULADD: b = T[I[1] - 1]; c = T[I[2] - 1]; r = cmuladd(a, b, c); break;
        case OP_MULSUB: b = T[I[1] - 1]; c = T[I[2] - 1]; r = cmulsub(a, b, c); break;
        case OP_SUBMUL: b = T[I[1] - 1]; c = T[I[2] - 1]; r = csubmul(a, b, c); break;
        case OP_ADD4: b = T[I[1] - 1]; c = T[I[2] - 1]; d = T[I[3] - 1]; r = cadd(cadd(cadd(a, b), c), d); break;
        case OP_MUL4: b = T[I[1] - 1]; c = T[I[2] - 1]; d = T[I[3] - 1]; r = cmul(cmul(cmul(a, b), c), d); break;
        case OP_MULMULADD: b = T[I[1] - 1]; c = T[I[2] - 1]; d = T[I[3] - 1]; r = cmuladd(a, b, cmul(c, d)); break;
        case OP_MULMULSUB: b = T[I[1] - 1]; c = T[I[2] - 1]; d = T[I[3] - 1]; r = csubmul(c, d, cmul(a, b)); break;
        default: r = C(NAN, NAN);
        }
        T[I[5] - 1] = r;
    }
}

static void tape_load_f64(const orc_tape* tp, cplx* T, const cplx* x, const cplx* p) {
    for (int k = 0; k < tp->nconst; k++) T[k] = C(tp->constants[2 * k], tp->constants[2 * k + 1]);
    for (int k = 0; k < tp->nparams; k++) T[tp->params_first - 1 + k] = p[k];
    for (int k = 0; k < tp->nvars; k++) T[tp->vars_first - 1 + k] = x[k];
}

/* evaluate! / evaluate_and_jacobian! of a System tape */
static void sys_eval(const orc_tape* tp, const cplx* x, const cplx* p, cplx* T, cplx* u, cplx* U, int n) {
    tape_load_f64(tp, T, x, p);
    tape_run_f64(tp, T);
    for (int i = 0; i < tp->output_dim; i++) u[i] = C(0, 0);
    for (int i = 0; i < tp->nu; i++) u[tp->u_assign[2 * i] - 1] = T[tp->u_assign[2 * i + 1] - 1];
    if (U) {
        for (int i = 0; i < n * n; i++) U[i] = C(0, 0);
        for (int i = 0; i < tp->nU; i++) U[tp->U_assign[2 * i] - 1] = T[tp->U_assign[2 * i + 1] - 1];
    }
}

/* evaluate! with x::Vector{ComplexDF64} (interpreted_system.jl eval_ComplexDF64 tape) */
static void sys_eval_dd(const orc_tape* tp, const cdd* x, const cplx* p, cdd* T, cdd* u) {
    for (int k = 0; k < tp->nconst; k++) T[k] = cdd_from(C(tp->constants[2 * k], tp->constants[2 * k + 1]));
    for (int k = 0; k < tp->nparams; k++) T[tp->params_first - 1 + k] = cdd_from(p[k]);
    for (int k = 0; k < tp->nvars; k++) T[tp->vars_first - 1 + k] = x[k];
    for (int i = 0; i < tp->ninstr; i++) {
        const int32_t* I = tp->instr + 6 * i;
        int op = I[4];
        if (op == OP_STOP) break;
        cdd a = T[I[0] - 1], b, c, d, r;
        switch (op) {
        case OP_CB: r = cdd_cb(a); break;
        case OP_INV: r = cdd_inv(a); break;
        case OP_INV_NOT_ZERO: r = (a.re.hi == 0 && a.re.lo == 0 && a.im.hi == 0 && a.im.lo == 0) ? a : cdd_inv(a); break;
        case OP_INVSQR: r = cdd_sqr(cdd_inv(a)); break;
        case OP_NEG: r = cdd_neg(a); break;
        case OP_SQR: r = cdd_sqr(a); break;
        case OP_IDENTITY: r = a; break;
        case OP_ADD: b = T[I[1] - 1]; r = cdd_add(a, b); break;
        case OP_SUB: b = T[I[1] - 1]; r = cdd_sub(a, b); break;
        case OP_MUL: b = T[I[1] - 1]; r = cdd_mul(a, b); break;
        case OP_DIV: b = T[I[1] - 1]; r = cdd_div(a, b); break;
        case OP_POW_INT: r = cdd_pow_int(a, I[1]); break;
        case OP_ADD3: b = T[I[1] - 1]; c = T[I[2] - 1]; r = cdd_add(cdd_add(a, b), c); break;
        case OP_MUL3: b = T[I[1] - 1]; c = T[I[2] - 1]; r = cdd_mul(cdd_mul(a, b), c); break;
        case OP_MULADD: b = T[I[1] - 1]; c = T[I[2] - 1]; r = cdd_add(cdd_mul(a, b), c); break;
        case OP_MULSUB: b = T[I[1] - 1]; c = T[I[2] - 1]; r = cdd_sub(cdd_mul(a, b), c); break;
        case OP_SUBMUL: b = T[I[1] - 1]; c = T[I[2] - 1]; r = cdd_sub(c, cdd_mul(a, b)); break;
        case OP_ADD4: b = T[I[1] - 1]; c = T[I[2] - 1]; d = T[I[3] - 1]; r = cdd_add(cdd_add(cdd_add(a, b), c), d); break;
        case OP_MUL4: b = T[I[1] - 1]; c = T[I[2] - 1]; d = T[I[3] - 1]; r = cdd_mul(cdd_mul(cdd_mul(a, b), c), d); break;
        case OP_MULMULADD: b = T[I[1] - 1]; c = T[I[2] - 1]; d = T[I[3] - 1]; r = cdd_add(cdd_mul(a, b), cdd_mul(c, d)); break;
        case OP_MULMULSUB: b = T[I[1] - 1]; c = T[I[2] - 1]; d = T[I[3] - 1]; r = cdd_sub(cdd_mul(a, b), cdd_mul(c, d)); break;
        default: { cdd nn = {{NAN, 0}, {NAN, 0}}; r = nn; }
        }
        T[I[5] - 1] = r;
    }
    cdd z = {{0, 0}, {0, 0}};
    for (int i = 0; i < tp->output_dim; i++) u[i] = z;
    for (int i = 0; i < tp->nu; i++) u[tp->u_assign[2 * i] - 1] = T[tp->u_assign[2 * i + 1] - 1];
}

/* ---- truncated Taylor series arithmetic, src/model_kit/taylor.jl:272-920 ---- */
#define SMAX 5 /* orders 0..4 */
typedef struct { cplx c[SMAX]; } ser;

static ser s_mul(const ser* x, const ser* y, int K) {
    ser r;
    for (int k = 0; k <= K; k++) {
        cplx c = cmul(x->c[0], y->c[k]);
        for (int j = 1; j <= k; j++) c = cmuladd(x->c[j], y->c[k - j], c);
        r.c[k] = c;
    }
    return r;
}
static ser s_add(const ser* x, const ser* y, int K) { ser r; for (int k = 0; k <= K; k++) r.c[k] = cadd(x->c[k], y->c[k]); return r; }
static ser s_sub(const ser* x, const ser* y, int K) { ser r; for (int k = 0; k <= K; k++) r.c[k] = csub(x->c[k], y->c[k]); return r; }
static ser s_sqr(const ser* x, int K) {
    ser r;
    for (int k = 0; k <= K; k++) {
        if (k == 0) { r.c[0] = c_sqr(x->c[0]); continue; }
        cplx w = cmul(x->c[0], x->c[k]);
        for (int j = 1; j <= (k - 1) / 2; j++) w = cmuladd(x->c[j], x->c[k - j], w);
        if (k % 2 == 0) r.c[k] = cadd(cscale(2.0, w), c_sqr(x->c[k / 2]));
        else r.c[k] = cadd(w, w);
    }
    return r;
}
static ser s_div(const ser* x, const ser* y, int K) {
    ser r;
    for (int k = 0; k <= K; k++) {
        cplx num = x->c[k];
        if (k > 0) {
            cplx s = cmul(r.c[0], y->c[k]);
            for (int j = 1; j < k; j++) s = cmuladd(r.c[j], y->c[k - j], s);
            num = csub(num, s);
        }
        r.c[k] = cdiv(num, y->c[0]);
    }
    return r;
}
static ser s_inv(const ser* x, int K) {
    ser r;
    for (int k = 0; k <= K; k++) {
        if (k == 0) { r.c[0] = cdiv(C(1, 0), x->c[0]); continue; }
        cplx s = cmul(r.c[0], x->c[k]);
        for (int j = 1; j < k; j++) s = cmuladd(r.c[j], x->c[k - j], s);
        r.c[k] = cdiv(cneg(s), x->c[0]);
    }
    return r;
}
/* taylor_op_pow_int: the reference's recurrence, including its x0 == 0 -> 0 rule */
static ser s_pow_int(const ser* x, long rr, int K) {
    ser w;
    if (x->c[0].re == 0 && x->c[0].im == 0) { for (int k = 0; k <= K; k++) w.c[k] = C(0, 0); return w; }
    w.c[0] = c_pow_int(x->c[0], rr);
    cplx u0inv = cinv(x->c[0]);
    for (int k = 1; k <= K; k++) {
        cplx s = cmul(w.c[k - 1], cscale((double)1, x->c[1]));
        for (int j = 2; j <= k; j++) s = cmuladd(w.c[k - j], cscale((double)j, x->c[j]), s);
        s = cscale((double)rr, s);
        cplx wt;
        if (k > 1) {
            cplx t = cmul(x->c[k - 1], cscale(1.0, w.c[1]));
            for (int j = 2; j <= k - 1; j++) t = cmuladd(x->c[k - j], cscale((double)j, w.c[j]), t);
            wt = cmul(u0inv, csub(s, t));
        } else {
            wt = cmul(u0inv, s);
        }
        w.c[k] = cdivr(wt, (double)k);
    }
    return w;
}

static void tape_run_ser(const orc_tape* tp, ser* T, int K) {
    for (int i = 0; i < tp->ninstr; i++) {
        const int32_t* I = tp->instr + 6 * i;
        int op = I[4];
        if (op == OP_STOP) break;
        const ser* a = &T[I[0] - 1];
        ser r, t1, t2;
        switch (op) {
        case OP_CB: t1 = s_sqr(a, K); r = s_mul(&t1, a, K); break;
        case OP_INV: r = s_inv(a, K); break;
        case OP_INV_NOT_ZERO: r = s_inv(a, K); break;
        case OP_INVSQR: t1 = s_sqr(a, K); r = s_inv(&t1, K); break;
        case OP_NEG: for (int k = 0; k <= K; k++) r.c[k] = cneg(a->c[k]); break;
        case OP_SQR: r = s_sqr(a, K); break;
        case OP_IDENTITY: r = *a; break;
        case OP_ADD: r = s_add(a, &T[I[1] - 1], K); break;
        case OP_SUB: r = s_sub(a, &T[I[1] - 1], K); break;
        case OP_MUL: r = s_mul(a, &T[I[1] - 1], K); break;
        case OP_DIV: r = s_div(a, &T[I[1] - 1], K); break;
        case OP_POW_INT: r = s_pow_int(a, I[1], K); break;
        case OP_ADD3: t1 = s_add(a, &T[I[1] - 1], K); r = s_add(&t1, &T[I[2] - 1], K); break;
        case OP_MUL3: t1 = s_mul(a, &T[I[1] - 1], K); r = s_mul(&t1, &T[I[2] - 1], K); break;
        case OP_MULADD: {
            const ser *y = &T[I[1] - 1], *z = &T[I[2] - 1];
            for (int k = 0; k <= K; k++) {
                cplx c = z->c[k];
                for (int j = 0; j <= k; j++) c = cmuladd(a->c[j], y->c[k - j], c);
                r.c[k] = c;
            }
            break;
        }
        case OP_MULSUB: {
            const ser *y = &T[I[1] - 1], *z = &T[I[2] - 1];
            for (int k = 0; k <= K; k++) {
                cplx c = cmulsub(a->c[0], y->c[k], z->c[k]);
                for (int j = 1; j <= k; j++) c = cmuladd(a->c[j], y->c[k - j], c);
                r.c[k] = c;
            }
            break;
        }
        case OP_SUBMUL: {
            const ser *y = &T[I[1] - 1], *z = &T[I[2] - 1];
            for (int k = 0; k <= K; k++) {
                cplx c = z->c[k];
                for (int j = 0; j <= k; j++) c = csubmul(a->c[j], y->c[k - j], c);
                r.c[k] = c;
            }
            break;
        }
        case OP_ADD4: t1 = s_add(a, &T[I[1] - 1], K); t2 = s_add(&T[I[2] - 1], &T[I[3] - 1], K); r = s_add(&t1, &t2, K); break;
        case OP_MUL4: t1 = s_mul(a, &T[I[1] - 1], K); t2 = s_mul(&T[I[2] - 1], &T[I[3] - 1], K); r = s_mul(&t1, &t2, K); break;
        case OP_MULMULADD: t1 = s_mul(a, &T[I[1] - 1], K); t2 = s_mul(&T[I[2] - 1], &T[I[3] - 1], K); r = s_add(&t1, &t2, K); break;
        case OP_MULMULSUB: t1 = s_mul(a, &T[I[1] - 1], K); t2 = s_mul(&T[I[2] - 1], &T[I[3] - 1], K); r = s_sub(&t1, &t2, K); break;
        default: for (int k = 0; k <= K; k++) r.c[k] = C(NAN, NAN);
        }
        T[I[5] - 1] = r;
    }
}

/* taylor!(u, Val(K), F, tx, p): series of the outputs up to order K.
 * tx: K vectors (orders 0..K-1) each n; p0/p1: parameter series (order 0 and 1, p1 may be NULL).
 * out: (K+1) x output_dim, out[k*m + i] is the order-k coefficient of output i. */
static void sys_taylor(const orc_tape* tp, int K, const cplx* const* tx, const cplx* p0, const cplx* p1, ser* T, cplx* out) {
    for (int k = 0; k < tp->tape_space; k++) for (int j = 0; j < SMAX; j++) T[k].c[j] = C(0, 0);
    for (int k = 0; k < tp->nconst; k++) T[k].c[0] = C(tp->constants[2 * k], tp->constants[2 * k + 1]);
    for (int k = 0; k < tp->nparams; k++) {
        T[tp->params_first - 1 + k].c[0] = p0[k];
        if (p1) T[tp->params_first - 1 + k].c[1] = p1[k];
    }
    for (int k = 0; k < tp->nvars; k++)
        for (int j = 0; j < K; j++) T[tp->vars_first - 1 + k].c[j] = tx[j][k];
    tape_run_ser(tp, T, K);
    int m = tp->output_dim;
    for (int k = 0; k <= K; k++) for (int i = 0; i < m; i++) out[k * m + i] = C(0, 0);
    for (int i = 0; i < tp->nu; i++) {
        int row = tp->u_assign[2 * i] - 1, slot = tp->u_assign[2 * i + 1] - 1;
        for (int k = 0; k <= K; k++) out[k * m + row] = T[slot].c[k];
    }
}

/* ================================ homotopy ================================ */
typedef struct {
    const orc_homotopy_desc* hd;
    int n;
    cplx gamma;
    cplx *Gp, *Fp;   /* fixed parameters of G and F */
    int npar;        /* parameter homotopy: number of parameters */
    cplx *p, *q, *pt, *dp;
    int tcache_valid;
    cplx tcache;
    /* buffers */
    cplx* T;
    cdd* Tdd;
    ser* Ts;
    cplx *u1, *U1, *u2, *U2;
    cdd *ud1, *ud2;
    cplx *ser1, *ser2;
} orc_H;

static int max_tape_space(const orc_homotopy_desc* hd) {
    int m = hd->target.eval_jac->tape_space;
    if (hd->target.eval->tape_space > m) m = hd->target.eval->tape_space;
    if (hd->kind == 0) {
        if (hd->start.eval_jac->tape_space > m) m = hd->start.eval_jac->tape_space;
        if (hd->start.eval->tape_space > m) m = hd->start.eval->tape_space;
    }
    return m;
}

static void H_init(orc_H* H, const orc_homotopy_desc* hd) {
    memset(H, 0, sizeof *H);
    H->hd = hd;
    int n = hd->n;
    H->n = n;
    H->gamma = C(hd->gamma_re, hd->gamma_im);
    int ts = max_tape_space(hd);
    H->T = calloc(ts, sizeof(cplx));
    H->Tdd = calloc(ts, sizeof(cdd));
    H->Ts = calloc(ts, sizeof(ser));
    H->u1 = calloc(n, sizeof(cplx)); H->u2 = calloc(n, sizeof(cplx));
    H->U1 = calloc(n * n, sizeof(cplx)); H->U2 = calloc(n * n, sizeof(cplx));
    H->ud1 = calloc(n, sizeof(cdd)); H->ud2 = calloc(n, sizeof(cdd));
    H->ser1 = calloc(SMAX * n, sizeof(cplx)); H->ser2 = calloc(SMAX * n, sizeof(cplx));
    int gp = hd->kind == 0 ? hd->start.eval_jac->nparams : 0;
    int fp = hd->target.eval_jac->nparams;
    H->Gp = calloc(gp + 1, sizeof(cplx));
    H->Fp = calloc(fp + 1, sizeof(cplx));
    for (int i = 0; i < gp; i++) H->Gp[i] = C(hd->start.fixed_params[2 * i], hd->start.fixed_params[2 * i + 1]);
    if (hd->kind == 0)
        for (int i = 0; i < fp; i++) H->Fp[i] = C(hd->target.fixed_params[2 * i], hd->target.fixed_params[2 * i + 1]);
    H->npar = hd->kind == 1 ? fp : 0;
    H->p = calloc(fp + 1, sizeof(cplx)); H->q = calloc(fp + 1, sizeof(cplx));
    H->pt = calloc(fp + 1, sizeof(cplx)); H->dp = calloc(fp + 1, sizeof(cplx));
}
static void H_free(orc_H* H) {
    free(H->T); free(H->Tdd); free(H->Ts); free(H->u1); free(H->u2); free(H->U1); free(H->U2);
    free(H->ud1); free(H->ud2); free(H->ser1); free(H->ser2); free(H->Gp); free(H->Fp);
    free(H->p); free(H->q); free(H->pt); free(H->dp);
}
/* parameters!(H, p, q) */
static void H_set_params(orc_H* H, const cplx* p, const cplx* q) {
    for (int i = 0; i < H->npar; i++) { H->p[i] = p[i]; H->q[i] = q[i]; }
    H->tcache_valid = 0;
}
/* tp!, src/homotopies/parameter_homotopy.jl:66-87 */
static void H_tp(orc_H* H, cplx t) {
    if (H->tcache_valid && H->tcache.re == t.re && H->tcache.im == t.im) return;
    for (int i = 0; i < H->npar; i++) {
        cplx pti;
        if (t.im == 0) {
            double tr = t.re;
            pti = cadd(cscale(tr, H->p[i]), cscale(1.0 - tr, H->q[i]));
        } else {
            cplx omt = C(1.0 - t.re, -t.im);
            pti = cadd(cmul(t, H->p[i]), cmul(omt, H->q[i]));
        }
        H->pt[i] = pti;
        H->dp[i] = csub(H->p[i], H->q[i]);
    }
    H->tcache = t;
    H->tcache_valid = 1;
}

/* H(x,t) in ComplexF64 */
static void H_eval(orc_H* H, const cplx* x, cplx t, cplx* u) {
    const orc_homotopy_desc* hd = H->hd;
    int n = H->n;
    if (hd->kind == 0) {
        sys_eval(hd->start.eval, x, H->Gp, H->T, u, NULL, n);
        sys_eval(hd->target.eval, x, H->Fp, H->T, H->u2, NULL, n);
        cplx ts = cmul(H->gamma, t), tt = C(1.0 - t.re, -t.im);
        for (int i = 0; i < n; i++) u[i] = cmuladd(ts, u[i], cmul(tt, H->u2[i]));
    } else {
        H_tp(H, t);
        sys_eval(hd->target.eval, x, H->pt, H->T, u, NULL, n);
    }
}
/* H(x,t) with x in double-double, rounded to ComplexF64 (straight_line_homotopy.jl:81-94) */
static void H_eval_dd(orc_H* H, const cdd* x, cplx t, cplx* u) {
    const orc_homotopy_desc* hd = H->hd;
    int n = H->n;
    if (hd->kind == 0) {
        sys_eval_dd(hd->start.eval, x, H->Gp, H->Tdd, H->ud1);
        sys_eval_dd(hd->target.eval, x, H->Fp, H->Tdd, H->ud2);
        cplx ts = cmul(H->gamma, t), tt = C(1.0 - t.re, -t.im);
        for (int i = 0; i < n; i++) u[i] = cdd_round(cdd_add(cdd_mul_c(ts, H->ud1[i]), cdd_mul_c(tt, H->ud2[i])));
    } else {
        H_tp(H, t);
        sys_eval_dd(hd->target.eval, x, H->pt, H->Tdd, H->ud1);
        for (int i = 0; i < n; i++) u[i] = cdd_round(H->ud1[i]);
    }
}
/* evaluate_and_jacobian!; U column-major */
static void H_eval_jac(orc_H* H, const cplx* x, cplx t, cplx* u, cplx* U) {
    const orc_homotopy_desc* hd = H->hd;
    int n = H->n;
    if (hd->kind == 0) {
        sys_eval(hd->start.eval_jac, x, H->Gp, H->T, u, U, n);
        sys_eval(hd->target.eval_jac, x, H->Fp, H->T, H->u2, H->U2, n);
        cplx ts = cmul(H->gamma, t), tt = C(1.0 - t.re, -t.im);
        for (int i = 0; i < n; i++) u[i] = cmuladd(ts, u[i], cmul(tt, H->u2[i]));
        for (int k = 0; k < n * n; k++) U[k] = cmuladd(ts, U[k], cmul(tt, H->U2[k]));
    } else {
        H_tp(H, t);
        sys_eval(hd->target.eval_jac, x, H->pt, H->T, u, U, n);
    }
}
/* taylor!(u, Val(K), H, tx, t): K-th coefficient of H(x(s), t+s); tx has K vectors */
static void H_taylor(orc_H* H, int K, const cplx* const* tx, cplx t, cplx* u) {
    const orc_homotopy_desc* hd = H->hd;
    int n = H->n;
    if (hd->kind == 0) {
        if (K == 1) { /* taylor_1!: γ G(x) - F(x) */
            sys_eval(hd->start.eval, tx[0], H->Gp, H->T, u, NULL, n);
            sys_eval(hd->target.eval, tx[0], H->Fp, H->T, H->u2, NULL, n);
            for (int i = 0; i < n; i++) u[i] = csub(cmul(H->gamma, u[i]), H->u2[i]);
            return;
        }
        sys_taylor(hd->start.eval, K, tx, H->Gp, NULL, H->Ts, H->ser1);
        sys_taylor(hd->target.eval, K, tx, H->Fp, NULL, H->Ts, H->ser2);
        cplx omt = C(1.0 - t.re, -t.im);
        for (int i = 0; i < n; i++) {
            cplx gK1 = H->ser1[(K - 1) * n + i], gK = H->ser1[K * n + i];
            cplx fK1 = H->ser2[(K - 1) * n + i], fK = H->ser2[K * n + i];
            cplx st = cmul(H->gamma, cadd(gK1, cmul(t, gK)));
            cplx tg = csub(cmul(omt, fK), fK1);
            u[i] = cadd(st, tg);
        }
    } else {
        H_tp(H, t);
        sys_taylor(hd->target.eval, K, tx, H->pt, H->dp, H->Ts, H->ser1);
        for (int i = 0; i < n; i++) u[i] = H->ser1[K * n + i];
    }
}

/* ============================== linear algebra ============================== */
typedef struct {
    int n;
    cplx *A, *LU;
    int* ipiv;
    double* row_scaling;
    int factorized, scaled;
    cplx *r, *dx, *work;
    cdd *xdd, *rdd;
    double* rwork;
} orc_ws;

static void ws_init(orc_ws* W, int n) {
    memset(W, 0, sizeof *W);
    W->n = n;
    W->A = calloc(n * n, sizeof(cplx)); W->LU = calloc(n * n, sizeof(cplx));
    W->ipiv = calloc(n, sizeof(int)); W->row_scaling = calloc(n, sizeof(double));
    for (int i = 0; i < n; i++) W->row_scaling[i] = 1.0;
    W->r = calloc(n, sizeof(cplx)); W->dx = calloc(n, sizeof(cplx)); W->work = calloc(n, sizeof(cplx));
    W->xdd = calloc(n, sizeof(cdd)); W->rdd = calloc(n, sizeof(cdd)); W->rwork = calloc(n, sizeof(double));
}
static void ws_free(orc_ws* W) {
    free(W->A); free(W->LU); free(W->ipiv); free(W->row_scaling); free(W->r); free(W->dx); free(W->work);
    free(W->xdd); free(W->rdd); free(W->rwork);
}
#define AT(M, i, j, n) (M)[(j) * (n) + (i)]

/* updated!, src/linear_algebra.jl:88-98 */
static void ws_updated(orc_ws* W) {
    W->factorized = 0; W->scaled = 0;
    memcpy(W->LU, W->A, sizeof(cplx) * W->n * W->n);
}
/* lu!, src/linear_algebra.jl:130-184 */
static void lu_factor(cplx* A, int* ipiv, int n) {
    for (int k = 0; k < n; k++) {
        int kp = k;
        double amax = cabs2(AT(A, k, k, n));
        for (int i = k + 1; i < n; i++) {
            double absi = cabs2(AT(A, i, k, n));
            if (absi > amax) { kp = i; amax = absi; }
        }
        ipiv[k] = kp;
        if (amax != 0) {
            if (k != kp)
                for (int i = 0; i < n; i++) { cplx tmp = AT(A, k, i, n); AT(A, k, i, n) = AT(A, kp, i, n); AT(A, kp, i, n) = tmp; }
            cplx Akk = AT(A, k, k, n);
            for (int i = k + 1; i < n; i++) AT(A, i, k, n) = cdiv(AT(A, i, k, n), Akk);
        }
        for (int j = k + 1; j < n; j++) {
            cplx Akj = AT(A, k, j, n);
            for (int i = k + 1; i < n; i++) AT(A, i, j, n) = csubmul(AT(A, i, k, n), Akj, AT(A, i, j, n));
        }
    }
}
/* lu_ldiv!, src/linear_algebra.jl:268-316 (in place on x) */
static void lu_solve(const cplx* A, const int* ipiv, int n, cplx* x) {
    for (int i = 0; i < n; i++) if (i != ipiv[i]) { cplx t = x[i]; x[i] = x[ipiv[i]]; x[ipiv[i]] = t; }
    for (int j = 0; j < n; j++) {
        cplx xj = x[j];
        for (int i = j + 1; i < n; i++) x[i] = csubmul(AT(A, i, j, n), xj, x[i]);
    }
    for (int j = n - 1; j >= 0; j--) {
        cplx xj = cdiv(x[j], AT(A, j, j, n));
        x[j] = xj;
        for (int i = 0; i < j; i++) x[i] = csubmul(AT(A, i, j, n), xj, x[i]);
    }
}
/* lu_ldiv_adj!, src/linear_algebra.jl:318-354 */
static void lu_solve_adj(const cplx* A, const int* ipiv, int n, cplx* x) {
    for (int j = 0; j < n; j++) {
        cplx z = x[j];
        for (int i = 0; i < j; i++) z = csubmul(cconj(AT(A, i, j, n)), x[i], z);
        x[j] = cdiv(z, cconj(AT(A, j, j, n)));
    }
    for (int j = n - 1; j >= 0; j--) {
        cplx z = x[j];
        for (int i = n - 1; i > j; i--) z = csubmul(cconj(AT(A, i, j, n)), x[i], z);
        x[j] = z;
    }
    for (int i = n - 1; i >= 0; i--) if (i != ipiv[i]) { cplx t = x[i]; x[i] = x[ipiv[i]]; x[ipiv[i]] = t; }
}
static void ws_factorize(orc_ws* W) { lu_factor(W->LU, W->ipiv, W->n); W->factorized = 1; }
/* LA.ldiv!(x, WS::MatrixWorkspace, b), src/linear_algebra.jl:389-408 */
static void ws_ldiv(orc_ws* W, cplx* x, const cplx* b) {
    int n = W->n;
    if (n == 1) { x[0] = cdiv(b[0], W->A[0]); return; }
    if (!W->factorized) ws_factorize(W);
    if (W->scaled) for (int i = 0; i < n; i++) x[i] = cscale(W->row_scaling[i], b[i]);
    else if (x != b) memcpy(x, b, sizeof(cplx) * n);
    lu_solve(W->LU, W->ipiv, n, x);
}
/* skeel_row_scaling!, src/linear_algebra.jl:432-459 */
static void skeel_row_scaling(double* d, const cplx* A, const double* c, int n, double threshold) {
    for (int i = 0; i < n; i++) d[i] = 0.0;
    for (int j = 0; j < n; j++) { double cj = c[j]; for (int i = 0; i < n; i++) d[i] += fast_abs(AT(A, i, j, n)) * cj; }
    double m = d[0];
    for (int i = 1; i < n; i++) m = jmax(m, d[i]);
    double s = threshold + m;
    for (int i = 0; i < n; i++) {
        int e; frexp(d[i], &e);
        if (!isfinite(d[i])) e = 0;
        if (e < s) d[i] = 1.0; else d[i] = ldexp(1.0, -e);
    }
}
/* apply_row_scaling!, src/linear_algebra.jl:489-497 */
static void apply_row_scaling(orc_ws* W) {
    int n = W->n;
    for (int j = 0; j < n; j++) for (int i = 0; i < n; i++) AT(W->LU, i, j, n) = cscale(W->row_scaling[i], AT(W->LU, i, j, n));
    W->scaled = 1;
}

/* norms: w == NULL is InfNorm, else WeightedNorm{InfNorm}, src/norm.jl:150-234 */
static double norm_inf(const cplx* x, const double* w, int n) {
    double dmax = w ? cabs2(cdivr(x[0], w[0])) : cabs2(x[0]);
    for (int i = 1; i < n; i++) dmax = max_fast(dmax, w ? cabs2(cdivr(x[i], w[i])) : cabs2(x[i]));
    double d = sqrt(dmax);
    if (isinf(d)) {
        dmax = w ? chypot(cdivr(x[0], w[0])) : chypot(x[0]);
        for (int i = 1; i < n; i++) dmax = jmax(dmax, w ? chypot(cdivr(x[i], w[i])) : chypot(x[i]));
        return dmax;
    }
    return d;
}
static double dist_inf(const cplx* x, const cplx* y, const double* w, int n) {
    double dmax = w ? cabs2(cdivr(csub(x[0], y[0]), w[0])) : cabs2(csub(x[0], y[0]));
    for (int i = 1; i < n; i++) dmax = max_fast(dmax, w ? cabs2(cdivr(csub(x[i], y[i]), w[i])) : cabs2(csub(x[i], y[i])));
    double d = sqrt(dmax);
    if (isinf(d)) {
        /* note: the reference's unweighted fallback mixes abs and abs2 (src/norm.jl:160-163) */
        if (w) {
            dmax = chypot(cdivr(csub(x[0], y[0]), w[0]));
            for (int i = 1; i < n; i++) dmax = jmax(dmax, chypot(cdivr(csub(x[i], y[i]), w[i])));
        } else {
            dmax = chypot(csub(x[0], y[0]));
            for (int i = 1; i < n; i++) dmax = jmax(dmax, cabs2(csub(x[i], y[i])));
        }
        return dmax;
    }
    return d;
}

/* residual!(r, A, x, b) in ComplexF64, src/linear_algebra.jl:505-519 */
static void residual_f64(cplx* r, const cplx* A, const cplx* x, const cplx* b, int n) {
    for (int i = 0; i < n; i++) r[i] = C(0, 0);
    for (int j = 0; j < n; j++) { cplx xj = x[j]; for (int i = 0; i < n; i++) r[i] = cmuladd(AT(A, i, j, n), xj, r[i]); }
    for (int i = 0; i < n; i++) r[i] = csub(r[i], b[i]);
}
/* same with x̄, r̄ in double-double */
static void residual_dd(cdd* r, const cplx* A, const cdd* x, const cplx* b, int n) {
    cdd z = {{0, 0}, {0, 0}};
    for (int i = 0; i < n; i++) r[i] = z;
    for (int j = 0; j < n; j++) { cdd xj = x[j]; for (int i = 0; i < n; i++) r[i] = cdd_add(r[i], cdd_mul_c(AT(A, i, j, n), xj)); }
    for (int i = 0; i < n; i++) r[i] = cdd_sub_c(r[i], b[i]);
}
/* mixed_precision_iterative_refinement!, src/linear_algebra.jl:528-544 */
static double mixed_refine(orc_ws* W, cplx* x, const cplx* b, const double* w) {
    int n = W->n;
    for (int i = 0; i < n; i++) W->xdd[i] = cdd_from(x[i]);
    residual_dd(W->rdd, W->A, W->xdd, b, n);
    for (int i = 0; i < n; i++) W->r[i] = cdd_round(W->rdd[i]);
    ws_ldiv(W, W->dx, W->r);
    for (int i = 0; i < n; i++) x[i] = csub(x[i], W->dx[i]);
    return norm_inf(W->dx, w, n) / norm_inf(x, w, n);
}
/* fixed_precision_iterative_refinement!, src/linear_algebra.jl:553-567 */
static double fixed_refine(orc_ws* W, cplx* x, const cplx* b, const double* w) {
    int n = W->n;
    residual_f64(W->r, W->A, x, b, n);
    ws_ldiv(W, W->dx, W->r);
    for (int i = 0; i < n; i++) x[i] = csub(x[i], W->dx[i]);
    return norm_inf(W->dx, w, n) / norm_inf(x, w, n);
}
/* iterative_refinement!, src/linear_algebra.jl:864-885 */
static double iterative_refinement(orc_ws* W, cplx* x, const cplx* b, const double* w, double tol, int max_iters) {
    double d = mixed_refine(W, x, b, w);
    for (int i = 2; i <= max_iters; i++) {
        double d2 = mixed_refine(W, x, b, w);
        if (d2 < tol) return d2;
        else if (d2 > 0.5 * d) return d2;
        d = d2;
    }
    return d;
}
/* Jacobian ldiv! with WeightedNorm: Skeel scaling on first factorization (src/linear_algebra.jl:846-862) */
static void jac_ldiv_w(orc_ws* W, cplx* x, const cplx* b, const double* weights) {
    if (!W->factorized) {
        skeel_row_scaling(W->row_scaling, W->A, weights, W->n, -30.0);
        apply_row_scaling(W);
    }
    ws_ldiv(W, x, b);
}
/* inverse_inf_norm_est, src/linear_algebra.jl:585-682 */
static double inverse_inf_norm_est(orc_ws* W, const double* dl, const double* dr) {
    if (!W->factorized) ws_factorize(W);
    int n = W->n;
    const double* rs = W->scaled ? W->row_scaling : NULL;
    cplx* y = W->work;
    double* x = W->rwork;
    for (int i = 0; i < n; i++) { x[i] = 1.0 / n; if (dr) x[i] /= dr[i]; }
    for (int i = 0; i < n; i++) y[i] = C(x[i], 0);
    lu_solve_adj(W->LU, W->ipiv, n, y);
    for (int i = 0; i < n; i++) { if (dl) y[i] = cdivr(y[i], dl[i]); if (rs) y[i] = cscale(rs[i], y[i]); }
    double g = 0.0;
    for (int i = 0; i < n; i++) g += fast_abs(y[i]);
    for (int i = 0; i < n; i++) {
        y[i] = cdivr(y[i], fast_abs(y[i]));
        if (dl) y[i] = cdivr(y[i], dl[i]);
        if (rs) y[i] = cdivr(y[i], rs[i]);
    }
    lu_solve(W->LU, W->ipiv, n, y);
    for (int i = 0; i < n; i++) x[i] = dr ? y[i].re / dr[i] : y[i].re;
    int k = 2;
    while (1) {
        int j = 0;
        double mx = fabs(x[0]);
        for (int i = 1; i < n; i++) if (fabs(x[i]) > mx) { j = i; mx = fabs(x[i]); }
        for (int i = 0; i < n; i++) x[i] = 0.0;
        x[j] = 1.0;
        if (dr) for (int i = 0; i < n; i++) x[i] /= dr[i];
        for (int i = 0; i < n; i++) y[i] = C(x[i], 0);
        lu_solve_adj(W->LU, W->ipiv, n, y);
        for (int i = 0; i < n; i++) { if (dl) y[i] = cdivr(y[i], dl[i]); if (rs) y[i] = cscale(rs[i], y[i]); }
        double gb = g;
        g = 0.0;
        for (int i = 0; i < n; i++) g += fast_abs(y[i]);
        if (g <= gb) { g = gb; break; }
        for (int i = 0; i < n; i++) {
            y[i] = cdivr(y[i], fast_abs(y[i]));
            if (dl) y[i] = cdivr(y[i], dl[i]);
            if (rs) y[i] = cdivr(y[i], rs[i]);
        }
        lu_solve(W->LU, W->ipiv, n, y);
        for (int i = 0; i < n; i++) x[i] = dr ? y[i].re / dr[i] : y[i].re;
        k += 1;
        double xn = fabs(x[0]);
        for (int i = 1; i < n; i++) xn = jmax(xn, fabs(x[i]));
        if (x[j] == xn || k > 2) break;
    }
    return nanmin(g, INFINITY);
}
/* inf_norm, src/linear_algebra.jl:684-707 */
static double inf_norm_ws(const orc_ws* W, const double* dl, const double* dr) {
    int n = W->n;
    double nrm = -INFINITY;
    for (int i = 0; i < n; i++) {
        double ni = 0.0;
        for (int j = 0; j < n; j++) ni += dr ? fast_abs(AT(W->A, i, j, n)) * dr[j] : fast_abs(AT(W->A, i, j, n));
        if (dl) ni *= dl[i];
        nrm = max_fast(nrm, ni);
    }
    return nrm;
}
/* LA.cond(WS, d_l, d_r), src/linear_algebra.jl:745-774 */
static double ws_cond(orc_ws* W, const double* dl, const double* dr) {
    if (W->n == 1) {
        double a = chypot(W->A[0]);
        if (!dl && !dr) return 1.0 / a;
        if (!dl) return 1.0 / (a * dr[0]);
        if (!dr) return 1.0 / (dl[0] * a);
        return 1.0 / (dl[0] * a * dr[0]);
    }
    return inverse_inf_norm_est(W, dl, dr) * inf_norm_ws(W, dl, dr);
}

/* numerical rank as LinearAlgebra.rank(J, rtol = 1e-14): singular values by
 * one-sided (Hestenes) Jacobi on a copy of J */
static int rank_rtol(const cplx* J, int n, double rtol) {
    cplx* A = malloc(sizeof(cplx) * n * n);
    memcpy(A, J, sizeof(cplx) * n * n);
    for (int sweep = 0; sweep < 60; sweep++) {
        double off = 0.0;
        for (int p = 0; p < n; p++)
            for (int q = p + 1; q < n; q++) {
                double app = 0, aqq = 0; cplx apq = C(0, 0);
                for (int i = 0; i < n; i++) {
                    app += cabs2(AT(A, i, p, n)); aqq += cabs2(AT(A, i, q, n));
                    apq = cadd(apq, cmul(cconj(AT(A, i, p, n)), AT(A, i, q, n)));
                }
                double g = sqrt(cabs2(apq));
                if (g == 0 || g <= 1e-300) continue;
                double rel = g / sqrt(app * aqq);
                if (rel > off) off = rel;
                if (rel < 1e-15) continue;
                double zeta = (aqq - app) / (2 * g);
                double tt = (zeta >= 0 ? 1.0 : -1.0) / (fabs(zeta) + sqrt(1 + zeta * zeta));
                double cs = 1 / sqrt(1 + tt * tt), sn = cs * tt;
                cplx ph = C(apq.re / g, apq.im / g);
                for (int i = 0; i < n; i++) {
                    cplx ap = AT(A, i, p, n), aq = AT(A, i, q, n);
                    cplx aqr = cmul(aq, cconj(ph));
                    AT(A, i, p, n) = csub(cscale(cs, ap), cscale(sn, aqr));
                    AT(A, i, q, n) = cmul(cadd(cscale(sn, ap), cscale(cs, aqr)), ph);
                }
            }
        if (off < 1e-15) break;
    }
    double smax = 0; double* s = malloc(sizeof(double) * n);
    for (int p = 0; p < n; p++) { double v = 0; for (int i = 0; i < n; i++) v += cabs2(AT(A, i, p, n)); s[p] = sqrt(v); if (s[p] > smax) smax = s[p]; }
    int r = 0;
    for (int p = 0; p < n; p++) if (s[p] > rtol * smax) r++;
    free(A); free(s);
    return r;
}

/* =============================== tracker state =============================== */
typedef struct {
    int n;
    const orc_options* opt;
    orc_H H;
    orc_ws J;
    /* TrackerState, src/tracker.jl:307-338 */
    cplx *x, *xh, *xb;
    /* SegmentStepper, src/utils.jl:300-309 */
    cplx seg_start, seg_target;
    double seg_absD; int seg_forward; double seg_s, seg_s1;
    double ds_prev, accuracy, omega, omega_prev, mu, tau, norm_dx0;
    int extended_prec, used_extended_prec, refined_extended_prec, keep_extended_prec, use_strict_beta_tau;
    double* w; /* weighted norm weights */
    double cond_J_xd;
    int code;
    long accepted, rejected, last_steps_failed, ext_accepted, ext_rejected;
    /* NewtonCorrector buffers */
    cplx *dx, *r;
    cdd* xdd;
    /* Predictor, src/predictor.jl:72-103 */
    int pred_hermite; int pred_order;
    double trust_region, local_error, cond_H_xd, ds_plain;
    cplx *tx0, *tx1, *tx2, *tx3; /* x, x¹, x², x³ */
    cplx pred_t, pred_prev_t;
    double tx_norm[4];
    cplx *xtemp, *u;
    cplx *prev_tx0, *prev_tx1;
    int winding;
    cplx pred_s, pred_prev_s;
    cplx *ty0, *ty1, *prev_ty0, *prev_ty1;
    long newton_iters;
    /* EndgameTrackerState, src/endgame_tracker.jl:177-215 */
    int eg_code, singular_endgame;
    double *val_x, *val_tx, *dval_x, *dval_tx, *vx2, *vx1, *vdx2, *vdx1, *lx2, *lx1, *ldx2, *ldx1;
    double logt2, logt1;
    int eg_winding; /* 0 = nothing */
    cplx* solution;
    double eg_accuracy, eg_cond;
    int eg_singular;
    long steps_eg, ext_steps_eg_start;
    int jtzf0, jtzf1;
    cplx* last_point; double last_t;
    double *row_scaling, *col_scaling;
    double *ai_starts, *ai_tols, *ai_abs_coords, *ai_conds;
    cplx* samples[3][2];
    double sample_times[3], sample_conds[3];
    double singular_start;
    long singular_steps;
    cplx *prediction, *prev_prediction;
    /* per-step record log */
    float* log_records;
    atomic_int* log_count;
    int log_capacity, path;
} orc_tracker;

static cplx* cvec(int n) { return calloc(n, sizeof(cplx)); }
static double* dvec(int n) { return calloc(n, sizeof(double)); }

static void tracker_create(orc_tracker* T, const orc_homotopy_desc* hd, const orc_options* opt) {
    memset(T, 0, sizeof *T);
    int n = hd->n;
    T->n = n; T->opt = opt;
    H_init(&T->H, hd);
    ws_init(&T->J, n);
    T->x = cvec(n); T->xh = cvec(n); T->xb = cvec(n);
    T->w = dvec(n); for (int i = 0; i < n; i++) T->w[i] = 1.0;
    T->dx = cvec(n); T->r = cvec(n); T->xdd = calloc(n, sizeof(cdd));
    T->tx0 = cvec(n); T->tx1 = cvec(n); T->tx2 = cvec(n); T->tx3 = cvec(n);
    T->xtemp = cvec(n); T->u = cvec(n); T->prev_tx0 = cvec(n); T->prev_tx1 = cvec(n);
    T->ty0 = cvec(n); T->ty1 = cvec(n); T->prev_ty0 = cvec(n); T->prev_ty1 = cvec(n);
    T->val_x = dvec(n); T->val_tx = dvec(n); T->dval_x = dvec(n); T->dval_tx = dvec(n);
    T->vx2 = dvec(n); T->vx1 = dvec(n); T->vdx2 = dvec(n); T->vdx1 = dvec(n);
    T->lx2 = dvec(n); T->lx1 = dvec(n); T->ldx2 = dvec(n); T->ldx1 = dvec(n);
    T->solution = cvec(n); T->last_point = cvec(n);
    T->row_scaling = dvec(n); T->col_scaling = dvec(n);
    T->ai_starts = dvec(n); T->ai_tols = dvec(n); T->ai_abs_coords = dvec(n); T->ai_conds = dvec(n);
    for (int k = 0; k < 3; k++) { T->samples[k][0] = cvec(n); T->samples[k][1] = cvec(n); }
    T->prediction = cvec(n); T->prev_prediction = cvec(n);
    T->pred_order = 4;
}
static void tracker_destroy(orc_tracker* T) {
    H_free(&T->H); ws_free(&T->J);
    void* ps[] = {T->x, T->xh, T->xb, T->w, T->dx, T->r, T->xdd, T->tx0, T->tx1, T->tx2, T->tx3, T->xtemp, T->u,
                  T->prev_tx0, T->prev_tx1, T->ty0, T->ty1, T->prev_ty0, T->prev_ty1, T->val_x, T->val_tx,
                  T->dval_x, T->dval_tx, T->vx2, T->vx1, T->vdx2, T->vdx1, T->lx2, T->lx1, T->ldx2, T->ldx1,
                  T->solution, T->last_point, T->row_scaling, T->col_scaling, T->ai_starts, T->ai_tols,
                  T->ai_abs_coords, T->ai_conds, T->prediction, T->prev_prediction};
    for (size_t i = 0; i < sizeof ps / sizeof ps[0]; i++) free(ps[i]);
    for (int k = 0; k < 3; k++) { free(T->samples[k][0]); free(T->samples[k][1]); }
}

/* ---- SegmentStepper, src/utils.jl:300-401 ---- */
static void seg_init(orc_tracker* T, cplx start, cplx target) {
    T->seg_start = start; T->seg_target = target;
    cplx d = csub(target, start);
    T->seg_absD = chypot(d);
    T->seg_forward = chypot(start) < chypot(target);
    T->seg_s = T->seg_s1 = T->seg_forward ? 0.0 : T->seg_absD;
}
static int seg_is_done(const orc_tracker* T) { return T->seg_forward ? T->seg_s == T->seg_absD : T->seg_s == 0.0; }
static void seg_step_success(orc_tracker* T) { T->seg_s = T->seg_s1; }
static void seg_propose(orc_tracker* T, double ds) {
    if (T->seg_forward) T->seg_s1 = jmin(T->seg_s + ds, T->seg_absD);
    else T->seg_s1 = jmax(T->seg_s - ds, 0.0);
}
static double seg_dist_to_target(const orc_tracker* T) { return T->seg_forward ? T->seg_absD - T->seg_s : T->seg_s; }
static double seg_ds(const orc_tracker* T) { return T->seg_forward ? T->seg_s1 - T->seg_s : T->seg_s - T->seg_s1; }
static cplx seg_t_helper(const orc_tracker* T, double s) {
    cplx st = T->seg_start, tg = T->seg_target; double D = T->seg_absD;
    if (T->seg_forward) {
        if (s == 0.0) return st;
        if (s == D) return tg;
        return cadd(st, cscale(s / D, csub(tg, st)));
    } else {
        if (s == D) return st;
        if (s == 0.0) return tg;
        return cadd(tg, cscale(s / D, csub(st, tg)));
    }
}
static cplx seg_t(const orc_tracker* T) { return seg_t_helper(T, T->seg_s); }
static cplx seg_t1(const orc_tracker* T) { return seg_t_helper(T, T->seg_s1); }
static cplx seg_dt(const orc_tracker* T) {
    if (T->seg_forward) return cscale((T->seg_s1 - T->seg_s) / T->seg_absD, csub(T->seg_target, T->seg_start));
    return cscale((T->seg_s - T->seg_s1) / T->seg_absD, csub(T->seg_target, T->seg_start));
}

/* ---- WeightedNorm init!/update!, src/norm.jl:101-136 ---- */
static void wnorm_init(orc_tracker* T, const cplx* x) {
    const orc_options* o = T->opt;
    double pn = norm_inf(x, NULL, T->n);
    for (int i = 0; i < T->n; i++) {
        double wi = fast_abs(x[i]);
        if (wi < o->scale_min * pn) wi = o->scale_min * pn;
        else if (wi > o->scale_max * pn) wi = o->scale_max * pn;
        T->w[i] = jmax(wi, o->scale_abs_min);
    }
}
static void wnorm_update(orc_tracker* T, const cplx* x) {
    const orc_options* o = T->opt;
    double nx = norm_inf(x, T->w, T->n);
    for (int i = 0; i < T->n; i++) {
        double wi = (fast_abs(x[i]) + T->w[i]) / 2;
        if (wi < o->scale_min * nx) wi = o->scale_min * nx;
        else if (wi > o->scale_max * nx) wi = o->scale_max * nx;
        if (isfinite(wi)) T->w[i] = jmax(wi, o->scale_abs_min);
    }
}

/* ============================== Newton corrector ============================== */
typedef struct { int code; double accuracy; int iters; double omega, theta, mu_low, norm_dx0; } newton_result;

static void eval_dd_at(orc_tracker* T, const cplx* x, cplx t, cplx* r) {
    for (int i = 0; i < T->n; i++) T->xdd[i] = cdd_from(x[i]);
    H_eval_dd(&T->H, T->xdd, t, r);
}

/* newton!, src/newton_corrector.jl:80-205.  xb = x̄ (output, all iterates alias it) */
static newton_result newton(orc_tracker* T, cplx* xb, const cplx* x0, cplx t, double mu, double omega,
                            int ext, int accurate_mu, int first_correction) {
    int n = T->n;
    double a = T->opt->a, ha = h_of(a);
    cplx *dx = T->dx, *r = T->r;
    if (xb != x0) memcpy(xb, x0, sizeof(cplx) * n);
    double mu_low = NAN, theta = NAN, ndx = NAN, ndx_prev = NAN, ndx0 = NAN;
    double abar = a;
    newton_result R;
    for (int i = 0; i <= 10; i++) {
        T->newton_iters++;
        H_eval_jac(&T->H, xb, t, r, T->J.A);
        if (ext) eval_dd_at(T, xb, t, r);
        ws_updated(&T->J);
        jac_ldiv_w(&T->J, dx, r, T->w);
        if (ext) iterative_refinement(&T->J, dx, r, T->w, abar * abar, 3);
        ndx = norm_inf(dx, T->w, n);
        if (isnan(ndx)) {
            R.code = NEWT_SINGULARITY; R.accuracy = norm_inf(dx, T->w, n); R.iters = i + 1;
            R.omega = omega; R.theta = theta; R.mu_low = mu_low; R.norm_dx0 = ndx0;
            return R;
        }
        for (int k = 0; k < n; k++) xb[k] = csub(xb[k], dx[k]);
        if (i == 0) ndx0 = ndx;
        if (i == 1) omega = 2 * ndx / (ndx_prev * ndx_prev);
        if (i >= 1) theta = ndx / ndx_prev;
        if ((i >= 1 && theta > abar) || (i == 0 && !first_correction && 0.125 * ndx0 * omega > ha)) {
            R.code = NEWT_TERMINATED; R.accuracy = norm_inf(dx, T->w, n); R.iters = i + 1;
            R.omega = omega; R.theta = theta; R.mu_low = mu_low; R.norm_dx0 = ndx0;
            return R;
        } else if (omega * (ndx * ndx) < 2 * mu * sqrt(1 - 2 * ha)) {
            H_eval_jac(&T->H, xb, t, r, T->J.A);
            ws_updated(&T->J);
            if (ext) {
                ws_ldiv(&T->J, dx, r);
                mu_low = norm_inf(dx, T->w, n);
                eval_dd_at(T, xb, t, r);
            }
            ws_ldiv(&T->J, dx, r);
            if (ext) iterative_refinement(&T->J, dx, r, T->w, abar * abar, 3);
            for (int k = 0; k < n; k++) xb[k] = csub(xb[k], dx[k]);
            double ndx1 = norm_inf(dx, T->w, n);
            if (isnan(ndx1)) {
                R.code = NEWT_SINGULARITY; R.accuracy = norm_inf(dx, T->w, n); R.iters = i + 1;
                R.omega = omega; R.theta = theta; R.mu_low = mu_low; R.norm_dx0 = ndx0;
                return R;
            } else if (ndx1 > sqrt(ndx)) {
                theta = ndx1 / ndx;
                R.code = NEWT_TERMINATED; R.accuracy = ndx1; R.iters = i + 2;
                R.omega = omega; R.theta = theta; R.mu_low = mu_low; R.norm_dx0 = ndx0;
                return R;
            }
            if (ndx1 > 2 * mu && ext) {
                eval_dd_at(T, xb, t, r);
                ws_ldiv(&T->J, dx, r);
                ndx = ndx1;
                ndx1 = norm_inf(dx, T->w, n);
                mu = ndx1;
            } else if (ndx1 > 2 * mu || accurate_mu) {
                H_eval(&T->H, xb, t, r);
                ws_ldiv(&T->J, dx, r);
                mu = norm_inf(dx, T->w, n);
            } else {
                mu = ndx1;
            }
            if (i == 0) {
                double wb = 2 * ndx / (ndx1 * ndx1);
                if (wb < omega) omega = wb; else omega *= 0.25;
            }
            R.code = NEWT_CONVERGED; R.accuracy = mu; R.iters = i + 2;
            R.omega = omega; R.theta = theta; R.mu_low = mu_low; R.norm_dx0 = ndx0;
            return R;
        }
        ndx_prev = ndx;
        if (i >= 1) abar *= abar;
    }
    R.code = NEWT_MAX_ITERS; R.accuracy = mu; R.iters = 11;
    R.omega = omega; R.theta = theta; R.mu_low = mu_low; R.norm_dx0 = ndx0;
    return R;
}

/* init_newton!, src/newton_corrector.jl:207-286 */
static int init_newton(orc_tracker* T, cplx* xb, const cplx* x0, cplx t, int ext, double* omega_out, double* mu_out) {
    int n = T->n;
    double a = T->opt->a;
    cplx *dx = T->dx, *r = T->r;
    H_eval_jac(&T->H, x0, t, r, T->J.A);
    if (ext) eval_dd_at(T, x0, t, r);
    ws_updated(&T->J);
    jac_ldiv_w(&T->J, dx, r, T->w);
    double v = norm_inf(dx, T->w, n) + DBL_EPSILON;
    int valid = 0;
    double omega = NAN, mu = NAN;
    double eps = sqrt(v);
    for (int k = 1; k <= 3; k++) {
        for (int i = 0; i < n; i++) xb[i] = C(x0[i].re + eps * T->w[i], x0[i].im);
        H_eval_jac(&T->H, xb, t, r, T->J.A);
        if (ext) eval_dd_at(T, xb, t, r);
        ws_updated(&T->J);
        jac_ldiv_w(&T->J, dx, r, T->w);
        for (int i = 0; i < n; i++) xb[i] = csub(xb[i], dx[i]);
        double ndx0 = norm_inf(dx, T->w, n);
        if (ext) eval_dd_at(T, xb, t, r);
        else H_eval(&T->H, xb, t, r);
        jac_ldiv_w(&T->J, dx, r, T->w);
        for (int i = 0; i < n; i++) xb[i] = csub(xb[i], dx[i]);
        double ndx1 = norm_inf(dx, T->w, n) + DBL_EPSILON;
        if (ndx1 < a * ndx0) {
            omega = 2 * ndx1 / (ndx0 * ndx0);
            mu = ndx1;
            double a7 = powi(a, 7);
            if (omega * mu > a7) {
                newton_result rr = newton(T, xb, xb, t, a7 / omega, omega, ext, 1, 0);
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

/* extended_prec_refinement_step!, src/newton_corrector.jl:55-78 */
static double ext_refine_step(orc_tracker* T, cplx* xb, const cplx* x, cplx t, int simple_newton_step) {
    int n = T->n;
    H_eval_jac(&T->H, x, t, T->r, T->J.A);
    eval_dd_at(T, x, t, T->r);
    ws_updated(&T->J);
    jac_ldiv_w(&T->J, T->dx, T->r, T->w);
    iterative_refinement(&T->J, T->dx, T->r, T->w, 1e-8, 3);
    for (int i = 0; i < n; i++) xb[i] = csub(x[i], T->dx[i]);
    if (simple_newton_step) {
        eval_dd_at(T, xb, t, T->r);
        jac_ldiv_w(&T->J, T->dx, T->r, T->w);
    }
    return norm_inf(T->dx, T->w, n);
}

/* ================================= predictor ================================= */
/* t_to_s_plane, src/predictor.jl:338-351 */
static cplx t_to_s_plane(cplx t, int m) {
    double r = fast_abs(t);
    if (t.im == 0 && t.re > 0) return C(nthroot(r, m), 0);
    double th = atan2(t.im, t.re);
    th = fmod(th, 2 * M_PI); if (th < 0) th += 2 * M_PI;
    double rr = nthroot(r, m);
    return C(rr * cos(th / m), rr * sin(th / m));
}
static cplx cpowi(cplx z, int n) { cplx r = C(1, 0); for (int i = 0; i < n; i++) r = cmul(r, z); return r; }

static void pred_init(orc_tracker* T) {
    T->cond_H_xd = 1.0;
    T->winding = 1;
    T->pred_t = T->pred_prev_t = C(NAN, 0);
    T->pred_s = T->pred_prev_s = C(NAN, 0);
    T->trust_region = T->local_error = NAN;
}

/* update!(predictor, H, x, t, J, norm, x̂), src/predictor.jl:158-284; xhat may be NULL */
static void pred_update(orc_tracker* T, const cplx* x, cplx t, const cplx* xhat) {
    int n = T->n;
    orc_ws* J = &T->J;
    cplx *u = T->u, *xt = T->xtemp;
    int m = T->winding;
    memcpy(T->prev_tx0, T->tx0, sizeof(cplx) * n);
    memcpy(T->prev_tx1, T->tx1, sizeof(cplx) * n);
    T->pred_prev_t = T->pred_t; T->pred_t = t;
    if (m > 1) { T->pred_prev_s = T->pred_s; T->pred_s = t_to_s_plane(t, m); }
    if (!xhat) T->local_error = NAN;
    else {
        double ds = fast_abs(csub(t, T->pred_prev_t));
        T->local_error = dist_inf(xhat, x, T->w, n) / powi(ds, T->pred_order);
    }
    memcpy(T->tx0, x, sizeof(cplx) * n);
    T->tx_norm[0] = norm_inf(x, T->w, n);
    if (m > 1) memcpy(T->ty0, x, sizeof(cplx) * n);

    const cplx* tx[4] = {T->tx0, T->tx1, T->tx2, T->tx3};
    H_taylor(&T->H, 1, tx, t, u);
    for (int i = 0; i < n; i++) u[i] = cneg(u[i]);
    ws_ldiv(J, xt, u);
    double delta = fixed_refine(J, xt, u, T->w);
    T->cond_H_xd = delta / DBL_EPSILON;
    if (delta > 1e-10) iterative_refinement(J, xt, u, NULL, 1e-10, 5);
    T->tx_norm[1] = norm_inf(xt, T->w, n);
    memcpy(T->tx1, xt, sizeof(cplx) * n);
    if (m > 1) {
        cplx mu = (m == 2) ? cscale(2.0, T->pred_s) : cscale((double)m, cpowi(T->pred_s, m - 1));
        for (int i = 0; i < n; i++) T->ty1[i] = cmul(mu, xt[i]);
        T->pred_hermite = 1;
        T->pred_order = 4;
        T->trust_region = T->tx_norm[0] / T->tx_norm[1];
        if (isnan(T->local_error)) { double q = T->tx_norm[1] / T->tx_norm[0]; T->local_error = q * q * q; }
        return;
    }
    H_taylor(&T->H, 2, tx, t, u);
    for (int i = 0; i < n; i++) u[i] = cneg(u[i]);
    ws_ldiv(J, xt, u);
    if (delta > 1e-10) iterative_refinement(J, xt, u, T->w, 1e-10, 4);
    T->tx_norm[2] = norm_inf(xt, T->w, n);
    memcpy(T->tx2, xt, sizeof(cplx) * n);

    H_taylor(&T->H, 3, tx, t, u);
    for (int i = 0; i < n; i++) u[i] = cneg(u[i]);
    ws_ldiv(J, xt, u);
    if (delta > 1e-4) iterative_refinement(J, xt, u, T->w, 1e-4, 3);
    T->tx_norm[3] = norm_inf(xt, T->w, n);
    memcpy(T->tx3, xt, sizeof(cplx) * n);

    double tau = INFINITY;
    for (int i = 0; i < n; i++) {
        double c1 = fast_abs(T->tx1[i]), c2 = fast_abs(T->tx2[i]), c3 = fast_abs(T->tx3[i]);
        double lam = jmax(1e-6, c1);
        c1 /= lam; c2 /= lam * lam; c3 /= lam * lam * lam;
        double tol = 1e-14 * jmax(jmax(c1, c2), c3);
        if (!((c1 <= tol && c2 <= tol && c3 <= tol) || c2 <= tol)) {
            double ti = (c2 / c3) / lam;
            if (ti < tau) tau = ti;
        }
    }
    if (!isfinite(tau)) tau = T->tx_norm[2] / T->tx_norm[3];
    if (!isfinite(tau)) {
        double mx = T->tx_norm[0];
        for (int k = 1; k < 4; k++) mx = jmax(mx, T->tx_norm[k]);
        tau = T->tx_norm[0] / mx;
    }
    T->pred_hermite = 0;
    T->pred_order = 4;
    T->trust_region = tau;
    if (isnan(T->local_error)) { double it = 1.0 / tau; double it2 = it * it; T->local_error = it2 * it2; }
}

/* cubic_hermite! with real nodes, src/predictor.jl:354-371 */
static void cubic_hermite_real(cplx* out, const cplx* a0, const cplx* a1, double t0, const cplx* b0, const cplx* b1,
                               double t1, double t, int n) {
    double s = (t - t0) / (t1 - t0);
    double oms = 1 - s;
    double h00 = (1 + 2 * s) * (oms * oms), h10 = (t - t0) * (oms * oms);
    double h01 = (s * s) * (3 - 2 * s), h11 = (t - t0) * s * (s - 1);
    for (int i = 0; i < n; i++)
        out[i] = cadd(cadd(cadd(cscale(h00, a0[i]), cscale(h10, a1[i])), cscale(h01, b0[i])), cscale(h11, b1[i]));
}
static void cubic_hermite_cplx(cplx* out, const cplx* a0, const cplx* a1, cplx t0, const cplx* b0, const cplx* b1,
                               cplx t1, cplx t, int n) {
    cplx s = cdiv(csub(t, t0), csub(t1, t0));
    cplx oms = csub(C(1, 0), s);
    cplx oms2 = cmul(oms, oms);
    cplx h00 = cmul(cadd(C(1, 0), cscale(2, s)), oms2), h10 = cmul(csub(t, t0), oms2);
    cplx h01 = cmul(cmul(s, s), csub(C(3, 0), cscale(2, s))), h11 = cmul(cmul(csub(t, t0), s), csub(s, C(1, 0)));
    for (int i = 0; i < n; i++)
        out[i] = cadd(cadd(cadd(cmul(h00, a0[i]), cmul(h10, a1[i])), cmul(h01, b0[i])), cmul(h11, b1[i]));
}

/* predict!, src/predictor.jl:286-329 */
static void predict(orc_tracker* T, cplx* xh, const cplx* x, cplx t, cplx dt) {
    int n = T->n;
    int m = T->winding;
    if (!T->pred_hermite) {
        double lam = T->trust_region, lam2 = lam * lam, lam3 = lam * lam * lam;
        for (int i = 0; i < n; i++) {
            cplx x0 = T->tx0[i], x1 = T->tx1[i], x2 = T->tx2[i], x3 = T->tx3[i];
            double c = fast_abs(x0), c1 = fast_abs(x1), c2 = fast_abs(x2), c3 = fast_abs(x3);
            double a1 = c1 * lam, a2 = c2 * lam2, a3 = c3 * lam3;
            double tau = 1e-12 * sqrt(c * c + a1 * a1 + a2 * a2 + a3 * a3);
            if (c3 * lam3 <= tau || c2 * lam2 <= tau) {
                xh[i] = cadd(x0, cmul(dt, cadd(x1, cmul(dt, x2))));
            } else {
                cplx di = csub(C(1, 0), cdiv(cmul(dt, x3), x2));
                xh[i] = cadd(x0, cmul(dt, cadd(x1, cdiv(cmul(dt, x2), di))));
            }
        }
    } else {
        cplx prev_s = t_to_s_plane(T->pred_prev_t, m);
        cplx s = t_to_s_plane(t, m);
        cplx s1 = t_to_s_plane(cadd(t, dt), m);
        cplx prev_sm, sm;
        if (m == 2) { prev_sm = cscale(2, prev_s); sm = cscale(2, s); }
        else { prev_sm = cscale((double)m, cpowi(prev_s, m - 1)); sm = cscale((double)m, cpowi(s, m - 1)); }
        for (int i = 0; i < n; i++) { T->prev_ty0[i] = T->prev_tx0[i]; T->prev_ty1[i] = cmul(prev_sm, T->prev_tx1[i]); }
        for (int i = 0; i < n; i++) { T->ty0[i] = T->tx0[i]; T->ty1[i] = cmul(sm, T->tx1[i]); }
        if (prev_s.im == 0 && s.im == 0 && s1.im == 0)
            cubic_hermite_real(xh, T->prev_ty0, T->prev_ty1, prev_s.re, T->ty0, T->ty1, s.re, s1.re, n);
        else
            cubic_hermite_cplx(xh, T->prev_ty0, T->prev_ty1, prev_s, T->ty0, T->ty1, s, s1, n);
    }
}

/* ================================== tracker ================================== */
static long ext_steps(const orc_tracker* T) { return T->ext_accepted + T->ext_rejected; }
static long tsteps(const orc_tracker* T) { return T->accepted + T->rejected; }

/* initial_step_size, src/tracker.jl:520-539 */
static double initial_step_size(orc_tracker* T) {
    const orc_options* o = T->opt;
    double a = o->beta_a * o->a;
    int p = T->pred_order;
    double e = T->local_error;
    if (isinf(e)) e = 1e5;
    double tau = T->trust_region;
    double ds1 = nthroot((sqrt(1 + 2 * h_of(a)) - 1) / (T->omega * e), p) / o->beta_omega_p;
    double ds2 = o->beta_tau * tau;
    double ds = nanmin(ds1, ds2);
    return jmin(jmin(ds, o->max_step_size), o->max_initial_step_size);
}

/* update_stepsize!, src/tracker.jl:541-588 */
static void update_stepsize(orc_tracker* T, const newton_result* res) {
    const orc_options* o = T->opt;
    double a = o->beta_a * o->a;
    int p = T->pred_order;
    double omega = jclamp(T->omega + 2 * (T->omega - T->omega_prev), T->omega, 8 * T->omega);
    double tau = T->tau;
    double ds;
    if (res->code == NEWT_CONVERGED) {
        double e = T->local_error;
        double ds1 = nthroot((sqrt(1 + 2 * h_of(a)) - 1) / (omega * e), p) / o->beta_omega_p;
        double ds2 = o->beta_tau * tau;
        if (T->use_strict_beta_tau || seg_dist_to_target(T) < ds2) ds2 = o->strict_beta_tau * tau;
        ds = jmin(nanmin(ds1, ds2), o->max_step_size);
        if (T->use_strict_beta_tau && seg_dist_to_target(T) < ds) ds *= o->strict_beta_tau;
        ds = jmin(ds, 10 * T->ds_prev);
        if (T->last_steps_failed > 0) ds = jmin(ds, T->ds_prev);
        T->ds_plain = ds;
    } else {
        int j = res->iters - 2;
        double Th = (j >= 0) ? nthroot(res->theta, 1 << j) : 1.0;
        double hT = h_of(Th), ha = h_of(0.5 * a);
        if (isnan(Th) || res->code == NEWT_SINGULARITY || isnan(res->accuracy) || res->iters == 1 || hT < ha)
            ds = 0.25 * seg_ds(T);
        else
            ds = nthroot((sqrt(1 + 2 * h_of(0.5 * a)) - 1) / (sqrt(1 + 2 * h_of(Th)) - 1), p) * seg_ds(T);
    }
    seg_propose(T, ds);
}

/* check_terminated!, src/tracker.jl:591-619 */
static void check_terminated(orc_tracker* T) {
    const orc_options* o = T->opt;
    double tol_acc;
    if (T->extended_prec || !o->extended_precision) tol_acc = powi(o->a, (1 << o->min_newton_iters) - 1) * h_of(o->a);
    else tol_acc = INFINITY;
    cplx t1 = seg_t1(T), t = seg_t(T);
    if (seg_is_done(T)) T->code = TC_SUCCESS;
    else if (tsteps(T) >= o->max_steps) T->code = TC_MAX_STEPS;
    else if (T->omega * T->mu > tol_acc) T->code = TC_ACCURACY_LIMIT;
    else if (seg_ds(T) < o->min_step_size) T->code = TC_STEP_TOO_SMALL;
    else if (fast_abs(csub(t1, t)) <= 2 * jeps(fast_abs(t))) T->code = TC_STEP_TOO_SMALL;
    else if (o->min_rel_step_size > 0 && !(t1.re == T->seg_target.re && t1.im == T->seg_target.im) &&
             fast_abs(csub(t1, t)) < fast_abs(t) * o->min_rel_step_size)
        T->code = TC_STEP_TOO_SMALL;
}

/* init!(tracker, x₁, t₁, t₀), src/tracker.jl:639-754 */
static int tracker_init(orc_tracker* T, const cplx* x1, cplx t1, cplx t0, double omega_in, double mu_in, int extp) {
    const orc_options* o = T->opt;
    int n = T->n;
    memcpy(T->x, x1, sizeof(cplx) * n);
    seg_init(T, t1, t0);
    T->ds_prev = 0.0;
    T->accuracy = DBL_EPSILON;
    T->omega = 1.0;
    T->keep_extended_prec = 0;
    T->use_strict_beta_tau = 0;
    wnorm_init(T, T->x);
    T->code = TC_TRACKING;
    T->accepted = T->rejected = T->ext_accepted = T->ext_rejected = 0;
    T->last_steps_failed = 0;
    cplx t = seg_t(T);
    int valid;
    double omega = omega_in, mu = mu_in;
    if (isnan(omega) || isnan(mu)) {
        valid = init_newton(T, T->xb, T->x, t, extp, &omega, &mu);
        if (!valid && !extp) {
            extp = 1;
            valid = init_newton(T, T->xb, T->x, t, 1, &omega, &mu);
        }
    } else valid = 1;
    T->used_extended_prec = T->extended_prec = extp;
    if (!isnan(omega)) T->omega = omega;
    if (valid) {
        T->accuracy = mu;
        T->mu = jmax(mu, DBL_EPSILON);
    } else {
        H_eval_jac(&T->H, T->x, t, T->r, T->J.A);
        int anynan = 0;
        for (int i = 0; i < n * n; i++) if (cisnan(T->J.A[i])) anynan = 1;
        int corank;
        if (anynan) { T->code = TC_INVALID_START; corank = 0; }
        else corank = n - rank_rtol(T->J.A, n, 1e-14);
        T->code = corank > 0 ? TC_INVALID_START_SINGULAR : TC_INVALID_START;
        return 0;
    }
    T->tau = INFINITY;
    H_eval_jac(&T->H, T->x, t, T->r, T->J.A);
    ws_updated(&T->J);
    pred_init(T);
    pred_update(T, T->x, t, NULL);
    T->tau = T->trust_region;
    double ds = initial_step_size(T);
    ds = jmax(jmin(ds, INFINITY), o->min_step_size);
    seg_propose(T, ds);
    T->omega_prev = T->omega;
    return T->code == TC_TRACKING;
}

/* init!(tracker, t₀), src/tracker.jl:756-766 */
static void tracker_init_t0(orc_tracker* T, cplx t0) {
    T->code = TC_TRACKING;
    seg_init(T, seg_t(T), t0);
    double ds = initial_step_size(T);
    seg_propose(T, ds);
    T->ds_prev = 0.0;
}

/* use_extended_precision!, src/tracker.jl:788-813 */
static void use_extended_precision(orc_tracker* T) {
    if (!T->opt->extended_precision) return;
    if (T->extended_prec) return;
    T->extended_prec = 1;
    T->used_extended_prec = 1;
    cplx t = seg_t(T);
    double mu = T->mu;
    for (int i = 0; i < 2; i++) mu = ext_refine_step(T, T->x, T->x, t, 0);
    T->mu = jmax(mu, DBL_EPSILON);
}
/* update_precision!, src/tracker.jl:768-786 */
static void update_precision(orc_tracker* T, double mu_low) {
    const orc_options* o = T->opt;
    if (!o->extended_precision) return;
    double a = o->a;
    double mu = T->mu, omega = T->omega;
    if (T->extended_prec && !T->keep_extended_prec && !isnan(mu_low) && mu_low > mu) {
        if (mu_low * omega < powi(a, 7) * h_of(a)) { T->extended_prec = 0; T->mu = mu_low; }
    } else if (mu * omega > powi(a, 5) * h_of(a)) {
        use_extended_precision(T);
    }
}
/* refine_current_solution!, src/tracker.jl:815-844 */
static double refine_current_solution(orc_tracker* T, double min_tol, int nsteps) {
    int n = T->n;
    cplx t = seg_t(T);
    double mu = T->accuracy;
    double mub = ext_refine_step(T, T->xb, T->x, t, 0);
    if (mub < mu) { memcpy(T->x, T->xb, sizeof(cplx) * n); mu = mub; }
    int k = 1;
    while (mu > min_tol && k <= nsteps) {
        mub = ext_refine_step(T, T->xb, T->x, t, 1);
        if (mub < mu) { memcpy(T->x, T->xb, sizeof(cplx) * n); mu = mub; }
        k++;
    }
    return mu;
}

/* Per-step features of the learned step-size proposer's training data (the GPU tracker
 * forms the same eight values after every step): log Δs, Newton contraction θ,
 * log(e Δs^4), t, log10 cond(H_x, ẋ), log ω, log μ, log τ; NaN -> 0, clamped to ±80. */
static float feat_clean(double v) { return isnan(v) ? 0.0f : (float)jclamp(v, -80.0, 80.0); }
static void step_features(orc_tracker* T, const newton_result* res, double ds0, float* f) {
    f[0] = feat_clean(hc_log(ds0));
    f[1] = feat_clean(res->theta);
    f[2] = feat_clean(hc_log(T->local_error * powi(ds0, 4)));
    f[3] = feat_clean(seg_t(T).re);
    f[4] = feat_clean(hc_log(T->cond_H_xd) / 2.302585092994046);
    f[5] = feat_clean(hc_log(T->omega));
    f[6] = feat_clean(hc_log(T->mu));
    f[7] = feat_clean(hc_log(T->tau));
}

/* step record: 0 path, 1 step index, 2-9 features, 10 accepted, 11 Newton iterations,
 * 12 log of the next proposed Δs, 13 flags (1 extended precision, 2 singular endgame),
 * 14 log10 accuracy, 15 log of the plain rule's next proposal before the segment end clamps it
 * (after a rejected step: same as 12) */
static void log_step(orc_tracker* T, const float* feat, const newton_result* res, int acc) {
    int idx = atomic_fetch_add(T->log_count, 1);
    if (idx >= T->log_capacity) return;
    float* r = T->log_records + (size_t)idx * 16;
    r[0] = (float)T->path;
    r[1] = (float)(T->accepted + T->rejected - 1);
    for (int i = 0; i < 8; i++) r[2 + i] = feat[i];
    r[10] = (float)acc;
    r[11] = (float)res->iters;
    r[12] = (float)hc_log(seg_ds(T));
    r[13] = (float)(T->extended_prec + 2 * T->singular_endgame);
    r[14] = (float)(hc_log(T->accuracy) / 2.302585092994046);
    r[15] = isnan(T->ds_plain) ? r[12] : (float)hc_log(T->ds_plain);
}

/* step!(tracker), src/tracker.jl:851-926; returns !last_step_failed */
static int tracker_step(orc_tracker* T) {
    const orc_options* o = T->opt;
    int n = T->n;
    cplx t = seg_t(T), dt = seg_dt(T), t1 = seg_t1(T);
    const double ds0 = seg_ds(T);
    predict(T, T->xh, T->x, t, dt);
    wnorm_update(T, T->xh);
    newton_result res = newton(T, T->xb, T->xh, t1, T->mu, T->omega, T->extended_prec, 0, T->accepted == 0);
    if (res.code == NEWT_CONVERGED) {
        memcpy(T->x, T->xb, sizeof(cplx) * n);
        T->ds_prev = seg_ds(T);
        seg_step_success(T);
        T->accuracy = res.accuracy;
        T->mu = jmax(res.accuracy, DBL_EPSILON);
        T->omega_prev = T->omega;
        T->omega = jmax(jmax(res.omega, 0.5 * T->omega), 0.1);
        update_precision(T, res.mu_low);
        if (seg_is_done(T) && o->extended_precision && T->accuracy > 1e-14) {
            T->accuracy = refine_current_solution(T, 1e-14, 3);
            T->refined_extended_prec = 1;
        }
        pred_update(T, T->x, seg_t(T), T->xh);
        T->tau = T->trust_region;
        T->accepted++;
        T->ext_accepted += T->extended_prec;
        T->last_steps_failed = 0;
    } else {
        T->rejected++;
        T->ext_rejected += T->extended_prec;
        T->last_steps_failed++;
    }
    T->norm_dx0 = res.norm_dx0;
    float feat[8];
    if (T->log_records) step_features(T, &res, ds0, feat);
    T->ds_plain = NAN;
    update_stepsize(T, &res);
    if (T->log_records) log_step(T, feat, &res, res.code == NEWT_CONVERGED);
    check_terminated(T);
    return !(T->last_steps_failed > 0);
}

/* ================================= valuation ================================= */
static void val_init(orc_tracker* T) {
    int n = T->n;
    for (int i = 0; i < n; i++) {
        T->val_x[i] = T->val_tx[i] = T->dval_x[i] = T->dval_tx[i] = 0;
        T->vx2[i] = T->vx1[i] = T->vdx2[i] = T->vdx1[i] = 0;
        T->lx2[i] = T->lx1[i] = T->ldx2[i] = T->ldx1[i] = 0;
    }
    T->logt2 = T->logt1 = NAN;
}
static double nu1(cplx x, cplx xd, double t) {
    double xx = x.re * x.re + x.im * x.im;
    double mu = x.re * xd.re + x.im * xd.im;
    return t * (mu / xx);
}
static void nu_nu1(cplx x, cplx xd, cplx x2, double t, double* v, double* v1) {
    double xx = x.re * x.re + x.im * x.im;
    double mu = x.re * xd.re + x.im * xd.im;
    double l = mu / xx;
    double mu1 = x.re * x2.re + xd.re * xd.re + x.im * x2.im + xd.im * xd.im;
    double l1 = mu1 / xx - 2 * (l * l);
    *v = t * l; *v1 = t * l1 + l;
}
/* finite_diff, src/valuation.jl:138-141 */
static double finite_diff(double v, double s, double v2, double s2, double v1, double s1) {
    double D1 = s - s1, D2 = s - s2, D12 = s1 - s2;
    return (D2 * v1) / (D12 * D1) - ((D12 + D2) * v2) / (D12 * D2) - (D12 * v) / (D1 * D2);
}
/* update!(val, pred, t), src/valuation.jl:82-124 */
static void val_update(orc_tracker* T, double t) {
    int n = T->n;
    double logt = hc_log(t);
    double logt2 = T->logt2, logt1 = T->logt1;
    int diff = T->winding > 1 && !isnan(logt2);
    for (int i = 0; i < n; i++) {
        cplx xi = T->tx0[i], xdi = T->tx1[i], x2i = T->tx2[i], x3i = T->tx3[i];
        double lxi = hc_log(fast_abs(xi)), ldxi = hc_log(fast_abs(xdi));
        if (diff) {
            double vi = nu1(xi, xdi, t);
            double dvi = finite_diff(vi, logt, T->vx2[i], logt2, T->vx1[i], logt1);
            T->val_x[i] = vi; T->dval_x[i] = dvi;
            double vdx = finite_diff(ldxi, logt, T->ldx2[i], logt2, T->ldx1[i], logt1);
            double dvdx = finite_diff(vdx, logt, T->vdx2[i], logt2, T->vdx1[i], logt1);
            T->val_tx[i] = vdx + 1; T->dval_tx[i] = dvdx;
        } else {
            double vi, v1;
            nu_nu1(xi, xdi, cscale(2, x2i), t, &vi, &v1);
            T->val_x[i] = vi; T->dval_x[i] = t * v1;
            T->vx1[i] = T->vx2[i]; T->vx2[i] = vi;
            double vdx, v1b;
            nu_nu1(xdi, cscale(2, x2i), cscale(6, x3i), t, &vdx, &v1b);
            T->val_tx[i] = vdx + 1; T->dval_tx[i] = t * v1b;
        }
        T->lx1[i] = T->lx2[i]; T->lx2[i] = lxi;
        T->ldx1[i] = T->ldx2[i]; T->ldx2[i] = ldxi;
    }
    T->logt1 = logt2; T->logt2 = logt;
}
/* at_infinity_tol!, src/valuation.jl:143-173 */
static void at_infinity_tol(orc_tracker* T, double finite_tol, int zero_is_finite) {
    for (int i = 0; i < T->n; i++) {
        double vx = T->val_x[i];
        double e = jmax(jmax(fabs(1.0 - T->val_tx[i] / vx), fabs(T->dval_x[i] / vx)), fabs(T->dval_tx[i] / T->val_tx[i]));
        if (isnan(e)) { T->ai_tols[i] = INFINITY; continue; }
        if (vx + e < -finite_tol) T->ai_tols[i] = e;
        else if (!zero_is_finite && vx - e > -finite_tol) T->ai_tols[i] = e;
        else T->ai_tols[i] = INFINITY;
    }
}
/* is_finite(val), src/valuation.jl:175-205 */
static int val_is_finite(orc_tracker* T, double finite_tol, int zero_is_finite, int max_winding) {
    double delta = 1.0 / max_winding;
    for (int i = 0; i < T->n; i++) {
        double vx = T->val_x[i];
        if (fabs(vx) < finite_tol) {
            if (!(fabs(T->dval_x[i]) < finite_tol) || T->val_tx[i] < 0.5 * delta) return 0;
        } else if (zero_is_finite && vx > (delta - finite_tol)) {
            double e = jmax(jmax(fabs(1.0 - T->val_tx[i] / vx), fabs(T->dval_x[i] / vx)), fabs(T->dval_tx[i] / T->val_tx[i]));
            if (!(e < finite_tol)) return 0;
        } else return 0;
    }
    return 1;
}
/* estimate_winding_number / check_winding_number, src/valuation.jl:207-228 */
static double check_winding_number(orc_tracker* T, int m) {
    double err = 0.0;
    for (int i = 0; i < T->n; i++) {
        double mv = m * T->val_tx[i];
        double ei = fabs(nearbyint(mv) - mv);
        err = ei > err ? ei : err;
    }
    return err;
}
static int estimate_winding_number(orc_tracker* T, int max_w, double* err_out) {
    int m = 1; double min_err = INFINITY;
    for (int k = 1; k <= max_w; k++) {
        double e = check_winding_number(T, k);
        if (e < min_err) { m = k; min_err = e; }
    }
    *err_out = min_err;
    return m;
}

/* ================================== endgame ================================== */
/* LA.cond(tracker, x, t, d_l, d_r), src/tracker.jl:509-514 */
static double tracker_cond_at(orc_tracker* T, const cplx* x, cplx t, const double* dl, const double* dr) {
    H_eval_jac(&T->H, x, t, T->r, T->J.A);
    ws_updated(&T->J);
    return ws_cond(&T->J, dl, dr);
}
/* row_scaling! for a square workspace, src/linear_algebra.jl:466-479 */
static void eg_row_scaling(orc_tracker* T) {
    skeel_row_scaling(T->row_scaling, T->J.A, T->col_scaling, T->n, T->opt->scaling_threshold);
}

/* init!(endgame_tracker, x, t₁), src/endgame_tracker.jl:260-294 */
static void eg_init(orc_tracker* T, const cplx* x, double t1) {
    int n = T->n;
    tracker_init(T, x, C(t1, 0), C(0, 0), NAN, NAN, 0);
    T->eg_code = tc_to_eg(T->code);
    T->singular_endgame = 0;
    T->jtzf0 = T->jtzf1 = 0;
    val_init(T);
    T->eg_winding = 0;
    for (int i = 0; i < n; i++) T->solution[i] = C(NAN, NAN);
    T->eg_accuracy = NAN; T->eg_cond = NAN; T->eg_singular = 0;
    T->steps_eg = 0;
    T->ext_steps_eg_start = INT64_MAX;
    for (int i = 0; i < n; i++) { T->row_scaling[i] = 1; T->col_scaling[i] = 1; T->ai_starts[i] = NAN; }
    T->singular_steps = 0;
    T->last_t = NAN;
}

/* tracking_stopped!, src/endgame_tracker.jl:695-721 */
static void tracking_stopped(orc_tracker* T) {
    const orc_options* o = T->opt;
    int n = T->n;
    T->eg_accuracy = T->accuracy;
    if (T->eg_code == EG_SUCCESS && T->eg_accuracy > 1e-14) refine_current_solution(T, 1e-14, o->refine_steps);
    memcpy(T->solution, T->x, sizeof(cplx) * n);
    T->eg_winding = 0;
    if (T->eg_code == EG_SUCCESS) {
        memcpy(T->col_scaling, T->w, sizeof(double) * n);
        eg_row_scaling(T);
        T->eg_cond = tracker_cond_at(T, T->solution, C(0, 0), T->row_scaling, T->col_scaling);
        T->eg_singular = T->eg_cond > o->sing_cond || T->eg_accuracy > o->sing_accuracy;
    }
}

/* add_sample!, src/endgame_tracker.jl:630-662 */
static void add_sample(orc_tracker* T, double t) {
    int n = T->n;
    int m = T->eg_winding;
    double s = nthroot(t, m);
    double mu = m * powi(s, m - 1);
    double kappa = ws_cond(&T->J, T->row_scaling, T->col_scaling);
    int slot;
    if (T->singular_steps <= 2) slot = (int)T->singular_steps;
    else {
        cplx *a = T->samples[0][0], *b = T->samples[0][1];
        T->samples[0][0] = T->samples[1][0]; T->samples[0][1] = T->samples[1][1];
        T->sample_times[0] = T->sample_times[1]; T->sample_conds[0] = T->sample_conds[1];
        T->samples[1][0] = T->samples[2][0]; T->samples[1][1] = T->samples[2][1];
        T->sample_times[1] = T->sample_times[2]; T->sample_conds[1] = T->sample_conds[2];
        T->samples[2][0] = a; T->samples[2][1] = b;
        slot = 2;
    }
    for (int i = 0; i < n; i++) { T->samples[slot][0][i] = T->tx0[i]; T->samples[slot][1][i] = cscale(mu, T->tx1[i]); }
    T->sample_times[slot] = s;
    T->sample_conds[slot] = kappa;
}
/* predict_endpoint!, src/endgame_tracker.jl:664-693 */
static double predict_endpoint(orc_tracker* T) {
    int n = T->n;
    if (T->singular_steps < 2) return INFINITY;
    if (T->singular_steps == 2)
        cubic_hermite_real(T->prediction, T->samples[0][0], T->samples[0][1], T->sample_times[0], T->samples[1][0],
                           T->samples[1][1], T->sample_times[1], 0.0, n);
    memcpy(T->prev_prediction, T->prediction, sizeof(cplx) * n);
    cubic_hermite_real(T->prediction, T->samples[1][0], T->samples[1][1], T->sample_times[1], T->samples[2][0],
                       T->samples[2][1], T->sample_times[2], 0.0, n);
    double p = T->sample_times[2] / T->sample_times[1];
    double p2 = p * p;
    double err = dist_inf(T->prediction, T->prev_prediction, NULL, n) / fabs(p2 * p2 - 1);
    double ns = norm_inf(T->prediction, NULL, n);
    if (ns > 1e-8) err /= ns;
    return err;
}
/* check_finite!, src/endgame_tracker.jl:402-422 */
static int check_finite(orc_tracker* T) {
    const orc_options* o = T->opt;
    if (!val_is_finite(T, o->val_finite_tol, !o->zero_is_at_infinity, o->max_winding_number)) return 0;
    double merr;
    int m = estimate_winding_number(T, o->max_winding_number, &merr);
    if (merr < o->val_finite_tol) {
        if (m == 1 && !T->jtzf0) return 0;
        T->eg_winding = m;
        return 1;
    }
    return 0;
}
/* check_at_infinity!, src/endgame_tracker.jl:424-499 */
static int check_at_infinity(orc_tracker* T) {
    const orc_options* o = T->opt;
    int n = T->n;
    if (!o->at_infinity_check) return 0;
    at_infinity_tol(T, o->val_finite_tol, !o->zero_is_at_infinity);
    double kappa = NAN;
    double t = seg_t(T).re;
    for (int i = 0; i < n; i++) {
        double toli = T->ai_tols[i];
        if (toli < o->val_at_infinity_tol) {
            if (isnan(T->ai_starts[i])) {
                int allnan = 1;
                for (int k = 0; k < n; k++) if (!isnan(T->ai_starts[k])) allnan = 0;
                if (allnan) {
                    memcpy(T->col_scaling, T->w, sizeof(double) * n);
                    eg_row_scaling(T);
                }
                kappa = ws_cond(&T->J, T->row_scaling, T->col_scaling);
                T->ai_conds[i] = kappa;
                T->ai_abs_coords[i] = fast_abs(T->x[i]);
                T->ai_starts[i] = t;
            } else {
                if (isnan(kappa)) kappa = ws_cond(&T->J, T->row_scaling, T->col_scaling);
                double vi = T->val_x[i];
                int at_zero = T->val_x[i] > 0;
                double cond_growth = kappa / T->ai_conds[i];
                double coord_growth = at_zero ? T->ai_abs_coords[i] / fast_abs(T->x[i]) : fast_abs(T->x[i]) / T->ai_abs_coords[i];
                if (coord_growth > jclamp(hc_pow(0.25, 4 * vi), 20, o->min_coord_growth) &&
                    (cond_growth > o->min_cond_growth || kappa > jmax(1e8, o->min_cond))) {
                    T->eg_code = at_zero ? EG_AT_ZERO : EG_AT_INFINITY;
                    return 1;
                }
            }
        } else if (!isnan(T->ai_starts[i])) {
            T->ai_starts[i] = NAN;
        }
    }
    return 0;
}
/* switch_to_singular!, src/endgame_tracker.jl:501-524 */
static void switch_to_singular(orc_tracker* T) {
    int n = T->n;
    T->singular_endgame = 1;
    double t = seg_t(T).re;
    T->singular_start = t;
    int allone = 1;
    for (int i = 0; i < n; i++) if (T->row_scaling[i] != 1.0) allone = 0;
    if (allone) {
        memcpy(T->col_scaling, T->w, sizeof(double) * n);
        eg_row_scaling(T);
    }
    add_sample(T, t);
    T->singular_steps = 0;
    T->winding = T->eg_winding;
    T->ai_conds[0] = T->sample_conds[0];
    T->keep_extended_prec = 1;
}
static void switch_to_regular(orc_tracker* T) {
    T->singular_endgame = 0;
    T->winding = 1;
    tracker_init_t0(T, C(0, 0));
}

/* singular_endgame_step!, src/endgame_tracker.jl:533-628 */
static int singular_endgame_step(orc_tracker* T) {
    const orc_options* o = T->opt;
    int n = T->n;
    double lam = 0.25;
    double t = seg_t(T).re;
    tracker_init_t0(T, C(lam * t, 0));
    int max_steps = 0;
    while (T->code == TC_TRACKING) {
        tracker_step(T);
        if ((T->steps_eg += 1) >= o->max_endgame_steps) { T->eg_code = EG_MAX_STEPS; max_steps = 1; break; }
        else if (ext_steps(T) - T->ext_steps_eg_start > o->max_endgame_extended_steps) {
            T->eg_code = EG_MAX_EXT_STEPS; max_steps = 1; break;
        }
    }
    T->singular_steps += 1;
    if (orc_debug) fprintf(stderr, "sing step %ld t=%.6e code=%d maxsteps=%d acc=%.3e ext_steps=%ld start=%ld\n", T->singular_steps, seg_t(T).re, T->code, max_steps, T->accuracy, ext_steps(T), T->ext_steps_eg_start);
    if (!max_steps) {
        if (T->code != TC_SUCCESS) {
            T->eg_code = tc_to_eg(T->code);
            tracking_stopped(T);
            return T->eg_code;
        }
        val_update(T, lam * t);
        double merr;
        int mh = estimate_winding_number(T, o->max_winding_number, &merr);
        if (orc_debug) fprintf(stderr, "  winding est %d err %.3e (have %d) valtx=%.4f\n", mh, merr, T->eg_winding, T->val_tx[0]);
        if (mh != T->eg_winding || merr > 0.1) { switch_to_regular(T); return T->eg_code; }
        add_sample(T, lam * t);
        if (!(T->singular_steps >= 2)) return T->eg_code;
        double acc = predict_endpoint(T);
        if (orc_debug) fprintf(stderr, "  predict acc %.3e state.acc %.3e pred=%.12f%+.12fi\n", acc, T->eg_accuracy, T->prediction[0].re, T->prediction[0].im);
        if (T->singular_steps == 2) {
            T->eg_accuracy = acc;
            memcpy(T->solution, T->prediction, sizeof(cplx) * n);
            return T->eg_code;
        } else if (acc < T->eg_accuracy && T->eg_accuracy > 1e-12) {
            T->eg_accuracy = acc;
            memcpy(T->solution, T->prediction, sizeof(cplx) * n);
            return T->eg_code;
        }
    }
    /* @label prediction */
    int m = T->eg_winding;
    double kappa = T->sample_conds[2];
    double zero_cond = 1.0 / (m + 1);
    /* Bool * Complex is a strong zero: copysign(0, y) per component */
    for (int i = 0; i < n; i++)
        T->solution[i] = (T->val_x[i] < zero_cond) ? T->prediction[i]
                                                   : C(copysign(0.0, T->prediction[i].re), copysign(0.0, T->prediction[i].im));
    double kappa0 = tracker_cond_at(T, T->solution, C(0, 0), T->row_scaling, T->col_scaling);
    double J0 = inf_norm_ws(&T->J, T->row_scaling, NULL);
    if (orc_debug) fprintf(stderr, "  prediction label: kappa=%.3e kappa0=%.3e J0=%.3e acc=%.3e m=%d\n", kappa, kappa0, J0, T->eg_accuracy, m);
    if (T->eg_accuracy < o->singular_min_accuracy &&
        (((m > 1 && kappa > o->min_cond && nanmax(kappa0, 1.0 / J0) > kappa) || (m == 1 && kappa0 > 1e12)) ||
         max_steps || (n == 1 && 1.0 / J0 < o->min_cond))) {
        T->eg_cond = jmax(kappa0, 1.0 / J0);
        T->eg_singular = 1;
        T->eg_code = EG_SUCCESS;
        return T->eg_code;
    } else if (!max_steps) {
        switch_to_regular(T);
        return T->eg_code;
    }
    return T->eg_code;
}

/* step!(endgame_tracker), src/endgame_tracker.jl:329-399 */
static int eg_step(orc_tracker* T) {
    const orc_options* o = T->opt;
    int n = T->n;
    if (T->steps_eg >= o->max_endgame_steps) return T->eg_code = EG_MAX_STEPS;
    else if (ext_steps(T) - T->ext_steps_eg_start > o->max_endgame_extended_steps) {
        int allfinite = 1;
        for (int i = 0; i < n; i++) if (cisnan(T->solution[i])) allfinite = 0;
        if (allfinite && T->eg_winding != 0 && T->eg_accuracy < o->singular_min_accuracy) {
            T->eg_cond = tracker_cond_at(T, T->solution, C(0, 0), T->row_scaling, T->col_scaling);
            T->eg_singular = 1;
            return T->eg_code = EG_SUCCESS;
        }
        return T->eg_code = EG_MAX_EXT_STEPS;
    }
    memcpy(T->last_point, T->x, sizeof(cplx) * n);
    T->last_t = seg_t(T).re;
    if (T->singular_endgame) return singular_endgame_step(T);
    cplx tp = seg_t1(T);
    int is_jump_to_zero = tp.re == 0 && tp.im == 0;
    int step_success = tracker_step(T);
    T->eg_code = tc_to_eg(T->code);
    if (orc_debug) fprintf(stderr, "eg t=%.6e ok=%d acc=%.3e ext=%d code=%d steps=%ld/%ld\n", seg_t(T).re, step_success, T->accuracy, T->extended_prec, T->code, T->accepted, T->rejected);
    if (T->eg_code != EG_TRACKING) { tracking_stopped(T); return T->eg_code; }
    T->jtzf0 = T->jtzf1; T->jtzf1 = is_jump_to_zero;
    double t = seg_t(T).re;
    if (!(t <= o->endgame_start)) return T->eg_code;
    if (T->steps_eg == 0) T->ext_steps_eg_start = ext_steps(T);
    T->steps_eg += 1;
    if (!step_success) return T->eg_code;
    val_update(T, t);
    if (orc_debug) fprintf(stderr, "  val_x=%.4f val_tx=%.4f dval_x=%.4f dval_tx=%.4f\n", T->val_x[0], T->val_tx[0], T->dval_x[0], T->dval_tx[0]);
    if (check_finite(T)) { if (orc_debug) fprintf(stderr, "  SWITCH TO SINGULAR m=%d\n", T->eg_winding); switch_to_singular(T); return T->eg_code; }
    else if (check_at_infinity(T)) return T->eg_code;
    return T->eg_code;
}

/* PathResult(endgame_tracker, x, path_number), src/endgame_tracker.jl:847-888 */
static void eg_path_result(orc_tracker* T, orc_path_result* R, double* sol, double* last, double* val) {
    const orc_options* o = T->opt;
    int n = T->n;
    double t;
    const cplx* s;
    if (T->eg_code == EG_SUCCESS) { t = 0.0; s = T->solution; H_eval(&T->H, T->solution, C(0, 0), T->r); }
    else { t = seg_t(T).re; s = T->x; H_eval(&T->H, T->x, seg_t(T), T->r); }
    R->code = T->eg_code;
    R->singular = T->eg_singular;
    R->winding_number = T->eg_winding;
    R->extended_precision = T->extended_prec;
    R->extended_precision_used = T->used_extended_prec;
    R->accepted_steps = (int)T->accepted;
    R->rejected_steps = (int)T->rejected;
    R->steps_eg = (int)T->steps_eg;
    R->valuation_valid = !(t > o->endgame_start);
    R->newton_iters = (int)T->newton_iters;
    R->learned_proposed = R->learned_accepted = 0;
    R->t = t;
    R->accuracy = T->eg_accuracy;
    R->residual = norm_inf(T->r, NULL, n);
    R->cond = T->eg_cond;
    R->omega = T->omega;
    R->mu = T->mu;
    R->last_t = T->last_t;
    for (int i = 0; i < n; i++) {
        sol[2 * i] = s[i].re; sol[2 * i + 1] = s[i].im;
        last[2 * i] = T->last_point[i].re; last[2 * i + 1] = T->last_point[i].im;
        val[i] = T->val_x[i];
    }
}

/* track(endgame_tracker, x, t₁), src/endgame_tracker.jl:902-911 */
static void eg_track(orc_tracker* T, const cplx* x, double t1, orc_path_result* R, double* sol, double* last, double* val) {
    T->newton_iters = 0;
    eg_init(T, x, t1);
    while (T->eg_code == EG_TRACKING) eg_step(T);
    eg_path_result(T, R, sol, last, val);
}

/* ============================ public oracle API ============================ */
static int check_desc(const orc_homotopy_desc* hd) {
    if (!hd || hd->n < 1) { snprintf(orc_err, sizeof orc_err, "bad homotopy"); return -1; }
    if (orc_tape_check(hd->target.eval_jac) || orc_tape_check(hd->target.eval)) return -1;
    if (hd->kind == 0 && (orc_tape_check(hd->start.eval_jac) || orc_tape_check(hd->start.eval))) return -1;
    return 0;
}

typedef struct {
    const orc_homotopy_desc* hd;
    const orc_batch* b;
    const orc_options* o;
    orc_path_result* res;
    double *sol, *last, *val;
    atomic_long next;
    float* log_records;
    int log_capacity;
    atomic_int log_count;
} job_t;

static void set_path_params(orc_tracker* T, const job_t* J, long k) {
    if (J->hd->kind != 1) return;
    int np = T->H.npar;
    int g = J->b->group_of_path ? J->b->group_of_path[k] : 0;
    cplx* p = malloc(sizeof(cplx) * (np + 1));
    cplx* q = malloc(sizeof(cplx) * (np + 1));
    for (int i = 0; i < np; i++) {
        p[i] = C(J->b->start_params[2 * i], J->b->start_params[2 * i + 1]);
        q[i] = C(J->b->target_params[2 * (g * np + i)], J->b->target_params[2 * (g * np + i) + 1]);
    }
    H_set_params(&T->H, p, q);
    free(p); free(q);
}

static void* worker(void* arg) {
    job_t* J = arg;
    orc_tracker T;
    tracker_create(&T, J->hd, J->o);
    T.log_records = J->log_records;
    T.log_capacity = J->log_capacity;
    T.log_count = &J->log_count;
    T.cond_H_xd = 1.0;
    int n = J->hd->n;
    cplx* x = cvec(n);
    for (;;) {
        long k = atomic_fetch_add(&J->next, 1);
        if (k >= J->b->npaths) break;
        for (int i = 0; i < n; i++) x[i] = C(J->b->starts[2 * (k * n + i)], J->b->starts[2 * (k * n + i) + 1]);
        set_path_params(&T, J, k);
        T.path = (int)k;
        eg_track(&T, x, J->b->t1, &J->res[k], J->sol + 2 * k * n, J->last + 2 * k * n, J->val + k * n);
    }
    free(x);
    tracker_destroy(&T);
    return NULL;
}

/* Batch solve over `nthreads` threads with a dynamic atomic queue (threaded_solve). */
int orc_track_batch_log(const orc_homotopy_desc* hd, const orc_batch* b, const orc_options* o, int nthreads,
                        orc_path_result* res, double* sol, double* last, double* val, float* log_records,
                        int log_capacity, int* log_count) {
    if (check_desc(hd)) return -1;
    if (nthreads < 1) nthreads = 1;
    job_t J = {hd, b, o, res, sol, last, val, 0, log_records, log_capacity, 0};
    pthread_t* th = malloc(sizeof(pthread_t) * nthreads);
    for (int i = 0; i < nthreads; i++) pthread_create(&th[i], NULL, worker, &J);
    for (int i = 0; i < nthreads; i++) pthread_join(th[i], NULL);
    free(th);
    if (log_count) *log_count = atomic_load(&J.log_count);
    return 0;
}
int orc_track_batch(const orc_homotopy_desc* hd, const orc_batch* b, const orc_options* o, int nthreads,
                    orc_path_result* res, double* sol, double* last, double* val) {
    return orc_track_batch_log(hd, b, o, nthreads, res, sol, last, val, NULL, 0, NULL);
}

/* Plain Tracker API: track(tracker, x, t₁, t₀) -> TrackerResult (src/tracker.jl:1023-1026).
 * out: code, accepted, rejected, extended_precision(_used); dout: t(re,im), accuracy, ω, μ, τ */
int orc_track_plain(const orc_homotopy_desc* hd, const orc_options* o, const double* p, const double* q, const double* x1,
                    double t1, double t0, double* xout, int32_t* iout, double* dout) {
    if (check_desc(hd)) return -1;
    orc_tracker T;
    tracker_create(&T, hd, o);
    int n = hd->n;
    if (hd->kind == 1) {
        int np = T.H.npar;
        cplx *pp = cvec(np + 1), *qq = cvec(np + 1);
        for (int i = 0; i < np; i++) { pp[i] = C(p[2 * i], p[2 * i + 1]); qq[i] = C(q[2 * i], q[2 * i + 1]); }
        H_set_params(&T.H, pp, qq);
        free(pp); free(qq);
    }
    cplx* x = cvec(n);
    for (int i = 0; i < n; i++) x[i] = C(x1[2 * i], x1[2 * i + 1]);
    tracker_init(&T, x, C(t1, 0), C(t0, 0), NAN, NAN, 0);
    while (T.code == TC_TRACKING) tracker_step(&T);
    for (int i = 0; i < n; i++) { xout[2 * i] = T.x[i].re; xout[2 * i + 1] = T.x[i].im; }
    iout[0] = T.code; iout[1] = (int)T.accepted; iout[2] = (int)T.rejected;
    iout[3] = T.extended_prec || T.refined_extended_prec;
    iout[4] = T.used_extended_prec || T.refined_extended_prec;
    cplx t = seg_t(&T);
    dout[0] = t.re; dout[1] = t.im; dout[2] = T.accuracy; dout[3] = T.omega; dout[4] = T.mu; dout[5] = T.tau;
    free(x);
    tracker_destroy(&T);
    return 0;
}

/* Predictor probe for the oracle's pinning tests: runs the tracker's predictor sequence
 * on given points of a path, exactly as step! uses it (src/tracker.jl:851-926):
 *   for k = 0..npts-1:  [k > 0: predict!(x̂, pred, H, x_{k-1}, t_{k-1}, t_k - t_{k-1})]
 *                       J <- factorization of H_x(x_k, t_k)  (the corrector leaves an unscaled one)
 *                       update!(pred, H, x_k, t_k, J, norm, k > 0 ? x̂ : nothing)
 *   then predict!(out, pred, H, x_last, t_last, dt).
 * xs: npts*n complex, ts: npts complex, w: n weights of the WeightedNorm (NULL: all 1).
 * Outputs: tx: 4n complex (x⁰..x³ after the last update), xh: n complex (final prediction),
 * xhats: npts*n complex (the prediction made before update k; zero for k = 0),
 * scal: npts*5 doubles per update (trust_region, local_error, cond_H_ẋ, hermite, order). */
int orc_predictor_probe(const orc_homotopy_desc* hd, const orc_options* o, const double* p, const double* q,
                        int winding, int npts, const double* xs, const double* ts, const double* w, double dt_re,
                        double dt_im, double* tx, double* xh, double* xhats, double* scal) {
    if (check_desc(hd) || npts < 1 || winding < 1) return -1;
    orc_tracker T;
    tracker_create(&T, hd, o);
    int n = hd->n;
    if (hd->kind == 1) {
        int np = T.H.npar;
        cplx *pp = cvec(np + 1), *qq = cvec(np + 1);
        for (int i = 0; i < np; i++) { pp[i] = C(p[2 * i], p[2 * i + 1]); qq[i] = C(q[2 * i], q[2 * i + 1]); }
        H_set_params(&T.H, pp, qq);
        free(pp); free(qq);
    }
    if (w) for (int i = 0; i < n; i++) T.w[i] = w[i];
    pred_init(&T);
    T.winding = winding;
    cplx *x = cvec(n), *xp = cvec(n), *xhat = cvec(n);
    cplx tprev = C(NAN, 0);
    for (int k = 0; k < npts; k++) {
        cplx t = C(ts[2 * k], ts[2 * k + 1]);
        for (int i = 0; i < n; i++) x[i] = C(xs[2 * (k * n + i)], xs[2 * (k * n + i) + 1]);
        if (k > 0) predict(&T, xhat, xp, tprev, csub(t, tprev));
        for (int i = 0; i < n; i++) {
            xhats[2 * (k * n + i)] = k > 0 ? xhat[i].re : 0.0;
            xhats[2 * (k * n + i) + 1] = k > 0 ? xhat[i].im : 0.0;
        }
        H_eval_jac(&T.H, x, t, T.r, T.J.A);
        ws_updated(&T.J);
        ws_factorize(&T.J);
        pred_update(&T, x, t, k > 0 ? xhat : NULL);
        scal[5 * k] = T.trust_region; scal[5 * k + 1] = T.local_error; scal[5 * k + 2] = T.cond_H_xd;
        scal[5 * k + 3] = T.pred_hermite; scal[5 * k + 4] = T.pred_order;
        memcpy(xp, x, sizeof(cplx) * n);
        tprev = t;
    }
    const cplx* txs[4] = {T.tx0, T.tx1, T.tx2, T.tx3};
    for (int k = 0; k < 4; k++)
        for (int i = 0; i < n; i++) { tx[2 * (k * n + i)] = txs[k][i].re; tx[2 * (k * n + i) + 1] = txs[k][i].im; }
    predict(&T, xhat, xp, tprev, C(dt_re, dt_im));
    for (int i = 0; i < n; i++) { xh[2 * i] = xhat[i].re; xh[2 * i + 1] = xhat[i].im; }
    free(x); free(xp); free(xhat);
    tracker_destroy(&T);
    return 0;
}

/* ---- unit-level entry points used by the oracle's own pinning tests ---- */
int orc_sys_eval(const orc_tape* ej, const double* x, const double* p, int n, double* u, double* U) {
    if (orc_tape_check(ej)) return -1;
    cplx* T = calloc(ej->tape_space, sizeof(cplx));
    cplx *xx = cvec(n), *pp = cvec(ej->nparams + 1), *uu = cvec(ej->output_dim), *UU = U ? cvec(n * n) : NULL;
    for (int i = 0; i < n; i++) xx[i] = C(x[2 * i], x[2 * i + 1]);
    for (int i = 0; i < ej->nparams; i++) pp[i] = C(p[2 * i], p[2 * i + 1]);
    sys_eval(ej, xx, pp, T, uu, UU, n);
    for (int i = 0; i < ej->output_dim; i++) { u[2 * i] = uu[i].re; u[2 * i + 1] = uu[i].im; }
    if (U) for (int i = 0; i < n * n; i++) { U[2 * i] = UU[i].re; U[2 * i + 1] = UU[i].im; }
    free(T); free(xx); free(pp); free(uu); free(UU);
    return 0;
}
/* out: (K+1) * output_dim complex; tx: K*n complex (order-major); p0/p1 may be NULL */
int orc_sys_taylor(const orc_tape* ev, int K, const double* tx, const double* p0, const double* p1, int n, double* out) {
    if (orc_tape_check(ev) || K < 0 || K >= SMAX) return -1;
    ser* T = calloc(ev->tape_space, sizeof(ser));
    cplx* txc = cvec(K * n + 1);
    const cplx* txp[SMAX];
    for (int k = 0; k < K; k++) { for (int i = 0; i < n; i++) txc[k * n + i] = C(tx[2 * (k * n + i)], tx[2 * (k * n + i) + 1]); txp[k] = txc + k * n; }
    int np = ev->nparams;
    cplx *q0 = cvec(np + 1), *q1 = cvec(np + 1);
    for (int i = 0; i < np; i++) { if (p0) q0[i] = C(p0[2 * i], p0[2 * i + 1]); if (p1) q1[i] = C(p1[2 * i], p1[2 * i + 1]); }
    int m = ev->output_dim;
    cplx* o = cvec((K + 1) * m);
    sys_taylor(ev, K, txp, q0, p1 ? q1 : NULL, T, o);
    for (int i = 0; i < (K + 1) * m; i++) { out[2 * i] = o[i].re; out[2 * i + 1] = o[i].im; }
    free(T); free(txc); free(q0); free(q1); free(o);
    return 0;
}
int orc_sys_eval_dd(const orc_tape* ev, const double* xhi, const double* xlo, const double* p, int n, double* uhi, double* ulo) {
    if (orc_tape_check(ev)) return -1;
    cdd* T = calloc(ev->tape_space, sizeof(cdd));
    cdd* xx = calloc(n, sizeof(cdd));
    cdd* uu = calloc(ev->output_dim, sizeof(cdd));
    cplx* pp = cvec(ev->nparams + 1);
    for (int i = 0; i < n; i++) xx[i] = CDD(DD(xhi[2 * i], xlo[2 * i]), DD(xhi[2 * i + 1], xlo[2 * i + 1]));
    for (int i = 0; i < ev->nparams; i++) pp[i] = C(p[2 * i], p[2 * i + 1]);
    sys_eval_dd(ev, xx, pp, T, uu);
    for (int i = 0; i < ev->output_dim; i++) {
        uhi[2 * i] = uu[i].re.hi; uhi[2 * i + 1] = uu[i].im.hi; ulo[2 * i] = uu[i].re.lo; ulo[2 * i + 1] = uu[i].im.lo;
    }
    free(T); free(xx); free(uu); free(pp);
    return 0;
}
/* DD scalar ops for pinning: op 0 add, 1 sub, 2 mul, 3 div */
void orc_dd_op(int op, double ah, double al, double bh, double bl, double* rh, double* rl) {
    dd a = DD(ah, al), b = DD(bh, bl), r;
    switch (op) {
    case 0: r = dd_add(a, b); break;
    case 1: r = dd_sub(a, b); break;
    case 2: r = dd_mul(a, b); break;
    default: r = dd_div(a, b);
    }
    *rh = r.hi; *rl = r.lo;
}
/* Solve A x = b with the reference's LU (optionally Skeel-scaled by weights) and
 * return the Higham cond estimate of diag(dl) A diag(dr). A column-major complex. */
int orc_linsolve(int n, const double* A, const double* b, const double* weights, double* x, double* cond_out,
                 const double* dl, const double* dr) {
    orc_ws W;
    ws_init(&W, n);
    for (int i = 0; i < n * n; i++) W.A[i] = C(A[2 * i], A[2 * i + 1]);
    cplx *bb = cvec(n), *xx = cvec(n);
    for (int i = 0; i < n; i++) bb[i] = C(b[2 * i], b[2 * i + 1]);
    ws_updated(&W);
    if (weights) jac_ldiv_w(&W, xx, bb, weights); else ws_ldiv(&W, xx, bb);
    for (int i = 0; i < n; i++) { x[2 * i] = xx[i].re; x[2 * i + 1] = xx[i].im; }
    if (cond_out) { ws_updated(&W); *cond_out = ws_cond(&W, dl, dr); }
    free(bb); free(xx);
    ws_free(&W);
    return 0;
}
int orc_rank(int n, const double* A, double rtol) {
    cplx* M = cvec(n * n);
    for (int i = 0; i < n * n; i++) M[i] = C(A[2 * i], A[2 * i + 1]);
    int r = rank_rtol(M, n, rtol);
    free(M);
    return r;
}
/* weighted inf-norm, norm.jl; w may be NULL */
double orc_norm(int n, const double* x, const double* w) {
    cplx* xx = cvec(n);
    for (int i = 0; i < n; i++) xx[i] = C(x[2 * i], x[2 * i + 1]);
    double r = norm_inf(xx, w, n);
    free(xx);
    return r;
}
/* SegmentStepper script for pinning (src/utils.jl:300-401): ops[k] < 0 is step_success!,
 * otherwise propose_step!(ops[k]).  After each op 12 doubles are written:
 * s, s', Δs, t, t', Δt (complex as re, im pairs), |Δ|, forward, is_done. */
int orc_segment_run(double sre, double sim, double tre, double tim, int nops, const double* ops, double* out) {
    orc_tracker* T = calloc(1, sizeof(orc_tracker));
    if (!T) return -1;
    seg_init(T, C(sre, sim), C(tre, tim));
    for (int k = 0; k < nops; k++) {
        if (ops[k] < 0) seg_step_success(T); else seg_propose(T, ops[k]);
        double* r = out + 14 * k;
        cplx t = seg_t(T), t1 = seg_t1(T), dt = seg_dt(T);
        r[0] = T->seg_s; r[1] = T->seg_s1; r[2] = seg_ds(T);
        r[3] = t.re; r[4] = t.im; r[5] = t1.re; r[6] = t1.im; r[7] = dt.re; r[8] = dt.im;
        r[9] = T->seg_absD; r[10] = T->seg_forward; r[11] = seg_is_done(T);
        r[12] = seg_dist_to_target(T); r[13] = 0;
    }
    free(T);
    return 0;
}
double orc_nthroot(double x, int N) { return nthroot(x, N); }
double orc_eps(double x) { return jeps(x); }
/* Valuation of a synthetic path: given Taylor data (x, x¹, x², x³) at t, returns val_x, val_tẋ (non-diff branch) */
void orc_valuation(int n, const double* tx, double t, double* val_x, double* val_tx) {
    for (int i = 0; i < n; i++) {
        cplx x = C(tx[2 * i], tx[2 * i + 1]), x1 = C(tx[2 * (n + i)], tx[2 * (n + i) + 1]);
        cplx x2 = C(tx[2 * (2 * n + i)], tx[2 * (2 * n + i) + 1]), x3 = C(tx[2 * (3 * n + i)], tx[2 * (3 * n + i) + 1]);
        double v, v1, w, w1;
        nu_nu1(x, x1, cscale(2, x2), t, &v, &v1);
        nu_nu1(x1, cscale(2, x2), cscale(6, x3), t, &w, &w1);
        val_x[i] = v; val_tx[i] = w + 1;
    }
}
size_t orc_sizeof(int which) {
    switch (which) {
    case 0: return sizeof(orc_tape);
    case 1: return sizeof(orc_system);
    case 2: return sizeof(orc_homotopy_desc);
    case 3: return sizeof(orc_options);
    case 4: return sizeof(orc_batch);
    case 5: return sizeof(orc_path_result);
    default: return 0;
    }
}
/* deterministic transcendental kernels, for pinning: 0 log, 1 exp, 2 pow, 3 cbrt */
double orc_math(int which, double x, double y) {
    switch (which) {
    case 0: return hc_log(x);
    case 1: return hc_exp(x);
    case 2: return hc_pow(x, y);
    default: return hc_cbrt(x);
    }
}
/* Track x₁ from t₁ to t₀ with the plain tracker, then update a fresh Valuation at the final t
 * (the reference's valuation tests, test/valuation_test.jl). */
int orc_track_plain_valuation(const orc_homotopy_desc* hd, const orc_options* o, const double* x1, double t1, double t0,
                              double* val_x, double* val_tx) {
    if (check_desc(hd)) return -1;
    orc_tracker T;
    tracker_create(&T, hd, o);
    int n = hd->n;
    cplx* x = cvec(n);
    for (int i = 0; i < n; i++) x[i] = C(x1[2 * i], x1[2 * i + 1]);
    tracker_init(&T, x, C(t1, 0), C(t0, 0), NAN, NAN, 0);
    while (T.code == TC_TRACKING) tracker_step(&T);
    val_init(&T);
    val_update(&T, seg_t(&T).re);
    for (int i = 0; i < n; i++) { val_x[i] = T.val_x[i]; val_tx[i] = T.val_tx[i]; }
    int code = T.code;
    free(x);
    tracker_destroy(&T);
    return code;
}

/* ============================ certification ============================
 * Krawczyk certification of approximate zeros (src/certification.jl:1029-1114
 * certify_solution, :1225-1287 ε_inflation_krawczyk) in the FP64 interval arithmetic of
 * src/interval_arithmetic.jl: every bound is the round-to-nearest result moved one ulp
 * outward (round_down = prevfloat, round_up = nextfloat, :57-61), products and quotients
 * by the case analysis of :118-234.  Interval muladd uses FMA (one rounding per bound).
 * Pinned by exact rational enclosure checks of every interval operation, the reference's
 * certification known answers (test/certification_test.jl) and the 3264 real conics. */
typedef struct { double lo, hi; } ival;
typedef struct { ival re, im; } icx;

static inline double r_up(double x) { return nextafter(x, INFINITY); }
static inline double r_dn(double x) { return nextafter(x, -INFINITY); }
static inline ival IV(double lo, double hi) { ival r = {lo, hi}; return r; }
static inline ival iround(double lo, double hi) { return IV(r_dn(lo), r_up(hi)); }
static inline int i_iszero(ival a) { return a.lo == 0 && a.hi == 0; }
static inline int i_in(double x, ival a) { return a.lo <= x && x <= a.hi; }
static inline double jmin_(double a, double b) { if (isnan(a) || isnan(b)) return NAN; return a < b ? a : b; }
static inline double i_mag(ival a) { return jmax(fabs(a.lo), fabs(a.hi)); }
static inline double i_mig(ival a) { return i_in(0.0, a) ? 0.0 : jmin_(fabs(a.lo), fabs(a.hi)); }
static inline double min_fast_(double x, double y) { return y < x ? y : x; }
static inline int i_interior(ival a, ival b) { return a.lo > b.lo && a.hi < b.hi; }

static ival i_add(ival a, ival b) { return iround(a.lo + b.lo, a.hi + b.hi); }
static ival i_add_r(ival a, double b) { return iround(a.lo + b, a.hi + b); }
static ival i_sub(ival a, ival b) { return iround(a.lo - b.hi, a.hi - b.lo); }
static ival i_sub_r(ival a, double b) { return iround(a.lo - b, a.hi - b); }
static ival i_neg(ival a) { return IV(-a.hi, -a.lo); }
static ival i_mul_r(double x, ival a) {
    if (i_iszero(a) || x == 0) return IV(0, 0);
    if (x >= 0.0) return iround(a.lo * x, a.hi * x);
    return iround(a.hi * x, a.lo * x);
}
static ival i_mul(ival a, ival b) {
    if (b.lo >= 0) {
        if (a.lo >= 0) return iround(a.lo * b.lo, a.hi * b.hi);
        if (a.hi <= 0) return iround(a.lo * b.hi, a.hi * b.lo);
        return iround(a.lo * b.hi, a.hi * b.hi);
    } else if (b.hi <= 0) {
        if (a.lo >= 0) return iround(a.hi * b.lo, a.lo * b.hi);
        if (a.hi <= 0) return iround(a.hi * b.hi, a.lo * b.lo);
        return iround(a.hi * b.lo, a.lo * b.lo);
    } else {
        if (a.lo > 0) return iround(a.hi * b.lo, a.hi * b.hi);
        if (a.hi < 0) return iround(a.lo * b.hi, a.lo * b.lo);
        return iround(jmin_(a.lo * b.hi, a.hi * b.lo), jmax(a.lo * b.lo, a.hi * b.hi));
    }
}
static ival i_muladd(ival a, ival b, ival c) {
    double lo = min_fast_(min_fast_(min_fast_(fma(a.lo, b.lo, c.lo), fma(a.lo, b.hi, c.lo)), fma(a.hi, b.lo, c.lo)),
                          fma(a.hi, b.hi, c.lo));
    double hi = max_fast(max_fast(max_fast(fma(a.lo, b.lo, c.hi), fma(a.lo, b.hi, c.hi)), fma(a.hi, b.lo, c.hi)),
                         fma(a.hi, b.hi, c.hi));
    return IV(r_dn(lo), r_up(hi));
}
static ival i_muladd_r(double a, ival b, ival c) {
    double lo = min_fast_(fma(a, b.lo, c.lo), fma(a, b.hi, c.lo));
    double hi = max_fast(fma(a, b.lo, c.hi), fma(a, b.hi, c.hi));
    return IV(r_dn(lo), r_up(hi));
}
static ival i_div(ival a, ival b) {
    if (i_iszero(a)) return IV(0, 0);
    if (b.lo > 0) {
        if (a.lo >= 0) return iround(a.lo / b.hi, a.hi / b.lo);
        if (a.hi <= 0) return iround(a.lo / b.lo, a.hi / b.hi);
        return iround(a.lo / b.lo, a.hi / b.lo);
    } else if (b.hi < 0) {
        if (a.lo >= 0) return iround(a.hi / b.hi, a.lo / b.lo);
        if (a.hi <= 0) return iround(a.hi / b.lo, a.lo / b.hi);
        return iround(a.hi / b.hi, a.lo / b.hi);
    }
    return IV(NAN, NAN);
}
static ival i_inv(ival a) {
    if (i_in(0.0, a)) return IV(NAN, NAN);
    return iround(1.0 / a.hi, 1.0 / a.lo);
}
static ival i_sqr(ival a) {
    if (a.lo >= 0) return iround(a.lo * a.lo, a.hi * a.hi);
    if (a.hi <= 0) return iround(a.hi * a.hi, a.lo * a.lo);
    double g = i_mig(a), m = i_mag(a);
    return iround(g * g, m * m);
}

static inline icx ICX(ival re, ival im) { icx z = {re, im}; return z; }
static inline icx ic_point(cplx z) { return ICX(IV(z.re, z.re), IV(z.im, z.im)); }
static inline icx ic_zero_(void) { return ICX(IV(0, 0), IV(0, 0)); }
static inline double icx_mag(icx z) { return jmax(i_mag(z.re), i_mag(z.im)); }
static icx icx_add(icx a, icx b) { return ICX(i_add(a.re, b.re), i_add(a.im, b.im)); }
static icx icx_sub(icx a, icx b) { return ICX(i_sub(a.re, b.re), i_sub(a.im, b.im)); }
static icx icx_neg(icx a) { return ICX(i_neg(a.re), i_neg(a.im)); }
static icx icx_mul(icx z, icx w) {
    return ICX(i_sub(i_mul(z.re, w.re), i_mul(z.im, w.im)), i_add(i_mul(z.re, w.im), i_mul(z.im, w.re)));
}
static icx icx_inv(icx b) {
    ival d = i_add(i_sqr(b.re), i_sqr(b.im));
    return ICX(i_div(b.re, d), i_div(i_neg(b.im), d));
}
static icx icx_div(icx a, icx b) {
    ival d = i_add(i_sqr(b.re), i_sqr(b.im));
    return ICX(i_div(i_add(i_mul(a.re, b.re), i_mul(a.im, b.im)), d), i_div(i_sub(i_mul(a.im, b.re), i_mul(a.re, b.im)), d));
}
/* muladd(z::Complex, w::IComplex, x::IComplex) and muladd on IComplex, :372-375 */
static icx icx_muladd_pt(cplx z, icx w, icx x) {
    return ICX(i_sub(i_muladd_r(z.re, w.re, x.re), i_mul_r(z.im, w.im)), i_muladd_r(z.re, w.im, i_muladd_r(z.im, w.re, x.im)));
}
static icx icx_muladd(icx z, icx w, icx x) {
    return ICX(i_sub(i_muladd(z.re, w.re, x.re), i_mul(z.im, w.im)), i_muladd(z.re, w.im, i_muladd(z.im, w.re, x.im)));
}
/* Base.power_by_squaring on IComplex */
static icx icx_pbs(icx x, long p) {
    if (p == 1) return x;
    if (p == 0) return ICX(IV(1, 1), IV(0, 0));
    if (p == 2) return icx_mul(x, x);
    int t = __builtin_ctzl((unsigned long)p) + 1;
    p >>= t;
    while (--t > 0) x = icx_mul(x, x);
    icx y = x;
    while (p > 0) {
        t = __builtin_ctzl((unsigned long)p) + 1;
        p >>= t;
        while (--t >= 0) x = icx_mul(x, x);
        y = icx_mul(y, x);
    }
    return y;
}

/* The generic operations (src/model_kit/operations.jl:183-248) on IComplex: one interval
 * execution of the tape. */
static void tape_run_ival(const orc_tape* tp, icx* T) {
    for (int i = 0; i < tp->ninstr; i++) {
        const int32_t* I = tp->instr + 6 * i;
        int op = I[4];
        if (op == OP_STOP) break;
        icx a = T[I[0] - 1], b, c, d, r;
        switch (op) {
        case OP_CB: r = icx_mul(icx_mul(a, a), a); break;
        case OP_INV: r = icx_inv(a); break;
        case OP_INV_NOT_ZERO: r = (i_iszero(a.re) && i_iszero(a.im)) ? a : icx_inv(a); break;
        case OP_INVSQR: b = icx_inv(a); r = icx_mul(b, b); break;
        case OP_NEG: r = icx_neg(a); break;
        case OP_SQR: r = icx_mul(a, a); break;
        case OP_IDENTITY: r = a; break;
        case OP_ADD: r = icx_add(a, T[I[1] - 1]); break;
        case OP_SUB: r = icx_sub(a, T[I[1] - 1]); break;
        case OP_MUL: r = icx_mul(a, T[I[1] - 1]); break;
        case OP_DIV: r = icx_div(a, T[I[1] - 1]); break;
        case OP_POW_INT: r = I[1] > 0 ? icx_pbs(a, I[1]) : icx_inv(icx_pbs(a, -(long)I[1])); break;
        case OP_ADD3: b = T[I[1] - 1]; c = T[I[2] - 1]; r = icx_add(icx_add(a, b), c); break;
        case OP_MUL3: b = T[I[1] - 1]; c = T[I[2] - 1]; r = icx_mul(icx_mul(a, b), c); break;
        case OP_MULADD: b = T[I[1] - 1]; c = T[I[2] - 1]; r = icx_add(icx_mul(a, b), c); break;
        case OP_MULSUB: b = T[I[1] - 1]; c = T[I[2] - 1]; r = icx_sub(icx_mul(a, b), c); break;
        case OP_SUBMUL: b = T[I[1] - 1]; c = T[I[2] - 1]; r = icx_sub(c, icx_mul(a, b)); break;
        case OP_ADD4: b = T[I[1] - 1]; c = T[I[2] - 1]; d = T[I[3] - 1]; r = icx_add(icx_add(icx_add(a, b), c), d); break;
        case OP_MUL4: b = T[I[1] - 1]; c = T[I[2] - 1]; d = T[I[3] - 1]; r = icx_mul(icx_mul(icx_mul(a, b), c), d); break;
        case OP_MULMULADD: b = T[I[1] - 1]; c = T[I[2] - 1]; d = T[I[3] - 1]; r = icx_add(icx_mul(a, b), icx_mul(c, d)); break;
        case OP_MULMULSUB: b = T[I[1] - 1]; c = T[I[2] - 1]; d = T[I[3] - 1]; r = icx_sub(icx_mul(a, b), icx_mul(c, d)); break;
        default: r = ICX(IV(NAN, NAN), IV(NAN, NAN));
        }
        T[I[5] - 1] = r;
    }
}
/* F (jac = 0: u) or ∂F/∂x (jac = 1: U, column-major) on the box X; parameters from the
 * packed box IP (re lo, re hi, im lo, im hi) or, for a fixed-parameter system, points. */
static void sys_eval_ival(const orc_tape* tp, const double* fixed, const icx* X, const double* IP, int n, int jac,
                          icx* T, icx* out) {
    for (int k = 0; k < tp->nconst; k++) T[k] = ic_point(C(tp->constants[2 * k], tp->constants[2 * k + 1]));
    for (int k = 0; k < tp->nparams; k++)
        T[tp->params_first - 1 + k] = fixed ? ic_point(C(fixed[2 * k], fixed[2 * k + 1]))
                                            : ICX(IV(IP[4 * k], IP[4 * k + 1]), IV(IP[4 * k + 2], IP[4 * k + 3]));
    for (int k = 0; k < tp->nvars; k++) T[tp->vars_first - 1 + k] = X[k];
    tape_run_ival(tp, T);
    int m = jac ? n * n : n, na = jac ? tp->nU : tp->nu;
    const int32_t* as = jac ? tp->U_assign : tp->u_assign;
    for (int i = 0; i < m; i++) out[i] = ic_zero_();
    for (int i = 0; i < na; i++) out[as[2 * i] - 1] = T[as[2 * i + 1] - 1];
}

/* Double-double balls for the extended retry: (m, r) holds every v with |v - m| <= r.
 * Midpoint errors: the accurate double-double sum errs by less than 3u²/(1-4u) of the exact
 * sum, the product (dd_mul, DWTimesDW1) by less than 7u² of the exact product (Joldes,
 * Muller, Popescu, ACM TOMS 44(2), 2017); the radius adds 8u² and 16u² of the magnitudes
 * plus 1e-300, rounded upward. */
typedef struct { dd m; double r; } dball;
typedef struct { dball re, im; } cball;
#define BALL_U2 1.2325951644078310e-32
static inline double dd_mag_(dd a) { return r_up(fabs(a.hi) + fabs(a.lo)); }
static dd dd_add_acc(dd x, dd y) {
    dd s = two_sum(x.hi, y.hi), t = two_sum(x.lo, y.lo);
    double c = s.lo + t.hi;
    dd v = quick_two_sum(s.hi, c);
    double w = t.lo + v.lo;
    return quick_two_sum(v.hi, w);
}
static inline dball DB(dd m, double r) { dball b = {m, r}; return b; }
static dball db_add(dball a, dball b) {
    dd m = dd_add_acc(a.m, b.m);
    double e = r_up(r_up(8.0 * BALL_U2 * dd_mag_(m)) + 1e-300);
    return DB(m, r_up(r_up(a.r + b.r) + e));
}
static dball db_neg(dball a) { return DB(DD(-a.m.hi, -a.m.lo), a.r); }
static dball db_sub(dball a, dball b) { return db_add(a, db_neg(b)); }
static dball db_mul(dball a, dball b) {
    dd m = dd_mul(a.m, b.m);
    double ma = dd_mag_(a.m), mb = dd_mag_(b.m);
    double e = r_up(r_up(16.0 * BALL_U2 * r_up(ma * mb)) + 1e-300);
    double r = r_up(r_up(r_up(ma * b.r) + r_up(mb * a.r)) + r_up(a.r * b.r));
    return DB(m, r_up(r + e));
}
static ival db_ival(dball a) {
    double t = a.m.hi + a.m.lo;
    return IV(r_dn(r_dn(t) - a.r), r_up(r_up(t) + a.r));
}
static inline cball CB(dball re, dball im) { cball z = {re, im}; return z; }
static cball cb_point(cplx z) { return CB(DB(DD(z.re, 0), 0), DB(DD(z.im, 0), 0)); }
static cball cb_nan_(void) { return CB(DB(DD(NAN, 0), INFINITY), DB(DD(NAN, 0), INFINITY)); }
static cball cb_add_(cball a, cball b) { return CB(db_add(a.re, b.re), db_add(a.im, b.im)); }
static cball cb_sub_(cball a, cball b) { return CB(db_sub(a.re, b.re), db_sub(a.im, b.im)); }
static cball cb_mul_(cball a, cball b) {
    return CB(db_sub(db_mul(a.re, b.re), db_mul(a.im, b.im)), db_add(db_mul(a.re, b.im), db_mul(a.im, b.re)));
}
static cball cb_pow_(cball x, long p) {
    if (p < 1) return cb_nan_();
    if (p == 1) return x;
    cball y = x;
    int have = 0;
    while (p > 0) {
        if (p & 1) { y = have ? cb_mul_(y, x) : x; have = 1; }
        p >>= 1;
        if (p) x = cb_mul_(x, x);
    }
    return y;
}
/* F at the double-double point x in balls; parameters from B (re hi, re lo, re rad, im hi, im lo, im rad)
 * or fixed points.  Quotients have no ball form: their results are unbounded. */
static void sys_eval_ball(const orc_tape* tp, const double* fixed, const cdd* x, const double* B, int n, int jac,
                          cball* T, cball* out) {
    for (int k = 0; k < tp->nconst; k++) T[k] = cb_point(C(tp->constants[2 * k], tp->constants[2 * k + 1]));
    for (int k = 0; k < tp->nparams; k++)
        T[tp->params_first - 1 + k] = fixed ? cb_point(C(fixed[2 * k], fixed[2 * k + 1]))
            : CB(DB(DD(B[6 * k], B[6 * k + 1]), B[6 * k + 2]), DB(DD(B[6 * k + 3], B[6 * k + 4]), B[6 * k + 5]));
    for (int k = 0; k < tp->nvars; k++) T[tp->vars_first - 1 + k] = CB(DB(x[k].re, 0), DB(x[k].im, 0));
    for (int i = 0; i < tp->ninstr; i++) {
        const int32_t* I = tp->instr + 6 * i;
        int op = I[4];
        if (op == OP_STOP) break;
        cball a = T[I[0] - 1], b, c, d, r;
        switch (op) {
        case OP_CB: r = cb_mul_(cb_mul_(a, a), a); break;
        case OP_NEG: r = CB(db_neg(a.re), db_neg(a.im)); break;
        case OP_SQR: r = cb_mul_(a, a); break;
        case OP_IDENTITY: r = a; break;
        case OP_ADD: r = cb_add_(a, T[I[1] - 1]); break;
        case OP_SUB: r = cb_sub_(a, T[I[1] - 1]); break;
        case OP_MUL: r = cb_mul_(a, T[I[1] - 1]); break;
        case OP_POW_INT: r = cb_pow_(a, I[1]); break;
        case OP_ADD3: b = T[I[1] - 1]; c = T[I[2] - 1]; r = cb_add_(cb_add_(a, b), c); break;
        case OP_MUL3: b = T[I[1] - 1]; c = T[I[2] - 1]; r = cb_mul_(cb_mul_(a, b), c); break;
        case OP_MULADD: b = T[I[1] - 1]; c = T[I[2] - 1]; r = cb_add_(cb_mul_(a, b), c); break;
        case OP_MULSUB: b = T[I[1] - 1]; c = T[I[2] - 1]; r = cb_sub_(cb_mul_(a, b), c); break;
        case OP_SUBMUL: b = T[I[1] - 1]; c = T[I[2] - 1]; r = cb_sub_(c, cb_mul_(a, b)); break;
        case OP_ADD4: b = T[I[1] - 1]; c = T[I[2] - 1]; d = T[I[3] - 1]; r = cb_add_(cb_add_(cb_add_(a, b), c), d); break;
        case OP_MUL4: b = T[I[1] - 1]; c = T[I[2] - 1]; d = T[I[3] - 1]; r = cb_mul_(cb_mul_(cb_mul_(a, b), c), d); break;
        case OP_MULMULADD: b = T[I[1] - 1]; c = T[I[2] - 1]; d = T[I[3] - 1]; r = cb_add_(cb_mul_(a, b), cb_mul_(c, d)); break;
        case OP_MULMULSUB: b = T[I[1] - 1]; c = T[I[2] - 1]; d = T[I[3] - 1]; r = cb_sub_(cb_mul_(a, b), cb_mul_(c, d)); break;
        default: r = cb_nan_();  /* inv, div: no ball form */
        }
        T[I[5] - 1] = r;
    }
    int m = jac ? n * n : n, na = jac ? tp->nU : tp->nu;
    const int32_t* as = jac ? tp->U_assign : tp->u_assign;
    for (int i = 0; i < m; i++) out[i] = cb_point(C(0, 0));
    for (int i = 0; i < na; i++) out[as[2 * i] - 1] = T[as[2 * i + 1] - 1];
}

/* Double-double LU with partial pivoting (pivots on the rounded entries), its solve, and
 * the inverse of the Jacobian at a double-double point rounded to FP64 (the role of
 * approx_inv! on the Arb Jacobian, src/certification.jl:1384-1391). */
static void lu_factor_dd(cdd* A, int* ipiv, int n) {
    for (int k = 0; k < n; k++) {
        int kp = k;
        double amax = cabs2(cdd_round(A[k + k * n]));
        for (int i = k + 1; i < n; i++) {
            double absi = cabs2(cdd_round(A[i + k * n]));
            if (absi > amax) { kp = i; amax = absi; }
        }
        ipiv[k] = kp;
        if (amax != 0) {
            if (k != kp)
                for (int j = 0; j < n; j++) { cdd t = A[k + j * n]; A[k + j * n] = A[kp + j * n]; A[kp + j * n] = t; }
            cdd akk = A[k + k * n];
            for (int i = k + 1; i < n; i++) A[i + k * n] = cdd_div(A[i + k * n], akk);
        }
        for (int j = k + 1; j < n; j++) {
            cdd akj = A[k + j * n];
            for (int i = k + 1; i < n; i++) A[i + j * n] = cdd_sub(A[i + j * n], cdd_mul(A[i + k * n], akj));
        }
    }
}
static void lu_solve_dd(const cdd* A, const int* ipiv, int n, cdd* x) {
    for (int i = 0; i < n; i++) if (i != ipiv[i]) { cdd t = x[i]; x[i] = x[ipiv[i]]; x[ipiv[i]] = t; }
    for (int j = 0; j < n; j++) {
        cdd xj = x[j];
        for (int i = j + 1; i < n; i++) x[i] = cdd_sub(x[i], cdd_mul(A[i + j * n], xj));
    }
    for (int j = n - 1; j >= 0; j--) {
        cdd xj = cdd_div(x[j], A[j + j * n]);
        x[j] = xj;
        for (int i = 0; i < j; i++) x[i] = cdd_sub(x[i], cdd_mul(A[i + j * n], xj));
    }
}
static void dd_inverse(const orc_system* F, int n, const cdd* xd, const double* BP, cball* TB, cplx* Cm) {
    cball* JB = malloc(sizeof(cball) * n * n);
    cdd* A = malloc(sizeof(cdd) * n * n);
    cdd* e = malloc(sizeof(cdd) * n);
    int* ipiv = malloc(sizeof(int) * n);
    sys_eval_ball(F->eval_jac, F->fixed_params, xd, BP, n, 1, TB, JB);
    for (int i = 0; i < n * n; i++) { A[i].re = JB[i].re.m; A[i].im = JB[i].im.m; }
    lu_factor_dd(A, ipiv, n);
    for (int j = 0; j < n; j++) {
        for (int i = 0; i < n; i++) { e[i].re = DD(i == j ? 1.0 : 0.0, 0.0); e[i].im = DD(0.0, 0.0); }
        lu_solve_dd(A, ipiv, n, e);
        for (int i = 0; i < n; i++) Cm[i + j * n] = cdd_round(e[i]);
    }
    free(JB); free(A); free(e); free(ipiv);
}

typedef struct orc_certificate {
    int32_t certified, real, complex_, newton_code, newton_iters, prec;
    double contraction, inflation;
} orc_certificate;

typedef struct { int certified, conj_inside, not_real; double bound, inflation; } kr_out;

/* Δx₀ = C r₀, x₀ = x̃ + [-ε, ε](1 + i), M = C J(x₀) - I, x₁ = (x̃ - Δx₀) - M (x₀ - x̃)
 * (ε_inflation_krawczyk, src/certification.jl:1225-1287); writes x₁ and x₀.  x̃ is a
 * double-double vector (low parts zero in the FP64 test, ext = 0); the extended test
 * encloses x̃ and x₀ - x̃ through balls and floors ε at 512 eps (1 + |x̃|). */
static kr_out krawczyk(const orc_system* F, int n, const cdd* xd, const cplx* Cm, const icx* R0, const double* IP,
                       int ext, icx* TI, double* box1, double* box0) {
    icx *D = malloc(sizeof(icx) * n), *X0 = malloc(sizeof(icx) * n), *dX = malloc(sizeof(icx) * n);
    icx *XT = malloc(sizeof(icx) * n), *JI = malloc(sizeof(icx) * n * n);
    kr_out o = {0, 1, 0, 0.0, 0.0};
    for (int i = 0; i < n; i++) {
        icx dd_ = ic_zero_();
        for (int k = 0; k < n; k++) dd_ = icx_muladd_pt(Cm[i + k * n], R0[k], dd_);
        D[i] = dd_;
        double hr = xd[i].re.hi, hi_ = xd[i].im.hi;
        if (!ext) {
            double eps_i = 512.0 * jmax(icx_mag(dd_), 2.220446049250313e-16);
            o.inflation = jmax(o.inflation, eps_i);
            X0[i] = ICX(IV(hr - eps_i, hr + eps_i), IV(hi_ - eps_i, hi_ + eps_i));
            dX[i] = ICX(i_sub_r(X0[i].re, hr), i_sub_r(X0[i].im, hi_));
            XT[i] = ICX(IV(hr, hr), IV(hi_, hi_));
        } else {
            double eps_i = 512.0 * jmax(icx_mag(dd_), 2.220446049250313e-16 * (1.0 + jmax(fabs(hr), fabs(hi_))));
            o.inflation = jmax(o.inflation, eps_i);
            X0[i] = ICX(IV(r_dn(r_dn(hr) - eps_i), r_up(r_up(hr) + eps_i)), IV(r_dn(r_dn(hi_) - eps_i), r_up(r_up(hi_) + eps_i)));
            dball xr = DB(xd[i].re, 0), xi = DB(xd[i].im, 0);
            dX[i] = ICX(IV(db_ival(db_sub(DB(DD(X0[i].re.lo, 0), 0), xr)).lo, db_ival(db_sub(DB(DD(X0[i].re.hi, 0), 0), xr)).hi),
                        IV(db_ival(db_sub(DB(DD(X0[i].im.lo, 0), 0), xi)).lo, db_ival(db_sub(DB(DD(X0[i].im.hi, 0), 0), xi)).hi));
            XT[i] = ICX(db_ival(xr), db_ival(xi));
        }
    }
    sys_eval_ival(F->eval_jac, F->fixed_params, X0, IP, n, 1, TI, JI);
    double bound = 0.0;
    int inside = 1;
    for (int i = 0; i < n; i++) {
        ival b = IV(0, 0);
        icx x1 = ic_zero_();
        for (int j = 0; j < n; j++) {
            icx m = ic_zero_();
            for (int k = 0; k < n; k++) m = icx_muladd_pt(Cm[i + k * n], JI[k + j * n], m);
            if (i == j) m.re = i_sub_r(m.re, 1.0);
            b = i_add_r(b, icx_mag(m));
            x1 = icx_muladd(m, dX[j], x1);
        }
        bound = jmax(bound, i_mag(b));
        x1 = icx_sub(icx_sub(XT[i], D[i]), x1);
        inside = inside && i_interior(x1.re, X0[i].re) && i_interior(x1.im, X0[i].im);
        o.conj_inside = o.conj_inside && i_interior(x1.re, X0[i].re) && i_interior(i_neg(x1.im), X0[i].im);
        o.not_real = o.not_real || !i_in(0.0, x1.im);
        double* o1 = box1 + 4 * i;
        o1[0] = x1.re.lo; o1[1] = x1.re.hi; o1[2] = x1.im.lo; o1[3] = x1.im.hi;
        double* o0 = box0 + 4 * i;
        o0[0] = X0[i].re.lo; o0[1] = X0[i].re.hi; o0[2] = X0[i].im.lo; o0[3] = X0[i].im.hi;
    }
    o.bound = 1.41422 * bound;
    o.certified = o.bound < 0.7071 && inside;
    free(D); free(X0); free(dX); free(XT); free(JI);
    return o;
}

/* x̃ ← x̃ - C mid(F(x̃)) in double-double with ball residuals (at most 10 steps, until
 * ‖C F(x̃)‖ stops decreasing; src/certification.jl:1303-1320); R0 encloses F at the final x̃. */
static void refine_dd(const orc_system* F, int n, cdd* xd, const cplx* Cm, const double* BP, cball* TB, icx* R0) {
    cball* rb = malloc(sizeof(cball) * n);
    cplx* dl = cvec(n);
    double acc = INFINITY;
    for (int it = 1; it <= 10; it++) {
        sys_eval_ball(F->eval, F->fixed_params, xd, BP, n, 0, TB, rb);
        for (int i = 0; i < n; i++) R0[i] = ICX(db_ival(rb[i].re), db_ival(rb[i].im));
        double nrm = 0.0;
        for (int i = 0; i < n; i++) {
            icx d = ic_zero_();
            for (int k = 0; k < n; k++) d = icx_muladd_pt(Cm[i + k * n], R0[k], d);
            nrm = jmax(nrm, icx_mag(d));
            dl[i] = C(0.5 * d.re.lo + 0.5 * d.re.hi, 0.5 * d.im.lo + 0.5 * d.im.hi);
        }
        if (!(nrm <= acc) || it == 10) break;
        acc = nrm;
        for (int i = 0; i < n; i++) {
            xd[i].re = dd_add_acc(xd[i].re, DD(-dl[i].re, 0));
            xd[i].im = dd_add_acc(xd[i].im, DD(-dl[i].im, 0));
        }
    }
    free(rb); free(dl);
}

static double cert_inf_norm(const cplx* x, int n) {
    double m = cabs2(x[0]);
    for (int i = 1; i < n; i++) m = max_fast(m, cabs2(x[i]));
    return sqrt(m);
}

/* One candidate: newton(F, x, p, InfNorm(); rtol = 8 eps, atol = 0, extended_precision =
 * true, max_iters = 8) (src/newton.jl:107-175), C = inverse of the last Jacobian, then the
 * ε-inflation Krawczyk test. */
static void certify_one(const orc_system* F, int n, const cplx* x0, const cplx* p, const double* IP, const double* BP,
                        int refine, int extended, int real_system, orc_certificate* out, double* box1, double* box0,
                        double* refined) {
    const orc_tape* ej = F->eval_jac;
    const orc_tape* ev = F->eval;
    int space = ej->tape_space > ev->tape_space ? ej->tape_space : ev->tape_space;
    cplx* T = malloc(sizeof(cplx) * (space + 1));
    cdd* Td = malloc(sizeof(cdd) * (space + 1));
    icx* TI = malloc(sizeof(icx) * (space + 1));
    cplx *x = cvec(n), *r = cvec(n), *LU = cvec(n * n), *Cm = cvec(n * n), *e = cvec(n);
    cdd *xe = malloc(sizeof(cdd) * n), *re = malloc(sizeof(cdd) * n);
    icx *X = malloc(sizeof(icx) * n), *R0 = malloc(sizeof(icx) * n);
    int* ipiv = malloc(sizeof(int) * n);
    memcpy(x, x0, sizeof(cplx) * n);
    int code = -1, iters = 0;
    if (refine) {
        double norm_x = cert_inf_norm(x, n), tol = 8.0 * 2.220446049250313e-16 * norm_x;
        double prev = NAN, ac = 1.0;
        code = 2;
        iters = 8;
        for (int it = 1; it <= 8; it++) {
            sys_eval(ej, x, p, T, r, LU, n);
            for (int i = 0; i < n; i++) xe[i] = cdd_from(x[i]);
            sys_eval_dd(ev, xe, p, Td, re);
            for (int i = 0; i < n; i++) r[i] = C(re[i].re.hi, re[i].im.hi);
            lu_factor(LU, ipiv, n);
            lu_solve(LU, ipiv, n, r);
            double nrm = cert_inf_norm(r, n);
            for (int i = 0; i < n; i++) x[i] = csub(x[i], r[i]);
            if (it > 1) {
                if (nrm < ac * prev) { prev = nrm; ac *= ac; }
                else { code = nrm < jmax(0.0, tol) ? 0 : 1; iters = it; break; }
            } else {
                prev = nrm;
            }
        }
    } else {
        sys_eval(ej, x, p, T, r, LU, n);
        lu_factor(LU, ipiv, n);
    }
    for (int j = 0; j < n; j++) {
        for (int i = 0; i < n; i++) e[i] = C(i == j ? 1.0 : 0.0, 0.0);
        lu_solve(LU, ipiv, n, e);
        for (int i = 0; i < n; i++) Cm[i + j * n] = e[i];
    }
    cdd* xd = malloc(sizeof(cdd) * n);
    for (int i = 0; i < n; i++) { X[i] = ic_point(x[i]); xd[i] = cdd_from(x[i]); }
    sys_eval_ival(ev, F->fixed_params, X, IP, n, 0, TI, R0);
    kr_out o = krawczyk(F, n, xd, Cm, R0, IP, 0, TI, box1, box0);
    int prec = 53;
    if (!o.certified && extended) {
        /* x̃ becomes a double-double point refined with ball residuals, then the test runs
         * again; if it still fails, C is recomputed from a double-double Jacobian and it runs
         * once more (extended_prec_certify_solution / arb_ε_inflation_krawczyk,
         * src/certification.jl:1116-1218, 1289-1394, in double-double). */
        int tspace = ej->tape_space > space ? ej->tape_space : space;
        cball* TB = malloc(sizeof(cball) * (tspace + 1));
        refine_dd(F, n, xd, Cm, BP, TB, R0);
        o = krawczyk(F, n, xd, Cm, R0, IP, 1, TI, box1, box0);
        prec = 106;
        if (!o.certified) {
            dd_inverse(F, n, xd, BP, TB, Cm);
            refine_dd(F, n, xd, Cm, BP, TB, R0);
            o = krawczyk(F, n, xd, Cm, R0, IP, 1, TI, box1, box0);
        }
        free(TB);
    }
    for (int i = 0; i < n; i++) { refined[2 * i] = xd[i].re.hi; refined[2 * i + 1] = xd[i].im.hi; }
    free(xd);
    int is_complex = o.certified && o.not_real;
    out->certified = o.certified;
    out->complex_ = is_complex;
    out->real = o.certified && o.conj_inside && real_system && !is_complex;
    out->newton_code = code;
    out->newton_iters = iters;
    out->prec = prec;
    out->contraction = o.bound;
    out->inflation = o.inflation;
    free(T); free(Td); free(TI); free(x); free(r); free(LU); free(Cm); free(e); free(xe); free(re);
    free(X); free(R0); free(ipiv);
}

typedef struct {
    const orc_system* F;
    int n, nsol, refine, extended, real_system, npp, np;
    const double *cand, *IP, *BP;
    const int32_t* group_of;
    const cplx* p;
    orc_certificate* out;
    double *box1, *box0, *refined;
    atomic_long next;
} cert_job_t;

static void* cert_worker(void* arg) {
    cert_job_t* J = arg;
    int n = J->n;
    cplx* x = cvec(n);
    for (;;) {
        long k = atomic_fetch_add(&J->next, 1);
        if (k >= J->nsol) break;
        for (int i = 0; i < n; i++) x[i] = C(J->cand[2 * (k * n + i)], J->cand[2 * (k * n + i) + 1]);
        int g = J->group_of ? J->group_of[k] : 0;
        certify_one(J->F, n, x, J->p + (size_t)g * J->npp, J->IP + (size_t)g * 4 * J->np, J->BP + (size_t)g * 6 * J->np,
                    J->refine, J->extended, J->real_system, &J->out[k], J->box1 + 4 * k * n, J->box0 + 4 * k * n,
                    J->refined + 2 * k * n);
    }
    free(x);
    return NULL;
}

/* Certify nsol candidates (the loop of _certify, src/certification.jl:779-975).  Parameter
 * groups as in hcg_certify: ngroups points (params), boxes (params_ival, NULL = the points)
 * and double-double balls (params_ball, NULL = the points); candidate k uses group
 * group_of[k] (NULL: group 0). */
int orc_certify(const orc_system* F, int n, int nsol, const double* cand, int ngroups, const double* params,
                const double* params_ival, const double* params_ball, const int32_t* group_of, int refine,
                int extended, int real_system, int nthreads, orc_certificate* out, double* box1, double* box0,
                double* refined) {
    if (!F || !F->eval || !F->eval_jac || n < 1) { snprintf(orc_err, sizeof orc_err, "bad system"); return -1; }
    if (orc_tape_check(F->eval) || orc_tape_check(F->eval_jac)) return -1;
    int np = F->fixed_params ? 0 : F->eval->nparams;
    if (np > 0 && (!params || ngroups < 1)) { snprintf(orc_err, sizeof orc_err, "params missing"); return -1; }
    if (np == 0) { ngroups = 1; group_of = NULL; }
    if (group_of)
        for (int k = 0; k < nsol; k++)
            if (group_of[k] < 0 || group_of[k] >= ngroups) { snprintf(orc_err, sizeof orc_err, "bad group"); return -1; }
    const double* fixed = F->fixed_params;
    int npp = F->eval->nparams;
    cplx* p = cvec(npp * ngroups + 1);
    double* IP = calloc(4 * (size_t)(np * ngroups + 1), sizeof(double));
    double* BP = calloc(6 * (size_t)(np * ngroups + 1), sizeof(double));
    for (int g = 0; g < ngroups; g++)
        for (int k = 0; k < npp; k++)
            p[g * npp + k] = fixed ? C(fixed[2 * k], fixed[2 * k + 1])
                                   : C(params[2 * (g * npp + k)], params[2 * (g * npp + k) + 1]);
    for (long q = 0; q < (long)np * ngroups; q++) {
        if (params_ival) memcpy(IP + 4 * q, params_ival + 4 * q, 4 * sizeof(double));
        else { IP[4 * q] = IP[4 * q + 1] = params[2 * q]; IP[4 * q + 2] = IP[4 * q + 3] = params[2 * q + 1]; }
        if (params_ball) memcpy(BP + 6 * q, params_ball + 6 * q, 6 * sizeof(double));
        else { BP[6 * q] = params[2 * q]; BP[6 * q + 3] = params[2 * q + 1]; }
    }
    if (nthreads < 1) nthreads = 1;
    cert_job_t J = {F, n, nsol, refine, extended, real_system, npp, np, cand, IP, BP, group_of, p, out, box1, box0,
                    refined, 0};
    pthread_t* th = malloc(sizeof(pthread_t) * nthreads);
    for (int i = 0; i < nthreads; i++) pthread_create(&th[i], NULL, cert_worker, &J);
    for (int i = 0; i < nthreads; i++) pthread_join(th[i], NULL);
    free(th); free(p); free(IP); free(BP);
    return 0;
}

/* One interval operation on (a, b) for the enclosure tests: 0 add, 1 sub, 2 mul, 3 div,
 * 4 sqr(a), 5 inv(a), 6 muladd(a, b, c), 7 mul by the real a.lo, 8 muladd(a.lo, b, c). */
void orc_ival_op(int op, const double* a, const double* b, const double* c, double* out) {
    ival A = IV(a[0], a[1]), B = IV(b[0], b[1]), Cc = IV(c[0], c[1]), R;
    switch (op) {
    case 0: R = i_add(A, B); break;
    case 1: R = i_sub(A, B); break;
    case 2: R = i_mul(A, B); break;
    case 3: R = i_div(A, B); break;
    case 4: R = i_sqr(A); break;
    case 5: R = i_inv(A); break;
    case 6: R = i_muladd(A, B, Cc); break;
    case 7: R = i_mul_r(a[0], B); break;
    case 8: R = i_muladd_r(a[0], B, Cc); break;
    default: R = IV(NAN, NAN);
    }
    out[0] = R.lo;
    out[1] = R.hi;
}

/* One ball operation for the enclosure tests: 0 add, 1 sub, 2 mul; balls as (hi, lo, radius). */
void orc_ball_op(int op, const double* a, const double* b, double* out) {
    dball A = DB(DD(a[0], a[1]), a[2]), B = DB(DD(b[0], b[1]), b[2]), R;
    R = op == 0 ? db_add(A, B) : (op == 1 ? db_sub(A, B) : db_mul(A, B));
    out[0] = R.m.hi;
    out[1] = R.m.lo;
    out[2] = R.r;
}
