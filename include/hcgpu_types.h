/*
 * hcgpu_types.h — the plain data types of the C ABI that the generated device code shares
 * with the host: tracker options, per-path records, the step-record layout, certificates and
 * the learned proposer's shape.  hipRTC compiles every generated tracker program against this
 * file (and not against hcgpu.h, whose function declarations device code does not need), so
 * changes to the entry points leave the cached code objects valid.  Included by hcgpu.h.
 */
#ifndef HCGPU_TYPES_H
#define HCGPU_TYPES_H
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

/* Options: a POD mirror of TrackerOptions + TrackerParameters (src/tracker.jl:45-140),
 * EndgameOptions (src/endgame_tracker.jl:47-72) and WeightedNormOptions
 * (src/norm.jl:36-40), plus GPU-only knobs.  hcg_default_options fills the
 * reference defaults. */
typedef struct hcg_options {
    /* TrackerOptions */
    int32_t max_steps;
    int32_t extended_precision;
    double max_step_size;
    double max_initial_step_size;
    double min_step_size;
    double min_rel_step_size;
    double terminate_cond;
    /* TrackerParameters */
    double a, beta_a, beta_omega_p, beta_tau, strict_beta_tau;
    int32_t min_newton_iters;
    /* EndgameOptions */
    int32_t max_endgame_steps;
    double endgame_start;
    int32_t max_endgame_extended_steps;
    int32_t max_winding_number;
    double min_cond, min_cond_growth, min_coord_growth;
    int32_t zero_is_at_infinity, at_infinity_check;
    int32_t only_nonsingular, refine_steps;
    double singular_min_accuracy;
    double val_finite_tol, val_at_infinity_tol;
    double sing_cond, sing_accuracy, scaling_threshold;
    /* WeightedNormOptions */
    double scale_min, scale_abs_min, scale_max;
    /* GPU-only knobs */
    int32_t learned_step;     /* 1: use the learned step-size proposer (plain rule stays the judge) */
    int32_t block_size;       /* 0: autotuned */
    int32_t blocks_per_cu;    /* 0: autotuned */
    int32_t log_steps;        /* 1: write per-step feature records into hcg_step_log */
    double learned_max_mult;  /* largest factor a learned proposal may lengthen the plain one by */
    double learned_min_t;     /* learned proposals only while |t| > learned_min_t */
    double learned_inf_t;     /* > 0: a path the network flags as heading to infinity (logit above
                                 learned_inf_logit) starts its endgame once t <= learned_inf_t instead
                                 of endgame_start; 0 = the flag is not used */
    double learned_inf_logit;
    int32_t refill_idle;         /* scheduling (no counterpart in the reference): finished lanes of a
                                    wave start their next paths together once at least this many of
                                    the wave's lanes are idle, or all are; 0 = default 8, 1 = at once */
    int32_t reserved_i;
} hcg_options;

/* Per-path record, PathResult fields (src/path_result.jl:76-98). */
typedef struct hcg_path_result {
    int32_t code;              /* EndgameTrackerCode ordinal */
    int32_t singular;
    int32_t winding_number;    /* 0 = nothing */
    int32_t extended_precision;
    int32_t extended_precision_used;
    int32_t accepted_steps;
    int32_t rejected_steps;
    int32_t steps_eg;
    int32_t valuation_valid;   /* 1 when t <= endgame_start (valuation reported) */
    int32_t newton_iters;      /* GPU extras */
    int32_t learned_proposed;
    int32_t learned_accepted;
    double t, accuracy, residual, cond, omega, mu, last_t;
} hcg_path_result;

/* Optional per-step record log (training data for the learned proposer; modelled on
 * path_info's per-step table, src/path_info.jl:3-84).  One record of
 * HCG_STEP_RECORD_FLOATS floats per tracker step, written after the step:
 *  0  path index (0-based position in the batch)
 *  1  step index of the path (accepted + rejected steps so far, minus 1)
 *  2-9  the 8 step features, each clamped to [-80, 80] (NaN written as 0):
 *       2 log Δs of this step, 3 Newton contraction θ, 4 log(local error · Δs⁴),
 *       5 t (real part), 6 log10 cond(H_x, ẋ), 7 log ω, 8 log μ, 9 log τ (trust region)
 *  10 1 if the step was accepted, else 0
 *  11 Newton iterations of the step
 *  12 log of the next Δs the tracker will try
 *  13 flags: 1 extended precision, 2 singular endgame, 4 the learned proposal was used
 *  14 log10 of the accuracy after the step
 *  15 log of the plain rule's next proposal before the segment-end clamp (after a
 *     rejected step: the same as 12) */
#define HCG_LOG_PATH 0
#define HCG_LOG_STEP 1
#define HCG_LOG_FEATURES 2 /* first of 8 */
#define HCG_LOG_ACCEPTED 10
#define HCG_LOG_NEWTON_ITERS 11
#define HCG_LOG_NEXT_LOG_DS 12
#define HCG_LOG_FLAGS 13
#define HCG_LOG_LOG10_ACCURACY 14
#define HCG_LOG_PLAIN_LOG_DS 15
typedef struct hcg_step_log {
    int32_t capacity;   /* max records */
    int32_t pad_;
    float* records;     /* capacity * HCG_STEP_RECORD_FLOATS */
    int32_t* count;     /* out: records written (may exceed capacity; extra are dropped) */
} hcg_step_log;
#define HCG_STEP_RECORD_FLOATS 16

/* Learned step-size proposer: 8 features -> 64 -> 64 (GELU) -> 2 outputs
 * (log step, at-infinity logit).  Weights are BF16 with FP32 accumulation. */
#define HCG_MLP_IN 8
#define HCG_MLP_HIDDEN 64
#define HCG_MLP_OUT 2
#define HCG_MLP_NPARAMS (HCG_MLP_IN * HCG_MLP_HIDDEN + HCG_MLP_HIDDEN + HCG_MLP_HIDDEN * HCG_MLP_HIDDEN + HCG_MLP_HIDDEN + HCG_MLP_HIDDEN * HCG_MLP_OUT + HCG_MLP_OUT)

/* Certificate of one candidate (see hcg_certify in hcgpu.h). */
typedef struct hcg_certificate {
    int32_t certified;    /* the Krawczyk test holds: exactly one zero in the box x₀, and it lies in x₁ */
    int32_t real;         /* that zero is real (conj(x₁) in the interior of x₀, F real on real points) */
    int32_t complex_;     /* some coordinate of x₁ excludes the real axis: that zero is not real */
    int32_t newton_code;  /* refinement: 0 success, 1 rejected, 2 max_iters, -1 not refined */
    int32_t newton_iters;
    int32_t prec;         /* 53: FP64 residual enclosure; 106: the double-double retry */
    double contraction;   /* 1.41422 · max_i Σ_j mag(M_ij) for M = C·J(x₀) − I; certified needs < 0.7071 */
    double inflation;     /* largest inflation radius ε_i of x₀ */
} hcg_certificate;

#ifdef __cplusplus
}
#endif
#endif
