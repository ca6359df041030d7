/*
 * hcgpu.h — C ABI of the MI355X batched path tracker (libhcgpu.so).
 *
 * This is the drop-in boundary that HomotopyContinuation.jl's batch path would
 * `ccall`.  It replaces, for a whole batch of paths, the per-path call
 *   track(tracker::EndgameTracker, x, t₁; path_number)      src/endgame_tracker.jl:902-911
 * issued by threaded_solve / serial_solve                      src/solve.jl:603-709
 * and the per-parameter loop of many_solve                     src/solve.jl:815-881.
 * Each path's output record mirrors PathResult(endgame_tracker) src/endgame_tracker.jl:847-888.
 *
 * All functions return 0 on success and a negative HCG_ERR_* code on failure;
 * hcg_last_error() then returns a thread-local message.  No C++ exception ever
 * crosses this boundary.  Per-path numerical failures are not errors: they are
 * path return codes (EndgameTrackerCode ordinals, src/endgame_tracker.jl:100-117).
 *
 * Ownership: the caller owns every host array for the duration of a call; the
 * library copies inputs to the device and results back.  The library owns device
 * buffers, hipRTC code objects and streams (freed by the *_destroy calls).
 * Calls are synchronous.  An hcg_ctx is not thread-safe: use one per host thread.
 */
#ifndef HCGPU_H
#define HCGPU_H
#ifndef __HIPCC_RTC__
#include <stddef.h>
#endif
#include <stdint.h>
#include "hcgpu_types.h"
#ifdef __cplusplus
extern "C" {
#endif

#define HCG_OK 0
#define HCG_ERR_INVALID_ARGUMENT (-1)
#define HCG_ERR_COMPILE (-2)
#define HCG_ERR_HIP (-3)
#define HCG_ERR_OUT_OF_MEMORY (-4)
#define HCG_ERR_UNSUPPORTED (-5)

typedef struct hcg_ctx hcg_ctx;
typedef struct hcg_program hcg_program;

/*
 * Straight-line program, a field-for-field mirror of ModelKit's InstructionSequence
 * (src/model_kit/instruction_sequence.jl:1-27).
 *  instr: ninstr x 6 int32 = (input1, input2, input3, input4, op, output), i.e. the
 *         isbits layout of `Instruction` (:1-5).  Indices are 1-based tape slots;
 *         op is the OpType ordinal (src/model_kit/operations.jl:5-49, OP_STOP = 0 ...
 *         OP_MULMULSUB = 32).  For OP_POW_INT input2 is the integer exponent (:178-180).
 *  constants: nconst complex numbers (re, im) occupying tape slots 1..nconst.
 *  params_first/vars_first: 1-based first tape slot of the parameter / variable block.
 *  t_index: tape slot of the continuation parameter, or 0 for a System tape.
 *  u_assign: nu pairs (i, k): u[i] = tape[k]   (u_assignments, :22)
 *  U_assign: nU pairs (l, k): U[l] = tape[k], l a 1-based column-major linear index
 *            (U_assignments, :232-237).
 *  tape_space: number of tape slots (tape_space_needed).
 */
typedef struct hcg_tape {
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
} hcg_tape;

/* One AbstractSystem: the evaluate_and_jacobian! tape (u and U) and the evaluate!
 * tape (u only; also used for the double-double residual and for taylor!).
 * fixed_params: nparams complex values for a FixedParameterSystem
 * (src/systems/fixed_parameter_system.jl:9-50) or NULL when the homotopy drives the
 * parameters (ParameterHomotopy). */
typedef struct hcg_system {
    const hcg_tape* eval_jac;
    const hcg_tape* eval;
    const double* fixed_params;
} hcg_system;

#define HCG_STRAIGHT_LINE 0 /* H = γ t G(x) + (1-t) F(x)  src/homotopies/straight_line_homotopy.jl:10-154 */
#define HCG_PARAMETER 1     /* H = F(x; t p + (1-t) q)     src/homotopies/parameter_homotopy.jl:10-102 */

typedef struct hcg_homotopy_desc {
    int32_t kind;
    int32_t n; /* number of variables = number of equations (square systems) */
    hcg_system start;  /* G (STRAIGHT_LINE only) */
    hcg_system target; /* F */
    double gamma_re, gamma_im;
} hcg_homotopy_desc;


/* A batch of paths for one program. */
typedef struct hcg_batch {
    int32_t npaths;
    int32_t ngroups;              /* number of target-parameter groups (PARAMETER) */
    const double* starts;         /* npaths * n complex, path-major */
    const double* start_params;   /* nparams complex: p (PARAMETER) */
    const double* target_params;  /* ngroups * nparams complex: q per group */
    const int32_t* group_of_path; /* npaths entries in [0, ngroups) or NULL (group 0) */
    double t1, t0;                /* track from t1 to t0 (EndgameTracker: 1 -> 0) */
} hcg_batch;

int hcg_ctx_create(int device, hcg_ctx** out);
void hcg_ctx_destroy(hcg_ctx* ctx);

/* Generates HIP source for the homotopy's tapes, compiles it with hipRTC for gfx950
 * (cached by source hash) and loads the tracker kernels.  Counterpart of
 * CompiledSystem/CompiledHomotopy code generation
 * (src/model_kit/compiled_system_homotopy.jl:15-58, 178-361). */
int hcg_program_create(hcg_ctx* ctx, const hcg_homotopy_desc* h, hcg_program** out);
void hcg_program_destroy(hcg_program* prog);
/* Generated HIP source of a program (for inspection); valid until destroy. */
const char* hcg_program_source(const hcg_program* prog);

void hcg_default_options(hcg_options* opts);

/* Track every path of `batch` to completion in one persistent launch and write one
 * hcg_path_result per path plus solution / last_point (npaths*n complex) and
 * valuation (npaths*n doubles).  Replaces threaded_solve's loop of
 * track(tracker, S[k]; path_number = k) (src/solve.jl:628-709). */
int hcg_track_batch(hcg_ctx* ctx, const hcg_program* prog, const hcg_batch* batch,
                    const hcg_options* opts, hcg_path_result* results, double* solutions,
                    double* last_points, double* valuations, hcg_step_log* log);

/* Same as hcg_track_batch, but every array reachable from `batch_dev` and every output
 * pointer is device memory (HBM) of the context's device: nothing is copied, for
 * callers that keep start solutions and results resident.  Out-of-range
 * group_of_path entries are rejected per path on the device (code terminated_unknown). */
int hcg_track_batch_device(hcg_ctx* ctx, const hcg_program* prog, const hcg_batch* batch_dev,
                           const hcg_options* opts, hcg_path_result* results_dev, double* solutions_dev,
                           double* last_points_dev, double* valuations_dev);

/* hcg_track_batch_device that allocates nothing.  It has no counterpart in the reference:
 * it is for callers that cannot allocate device memory (code under HIP graph capture, tight
 * serving loops, harnesses that forbid hipMalloc).  The launch state (work-queue head, the
 * four launch counters, the options record) lives in one fixed 1 KiB device array of the
 * library, one per GPU, held until the process ends; a kernel on the context's stream clears
 * it at each call.  All calls on one GPU share that array, so calls that overlap on one GPU
 * race.  Inputs, outputs and limits are those of hcg_track_batch_device (the array's size
 * does not depend on the batch). */
int hcg_track_batch_device_noalloc(hcg_ctx* ctx, const hcg_program* prog, const hcg_batch* batch_dev,
                                   const hcg_options* opts, hcg_path_result* results_dev, double* solutions_dev,
                                   double* last_points_dev, double* valuations_dev);

/* Steady-state relaunch of a device-resident batch as one HIP graph.  No counterpart in the
 * reference, which re-enters its per-path loop for every batch (many_solve,
 * src/solve.jl:815-881); this is for callers that track batch after batch through the same
 * buffers (one parameter set per batch, a serving loop).  hcg_graph_create captures what
 * hcg_track_batch_device enqueues for these arguments: a kernel that clears the graph's own
 * launch state (work-queue head, counters, options; no memset node) and the tracker kernel.
 * The graph keeps every pointer and the options as given: the caller rewrites the same device
 * buffers (new starts, new target parameters, same path count) and calls hcg_graph_launch, which
 * replays both kernels on the context's stream and returns when they are done (kernel time and
 * counters as for hcg_last_kernel_stats / hcg_last_counters).  Results equal those of
 * hcg_track_batch_device on the same inputs.  Destroy the graph before its program and
 * context. */
typedef struct hcg_graph hcg_graph;
int hcg_graph_create(hcg_ctx* ctx, const hcg_program* prog, const hcg_batch* batch_dev, const hcg_options* opts,
                     hcg_path_result* results_dev, double* solutions_dev, double* last_points_dev,
                     double* valuations_dev, hcg_graph** out);
int hcg_graph_launch(hcg_graph* g);
void hcg_graph_destroy(hcg_graph* g);

/* Several homotopies (1..8, any sizes and kinds) compiled into one program: one
 * persistent launch tracks the paths of all of them from a single work queue.  This is the
 * batch form of calling threaded_solve (src/solve.jl:628-709) once per system; there is
 * no single reference call it replaces. */
int hcg_program_create_multi(hcg_ctx* ctx, const hcg_homotopy_desc* const* hs, int32_t nsys, hcg_program** out);
int hcg_compile_multi_only(const hcg_homotopy_desc* const* hs, int32_t nsys, size_t* code_size);

/* Build choices of a program (no counterpart in the reference; picked by the autotuner).
 * waves_per_eu: waves per SIMD the tracker kernel is compiled for, i.e. a register budget
 * of 512 / waves_per_eu per lane (1, 2, 4 or 8; 0 = the library default; per-lane layout).
 * layout: HCG_LAYOUT_LANE (one path per lane, the tapes compiled with hipRTC) or
 * HCG_LAYOUT_GROUP (one path per group of 16 lanes, 32 lanes for 17 to 24 variables, each
 * lane holding one row of the Jacobian and one entry of every vector in registers; the tapes
 * are scheduled for a precompiled kernel, so no compilation happens). */
#define HCG_LAYOUT_LANE 0
#define HCG_LAYOUT_GROUP 1
typedef struct hcg_build_options {
    int32_t waves_per_eu;
    int32_t layout;
    int32_t reserved[6];
} hcg_build_options;
/* General form of hcg_program_create / hcg_program_create_multi (multi = 1: one kernel over
 * all nsys homotopies; multi = 0 requires nsys = 1); bo may be NULL. */
int hcg_program_create_ex(hcg_ctx* ctx, const hcg_homotopy_desc* const* hs, int32_t nsys, int32_t multi,
                          const hcg_build_options* bo, hcg_program** out);
int hcg_compile_ex(const hcg_homotopy_desc* const* hs, int32_t nsys, int32_t multi, const hcg_build_options* bo,
                   size_t* code_size);
/* batches[s] describes the paths of system s.  Outputs are concatenated in system order:
 * results: sum_s npaths_s records; solutions and last_points: sum_s npaths_s * n_s complex;
 * valuations: sum_s npaths_s * n_s doubles. */
int hcg_track_multi(hcg_ctx* ctx, const hcg_program* prog, const hcg_batch* batches, const hcg_options* opts,
                    hcg_path_result* results, double* solutions, double* last_points, double* valuations);

/* Time of the last launch's tracker kernel (ms, HIP events on its stream), and the
 * device-side counters of that launch. */
int hcg_last_kernel_stats(hcg_ctx* ctx, double* kernel_ms, int64_t* total_steps,
                          int64_t* total_newton_iters);
/* Device counters of the last launch: [0] tracker steps, [1] Newton iterations,
 * [2] counted FP64 flops, [3] accepted learned proposals. */
int hcg_last_counters(hcg_ctx* ctx, uint64_t* out4);
/* Resource use of a program's kernel from its code object (no counterpart in the reference:
 * the occupancy evidence of the measurement): out6 = [VGPRs per lane, private segment bytes
 * per lane, static LDS bytes per workgroup, waves per SIMD the kernel was built for (lane-group
 * layout: LDS of a plain launch, to which the learned step adds its network's BF16 weights,
 * and the workgroups that fit on a CU with it, each one wave per SIMD),
 * workgroup size, compute units of the device]. */
int hcg_program_kernel_info(const hcg_program* prog, int32_t* out6);
/* The lane-group layout's slot schedules of nsys homotopies, built and checked against their
 * tapes without a device (no counterpart in the reference; the scheduling evidence of the
 * layout): out[0] = kernel size class (4, 8, 12, 16), out[1] = LDS bytes per path, then per
 * system six entries: eval_jac slots, LDS registers, instructions; eval slots, registers,
 * instructions.  out holds 2 + 6 nsys entries. */
int hcg_group_schedule_info(const hcg_homotopy_desc* const* hs, int32_t nsys, int32_t* out);
/* Generated HIP source and static flop counts (eval, eval_jac, eval_dd, taylor1..3) of a
 * homotopy without compiling it; no device needed. */
int hcg_generate(const hcg_homotopy_desc* h, char* buf, size_t bufsz, size_t* needed, long* flops6);
/* Generate and hipRTC-compile (filling the code-object cache); no device needed. */
int hcg_compile_only(const hcg_homotopy_desc* h, size_t* code_size);

/* Load FP32 master weights (HCG_MLP_NPARAMS floats, layout W1[64][8] b1 W2[64][64] b2 W3[2][64] b3)
 * into the context; the tracker rounds them to BF16. */
int hcg_mlp_load(hcg_ctx* ctx, const float* params);
/* Forward + backward on the device for one batch: writes loss and gradients
 * (HCG_MLP_NPARAMS floats).  x: n*8, y: n*2 targets, w: n*2 per-output weights. */
int hcg_mlp_forward_backward(hcg_ctx* ctx, const float* params, int32_t n, const float* x,
                             const float* y, const float* w, float* out_pred, float* loss,
                             float* grad);
/* Train with Adam for `epochs` over (x, y, w) on the device; params in/out. */
int hcg_mlp_train(hcg_ctx* ctx, float* params, int32_t n, const float* x, const float* y,
                  const float* w, int32_t epochs, int32_t batch, float lr, uint64_t seed,
                  float* final_loss);

/* ---------------- certification ----------------
 * Interval-arithmetic Krawczyk certification of approximate zeros of a square system
 * F(x; p), one candidate per GPU lane.  Replaces the per-candidate loop of _certify
 * (src/certification.jl:779-975) calling certify_solution (:1029-1114) with the FP64
 * ε-inflation test ε_inflation_krawczyk (:1225-1287), in the interval arithmetic of
 * src/interval_arithmetic.jl.  The reference retries candidates that fail in FP64 with
 * Arb balls of growing precision (extended_prec_certify_solution, :1116-1218); here a
 * failed candidate is retried once with the residual F(x̃) enclosed in double-double
 * balls (about 106 bits), the rest of the test staying in FP64 intervals.  Candidates that
 * fail both are reported as not certified.  Duplicate detection among certified boxes
 * (DistinctSolutionCertificates, :300-427) is host work on the returned boxes. */

/* Compile a certifier for the system F (its eval_jac and eval tapes; F.fixed_params fixes
 * the parameters, else they are given per hcg_certify call).  Counterpart of
 * CertificationCache(F) (src/certification.jl:631-671), which builds the IComplexF64
 * interpreters of F and of its Jacobian. */
int hcg_certifier_create(hcg_ctx* ctx, const hcg_system* F, int32_t n, hcg_program** out);
int hcg_certifier_compile_only(const hcg_system* F, int32_t n, size_t* code_size);
/* Certify nsol candidates (nsol * n complex).  Parameters come in ngroups groups, candidate
 * k belonging to group group_of_candidate[k] (NULL: all to group 0): params holds ngroups
 * parameter points (nparams complex each) used by the Newton refinement; params_ival their
 * enclosures (nparams x (re lo, re hi, im lo, im hi) each), NULL = the points themselves;
 * params_ball double-double enclosures for the extended retry (nparams x (re hi, re lo,
 * re radius, im hi, im lo, im radius) each), NULL = the points.  A system without free
 * parameters takes ngroups = 0 and NULL for all four.  refine = 1 refines each candidate by
 * Newton's method first (the reference's default); extended = 1 enables the double-double
 * retry.  real_system = 1 if F(·; p) maps real points to real values for every group
 * (is_real(F), src/model_kit/abstract_system_homotopy.jl:85-89, decided by the caller).
 * Outputs: out[nsol]; box1 / box0: nsol * n * 4 doubles (re lo, re hi, im lo, im hi) of x₁
 * and x₀; refined: nsol * n complex. */
int hcg_certify(hcg_ctx* ctx, const hcg_program* cert, int32_t nsol, const double* candidates, int32_t ngroups,
                const double* params, const double* params_ival, const double* params_ball,
                const int32_t* group_of_candidate, int32_t refine, int32_t extended, int32_t real_system,
                hcg_certificate* out, double* box1, double* box0, double* refined);

const char* hcg_last_error(void);

#ifdef __cplusplus
}
#endif
#endif
