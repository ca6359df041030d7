"""
    HCGPU

Julia side of the C ABI of `libhcgpu.so` (`include/hcgpu.h`, `include/hcgpu_types.h`): the
struct mirrors, one `ccall` wrapper per entry point, and a GPU path tracker that plugs into
HomotopyContinuation.jl's batch seam.

The seam is `solve(S::Solver, starts)` (src/solve.jl:547-570), whose `threaded_solve`
(src/solve.jl:628-709) calls `track(tracker, x; path_number)` (src/endgame_tracker.jl:902-911)
once per start solution.  `solve(S::Solver{GPUTracker}, starts)` below replaces that loop by
one `hcg_track_batch` call and builds the same `PathResult`s (src/endgame_tracker.jl:847-888)
and `Result` (src/result.jl:102-111).  `solve_many` replaces `many_solve`'s loop over target
parameters (src/solve.jl:815-881) by one launch over all targets.

Usage:

    using HomotopyContinuation, HCGPU
    S, starts = solver_startsolutions(F; start_system = :total_degree)
    result = solve(HCGPU.gpu_solver(S), starts)

The struct layouts and the `ccall` signatures are checked against the C headers by
`tests/test_julia_glue.py` (field names, order and types; entry-point names, arity and
argument types), since no Julia toolchain runs in the build's CI.
"""
module HCGPU

using HomotopyContinuation
const HC = HomotopyContinuation
const MK = HomotopyContinuation.ModelKit
import HomotopyContinuation: solve, track

const libhcgpu = get(ENV, "HCGPU_LIB",
                     normpath(joinpath(@__DIR__, "..", "homotopycontinuation.jl_amd", "_build", "libhcgpu.so")))

# ---------------------------------------------------------------------------------------------
# Constants (hcgpu.h, hcgpu_types.h)
# ---------------------------------------------------------------------------------------------
const HCG_OK = 0
const HCG_ERR_INVALID_ARGUMENT = -1
const HCG_ERR_COMPILE = -2
const HCG_ERR_HIP = -3
const HCG_ERR_OUT_OF_MEMORY = -4
const HCG_ERR_UNSUPPORTED = -5
const HCG_STRAIGHT_LINE = 0
const HCG_PARAMETER = 1
const HCG_LAYOUT_LANE = 0
const HCG_LAYOUT_GROUP = 1
const HCG_LOG_PATH = 0
const HCG_LOG_STEP = 1
const HCG_LOG_FEATURES = 2
const HCG_LOG_ACCEPTED = 10
const HCG_LOG_NEWTON_ITERS = 11
const HCG_LOG_NEXT_LOG_DS = 12
const HCG_LOG_FLAGS = 13
const HCG_LOG_LOG10_ACCURACY = 14
const HCG_LOG_PLAIN_LOG_DS = 15
const HCG_STEP_RECORD_FLOATS = 16
const HCG_MLP_IN = 8
const HCG_MLP_HIDDEN = 64
const HCG_MLP_OUT = 2
const HCG_MLP_NPARAMS = HCG_MLP_IN * HCG_MLP_HIDDEN + HCG_MLP_HIDDEN + HCG_MLP_HIDDEN * HCG_MLP_HIDDEN +
                        HCG_MLP_HIDDEN + HCG_MLP_HIDDEN * HCG_MLP_OUT + HCG_MLP_OUT

# ---------------------------------------------------------------------------------------------
# Struct mirrors, field for field (C name hcg_xxx_yyy ↔ Julia HcgXxxYyy)
# ---------------------------------------------------------------------------------------------
struct HcgTape
    instr::Ptr{Int32}
    constants::Ptr{Float64}
    u_assign::Ptr{Int32}
    U_assign::Ptr{Int32}
    ninstr::Int32
    nconst::Int32
    params_first::Int32
    nparams::Int32
    vars_first::Int32
    nvars::Int32
    t_index::Int32
    nu::Int32
    nU::Int32
    output_dim::Int32
    tape_space::Int32
    pad_::Int32
end

struct HcgSystem
    eval_jac::Ptr{HcgTape}
    eval::Ptr{HcgTape}
    fixed_params::Ptr{Float64}
end

struct HcgHomotopyDesc
    kind::Int32
    n::Int32
    start::HcgSystem
    target::HcgSystem
    gamma_re::Float64
    gamma_im::Float64
end

struct HcgBatch
    npaths::Int32
    ngroups::Int32
    starts::Ptr{Float64}
    start_params::Ptr{Float64}
    target_params::Ptr{Float64}
    group_of_path::Ptr{Int32}
    t1::Float64
    t0::Float64
end

struct HcgOptions
    max_steps::Int32
    extended_precision::Int32
    max_step_size::Float64
    max_initial_step_size::Float64
    min_step_size::Float64
    min_rel_step_size::Float64
    terminate_cond::Float64
    a::Float64
    beta_a::Float64
    beta_omega_p::Float64
    beta_tau::Float64
    strict_beta_tau::Float64
    min_newton_iters::Int32
    max_endgame_steps::Int32
    endgame_start::Float64
    max_endgame_extended_steps::Int32
    max_winding_number::Int32
    min_cond::Float64
    min_cond_growth::Float64
    min_coord_growth::Float64
    zero_is_at_infinity::Int32
    at_infinity_check::Int32
    only_nonsingular::Int32
    refine_steps::Int32
    singular_min_accuracy::Float64
    val_finite_tol::Float64
    val_at_infinity_tol::Float64
    sing_cond::Float64
    sing_accuracy::Float64
    scaling_threshold::Float64
    scale_min::Float64
    scale_abs_min::Float64
    scale_max::Float64
    learned_step::Int32
    block_size::Int32
    blocks_per_cu::Int32
    log_steps::Int32
    learned_max_mult::Float64
    learned_min_t::Float64
    learned_inf_t::Float64
    learned_inf_logit::Float64
    refill_idle::Int32
    reserved_i::Int32
end

struct HcgPathResult
    code::Int32
    singular::Int32
    winding_number::Int32
    extended_precision::Int32
    extended_precision_used::Int32
    accepted_steps::Int32
    rejected_steps::Int32
    steps_eg::Int32
    valuation_valid::Int32
    newton_iters::Int32
    learned_proposed::Int32
    learned_accepted::Int32
    t::Float64
    accuracy::Float64
    residual::Float64
    cond::Float64
    omega::Float64
    mu::Float64
    last_t::Float64
end

struct HcgStepLog
    capacity::Int32
    pad_::Int32
    records::Ptr{Float32}
    count::Ptr{Int32}
end

struct HcgCertificate
    certified::Int32
    real::Int32
    complex_::Int32
    newton_code::Int32
    newton_iters::Int32
    prec::Int32
    contraction::Float64
    inflation::Float64
end

struct HcgBuildOptions
    waves_per_eu::Int32
    layout::Int32
    reserved::NTuple{6,Int32}
end

# ---------------------------------------------------------------------------------------------
# Entry points: one ccall each, same names, argument order and meaning as the C declarations
# ---------------------------------------------------------------------------------------------
const Ctx = Ptr{Cvoid}       # hcg_ctx*
const Prog = Ptr{Cvoid}      # hcg_program*
const Graph = Ptr{Cvoid}     # hcg_graph*

hcg_ctx_create(device, out) = ccall((:hcg_ctx_create, libhcgpu), Cint, (Cint, Ptr{Ptr{Cvoid}}), device, out)
hcg_ctx_destroy(ctx) = ccall((:hcg_ctx_destroy, libhcgpu), Cvoid, (Ptr{Cvoid},), ctx)
hcg_program_create(ctx, h, out) =
    ccall((:hcg_program_create, libhcgpu), Cint, (Ptr{Cvoid}, Ptr{HcgHomotopyDesc}, Ptr{Ptr{Cvoid}}), ctx, h, out)
hcg_program_destroy(prog) = ccall((:hcg_program_destroy, libhcgpu), Cvoid, (Ptr{Cvoid},), prog)
hcg_program_source(prog) = ccall((:hcg_program_source, libhcgpu), Cstring, (Ptr{Cvoid},), prog)
hcg_default_options(opts) = ccall((:hcg_default_options, libhcgpu), Cvoid, (Ptr{HcgOptions},), opts)
hcg_track_batch(ctx, prog, batch, opts, results, solutions, last_points, valuations, log) =
    ccall((:hcg_track_batch, libhcgpu), Cint,
          (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{HcgBatch}, Ptr{HcgOptions}, Ptr{HcgPathResult}, Ptr{Float64},
           Ptr{Float64}, Ptr{Float64}, Ptr{HcgStepLog}),
          ctx, prog, batch, opts, results, solutions, last_points, valuations, log)
hcg_track_batch_device(ctx, prog, batch_dev, opts, results_dev, solutions_dev, last_points_dev, valuations_dev) =
    ccall((:hcg_track_batch_device, libhcgpu), Cint,
          (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{HcgBatch}, Ptr{HcgOptions}, Ptr{HcgPathResult}, Ptr{Float64},
           Ptr{Float64}, Ptr{Float64}),
          ctx, prog, batch_dev, opts, results_dev, solutions_dev, last_points_dev, valuations_dev)
hcg_track_batch_device_noalloc(ctx, prog, batch_dev, opts, results_dev, solutions_dev, last_points_dev,
                               valuations_dev) =
    ccall((:hcg_track_batch_device_noalloc, libhcgpu), Cint,
          (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{HcgBatch}, Ptr{HcgOptions}, Ptr{HcgPathResult}, Ptr{Float64},
           Ptr{Float64}, Ptr{Float64}),
          ctx, prog, batch_dev, opts, results_dev, solutions_dev, last_points_dev, valuations_dev)
hcg_graph_create(ctx, prog, batch_dev, opts, results_dev, solutions_dev, last_points_dev, valuations_dev, out) =
    ccall((:hcg_graph_create, libhcgpu), Cint,
          (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{HcgBatch}, Ptr{HcgOptions}, Ptr{HcgPathResult}, Ptr{Float64},
           Ptr{Float64}, Ptr{Float64}, Ptr{Ptr{Cvoid}}),
          ctx, prog, batch_dev, opts, results_dev, solutions_dev, last_points_dev, valuations_dev, out)
hcg_graph_launch(g) = ccall((:hcg_graph_launch, libhcgpu), Cint, (Ptr{Cvoid},), g)
hcg_graph_destroy(g) = ccall((:hcg_graph_destroy, libhcgpu), Cvoid, (Ptr{Cvoid},), g)
hcg_program_create_multi(ctx, hs, nsys, out) =
    ccall((:hcg_program_create_multi, libhcgpu), Cint,
          (Ptr{Cvoid}, Ptr{Ptr{HcgHomotopyDesc}}, Int32, Ptr{Ptr{Cvoid}}), ctx, hs, nsys, out)
hcg_compile_multi_only(hs, nsys, code_size) =
    ccall((:hcg_compile_multi_only, libhcgpu), Cint, (Ptr{Ptr{HcgHomotopyDesc}}, Int32, Ptr{Csize_t}),
          hs, nsys, code_size)
hcg_program_create_ex(ctx, hs, nsys, multi, bo, out) =
    ccall((:hcg_program_create_ex, libhcgpu), Cint,
          (Ptr{Cvoid}, Ptr{Ptr{HcgHomotopyDesc}}, Int32, Int32, Ptr{HcgBuildOptions}, Ptr{Ptr{Cvoid}}),
          ctx, hs, nsys, multi, bo, out)
hcg_compile_ex(hs, nsys, multi, bo, code_size) =
    ccall((:hcg_compile_ex, libhcgpu), Cint,
          (Ptr{Ptr{HcgHomotopyDesc}}, Int32, Int32, Ptr{HcgBuildOptions}, Ptr{Csize_t}), hs, nsys, multi, bo,
          code_size)
hcg_track_multi(ctx, prog, batches, opts, results, solutions, last_points, valuations) =
    ccall((:hcg_track_multi, libhcgpu), Cint,
          (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{HcgBatch}, Ptr{HcgOptions}, Ptr{HcgPathResult}, Ptr{Float64},
           Ptr{Float64}, Ptr{Float64}),
          ctx, prog, batches, opts, results, solutions, last_points, valuations)
hcg_last_kernel_stats(ctx, kernel_ms, total_steps, total_newton_iters) =
    ccall((:hcg_last_kernel_stats, libhcgpu), Cint, (Ptr{Cvoid}, Ptr{Float64}, Ptr{Int64}, Ptr{Int64}),
          ctx, kernel_ms, total_steps, total_newton_iters)
hcg_last_counters(ctx, out4) =
    ccall((:hcg_last_counters, libhcgpu), Cint, (Ptr{Cvoid}, Ptr{UInt64}), ctx, out4)
hcg_program_kernel_info(prog, out6) =
    ccall((:hcg_program_kernel_info, libhcgpu), Cint, (Ptr{Cvoid}, Ptr{Int32}), prog, out6)
hcg_group_schedule_info(hs, nsys, out) =
    ccall((:hcg_group_schedule_info, libhcgpu), Cint, (Ptr{Ptr{HcgHomotopyDesc}}, Int32, Ptr{Int32}), hs, nsys, out)
hcg_generate(h, buf, bufsz, needed, flops6) =
    ccall((:hcg_generate, libhcgpu), Cint, (Ptr{HcgHomotopyDesc}, Ptr{UInt8}, Csize_t, Ptr{Csize_t}, Ptr{Clong}),
          h, buf, bufsz, needed, flops6)
hcg_compile_only(h, code_size) =
    ccall((:hcg_compile_only, libhcgpu), Cint, (Ptr{HcgHomotopyDesc}, Ptr{Csize_t}), h, code_size)
hcg_mlp_load(ctx, params) = ccall((:hcg_mlp_load, libhcgpu), Cint, (Ptr{Cvoid}, Ptr{Float32}), ctx, params)
hcg_mlp_forward_backward(ctx, params, n, x, y, w, out_pred, loss, grad) =
    ccall((:hcg_mlp_forward_backward, libhcgpu), Cint,
          (Ptr{Cvoid}, Ptr{Float32}, Int32, Ptr{Float32}, Ptr{Float32}, Ptr{Float32}, Ptr{Float32}, Ptr{Float32},
           Ptr{Float32}),
          ctx, params, n, x, y, w, out_pred, loss, grad)
hcg_mlp_train(ctx, params, n, x, y, w, epochs, batch, lr, seed, final_loss) =
    ccall((:hcg_mlp_train, libhcgpu), Cint,
          (Ptr{Cvoid}, Ptr{Float32}, Int32, Ptr{Float32}, Ptr{Float32}, Ptr{Float32}, Int32, Int32, Float32, UInt64,
           Ptr{Float32}),
          ctx, params, n, x, y, w, epochs, batch, lr, seed, final_loss)
hcg_certifier_create(ctx, F, n, out) =
    ccall((:hcg_certifier_create, libhcgpu), Cint, (Ptr{Cvoid}, Ptr{HcgSystem}, Int32, Ptr{Ptr{Cvoid}}),
          ctx, F, n, out)
hcg_certifier_compile_only(F, n, code_size) =
    ccall((:hcg_certifier_compile_only, libhcgpu), Cint, (Ptr{HcgSystem}, Int32, Ptr{Csize_t}), F, n, code_size)
hcg_certify(ctx, cert, nsol, candidates, ngroups, params, params_ival, params_ball, group_of_candidate, refine,
            extended, real_system, out, box1, box0, refined) =
    ccall((:hcg_certify, libhcgpu), Cint,
          (Ptr{Cvoid}, Ptr{Cvoid}, Int32, Ptr{Float64}, Int32, Ptr{Float64}, Ptr{Float64}, Ptr{Float64},
           Ptr{Int32}, Int32, Int32, Int32, Ptr{HcgCertificate}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}),
          ctx, cert, nsol, candidates, ngroups, params, params_ival, params_ball, group_of_candidate, refine,
          extended, real_system, out, box1, box0, refined)
hcg_last_error() = ccall((:hcg_last_error, libhcgpu), Cstring, ())

struct HcgpuError <: Exception
    code::Int
    msg::String
end
Base.showerror(io::IO, e::HcgpuError) = print(io, "libhcgpu error ", e.code, ": ", e.msg)
check(rc) = rc == HCG_OK ? nothing : throw(HcgpuError(rc, unsafe_string(hcg_last_error())))

# ---------------------------------------------------------------------------------------------
# Tapes: ModelKit.InstructionSequence → hcg_tape (src/model_kit/instruction_sequence.jl:1-27)
# ---------------------------------------------------------------------------------------------
"""
The arrays an `HcgTape` points into, kept alive with it.  `Instruction` is isbits
`(NTuple{4,Int32}, OpType::Int32, Int32)`, so the instruction vector is the C `instr` array
as is (6 int32 per instruction, 1-based slots, the trailing `OP_STOP` included).
"""
mutable struct TapeData
    instr::Vector{Int32}
    constants::Vector{Float64}
    u_assign::Vector{Int32}
    U_assign::Vector{Int32}
    tape::Base.RefValue{HcgTape}
end

function pairs_i32(a::Vector{Tuple{Int,Int}})
    v = Vector{Int32}(undef, 2 * length(a))
    for (k, (i, j)) in enumerate(a)
        v[2k-1] = i
        v[2k] = j
    end
    v
end

function TapeData(seq::MK.InstructionSequence)
    instr = collect(reinterpret(Int32, seq.instructions))
    constants = Vector{Float64}(undef, 2 * length(seq.constants))
    for (k, c) in enumerate(seq.constants)
        z = ComplexF64(c)
        constants[2k-1] = real(z)
        constants[2k] = imag(z)
    end
    u = pairs_i32(seq.u_assignments)
    U = pairs_i32(seq.U_assignments)
    d = TapeData(instr, constants, u, U, Ref{HcgTape}())
    d.tape[] = HcgTape(pointer(d.instr), isempty(constants) ? C_NULL : pointer(d.constants),
                       isempty(u) ? C_NULL : pointer(d.u_assign), isempty(U) ? C_NULL : pointer(d.U_assign),
                       length(seq.instructions), length(seq.constants),
                       isempty(seq.parameters_range) ? 0 : first(seq.parameters_range), length(seq.parameters_range),
                       first(seq.variables_range), length(seq.variables_range),
                       something(seq.continuation_parameter_index, 0), length(seq.u_assignments),
                       length(seq.U_assignments), seq.output_dim, seq.tape_space_needed, 0)
    d
end

"""The tapes of one `AbstractSystem`: `evaluate_and_jacobian!` (u and U) and `evaluate!`."""
mutable struct SystemData
    eval_jac::TapeData
    eval::TapeData
    fixed_params::Union{Nothing,Vector{Float64}}
    sys::Base.RefValue{HcgSystem}
end

model_system(F::MK.System) = F
model_system(F::HC.ModelKit.CompiledSystem) = F.system
model_system(F::HC.ModelKit.InterpretedSystem) = F.system
model_system(F) = throw(ArgumentError("libhcgpu takes ModelKit systems (compiled or interpreted), got $(typeof(F))"))

function interleave(v::AbstractVector)
    out = Vector{Float64}(undef, 2 * length(v))
    for (k, z) in enumerate(v)
        out[2k-1] = real(ComplexF64(z))
        out[2k] = imag(ComplexF64(z))
    end
    out
end

function SystemData(F::MK.AbstractSystem; fixed = nothing)
    if F isa HC.FixedParameterSystem   # src/systems/fixed_parameter_system.jl:9-33
        return SystemData(F.system; fixed = F.parameters)
    end
    sys = model_system(F)
    ej = TapeData(MK.jacobian_instruction_sequence(sys))
    e = TapeData(MK.instruction_sequence(sys))
    fp = fixed === nothing ? nothing : interleave(fixed)
    d = SystemData(ej, e, fp, Ref{HcgSystem}())
    d.sys[] = HcgSystem(Base.unsafe_convert(Ptr{HcgTape}, ej.tape), Base.unsafe_convert(Ptr{HcgTape}, e.tape),
                        fp === nothing ? C_NULL : pointer(fp))
    d
end

"""Descriptor of a straight-line or parameter homotopy and everything it points into."""
mutable struct HomotopyData
    start::Union{Nothing,SystemData}
    target::SystemData
    p::Union{Nothing,Vector{Float64}}     # ParameterHomotopy: start parameters
    q::Union{Nothing,Vector{Float64}}     # ParameterHomotopy: target parameters
    desc::Base.RefValue{HcgHomotopyDesc}
end

const NULL_SYSTEM = HcgSystem(C_NULL, C_NULL, C_NULL)

function HomotopyData(H::HC.StraightLineHomotopy)   # src/homotopies/straight_line_homotopy.jl:10-68
    G = SystemData(H.start)
    F = SystemData(H.target)
    d = HomotopyData(G, F, nothing, nothing, Ref{HcgHomotopyDesc}())
    d.desc[] = HcgHomotopyDesc(HCG_STRAIGHT_LINE, size(H, 2), G.sys[], F.sys[], real(H.γ), imag(H.γ))
    d
end

function HomotopyData(H::HC.ParameterHomotopy)      # src/homotopies/parameter_homotopy.jl:10-64
    F = SystemData(H.F)
    d = HomotopyData(nothing, F, interleave(H.p), interleave(H.q), Ref{HcgHomotopyDesc}())
    d.desc[] = HcgHomotopyDesc(HCG_PARAMETER, size(H, 2), NULL_SYSTEM, F.sys[], 0.0, 0.0)
    d
end

HomotopyData(H) = throw(ArgumentError("libhcgpu tracks straight-line and parameter homotopies, got $(typeof(H))"))

# ---------------------------------------------------------------------------------------------
# Options: TrackerOptions + TrackerParameters (src/tracker.jl:45-140), EndgameOptions
# (src/endgame_tracker.jl:47-72), WeightedNormOptions (src/norm.jl:36-40)
# ---------------------------------------------------------------------------------------------
function with_fields(o::T; kw...) where {T}
    T((haskey(kw, f) ? convert(fieldtype(T, f), kw[f]) : getfield(o, f) for f in fieldnames(T))...)
end

function default_options()
    r = Ref{HcgOptions}()
    hcg_default_options(r)
    r[]
end

function HcgOptions(T::HC.EndgameTracker; learned_step::Bool = false)
    to = T.tracker.options
    tp = to.parameters
    eo = T.options
    wo = T.tracker.state.norm.options
    with_fields(default_options();
        max_steps = to.max_steps, extended_precision = to.extended_precision,
        max_step_size = to.max_step_size, max_initial_step_size = to.max_initial_step_size,
        min_step_size = to.min_step_size, min_rel_step_size = to.min_rel_step_size,
        terminate_cond = to.terminate_cond,
        a = tp.a, beta_a = tp.β_a, beta_omega_p = tp.β_ω_p, beta_tau = tp.β_τ,
        strict_beta_tau = tp.strict_β_τ, min_newton_iters = tp.min_newton_iters,
        max_endgame_steps = eo.max_endgame_steps, endgame_start = eo.endgame_start,
        max_endgame_extended_steps = eo.max_endgame_extended_steps,
        max_winding_number = eo.max_winding_number, min_cond = eo.min_cond,
        min_cond_growth = eo.min_cond_growth, min_coord_growth = eo.min_coord_growth,
        zero_is_at_infinity = eo.zero_is_at_infinity, at_infinity_check = eo.at_infinity_check,
        only_nonsingular = eo.only_nonsingular, refine_steps = eo.refine_steps,
        singular_min_accuracy = eo.singular_min_accuracy, val_finite_tol = eo.val_finite_tol,
        val_at_infinity_tol = eo.val_at_infinity_tol, sing_cond = eo.sing_cond,
        sing_accuracy = eo.sing_accuracy, scaling_threshold = eo.scaling_threshold,
        scale_min = wo.scale_min, scale_abs_min = wo.scale_abs_min, scale_max = wo.scale_max,
        learned_step = learned_step)
end

# ---------------------------------------------------------------------------------------------
# GPUTracker: an AbstractPathTracker whose batch entry is one launch
# ---------------------------------------------------------------------------------------------
mutable struct GPUTracker <: HC.AbstractPathTracker
    tracker::HC.EndgameTracker   # the CPU tracker it stands in for (homotopy, options)
    ctx::Ptr{Cvoid}
    prog::Ptr{Cvoid}
    homotopy::HomotopyData
    options::HcgOptions
end

"""
    GPUTracker(T::EndgameTracker; device = 0, layout = :auto, learned_weights = nothing)

Compile `T`'s homotopy for the GPU (`hcg_program_create_ex`): `layout = :lane` generates and
compiles the tapes (one path per lane), `:group` schedules them for the precompiled lane-group
kernel (one path per 16 lanes, 32 lanes for 17 to 24 variables); `:auto` takes `:lane` up to
12 variables, where the per-lane kernel factors the Jacobian in registers, and `:group` past
that.  `learned_weights`
(`HCG_MLP_NPARAMS` Float32) turns on the learned step proposer.
"""
function GPUTracker(T::HC.EndgameTracker; device::Integer = 0, layout::Symbol = :auto,
                    learned_weights::Union{Nothing,Vector{Float32}} = nothing)
    hd = HomotopyData(T.tracker.homotopy)
    if layout === :auto
        layout = hd.desc[].n > 12 ? :group : :lane
    end
    ctx = Ref{Ptr{Cvoid}}(C_NULL)
    check(hcg_ctx_create(device, ctx))
    prog = Ref{Ptr{Cvoid}}(C_NULL)
    bo = Ref(HcgBuildOptions(0, layout === :group ? HCG_LAYOUT_GROUP : HCG_LAYOUT_LANE, ntuple(_ -> Int32(0), 6)))
    descs = Ptr{HcgHomotopyDesc}[Base.unsafe_convert(Ptr{HcgHomotopyDesc}, hd.desc)]
    GC.@preserve hd descs begin
        rc = hcg_program_create_ex(ctx[], descs, 1, 0, bo, prog)
    end
    rc == HCG_OK || (hcg_ctx_destroy(ctx[]); check(rc))
    learned = learned_weights !== nothing
    if learned
        length(learned_weights) == HCG_MLP_NPARAMS || throw(ArgumentError("expected $HCG_MLP_NPARAMS weights"))
        check(hcg_mlp_load(ctx[], learned_weights))
    end
    G = GPUTracker(T, ctx[], prog[], hd, HcgOptions(T; learned_step = learned))
    finalizer(G) do g
        g.prog != C_NULL && hcg_program_destroy(g.prog)
        g.ctx != C_NULL && hcg_ctx_destroy(g.ctx)
    end
    G
end

"""
    gpu_solver(S::Solver; kwargs...)

The same solver with its tracker replaced by a `GPUTracker` (seed and start system kept).
"""
gpu_solver(S::HC.Solver{<:HC.EndgameTracker}; kwargs...) =
    HC.Solver([GPUTracker(S.trackers[1]; kwargs...)], S.seed, HC.SolveStats(), S.start_system)

const CODES = instances(HC.EndgameTrackerCode.codes)

"""Raw results of one launch: per-path records and the n × npaths arrays (column = path)."""
struct BatchOutput
    records::Vector{HcgPathResult}
    solutions::Matrix{ComplexF64}
    last_points::Matrix{ComplexF64}
    valuations::Matrix{Float64}
end

function to_complex(v::Vector{Float64}, n::Int, m::Int)
    X = Matrix{ComplexF64}(undef, n, m)
    for j = 1:m, i = 1:n
        k = 2 * ((j - 1) * n + (i - 1))
        X[i, j] = complex(v[k+1], v[k+2])
    end
    X
end

"""
    track_batch(G::GPUTracker, starts; targets = nothing, t₁ = 1.0) -> BatchOutput

Every start solution tracked from t₁ to 0 in one launch (`hcg_track_batch`).  For a parameter
homotopy, `targets` (a vector of parameter vectors) replaces the target by every entry in turn:
path (k, j) is start k towards target j, ordered target-major as `many_solve` does.
"""
function track_batch(G::GPUTracker, starts::AbstractVector; targets = nothing, t₁::Real = 1.0)
    n = Int(G.homotopy.desc[].n)
    S = collect(starts)
    ns = length(S)
    hd = G.homotopy
    ngroups = 1
    q = hd.q
    group_of_path = Int32[]
    if targets !== nothing
        hd.p === nothing && throw(ArgumentError("targets need a parameter homotopy"))
        ngroups = length(targets)
        q = reduce(vcat, interleave.(targets))
        group_of_path = Int32[j - 1 for j = 1:ngroups for _ = 1:ns]
    end
    npaths = ns * ngroups
    x = Vector{Float64}(undef, 2 * n * npaths)
    for j = 1:ngroups, k = 1:ns
        s = S[k]
        length(s) == n || throw(ArgumentError("start solution $k has length $(length(s)), expected $n"))
        base = 2 * n * ((j - 1) * ns + (k - 1))
        for i = 1:n
            x[base+2i-1] = real(ComplexF64(s[i]))
            x[base+2i] = imag(ComplexF64(s[i]))
        end
    end
    results = Vector{HcgPathResult}(undef, npaths)
    sols = Vector{Float64}(undef, 2 * n * npaths)
    last = Vector{Float64}(undef, 2 * n * npaths)
    val = Vector{Float64}(undef, n * npaths)
    opts = Ref(G.options)
    GC.@preserve x q group_of_path hd begin
        batch = Ref(HcgBatch(npaths, ngroups, pointer(x),
                             hd.p === nothing ? C_NULL : pointer(hd.p),
                             q === nothing ? C_NULL : pointer(q),
                             isempty(group_of_path) ? C_NULL : pointer(group_of_path),
                             Float64(t₁), 0.0))
        check(hcg_track_batch(G.ctx, G.prog, batch, opts, results, sols, last, val, C_NULL))
    end
    BatchOutput(results, to_complex(sols, n, npaths), to_complex(last, n, npaths), reshape(val, n, npaths))
end

"""PathResult of path `i` of a launch, as `PathResult(endgame_tracker, x, path_number)` builds it."""
function path_result(out::BatchOutput, i::Int, start_solution, path_number)
    r = out.records[i]
    HC.PathResult(
        return_code = Symbol(CODES[r.code+1]),
        solution = out.solutions[:, i],
        t = r.t,
        accuracy = r.accuracy,
        residual = r.residual,
        singular = r.singular != 0,
        condition_jacobian = r.cond,
        winding_number = r.winding_number == 0 ? nothing : Int(r.winding_number),
        last_path_point = (out.last_points[:, i], r.last_t),
        valuation = r.valuation_valid != 0 ? out.valuations[:, i] : nothing,
        start_solution = start_solution,
        path_number = path_number,
        ω = r.omega,
        μ = r.mu,
        extended_precision = r.extended_precision != 0,
        accepted_steps = r.accepted_steps,
        rejected_steps = r.rejected_steps,
        steps_eg = r.steps_eg,
        extended_precision_used = r.extended_precision_used != 0,
    )
end

"""
    solve(S::Solver{GPUTracker}, starts; kwargs...)

`threaded_solve`'s loop (src/solve.jl:628-709) as one launch: `Result` of every path, with the
solver's statistics updated path by path as the loop does.
"""
function solve(S::HC.Solver{GPUTracker}, starts::AbstractArray; show_progress::Bool = false, kwargs...)
    HC.init!(S.stats)
    G = S.trackers[1]
    out = track_batch(G, vec(starts))
    path_results = Vector{HC.PathResult}(undef, length(starts))
    for (k, s) in enumerate(starts)
        r = path_result(out, k, s, k)
        path_results[k] = r
        HC.update!(S.stats, r)
    end
    HC.Result(path_results; seed = S.seed, start_system = S.start_system)
end

"""
    solve_many(S::Solver{GPUTracker}, starts, targets) -> Vector{Result}

`many_solve`'s loop over target parameters (src/solve.jl:815-881) as one launch: one `Result`
per target.
"""
function solve_many(S::HC.Solver{GPUTracker}, starts::AbstractArray, targets::AbstractVector)
    G = S.trackers[1]
    ns = length(starts)
    out = track_batch(G, vec(starts); targets = targets)
    map(1:length(targets)) do j
        prs = [path_result(out, (j - 1) * ns + k, starts[k], k) for k = 1:ns]
        HC.Result(prs; seed = S.seed, start_system = S.start_system)
    end
end

"""`track(tracker, x, t₁; path_number)` (src/endgame_tracker.jl:902-911) for one path."""
function track(G::GPUTracker, x, t₁::Real = 1.0; path_number::Union{Nothing,Int} = nothing, kwargs...)
    out = track_batch(G, [x]; t₁ = t₁)
    path_result(out, 1, x, path_number)
end

end # module
