"""ctypes mirror of include/hcgpu.h and the loader of libhcgpu.so.

The loader never falls back to anything: if the HIP library is missing, calls
raise `HcgpuError` (the product path has no CPU fallback).
"""
from __future__ import annotations

import ctypes
import os
from pathlib import Path

from .tape import hcg_tape

PKG_ROOT = Path(__file__).resolve().parent.parent  # homotopycontinuation.jl_amd/
LIB_PATH = PKG_ROOT / "_build" / "libhcgpu.so"

HCG_STRAIGHT_LINE = 0
HCG_PARAMETER = 1
HCG_OK, HCG_ERR_INVALID_ARGUMENT, HCG_ERR_COMPILE, HCG_ERR_HIP, HCG_ERR_OUT_OF_MEMORY, HCG_ERR_UNSUPPORTED = 0, -1, -2, -3, -4, -5
HCG_STEP_RECORD_FLOATS = 16
HCG_MLP_IN, HCG_MLP_HIDDEN, HCG_MLP_OUT = 8, 64, 2
HCG_MLP_NPARAMS = (HCG_MLP_IN * HCG_MLP_HIDDEN + HCG_MLP_HIDDEN + HCG_MLP_HIDDEN * HCG_MLP_HIDDEN
                   + HCG_MLP_HIDDEN + HCG_MLP_HIDDEN * HCG_MLP_OUT + HCG_MLP_OUT)

# EndgameTrackerCode ordinals, src/endgame_tracker.jl:100-117
RETURN_CODES = [
    "tracking", "success", "at_infinity", "at_zero", "terminated_accuracy_limit",
    "terminated_invalid_startvalue", "terminated_invalid_startvalue_singular_jacobian",
    "terminated_ill_conditioned", "terminated_max_steps", "terminated_max_extended_steps",
    "terminated_max_winding_number", "terminated_step_size_too_small", "terminated_unknown",
    "post_check_failed", "excess_solution", "polyhedral_failed",
]
# TrackerCode ordinals, src/tracker.jl:166-176
TRACKER_CODES = [
    "tracking", "success", "terminated_max_steps", "terminated_accuracy_limit",
    "terminated_ill_conditioned", "terminated_invalid_startvalue",
    "terminated_invalid_startvalue_singular_jacobian", "terminated_step_size_too_small",
    "terminated_unknown",
]


class HcgpuError(RuntimeError):
    pass


class hcg_system(ctypes.Structure):
    _fields_ = [
        ("eval_jac", ctypes.POINTER(hcg_tape)),
        ("eval", ctypes.POINTER(hcg_tape)),
        ("fixed_params", ctypes.POINTER(ctypes.c_double)),
    ]


class hcg_homotopy_desc(ctypes.Structure):
    _fields_ = [
        ("kind", ctypes.c_int32),
        ("n", ctypes.c_int32),
        ("start", hcg_system),
        ("target", hcg_system),
        ("gamma_re", ctypes.c_double),
        ("gamma_im", ctypes.c_double),
    ]


_I, _D = ctypes.c_int32, ctypes.c_double


class hcg_options(ctypes.Structure):
    _fields_ = [
        ("max_steps", _I), ("extended_precision", _I),
        ("max_step_size", _D), ("max_initial_step_size", _D), ("min_step_size", _D),
        ("min_rel_step_size", _D), ("terminate_cond", _D),
        ("a", _D), ("beta_a", _D), ("beta_omega_p", _D), ("beta_tau", _D), ("strict_beta_tau", _D),
        ("min_newton_iters", _I),
        ("max_endgame_steps", _I),
        ("endgame_start", _D),
        ("max_endgame_extended_steps", _I), ("max_winding_number", _I),
        ("min_cond", _D), ("min_cond_growth", _D), ("min_coord_growth", _D),
        ("zero_is_at_infinity", _I), ("at_infinity_check", _I), ("only_nonsingular", _I), ("refine_steps", _I),
        ("singular_min_accuracy", _D), ("val_finite_tol", _D), ("val_at_infinity_tol", _D),
        ("sing_cond", _D), ("sing_accuracy", _D), ("scaling_threshold", _D),
        ("scale_min", _D), ("scale_abs_min", _D), ("scale_max", _D),
        ("learned_step", _I), ("block_size", _I), ("blocks_per_cu", _I), ("log_steps", _I),
        ("learned_max_mult", _D), ("learned_min_t", _D), ("learned_inf_t", _D), ("learned_inf_logit", _D),
        ("refill_idle", _I), ("reserved_i", _I),
    ]


def default_options(**overrides) -> hcg_options:
    """Reference defaults: TrackerOptions / TrackerParameters (src/tracker.jl:45-140),
    EndgameOptions (src/endgame_tracker.jl:47-72), WeightedNormOptions (src/norm.jl:36-40)."""
    o = hcg_options()
    o.max_steps = 10_000
    o.extended_precision = 1
    o.max_step_size = float("inf")
    o.max_initial_step_size = float("inf")
    o.min_step_size = 1e-48
    o.min_rel_step_size = 0.0
    o.terminate_cond = 1e13
    o.a, o.beta_a, o.beta_omega_p, o.beta_tau = 0.125, 1.0, 3.0, 0.4
    o.strict_beta_tau = min(0.75 * 0.4, 0.4)
    o.min_newton_iters = 2
    o.max_endgame_steps = 2000
    o.endgame_start = 0.1
    o.max_endgame_extended_steps = 400
    o.max_winding_number = 6
    o.min_cond, o.min_cond_growth, o.min_coord_growth = 1e6, 1e4, 100.0
    o.zero_is_at_infinity, o.at_infinity_check, o.only_nonsingular = 0, 1, 0
    o.refine_steps = 3
    o.singular_min_accuracy = 1e-6
    o.val_finite_tol, o.val_at_infinity_tol = 0.05, 0.01
    o.sing_cond, o.sing_accuracy, o.scaling_threshold = 1e14, 1e-12, -30.0
    o.scale_min, o.scale_abs_min, o.scale_max = 1e-4, 1e-6, 2.0 ** 511
    o.learned_step, o.block_size, o.blocks_per_cu, o.log_steps = 0, 0, 0, 0
    o.learned_max_mult, o.learned_min_t = 4.0, 0.0
    o.learned_inf_t, o.learned_inf_logit = 0.0, 4.0
    o.refill_idle = 0
    presets = {"default": (3.0, 0.4), "fast": (2.0, 0.75), "conservative": (4.0, 0.25)}
    params = overrides.pop("parameters", None)
    if params is not None:
        o.beta_omega_p, o.beta_tau = presets[params]
        o.strict_beta_tau = min(0.75 * o.beta_tau, 0.4)
    for k, v in overrides.items():
        if not hasattr(o, k):
            raise TypeError(f"unknown option {k}")
        setattr(o, k, v)
    return o


class hcg_build_options(ctypes.Structure):
    _fields_ = [("waves_per_eu", _I), ("layout", _I), ("reserved", _I * 6)]


class hcg_batch(ctypes.Structure):
    _fields_ = [
        ("npaths", _I), ("ngroups", _I),
        ("starts", ctypes.POINTER(_D)),
        ("start_params", ctypes.POINTER(_D)),
        ("target_params", ctypes.POINTER(_D)),
        ("group_of_path", ctypes.POINTER(_I)),
        ("t1", _D), ("t0", _D),
    ]


class hcg_path_result(ctypes.Structure):
    _fields_ = [
        ("code", _I), ("singular", _I), ("winding_number", _I), ("extended_precision", _I),
        ("extended_precision_used", _I), ("accepted_steps", _I), ("rejected_steps", _I),
        ("steps_eg", _I), ("valuation_valid", _I), ("newton_iters", _I),
        ("learned_proposed", _I), ("learned_accepted", _I),
        ("t", _D), ("accuracy", _D), ("residual", _D), ("cond", _D), ("omega", _D), ("mu", _D),
        ("last_t", _D),
    ]


class hcg_certificate(ctypes.Structure):
    _fields_ = [
        ("certified", _I), ("real", _I), ("complex_", _I), ("newton_code", _I), ("newton_iters", _I),
        ("prec", _I), ("contraction", _D), ("inflation", _D),
    ]


class hcg_step_log(ctypes.Structure):
    _fields_ = [
        ("capacity", _I), ("pad_", _I),
        ("records", ctypes.POINTER(ctypes.c_float)),
        ("count", ctypes.POINTER(_I)),
    ]


_lib = None


def lib():
    """Load libhcgpu.so (built in-tree by __graft_entry__.build())."""
    global _lib
    if _lib is not None:
        return _lib
    path = os.environ.get("HCGPU_LIB", str(LIB_PATH))
    if not os.path.exists(path):
        raise HcgpuError(f"libhcgpu.so not built ({path}); run __graft_entry__.build()")
    # The library binds whichever HIP runtime and hipRTC the process already holds under their
    # sonames (torch's bundled copies when torch was imported first, else /opt/rocm's).  It does
    # not import torch itself: the code-object cache is compiled by __graft_entry__.build() with
    # /opt/rocm's hipRTC, whose code is faster than torch's older bundled compiler's.  Code that
    # uses torch device buffers (hcgpu.device) imports torch before this library is loaded.
    L = ctypes.CDLL(path)
    P = ctypes.POINTER
    L.hcg_ctx_create.argtypes = [ctypes.c_int, P(ctypes.c_void_p)]
    L.hcg_ctx_destroy.argtypes = [ctypes.c_void_p]
    L.hcg_ctx_destroy.restype = None
    L.hcg_program_create.argtypes = [ctypes.c_void_p, P(hcg_homotopy_desc), P(ctypes.c_void_p)]
    L.hcg_program_destroy.argtypes = [ctypes.c_void_p]
    L.hcg_program_destroy.restype = None
    L.hcg_program_source.argtypes = [ctypes.c_void_p]
    L.hcg_program_source.restype = ctypes.c_char_p
    L.hcg_default_options.argtypes = [P(hcg_options)]
    L.hcg_default_options.restype = None
    L.hcg_track_batch.argtypes = [ctypes.c_void_p, ctypes.c_void_p, P(hcg_batch), P(hcg_options),
                                  P(hcg_path_result), P(_D), P(_D), P(_D), P(hcg_step_log)]
    L.hcg_last_kernel_stats.argtypes = [ctypes.c_void_p, P(_D), P(ctypes.c_int64), P(ctypes.c_int64)]
    L.hcg_mlp_load.argtypes = [ctypes.c_void_p, P(ctypes.c_float)]
    L.hcg_mlp_forward_backward.argtypes = [ctypes.c_void_p, P(ctypes.c_float), _I, P(ctypes.c_float),
                                           P(ctypes.c_float), P(ctypes.c_float), P(ctypes.c_float),
                                           P(ctypes.c_float), P(ctypes.c_float)]
    L.hcg_mlp_train.argtypes = [ctypes.c_void_p, P(ctypes.c_float), _I, P(ctypes.c_float), P(ctypes.c_float),
                                P(ctypes.c_float), _I, _I, ctypes.c_float, ctypes.c_uint64, P(ctypes.c_float)]
    L.hcg_last_error.restype = ctypes.c_char_p
    L.hcg_compile_only.argtypes = [P(hcg_homotopy_desc), P(ctypes.c_size_t)]
    L.hcg_generate.argtypes = [P(hcg_homotopy_desc), ctypes.c_char_p, ctypes.c_size_t, P(ctypes.c_size_t),
                               P(ctypes.c_long)]
    L.hcg_last_counters.argtypes = [ctypes.c_void_p, P(ctypes.c_uint64)]
    L.hcg_program_kernel_info.argtypes = [ctypes.c_void_p, P(ctypes.c_int32)]
    L.hcg_group_schedule_info.argtypes = [ctypes.c_void_p, ctypes.c_int32, P(ctypes.c_int32)]
    L.hcg_certifier_create.argtypes = [ctypes.c_void_p, P(hcg_system), _I, P(ctypes.c_void_p)]
    L.hcg_certifier_compile_only.argtypes = [P(hcg_system), _I, P(ctypes.c_size_t)]
    L.hcg_certify.argtypes = [ctypes.c_void_p, ctypes.c_void_p, _I, P(_D), _I, P(_D), P(_D), P(_D), P(_I), _I, _I,
                              _I, P(hcg_certificate), P(_D), P(_D), P(_D)]
    L.hcg_track_batch_device.argtypes = [ctypes.c_void_p, ctypes.c_void_p, P(hcg_batch), P(hcg_options),
                                         ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]
    L.hcg_graph_create.argtypes = [ctypes.c_void_p, ctypes.c_void_p, P(hcg_batch), P(hcg_options), ctypes.c_void_p,
                                   ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, P(ctypes.c_void_p)]
    L.hcg_graph_launch.argtypes = [ctypes.c_void_p]
    L.hcg_graph_destroy.argtypes = [ctypes.c_void_p]
    L.hcg_graph_destroy.restype = None
    _lib = L
    return L


def check(rc: int):
    if rc != 0:
        msg = lib().hcg_last_error()
        raise HcgpuError(f"hcgpu error {rc}: {msg.decode() if msg else ''}")


HCGPU_SYMBOLS = [
    "hcg_ctx_create", "hcg_ctx_destroy", "hcg_program_create", "hcg_program_destroy", "hcg_program_source",
    "hcg_default_options", "hcg_track_batch", "hcg_last_kernel_stats", "hcg_mlp_load",
    "hcg_mlp_forward_backward", "hcg_mlp_train", "hcg_last_error", "hcg_compile_only", "hcg_generate",
    "hcg_last_counters", "hcg_track_batch_device", "hcg_program_create_multi", "hcg_compile_multi_only",
    "hcg_track_multi", "hcg_program_create_ex", "hcg_compile_ex", "hcg_certifier_create",
    "hcg_certifier_compile_only", "hcg_certify", "hcg_program_kernel_info", "hcg_track_batch_device_noalloc",
    "hcg_group_schedule_info", "hcg_graph_create", "hcg_graph_launch", "hcg_graph_destroy",
]
