"""solve / EndgameTracker / track on the GPU: the host mirror of the reference API.

solve(F; start_system = :total_degree, gamma)             src/solve.jl:447-526
solve(F; start_parameters, target_parameters)             src/solve.jl:473-481, 753-783
solver_startsolutions                                     src/solve.jl:117-194
EndgameTracker / track(tracker, x; path_number)           src/endgame_tracker.jl:236-254, 902-911
many_solve (all target parameters in one launch here)     src/solve.jl:815-881

Every path runs on the GPU (libhcgpu.so); there is no CPU fallback.
"""
from __future__ import annotations

from typing import Optional, Sequence

import numpy as np

from . import abi
from .gpu import GpuContext, Program, default_context, track_batch
from .homotopies import ParameterHomotopy, StraightLineHomotopy, total_degree
from .polys import System
from .result import PathResult, Result, path_results_from_batch

_TRACKER_KEYS = {"max_steps", "max_step_size", "max_initial_step_size", "min_step_size", "min_rel_step_size",
                 "extended_precision", "terminate_cond", "parameters"}
_PARAM_KEYS = {"a", "beta_a", "beta_omega_p", "beta_tau", "strict_beta_tau", "min_newton_iters"}
_ENDGAME_KEYS = {"endgame_start", "max_endgame_steps", "max_endgame_extended_steps", "min_cond", "min_cond_growth",
                 "min_coord_growth", "zero_is_at_infinity", "at_infinity_check", "only_nonsingular",
                 "singular_min_accuracy", "max_winding_number", "val_finite_tol", "val_at_infinity_tol",
                 "sing_cond", "sing_accuracy", "scaling_threshold", "refine_steps"}


def make_options(tracker_options=None, endgame_options=None, **gpu) -> abi.hcg_options:
    kw = {}
    for src, keys in ((tracker_options or {}, _TRACKER_KEYS | _PARAM_KEYS), (endgame_options or {}, _ENDGAME_KEYS)):
        for k, v in dict(src).items():
            if k == "automatic_differentiation":
                continue  # accepted and unused, as in the reference (src/tracker.jl:95-113)
            if k not in keys:
                raise TypeError(f"unknown option {k}")
            if isinstance(v, bool):
                v = int(v)
            kw[k] = v
    kw.update(gpu)
    o = abi.default_options(**kw)
    o.min_rel_step_size = 0.0  # EndgameTracker forces this (src/endgame_tracker.jl:270)
    return o


class EndgameTracker:
    """A homotopy compiled for the GPU plus options.  track() runs single paths,
    track_batch() whole batches (one persistent launch)."""

    def __init__(self, homotopy, tracker_options=None, options=None, ctx: Optional[GpuContext] = None, **gpu):
        self.homotopy = homotopy
        self.ctx = ctx or default_context()
        self.options = make_options(tracker_options, options, **gpu)
        self._program: Optional[Program] = None

    @property
    def program(self) -> Program:
        if self._program is None:
            self._program = self.ctx.program(self.homotopy)
        return self._program

    def track_batch(self, starts, target_groups=None, group_of_path=None, log_capacity=0):
        return track_batch(self.program, starts, self.options, target_groups, group_of_path, log_capacity)


def track(tracker: EndgameTracker, x, t1: float = 1.0, path_number: Optional[int] = None) -> PathResult:
    if t1 != 1.0:
        raise NotImplementedError("the GPU endgame tracker tracks from t = 1")
    out = tracker.track_batch(np.asarray([x], dtype=np.complex128))
    r = path_results_from_batch(out, [x])[0]
    r.path_number = path_number
    return r


def polyhedral_failed_results(failed, first_path_number=1):
    """PathResults with return code `polyhedral_failed` for toric-stage paths that did not reach
    the generic system, given as (start solution, last point, s) (src/polyhedral.jl:491-511)."""
    out = []
    for k, (start, point, s) in enumerate(failed):
        out.append(PathResult(return_code="polyhedral_failed", solution=np.array(point), t=float(s),
                              accuracy=float("nan"), residual=float("nan"), singular=False,
                              condition_jacobian=float("nan"), winding_number=None, extended_precision=False,
                              path_number=first_path_number + k, start_solution=np.array(start),
                              last_path_point=(np.array(point), float(s)), valuation=None, ω=float("nan"),
                              μ=float("nan"), accepted_steps=0, rejected_steps=0, steps_eg=0,
                              extended_precision_used=False, newton_iters=0))
    return out


class Solver:
    """Solver(tracker; seed, start_system), src/solve.jl:41-51."""

    def __init__(self, tracker: EndgameTracker, seed=None, start_system=None, failed_starts=()):
        self.tracker = tracker
        self.seed = seed
        self.start_system = start_system
        # paths that failed before reaching this tracker's start system (the toric stage of a
        # polyhedral start system): reported as `polyhedral_failed` path results
        self.failed_starts = list(failed_starts)

    def solve(self, starts, target_parameters=None):
        starts = np.asarray(starts, dtype=np.complex128)
        H = self.tracker.homotopy
        if target_parameters is None:
            out = self.tracker.track_batch(starts)
            prs = path_results_from_batch(out, starts)
            prs += polyhedral_failed_results(self.failed_starts, len(prs) + 1)
            return Result(prs, seed=self.seed, start_system=self.start_system)
        # many parameters: every (target, start) pair in one launch
        Qorig = np.asarray(target_parameters, dtype=np.complex128).reshape(-1, H.original.nparams)
        Q = H.target_matrix(Qorig)
        S = len(starts)
        all_starts = np.tile(starts, (len(Q), 1))
        groups = np.repeat(np.arange(len(Q), dtype=np.int32), S)
        out = self.tracker.track_batch(all_starts, target_groups=Q, group_of_path=groups)
        prs = path_results_from_batch(out, all_starts)
        results = []
        for g in range(len(Q)):
            chunk = prs[g * S:(g + 1) * S]
            for k, r in enumerate(chunk):
                r.path_number = k + 1
            results.append((Result(chunk, seed=self.seed), Qorig[g]))
        return results


def solver_startsolutions(F: System, starts=None, *, start_system="total_degree", gamma=None, seed=None,
                          start_parameters=None, target_parameters=None, tracker_options=None,
                          endgame_options=None, lift_affine=False, **gpu):
    if seed is None:
        seed = int(np.random.default_rng().integers(0, 2**32))
    if start_parameters is not None:
        if target_parameters is None:
            raise ValueError("target_parameters required")
        q = np.asarray(target_parameters, dtype=np.complex128)
        q0 = q if q.ndim == 1 else q[0]
        H = ParameterHomotopy(F, start_parameters, q0, lift_affine=lift_affine)
        T = EndgameTracker(H, tracker_options, endgame_options, **gpu)
        return Solver(T, seed=seed), starts
    if start_system == "polyhedral":
        # src/solve.jl: solver_startsolutions(F; start_system = :polyhedral, only_non_zero):
        # the toric stage runs here (one launch per mixed cell); the returned solver tracks
        # the coefficient homotopy from the generic system to F, with the endgame
        from .polyhedral import polyhedral
        from .polyhedral import _gpu_toric_tracker
        P = polyhedral(F if target_parameters is None else F.fix_parameters(target_parameters),
                       only_non_zero=gpu.pop("only_non_zero", None), only_torus=gpu.pop("only_torus", None), seed=seed)
        ctx = gpu.get("ctx")
        S, _ = P.start_solutions(tracker=_gpu_toric_tracker(ctx))
        T = EndgameTracker(P.homotopy(), tracker_options, endgame_options, **gpu)
        return Solver(T, seed=seed, start_system="polyhedral", failed_starts=P.failed), S
    if start_system != "total_degree":
        raise NotImplementedError(f"start_system={start_system!r}: use 'total_degree' or 'polyhedral'")
    H, S = total_degree(F, gamma=gamma, seed=seed, target_parameters=target_parameters)
    T = EndgameTracker(H, tracker_options, endgame_options, **gpu)
    return Solver(T, seed=seed, start_system="total_degree"), S


def solve(F: System, starts=None, **kw):
    """solve(F; start_system = "total_degree", gamma, seed) -> Result, or
    solve(F, starts; start_parameters = p, target_parameters = q) -> Result /
    [(Result, q_k)] when q is a list of parameter vectors.  A monodromy result as `starts`
    supplies the start solutions and start parameters (its solutions and parameters,
    src/monodromy.jl:351, 391; used so in test/monodromy_test.jl:347-356)."""
    if starts is not None and hasattr(starts, "parameters") and hasattr(starts, "new_per_loop"):
        kw.setdefault("start_parameters", starts.parameters)
        starts = starts.solutions
    tp = kw.get("target_parameters")
    many = kw.get("start_parameters") is not None and tp is not None and np.asarray(tp).ndim == 2
    solver, S = solver_startsolutions(F, starts, **kw)
    if kw.get("start_parameters") is not None:
        if many:
            return solver.solve(S, target_parameters=tp)
        return solver.solve(S)
    return solver.solve(S)
