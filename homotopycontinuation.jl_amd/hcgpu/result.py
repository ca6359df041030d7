"""PathResult / Result: the output contract of solve (host side).

PathResult fields      src/path_result.jl:76-98; predicates :220-298
Result + multiplicity  src/result.jl:38-111, src/unique_points.jl:314-340
counting API           src/result.jl:627-841
"""
from __future__ import annotations

from typing import Dict, List, Optional, Sequence

import numpy as np

from .abi import RETURN_CODES


class PathResult:
    __slots__ = ("return_code", "solution", "t", "accuracy", "residual", "singular", "multiplicity",
                 "condition_jacobian", "winding_number", "extended_precision", "path_number", "start_solution",
                 "last_path_point", "valuation", "ω", "μ", "accepted_steps", "rejected_steps", "steps_eg",
                 "extended_precision_used", "newton_iters")

    def __init__(self, **kw):
        self.multiplicity = 1
        for k, v in kw.items():
            setattr(self, k, v)

    # predicates, src/path_result.jl:220-265
    def is_success(self):
        return self.return_code == "success"

    def is_at_infinity(self):
        return self.return_code in ("at_infinity", "at_zero")

    def is_excess_solution(self):
        return self.return_code == "excess_solution"

    def is_failed(self):
        return not (self.is_at_infinity() or self.is_success() or self.is_excess_solution())

    def is_finite(self):
        return self.is_success()

    def is_singular(self):
        return bool(self.singular)

    def is_nonsingular(self):
        return not self.is_singular() and self.is_success()

    def is_real(self, atol: float = 1e-6, rtol: float = 0.0):
        m = float(np.max(np.abs(np.imag(self.solution))))
        if m < atol:
            return True
        if rtol == 0:
            return False
        return m < rtol * float(np.sum(np.abs(self.solution)))

    def steps(self):
        return self.accepted_steps + self.rejected_steps

    def __repr__(self):
        return f"PathResult(return_code={self.return_code}, path_number={self.path_number})"


def path_results_from_batch(out, starts, offset: int = 0) -> List[PathResult]:
    """PathResult(endgame_tracker, start, path_number) for each path of a batch
    (src/endgame_tracker.jl:847-888); path numbers are 1-based."""
    res = []
    for k in range(len(out.code)):
        w = int(out.winding_number[k])
        res.append(PathResult(
            return_code=RETURN_CODES[int(out.code[k])],
            solution=out.solution[k].copy(),
            t=float(out.t[k]),
            accuracy=float(out.accuracy[k]),
            residual=float(out.residual[k]),
            singular=bool(out.singular[k]),
            condition_jacobian=float(out.cond[k]),
            winding_number=w if w != 0 else None,
            extended_precision=bool(out.extended_precision[k]),
            path_number=offset + k + 1,
            start_solution=np.asarray(starts[k]).copy(),
            last_path_point=(out.last_point[k].copy(), float(out.last_t[k])),
            valuation=out.valuation[k].copy() if out.valuation_valid[k] else None,
            ω=float(out.omega[k]),
            μ=float(out.mu[k]),
            accepted_steps=int(out.accepted_steps[k]),
            rejected_steps=int(out.rejected_steps[k]),
            steps_eg=int(out.steps_eg[k]),
            extended_precision_used=bool(out.extended_precision_used[k]),
            newton_iters=int(out.newton_iters[k]),
        ))
    return res


def multiplicities(points: Sequence[np.ndarray], atol: float = 1e-14, rtol: float = 1e-8) -> List[List[int]]:
    """Clusters of points within max(atol, rtol * ‖v‖) in the Euclidean norm
    (_multiplicities, src/unique_points.jl:314-340); returns clusters of size >= 2
    as lists of indices, the first index being the cluster's representative."""
    reps: List[int] = []
    mults: Dict[int, List[int]] = {}
    for i, v in enumerate(points):
        v = np.asarray(v)
        tol = max(atol, rtol * float(np.linalg.norm(v)))
        found = None
        for r in reps:
            if float(np.linalg.norm(points[r] - v)) <= tol:
                found = r
                break
        if found is None:
            reps.append(i)
        else:
            mults.setdefault(found, [found]).append(i)
    return list(mults.values())


class Result:
    """Result(path_results; seed, start_system), src/result.jl:95-111."""

    def __init__(self, path_results: List[PathResult], seed=None, start_system=None, tracked_paths=None):
        self.path_results = path_results
        self.seed = seed
        self.start_system = start_system
        self.tracked_paths = len(path_results) if tracked_paths is None else tracked_paths
        sing = [r for r in path_results if r.is_singular()]
        clusters = multiplicities([r.solution for r in sing])
        self.multiplicities: Dict[int, List[List[int]]] = {}
        self.multiple_indicator = set()
        for cl in clusters:
            pns = [sing[i].path_number for i in cl]
            self.multiplicities.setdefault(len(cl), []).append(pns)
            for pn in pns[1:]:
                self.multiple_indicator.add(pn)
        by_pn = {r.path_number: r for r in path_results}
        for k, cls in self.multiplicities.items():
            for cl in cls:
                for pn in cl:
                    r = by_pn[pn]
                    r.multiplicity = max(k, r.winding_number or 1)

    def __len__(self):
        return len(self.path_results)

    def __iter__(self):
        return iter(self.path_results)

    def __getitem__(self, i):
        return self.path_results[i]

    def is_multiple_result(self, r: PathResult):
        return r.path_number in self.multiple_indicator

    # nresults, src/result.jl:627-658
    def nresults(self, only_real=False, real_atol=1e-6, real_rtol=0.0, only_nonsingular=False,
                 only_singular=False, only_finite=True, multiple_results=False):
        return sum(1 for r in self.path_results
                   if (not only_real or r.is_real(real_atol, real_rtol))
                   and (not only_nonsingular or r.is_nonsingular())
                   and (not only_singular or r.is_singular())
                   and (not only_finite or r.is_finite())
                   and (multiple_results or not self.is_multiple_result(r)))

    def results(self, only_real=False, real_atol=1e-6, real_rtol=0.0, only_nonsingular=False,
                only_singular=False, only_finite=True, multiple_results=False):
        return [r for r in self.path_results
                if (not only_real or r.is_real(real_atol, real_rtol))
                and (not only_nonsingular or r.is_nonsingular())
                and (not only_singular or r.is_singular())
                and (not only_finite or r.is_finite())
                and (multiple_results or not self.is_multiple_result(r))]

    def solutions(self, only_nonsingular=True, **kw):
        return [r.solution for r in self.results(only_nonsingular=only_nonsingular, **kw)]

    def nsolutions(self, only_nonsingular=True, **kw):
        return self.nresults(only_nonsingular=only_nonsingular, **kw)

    def nsingular(self, counting_multiplicities=False, **kw):
        return self.nresults(only_singular=True, multiple_results=counting_multiplicities, **kw)

    def nat_infinity(self):
        return sum(1 for r in self.path_results if r.is_at_infinity())

    def nexcess_solutions(self):
        return sum(1 for r in self.path_results if r.is_excess_solution())

    def nfailed(self):
        return sum(1 for r in self.path_results if r.is_failed())

    def nnonsingular(self):
        return sum(1 for r in self.path_results if r.is_nonsingular())

    def nreal(self, **kw):
        return self.nresults(only_real=True, **kw)

    def ntracked(self):
        return self.tracked_paths

    def statistics(self):
        return dict(total=len(self), nonsingular=self.nnonsingular(), singular=self.nsingular(),
                    at_infinity=self.nat_infinity(), failed=self.nfailed(), real=self.nreal())
