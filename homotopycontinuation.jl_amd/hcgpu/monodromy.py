"""Monodromy: all solutions of a parameterized system at one parameter value, from one
start pair, by tracking every known solution around random loops in parameter space.

Follows the reference's monodromy_solve (src/monodromy.jl): a loop p -> p1 -> p2 -> p with
p1, p2 drawn by `independent_normal` (MonodromyLoop, :103-117); every known solution is
tracked around each new loop (track, :1514-1597); endpoints back at p that are not known
yet are added; the search stops when `target_solutions_count` solutions are known or
`max_loops_no_progress` loops in a row found nothing new (:1151-1160).  A start pair comes
from a random x and a linear solve for p when the system is linear in its parameters
(find_start_pair_linear_in_params, :855-914), else from Newton's method in (x, p)
(:827-853).

Each loop edge is ONE batched launch over all known solutions (a parameter homotopy from
the edge's start to its end parameters), on the GPU by default.  New solutions are
recognized by their distance to the known ones (relative ∞-norm below `dedup_rtol`).
"""
from __future__ import annotations

from dataclasses import dataclass, field
from typing import Callable, List, Optional, Sequence, Tuple

import numpy as np

from .homotopies import ParameterHomotopy
from .polys import System


def independent_normal(rng: np.random.Generator, p: np.ndarray) -> np.ndarray:
    """Complex Gaussian parameters, re and im iid N(0, 1/2) (as Julia's randn(ComplexF64))."""
    return (rng.standard_normal(p.shape) + 1j * rng.standard_normal(p.shape)) / np.sqrt(2)


def _linear_start_pair(F: System, rng: np.random.Generator) -> Optional[Tuple[np.ndarray, np.ndarray]]:
    """find_start_pair_linear_in_params (src/monodromy.jl:855-880): x0 random, p0 the
    least-squares solution of the linear system F(x0; p) = 0; None when F is not linear in
    its parameters or the linear system has more equations than parameters."""
    n, m = F.nvars, F.nparams
    x0 = (rng.standard_normal(n) + 1j * rng.standard_normal(n)) / np.sqrt(2)
    A = np.zeros((len(F.polys), m), dtype=np.complex128)
    b = np.zeros(len(F.polys), dtype=np.complex128)
    for i, f in enumerate(F.polys):
        for e, c in f.terms.items():
            xe, pe = e[:n], e[n:]
            if sum(pe) > 1:
                return None
            v = c * np.prod(x0 ** np.asarray(xe))
            if sum(pe) == 0:
                b[i] -= v
            else:
                A[i, int(np.argmax(pe))] += v
    if len(F.polys) > m or not np.all(np.any(A != 0, axis=1)):
        return None  # more equations than parameters, or an equation without parameters
    if not np.any(b):
        if len(F.polys) == m:
            return None  # only p = 0 (the reference does not take 0 as a parameter)
        _, _, vh = np.linalg.svd(A)
        N = vh[len(F.polys):].conj().T
        p0 = N @ ((rng.standard_normal(N.shape[1]) + 1j * rng.standard_normal(N.shape[1])) / np.sqrt(2))
    else:
        p0 = np.linalg.lstsq(A, b, rcond=None)[0]
    return x0, p0


def _newton_start_pair(F: System, rng: np.random.Generator, max_tries: int = 1000, rtol: float = 1e-12,
                       max_iters: int = 50) -> Optional[Tuple[np.ndarray, np.ndarray]]:
    """find_start_pair(F::AbstractSystem) (src/monodromy.jl:827-853): Newton's method in
    variable and parameter space from random (x, p), with minimum-norm steps for the
    underdetermined system, restarted from new random points until it converges."""
    n, m = F.nvars, F.nparams
    N = n + m
    J = [[f.diff(k) for k in range(N)] for f in F.polys]
    deg = max(max(sum(e) for e in f.terms) for f in F.polys)
    for _ in range(max_tries):
        z = (rng.standard_normal(N) + 1j * rng.standard_normal(N)) / np.sqrt(2)
        prev = np.inf
        for _ in range(max_iters):
            zl = list(z)
            r = np.array([f(zl) for f in F.polys], dtype=np.complex128)
            Jz = np.array([[d(zl) for d in row] for row in J], dtype=np.complex128)
            dz = np.linalg.lstsq(Jz, r, rcond=None)[0]
            z = z - dz
            step = np.abs(dz).max()
            if step <= rtol * max(1.0, np.abs(z).max()):
                # a stationary point of |F| that is not a zero (rank-deficient Jacobian) is
                # rejected, as the reference's second Newton run at fixed p would reject it
                zl = list(z)
                res = max(abs(f(zl)) for f in F.polys)
                if res <= 1e-9 * max(1.0, np.abs(z).max()) ** deg:
                    return z[:n], z[n:]
                break
            if not np.isfinite(step) or (step > prev and step > 1e-6):
                break
            prev = step
    return None


def find_start_pair(F: System, rng: np.random.Generator) -> Tuple[np.ndarray, np.ndarray]:
    """(x0, p0) with F(x0; p0) = 0 (find_start_pair, src/monodromy.jl:814-853): the linear
    solve when F is linear in its parameters, else Newton's method in (x, p)."""
    if F.nparams == 0:
        raise ValueError("the system has no parameters")
    pair = _linear_start_pair(F, rng)
    if pair is None:
        pair = _newton_start_pair(F, rng)
    if pair is None:
        raise ValueError("no start pair found")
    return pair


@dataclass
class MonodromyResult:
    solutions: np.ndarray                 # (nsolutions, n) at `parameters`
    parameters: np.ndarray
    returncode: str                       # "success" (target reached) or "heuristic_stop"
    loops: int = 0
    tracked_paths: int = 0
    new_per_loop: List[int] = field(default_factory=list)

    def nsolutions(self) -> int:
        return len(self.solutions)


def _new_points(known: np.ndarray, X: np.ndarray, rtol: float) -> np.ndarray:
    """Rows of X farther than rtol (relative ∞-norm) from every row of known and from each
    other (first occurrence kept)."""
    out = []
    pool = known
    for x in X:
        if len(pool):
            d = np.abs(pool - x).max(axis=1) / np.maximum(1.0, np.abs(x).max())
            if d.min() <= rtol:
                continue
        out.append(x)
        pool = np.vstack([pool, x[None, :]]) if len(pool) else x[None, :]
    return np.array(out).reshape(-1, X.shape[1]) if out else np.zeros((0, X.shape[1]), dtype=np.complex128)


# A tracker for one loop edge: (homotopy template, starts, start params, end params) ->
# (return codes, endpoints); parameters are given in the template's tape space.
EdgeTracker = Callable[[ParameterHomotopy, np.ndarray, np.ndarray, np.ndarray], Tuple[np.ndarray, np.ndarray]]


def gpu_edge_tracker(ctx=None, options=None) -> EdgeTracker:
    """Loop edges tracked by libhcgpu (one launch per edge)."""
    from .gpu import default_context, track_batch
    from .solve import make_options
    ctx = ctx or default_context()
    progs = {}

    def track(H, starts, p, q):
        prog = progs.get(id(H))
        if prog is None:
            prog = progs[id(H)] = ctx.program(H)
        out = track_batch(prog, starts, options or make_options(), target_groups=q[None, :], start_params=p)
        return out.code, out.solution

    return track


def monodromy_solve(F: System, start_solutions: Optional[Sequence[Sequence[complex]]] = None,
                    start_parameters: Optional[Sequence[complex]] = None, *, target_solutions_count: Optional[int] = None,
                    max_loops_no_progress: int = 5, max_loops: int = 100, seed: int = 0, lift_affine: bool = False,
                    dedup_rtol: float = 1e-8, tracker: Optional[EdgeTracker] = None) -> MonodromyResult:
    """monodromy_solve(F, [x0], p0; target_solutions_count, max_loops_no_progress, seed)."""
    rng = np.random.default_rng(seed)
    if start_solutions is None:
        x0, p0 = find_start_pair(F, rng)
        known = x0[None, :]
    else:
        if start_parameters is None:
            raise ValueError("start_parameters are needed with start_solutions")
        p0 = np.asarray(start_parameters, dtype=np.complex128)
        known = np.asarray(start_solutions, dtype=np.complex128).reshape(-1, F.nvars)
    H = ParameterHomotopy(F, p0, [p0], lift_affine=lift_affine)
    tracker = tracker or gpu_edge_tracker()
    res = MonodromyResult(known, p0, "heuristic_stop")
    no_change = 0
    for loop in range(max_loops):
        if target_solutions_count is not None and len(known) >= target_solutions_count:
            break
        p1, p2 = independent_normal(rng, p0), independent_normal(rng, p0)
        X = known
        for pa, pb in ((p0, p1), (p1, p2), (p2, p0)):
            if not len(X):
                break
            codes, sols = tracker(H, X, H.target_matrix(pa)[0], H.target_matrix(pb)[0])
            res.tracked_paths += len(X)
            X = sols[np.asarray(codes) == 1]
        new = _new_points(known, X, dedup_rtol)
        known = np.vstack([known, new]) if len(new) else known
        res.loops += 1
        res.new_per_loop.append(len(new))
        no_change = 0 if len(new) else no_change + 1
        if target_solutions_count is None and no_change >= max_loops_no_progress:
            break
    res.solutions = known
    if target_solutions_count is not None and len(known) >= target_solutions_count:
        res.returncode = "success"
    return res
