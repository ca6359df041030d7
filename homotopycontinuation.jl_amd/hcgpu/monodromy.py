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
                    dedup_rtol: float = 1e-8, tracker: Optional[EdgeTracker] = None,
                    parameter_sampler: Optional[Callable[[np.random.Generator, np.ndarray], np.ndarray]] = None
                    ) -> MonodromyResult:
    """monodromy_solve(F, [x0], p0; target_solutions_count, max_loops_no_progress, seed,
    parameter_sampler): parameter_sampler(rng, p0) draws the loop's intermediate parameters
    (default independent_normal)."""
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
        sample = parameter_sampler or independent_normal
        p1, p2 = np.asarray(sample(rng, p0), dtype=np.complex128), np.asarray(sample(rng, p0), dtype=np.complex128)
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


def _substituted_system(F: System) -> System:
    """[F(x, p + λ v); (a·x − 1) λ + t] in variables [x; λ] and parameters [t; p; v; a]
    (the verify_system of verify_solution_completeness, src/monodromy.jl:1711-1715)."""
    from .polys import variables
    n, m = F.nvars, F.nparams
    xs, ps = variables(n + 1, 1 + 2 * m + n)
    x, lam = xs[:n], xs[n]
    t, p, v, a = ps[0], ps[1:1 + m], ps[1 + m:1 + 2 * m], ps[1 + 2 * m:]
    subs = list(x) + [p[k] + lam * v[k] for k in range(m)]
    eqs = []
    for f in F.polys:
        acc = None
        for e, c in f.terms.items():
            term = None
            for sym, k in zip(subs, e):
                if k:
                    fac = sym ** k
                    term = fac if term is None else term * fac
            term = c if term is None else term * c
            acc = term if acc is None else acc + term
        eqs.append(acc)
    lin = a[0] * x[0]
    for i in range(1, n):
        lin = lin + a[i] * x[i]
    eqs.append((lin - 1) * lam + t)
    return System(eqs, n + 1, 1 + 2 * m + n)


def verify_solution_completeness(F: System, sols, q, *, trace_tol: float = 1e-14, seed: int = 0,
                                 tracker: Optional[EdgeTracker] = None, monodromy_options=None) -> Optional[bool]:
    """The trace test of verify_solution_completeness (src/monodromy.jl:1694-1806): True if
    `sols` are all the isolated solutions of F(x; q), False if the trace test fails, None if
    a solution was lost on the way.  Every homotopy runs as a batched launch (tracker)."""
    rng = np.random.default_rng(seed)
    n, m = F.nvars, F.nparams
    sols = np.asarray(sols, dtype=np.complex128).reshape(-1, n)
    q = np.asarray(q, dtype=np.complex128)
    tracker = tracker or gpu_edge_tracker()
    V = _substituted_system(F)
    # start witnesses: n solutions moved to random parameters qq, on the hyperplane a0·x = 1
    qq = (rng.standard_normal(m) + 1j * rng.standard_normal(m)) / np.sqrt(2)
    HF = ParameterHomotopy(F, q, [q])
    codes, X = tracker(HF, sols[:min(n, len(sols))], q, qq)
    Y = X[np.asarray(codes) == 1]
    a0 = np.linalg.lstsq(Y, np.ones(len(Y), dtype=np.complex128), rcond=None)[0]
    base = np.concatenate([q, qq - q, a0])
    start = np.concatenate([[0.0], base]).astype(np.complex128)
    Y1 = np.hstack([Y, np.ones((len(Y), 1), dtype=np.complex128)])
    # more witnesses on the λ ≠ 0 component by monodromy with t = 0
    sampler = lambda r, p: np.concatenate([[0.0], (r.standard_normal(len(p) - 1) +  # noqa: E731
                                                  1j * r.standard_normal(len(p) - 1)) / np.sqrt(2)])
    mres = monodromy_solve(V, Y1, start, seed=seed + 1, tracker=tracker, parameter_sampler=sampler,
                           **(monodromy_options or {}))
    S = np.vstack([np.hstack([sols, np.zeros((len(sols), 1), dtype=np.complex128)]), mres.solutions])
    gamma = complex(rng.standard_normal(), rng.standard_normal()) / np.sqrt(2)
    HV = ParameterHomotopy(V, start, [start])
    mid = start.copy()
    mid[0] = 0.5 * gamma
    end = start.copy()
    end[0] = gamma
    c1, S1 = tracker(HV, S, start, mid)
    if (np.asarray(c1) == 1).sum() != len(S):
        return None
    c2, S2 = tracker(HV, S1, mid, end)
    if (np.asarray(c2) == 1).sum() != len(S):
        return None
    M = np.vstack([np.stack([S.sum(axis=0), S1.sum(axis=0), S2.sum(axis=0)], axis=1), np.ones((1, 3))])
    sv = np.linalg.svd(M, compute_uv=False)
    return bool(sv[2] / sv[0] < trace_tol)
