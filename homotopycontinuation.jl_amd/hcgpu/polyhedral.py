"""Polyhedral start systems (Huber-Sturmfels): start solutions counted by the mixed volume.

Follows `polyhedral(F; only_non_zero)` and `PolyhedralStartSolutionsIterator`
(src/polyhedral.jl:1-60, 186-260): a random integer lifting of the supports, the fine
mixed cells of the lifted Minkowski sum, for each cell the solutions of its binomial
system, then two homotopies:

1. the toric homotopy of each cell, h_i(x, s) = sum_a c_ia x^a s^(e_ia) with generic
   coefficients c, from s = 0 (only the cell's two terms per equation remain: the binomial
   system) to s = 1 (the generic system G(x; c));
2. the coefficient homotopy from G(x; c) to the target F (a parameter homotopy over the
   coefficient vector, with the endgame).

Both are parameter homotopies of the GPU tracker: stage 1 is a system with one parameter
s that appears with integer powers.  Mixed cells come from MixedSubdivisions.jl in the
reference (an external package, not in the reference tree); here they are enumerated
exactly: every choice of one pair of points per support, its lifted normal solved in
integers, kept when the pair attains the minimum alone in every support.  The cost is the
product of the pair counts, fine for the supports of small systems (cyclic-5: 15^4 pairs).

Without `only_non_zero`, the origin is added to every support that lacks it (coefficient
zero in the target), so the start count bounds all isolated solutions, not only those in
the torus (src/polyhedral.jl:200-225).
"""
from __future__ import annotations

import itertools
import math
from dataclasses import dataclass
from typing import List, Optional, Sequence, Tuple

import numpy as np

from .homotopies import ParameterHomotopy
from .polys import Poly, System

_CHUNK = 1 << 16
# exact enumeration visits every tuple of one point pair per support; beyond this many tuples
# it is refused with an error instead of running for hours
MAX_PAIR_TUPLES = 50_000_000


class EnumerationTooLarge(RuntimeError):
    """The supports have too many point-pair tuples for the exact mixed-cell enumeration."""


def _supports(F: System, only_non_zero: bool):
    if F.nparams:
        raise ValueError("polyhedral: fix the parameters first (F.fix_parameters(p))")
    n = F.nvars
    if len(F.polys) != n:
        raise ValueError("polyhedral: square systems only")
    A, C = [], []
    for f in F.polys:
        exps = [tuple(e[:n]) for e in f.terms]
        coef = [f.terms[e] for e in f.terms]
        if not only_non_zero and (0,) * n not in exps:
            exps.append((0,) * n)
            coef.append(0.0)
        A.append(np.array(exps, dtype=np.int64).reshape(-1, n))
        C.append(np.array(coef, dtype=np.complex128))
    return A, C


@dataclass
class MixedCell:
    """One fine mixed cell: the pair (j0, j1) of each support, |det| of the edge matrix,
    and the integer exponents of s in each term of the cell's toric homotopy."""
    pairs: np.ndarray        # (n, 2) term indices
    volume: int              # |det V|, the cell's number of start solutions
    V: np.ndarray            # (n, n) rows a_j1 - a_j0
    W: np.ndarray            # (n, n) integer, V @ W = volume * I
    s_exponents: List[np.ndarray]


def mixed_cells(A: Sequence[np.ndarray], lifting: Sequence[np.ndarray]) -> List[MixedCell]:
    """Fine mixed cells of the lifted supports (lower facets with inner normal (g, 1)).
    Raises ValueError when the lifting is not generic (a facet with more than two points
    of one support), so the caller can draw another."""
    n = len(A)
    P = [np.array(list(itertools.combinations(range(len(a)), 2)), dtype=np.int64).reshape(-1, 2) for a in A]
    if any(len(p) == 0 for p in P):
        return []
    sizes = [len(p) for p in P]
    total = math.prod(sizes)  # Python integers: no int64 overflow for large supports
    if total > MAX_PAIR_TUPLES:
        raise EnumerationTooLarge(
            f"mixed cells: {total:.3g} point-pair tuples exceed the exact enumeration's limit of "
            f"{MAX_PAIR_TUPLES:.3g}; use start_system='total_degree' for this system")
    cells = []
    for lo in range(0, total, _CHUNK):
        idx = np.stack(np.unravel_index(np.arange(lo, min(total, lo + _CHUNK)), sizes), axis=1)
        m = len(idx)
        V = np.empty((m, n, n))
        rhs = np.empty((m, n))
        J0 = np.empty((m, n), dtype=np.int64)
        J1 = np.empty((m, n), dtype=np.int64)
        for i in range(n):
            j0, j1 = P[i][idx[:, i], 0], P[i][idx[:, i], 1]
            J0[:, i], J1[:, i] = j0, j1
            V[:, i, :] = A[i][j1] - A[i][j0]
            rhs[:, i] = lifting[i][j0] - lifting[i][j1]
        det = np.linalg.det(V)
        ok = np.abs(det) > 0.5
        if not ok.any():
            continue
        V, rhs, det, J0, J1 = V[ok], rhs[ok], det[ok], J0[ok], J1[ok]
        g = np.linalg.solve(V, rhs[..., None])[..., 0]
        keep = np.ones(len(V), dtype=bool)
        for i in range(n):
            val = A[i] @ g.T + lifting[i][:, None]                 # (|A_i|, m)
            beta = val[J0[:, i], np.arange(len(V))]
            d = val - beta
            tol = 1e-7 * (1 + np.abs(val).max(axis=0))
            keep &= (d >= -tol).all(axis=0)
            ties = (np.abs(d) <= tol).sum(axis=0)
            if (keep & (ties > 2)).any():
                raise ValueError("lifting not generic")
        for k in np.flatnonzero(keep):
            cells.append(_exact_cell(A, lifting, np.stack([J0[k], J1[k]], axis=1), int(round(abs(det[k])))))
    return cells


def _exact_cell(A, lifting, pairs, vol) -> MixedCell:
    n = len(A)
    V = np.stack([A[i][pairs[i, 1]] - A[i][pairs[i, 0]] for i in range(n)]).astype(np.int64)
    W = np.rint(vol * np.linalg.inv(V.astype(np.float64))).astype(np.int64)
    if not np.array_equal(V @ W, vol * np.eye(n, dtype=np.int64)):
        raise ValueError("edge matrix too large for exact integer inversion")
    rhs = np.array([lifting[i][pairs[i, 0]] - lifting[i][pairs[i, 1]] for i in range(n)], dtype=np.int64)
    gv = W @ rhs                                                     # vol * g, exact
    num = []
    for i in range(n):
        val = A[i] @ gv + vol * lifting[i]                           # vol * (<a, g> + w(a))
        e = val - val[pairs[i, 0]]
        if (e < 0).any() or (e[pairs[i]] != 0).any() or (np.count_nonzero(e == 0) != 2):
            raise ValueError("cell check failed in exact arithmetic")
        num.append(e)
    g = int(np.gcd.reduce(np.concatenate(num + [np.array([vol])])))
    return MixedCell(pairs, vol, V, W, [e // g for e in num])


def binomial_solutions(cell: MixedCell, A, C) -> np.ndarray:
    """All |det V| solutions of x^(v_i) = b_i, b_i = -c_i,j0 / c_i,j1: log x = V^-1 (log b +
    2 pi i k), k running over Z^n / V Z^n, enumerated as the group generated by the columns
    of W = |det V| V^-1 modulo |det V|."""
    n, d = len(A), cell.volume
    b = np.array([-C[i][cell.pairs[i, 0]] / C[i][cell.pairs[i, 1]] for i in range(n)])
    z0 = np.linalg.solve(cell.V.astype(np.float64), np.log(b))
    seen = {tuple([0] * n)}
    frontier = [np.zeros(n, dtype=np.int64)]
    gens = [cell.W[:, j] % d for j in range(n)]
    while frontier:
        nxt = []
        for y in frontier:
            for w in gens:
                z = tuple((y + w) % d)
                if z not in seen:
                    seen.add(z)
                    nxt.append(np.array(z))
        frontier = nxt
    if len(seen) != d:
        raise ValueError("binomial root count differs from the cell volume")
    Y = np.array(sorted(seen), dtype=np.float64)
    return np.exp(z0[None, :] + 2j * np.pi * Y / d)


class PolyhedralStart:
    """Start data of the polyhedral homotopy of F: lifting, mixed cells, generic
    coefficients, and the two stages (toric per cell, then coefficients)."""

    def __init__(self, F: System, only_non_zero: bool = False, seed: Optional[int] = None, lifting_range: int = 32,
                 liftings: int = 8):
        self.F = F
        self.n = F.nvars
        self.only_non_zero = only_non_zero
        rng = np.random.default_rng(seed)
        self.A, self.target_coefficients = _supports(F, only_non_zero)
        # Among `liftings` generic liftings keep the one whose toric homotopies have the lowest
        # largest power of s: high powers make the toric paths steep near s = 1 (the reason the
        # reference splits such paths in two stages, src/polyhedral.jl:436-490).
        best = None
        attempt = 0
        while (best is None or attempt < liftings) and attempt < 60 + liftings:
            r = lifting_range * 4 ** (attempt // 10)  # small liftings keep the powers low; widen on ties
            attempt += 1
            lifting = [rng.integers(0, r, len(a)).astype(np.int64) for a in self.A]
            try:
                cells = mixed_cells(self.A, lifting)
            except EnumerationTooLarge:
                raise
            except ValueError:
                continue
            top = max((int(max(e.max() for e in c.s_exponents)) for c in cells), default=0)
            if best is None or top < best[0]:
                best = (top, lifting, cells)
        if best is None:
            raise RuntimeError("no generic lifting found")
        self.max_power, self.lifting, self.cells = best
        self.failed = []  # toric paths that did not reach s = 1 (filled by start_solutions)
        self.start_coefficients = [(rng.standard_normal(len(a)) + 1j * rng.standard_normal(len(a))) / np.sqrt(2)
                                   for a in self.A]

    @property
    def mixed_volume(self) -> int:
        return sum(c.volume for c in self.cells)

    def coefficient_system(self) -> System:
        """G(x; c) = sum_a c_ia x^a, one parameter per term."""
        n = self.n
        npar = sum(len(a) for a in self.A)
        polys, k = [], 0
        for a in self.A:
            t = {}
            for row in a:
                e = [0] * (n + npar)
                e[:n] = row
                e[n + k] = 1
                t[tuple(e)] = 1.0
                k += 1
            polys.append(Poly(t, n + npar))
        return System(polys, n, npar)

    def start_parameters(self) -> np.ndarray:
        return np.concatenate(self.start_coefficients)

    def target_parameters(self) -> np.ndarray:
        return np.concatenate(self.target_coefficients)

    def toric_system(self, cell: MixedCell) -> System:
        """h_i(x, s) = sum_a c_ia x^a s^(e_ia): at s = 0 the cell's binomial system."""
        n = self.n
        polys = []
        for i, a in enumerate(self.A):
            t = {}
            for j, row in enumerate(a):
                t[tuple(row) + (int(cell.s_exponents[i][j]),)] = self.start_coefficients[i][j]
            polys.append(Poly(t, n + 1))
        return System(polys, n, 1)

    def start_solutions(self, tracker=None) -> Tuple[np.ndarray, int]:
        """Stage 1 for every cell (one launch per cell): solutions of G(x; c) and the number
        of toric paths that failed.  `tracker(H, starts, p, q)` returns (codes, endpoints) or
        (codes, endpoints, final t); the failed paths are kept in self.failed as
        (start solution, last point, t) for their `polyhedral_failed` path results."""
        tracker = tracker or _gpu_toric_tracker()
        out = []
        self.failed = []
        for cell in self.cells:
            X = binomial_solutions(cell, self.A, self.start_coefficients)
            H = ParameterHomotopy(self.toric_system(cell), [0.0], [1.0])
            r = tracker(H, X, np.array([0.0 + 0j]), np.array([1.0 + 0j]))
            codes, sols = np.asarray(r[0]), r[1]
            ts = np.asarray(r[2]) if len(r) > 2 else np.full(len(X), np.nan)
            ok = codes == 1
            for k in np.flatnonzero(~ok):
                self.failed.append((X[k], sols[k], float(ts[k])))
            out.append(sols[ok])
        S = np.vstack(out) if out else np.zeros((0, self.n), dtype=np.complex128)
        return S, len(self.failed)

    def homotopy(self) -> ParameterHomotopy:
        """Stage 2: the coefficient homotopy from G(x; c) to F."""
        return ParameterHomotopy(self.coefficient_system(), self.start_parameters(), self.target_parameters())


def _gpu_toric_tracker(ctx=None):
    """Toric stage on the GPU: (codes, endpoints, final t) of one launch."""
    from .gpu import default_context, track_batch
    from .solve import make_options
    ctx = ctx or default_context()

    def track(H, starts, p, q):
        out = track_batch(ctx.program(H), starts, make_options(), target_groups=q[None, :], start_params=p)
        return out.code, out.solution, 1.0 - out.t  # s = 1 - t runs from 0 to 1

    return track


def _non_zero(only_non_zero: Optional[bool], only_torus: Optional[bool]) -> bool:
    """only_torus is only the default of only_non_zero (src/polyhedral.jl:303-304, 364-365):
    an explicitly given only_non_zero wins."""
    if only_non_zero is not None:
        return bool(only_non_zero)
    return bool(only_torus) if only_torus is not None else False


def polyhedral(F: System, only_non_zero: Optional[bool] = None, only_torus: Optional[bool] = None,
               seed: Optional[int] = None) -> PolyhedralStart:
    """polyhedral(F; only_non_zero) (src/polyhedral.jl:186-260); `only_torus` is the older
    name of the same switch and only its default (test/polyhedral_test.jl:6-10)."""
    return PolyhedralStart(F, only_non_zero=_non_zero(only_non_zero, only_torus), seed=seed)


def mixed_volume(F: System, seed: Optional[int] = 0) -> int:
    """Mixed volume of the supports of F (mixed_volume, src/polyhedral.jl:1)."""
    return PolyhedralStart(F, only_non_zero=True, seed=seed).mixed_volume


def paths_to_track(F: System, start_system: str = "polyhedral", only_non_zero: Optional[bool] = None,
                   only_torus: Optional[bool] = None) -> int:
    """paths_to_track (src/polyhedral.jl:300, src/total_degree.jl:481): Bezout number, or the mixed volume of the supports
    (with the origin added unless only_non_zero)."""
    if start_system == "total_degree":
        return math.prod(F.degrees())
    return PolyhedralStart(F, only_non_zero=_non_zero(only_non_zero, only_torus), seed=0).mixed_volume
