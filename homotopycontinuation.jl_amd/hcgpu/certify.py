"""Certification of approximate zeros: certify(F, solutions, p) -> CertificationResult.

Host side of src/certification.jl.  The per-candidate work (Newton refinement, the
approximate inverse of the Jacobian and the ε-inflation Krawczyk test in FP64 interval
arithmetic, certify_solution / ε_inflation_krawczyk, :1029-1114 and :1225-1287) runs in one
libhcgpu launch over all candidates (hcg_certify).  This module supplies what the reference
does around it:
  * parameters as enclosures: Fractions (and ints) become intervals that contain the exact
    rational, as Interval{Float64}(::Rational) does (src/interval_arithmetic.jl:24-25);
  * is_real(F) (src/model_kit/abstract_system_homotopy.jl:85-89): F at a random real point
    (at the parameter point) has real values;
  * duplicates: a certified box that overlaps the box of an earlier distinct certificate is
    a duplicate of it (add_certificate!, src/certification.jl:351-375), in candidate order;
  * CertificationResult and its counts (:436-553), is_positive (:141-163), save (:586-602).
Candidates that fail the FP64 test are retried once with the residual F(x̃) enclosed in
double-double balls (prec 106); the reference retries them with Arb balls of growing
precision instead (extended_prec_certify_solution, :1116-1218).
"""
from __future__ import annotations

import math
from dataclasses import dataclass
from fractions import Fraction
from typing import List, Optional, Sequence, Tuple

import numpy as np

from .polys import System


@dataclass
class SolutionCertificate:
    """One candidate's certificate (SolutionCertificate, src/certification.jl:69-81).
    I: the certified box x₁ and I0: the inflated box x₀, as (n, 4) arrays of
    (re lo, re hi, im lo, im hi); a certified box holds exactly one zero of F."""
    solution_candidate: np.ndarray
    certified: bool
    real: bool
    complex: bool
    index: int                       # 1-based position among the candidates
    I: np.ndarray
    I0: np.ndarray
    solution: Optional[np.ndarray]   # midpoint of I (None unless certified)
    newton_code: int = -1
    contraction: float = math.nan
    prec: int = 53                   # 53: FP64 residual enclosure, 106: the double-double retry

    def is_certified(self) -> bool:
        return self.certified

    def is_real(self) -> bool:
        return self.real

    def is_complex(self) -> bool:
        return self.complex

    def is_positive(self, i: Optional[int] = None) -> bool:
        """Certified, real, and the real part of coordinate i (1-based; all if None) is
        certainly positive."""
        if not (self.certified and self.real):
            return False
        lo = self.I[:, 0]
        return bool(np.all(lo > 0)) if i is None else bool(lo[i - 1] > 0)


class CertificationResult:
    """Certificates in candidate order and duplicates (lists of 1-based candidate indices, the
    distinct representative first) (CertificationResult, src/certification.jl:436-441).  The
    per-candidate arrays of the launch are kept as arrays (`certified`, `real`, `complex`,
    `prec`, boxes `box1` / `box0`); SolutionCertificate objects are built on first access."""

    def __init__(self, S: np.ndarray, out, duplicates: List[List[int]], kernel_ms: float = 0.0):
        self.candidates = S
        self.certified = out.certified.astype(bool)
        self.real = out.real.astype(bool)
        self.complex = out.complex_.astype(bool)
        self.prec = out.prec.astype(np.int32)
        self.newton_code = out.newton_code
        self.contraction = out.contraction
        self.box1, self.box0 = out.box1, out.box0
        self.duplicates = duplicates
        self.kernel_ms = kernel_ms
        self._certs: Optional[List[SolutionCertificate]] = None

    def _cert(self, k: int) -> SolutionCertificate:
        ok = bool(self.certified[k])
        b1 = self.box1[k]
        mid = ((b1[:, 0] + b1[:, 1]) / 2 + 1j * (b1[:, 2] + b1[:, 3]) / 2) if ok else None
        return SolutionCertificate(solution_candidate=self.candidates[k], certified=ok, real=bool(self.real[k]),
                                   complex=bool(self.complex[k]), index=k + 1, I=b1, I0=self.box0[k], solution=mid,
                                   newton_code=int(self.newton_code[k]), contraction=float(self.contraction[k]),
                                   prec=int(self.prec[k]))

    @property
    def certificates(self) -> List[SolutionCertificate]:
        if self._certs is None:
            self._certs = [self._cert(k) for k in range(len(self.candidates))]
        return self._certs

    def _dups(self, flags) -> int:
        return sum(len(d) - 1 for d in self.duplicates if flags[d[0] - 1])

    def ncertified(self) -> int:
        return int(self.certified.sum())

    def nreal_certified(self) -> int:
        return int((self.certified & self.real).sum())

    def ncomplex_certified(self) -> int:
        return int((self.certified & self.complex).sum())

    def ndistinct_certified(self) -> int:
        return self.ncertified() - sum(len(d) - 1 for d in self.duplicates)

    def ndistinct_real_certified(self) -> int:
        return self.nreal_certified() - self._dups(self.real)

    def ndistinct_complex_certified(self) -> int:
        return self.ncomplex_certified() - self._dups(self.complex)

    def distinct_certificates(self) -> List[SolutionCertificate]:
        drop = {k for d in self.duplicates for k in d[1:]}
        return [c for c in self.certificates if c.index not in drop]

    def distinct_solutions(self) -> List[np.ndarray]:
        return [c.solution for c in self.distinct_certificates() if c.certified]

    def __str__(self):
        return ("CertificationResult\n===================\n"
                f"• {len(self.candidates)} solution candidates given\n"
                f"• {self.ncertified()} certified solution intervals ({self.nreal_certified()} real, "
                f"{self.ncomplex_certified()} complex)\n"
                f"• {self.ndistinct_certified()} distinct certified solution intervals "
                f"({self.ndistinct_real_certified()} real, {self.ndistinct_complex_certified()} complex)")

    def save(self, filename: str):
        """Text file: the summary, then per certificate its index, flags and the box x₁."""
        with open(filename, "w") as fh:
            fh.write(str(self) + "\n\n")
            for c in self.certificates:
                fh.write(f"index {c.index}: certified={c.certified} real={c.real} complex={c.complex}\n")
                for row in c.I:
                    fh.write(f"  [{row[0]!r}, {row[1]!r}] + [{row[2]!r}, {row[3]!r}]im\n")


def _up_float(f: Fraction) -> float:
    """The smallest double >= f."""
    d = float(f)
    return d if Fraction(d) >= f else math.nextafter(d, math.inf)


def _part(x):
    """(point, lo, hi, dd hi, dd lo, dd radius) of one real parameter: a float is exact; a
    Fraction (or int) becomes its nearest double, the interval quotient of numerator and
    denominator (Interval{Float64}(::Rational), src/interval_arithmetic.jl:24-25), and a
    double-double ball around the exact rational."""
    if isinstance(x, (Fraction, int)) and not isinstance(x, bool):
        f = Fraction(x)
        point = float(f)
        hi = point
        lo = float(f - Fraction(hi))
        rad = _up_float(abs(f - Fraction(hi) - Fraction(lo)))
        if f.denominator == 1 and Fraction(point) == f:
            return point, point, point, hi, lo, rad
        num, den = float(f.numerator), float(f.denominator)
        if Fraction(num) == f.numerator and Fraction(den) == f.denominator:
            q = num / den
            return point, math.nextafter(q, -math.inf), math.nextafter(q, math.inf), hi, lo, rad
        # numerator or denominator not a double: the tightest enclosure of the exact value
        d = Fraction(point)
        blo = point if d <= f else math.nextafter(point, -math.inf)
        bhi = point if d >= f else math.nextafter(point, math.inf)
        return point, blo, bhi, hi, lo, rad
    x = float(x)
    return x, x, x, x, 0.0, 0.0


def _enclose(v):
    if isinstance(v, complex):
        re, im = _part(v.real), _part(v.imag)
    elif isinstance(v, tuple) and len(v) == 2:  # (re, im), each a float or a Fraction
        re, im = _part(v[0]), _part(v[1])
    else:
        re, im = _part(v), _part(0.0)
    return complex(re[0], im[0]), (re[1], re[2], im[1], im[2]), (re[3], re[4], re[5], im[3], im[4], im[5])


def parameter_boxes(p: Sequence) -> Tuple[np.ndarray, np.ndarray, np.ndarray]:
    """Parameter point (complex), FP64 box (nparams x 4: re lo, re hi, im lo, im hi) and
    double-double ball (nparams x 6: re hi, re lo, re radius, im hi, im lo, im radius) of a
    parameter vector whose entries are complex/floats (exact) or Fractions (enclosed)."""
    if not len(p):
        return np.zeros(0, np.complex128), np.zeros((0, 4)), np.zeros((0, 6))
    pts, boxes, balls = zip(*[_enclose(v) for v in p])
    return (np.asarray(pts, dtype=np.complex128), np.asarray(boxes, dtype=np.float64).reshape(-1, 4),
            np.asarray(balls, dtype=np.float64).reshape(-1, 6))


def is_real_system(F: System, p: Optional[np.ndarray] = None, seed: int = 0) -> bool:
    """is_real(F): F at a random real point has real values (at the parameter point p)."""
    rng = np.random.default_rng(seed)
    x = list(rng.standard_normal(F.nvars))
    args = x + (list(p) if p is not None and len(p) else [])
    return all(complex(f(args)).imag == 0 for f in F.polys)


def duplicate_groups(certified: np.ndarray, box: np.ndarray) -> List[List[int]]:
    """Candidates whose certified box overlaps the box of an earlier distinct certified
    candidate, grouped by that representative (1-based indices, representative first).
    box: (nsol, n, 4).  Boxes overlap when every coordinate's real and imaginary intervals
    intersect."""
    idx = np.flatnonzero(certified)
    if len(idx) < 2:
        return []
    B = box[idx]
    key_lo, key_hi = B[:, 0, 0], B[:, 0, 1]
    width = float(np.max(key_hi - key_lo)) if len(B) else 0.0
    order = np.argsort(key_lo, kind="stable")
    lo_sorted = key_lo[order]
    # candidates of each box: key_lo in [lo - width, hi]
    start = np.searchsorted(lo_sorted, key_lo - width, side="left")
    stop = np.searchsorted(lo_sorted, key_hi, side="right")
    rep = {}          # position in idx -> representative position
    groups = {}
    for a in range(len(idx)):
        best = None
        for s in order[start[a]:stop[a]]:
            if s >= a or s in rep:
                continue  # only earlier distinct certificates
            if (np.all(B[s, :, 0] <= B[a, :, 1]) and np.all(B[a, :, 0] <= B[s, :, 1])
                    and np.all(B[s, :, 2] <= B[a, :, 3]) and np.all(B[a, :, 2] <= B[s, :, 3])):
                best = s if best is None else min(best, s)
        if best is not None:
            rep[a] = best
            groups.setdefault(best, [int(idx[best]) + 1]).append(int(idx[a]) + 1)
    return [groups[k] for k in sorted(groups)]


def _candidates(X, n) -> np.ndarray:
    if hasattr(X, "path_results") or hasattr(X, "results"):
        S = X.solutions(only_nonsingular=False)
        return np.asarray(S, dtype=np.complex128).reshape(-1, n)
    if hasattr(X, "solutions") and isinstance(getattr(X, "solutions"), np.ndarray):  # MonodromyResult
        return np.asarray(X.solutions, dtype=np.complex128).reshape(-1, n)
    if hasattr(X, "solution"):  # one PathResult
        X = [X]
    if isinstance(X, (list, tuple)) and X and hasattr(X[0], "solution"):  # PathResults
        X = [r.solution for r in X]
    return np.asarray(X, dtype=np.complex128).reshape(-1, n)


def certify(F: System, X, p: Optional[Sequence] = None, *, target_parameters: Optional[Sequence] = None,
            refine: bool = True, extended: bool = True, check_distinct: bool = True, ctx=None,
            certifier=None) -> CertificationResult:
    """certify(F, solutions, [p]): certify the candidates X (an array of points, a Result,
    PathResults, or a MonodromyResult, whose parameters are used when p is not given) as
    zeros of F(x; p), on the GPU in one launch (src/certification.jl:1411-1560)."""
    n = F.nvars
    if len(F) != n:
        raise ValueError("We can only certify solutions to square systems.")
    if p is None:
        p = target_parameters
    if p is None and hasattr(X, "parameters") and F.nparams:
        p = list(X.parameters)
    if p is None and F.nparams:
        raise ValueError("The given system expects parameters but none are given.")
    S = _candidates(X, n)
    pts, boxes, balls = parameter_boxes(list(p) if p is not None else [])
    if F.nparams and len(pts) != F.nparams:
        raise ValueError(f"expected {F.nparams} parameters, got {len(pts)}")
    real_system = is_real_system(F, pts)
    if certifier is None:
        from .gpu import default_context
        certifier = (ctx or default_context()).certifier(F)
    has = F.nparams > 0
    out = certifier.certify(S, pts if has else None, boxes if has else None, balls if has else None, refine=refine,
                            extended=extended, real_system=real_system)
    return assemble(S, out, check_distinct=check_distinct, kernel_ms=certifier.ctx.last_kernel_ms())


def certify_groups(F: System, S, parameter_sets: Sequence[Sequence], group_of, *, refine: bool = True,
                   extended: bool = True, check_distinct: bool = True, ctx=None, certifier=None) -> CertificationResult:
    """Certify candidates of F(x; p_g) for several parameter values p_g in ONE launch:
    candidate k is checked against parameter set group_of[k].  Duplicates are looked for
    within each group.  This is the batch form of calling certify once per parameter value
    (e.g. on every target of a many-parameter solve)."""
    n = F.nvars
    if len(F) != n or not F.nparams:
        raise ValueError("a square system with parameters is needed")
    S = np.asarray(S, dtype=np.complex128).reshape(-1, n)
    G = np.asarray(group_of, dtype=np.int32).reshape(len(S))
    enc = [parameter_boxes(list(p)) for p in parameter_sets]
    pts = np.stack([e[0] for e in enc])
    boxes = np.stack([e[1] for e in enc])
    balls = np.stack([e[2] for e in enc])
    real_system = all(is_real_system(F, q) for q in pts)
    if certifier is None:
        from .gpu import default_context
        certifier = (ctx or default_context()).certifier(F)
    out = certifier.certify(S, pts, boxes, balls, refine=refine, extended=extended, real_system=real_system, group_of=G)
    res = assemble(S, out, check_distinct=False, kernel_ms=certifier.ctx.last_kernel_ms())
    if check_distinct:
        dups = []
        for g in np.unique(G):
            idx = np.flatnonzero(G == g)
            for grp in duplicate_groups(out.certified[idx].astype(bool), out.box1[idx]):
                dups.append([int(idx[k - 1]) + 1 for k in grp])
        res.duplicates = sorted(dups)
    return res


def assemble(S: np.ndarray, out, check_distinct: bool = True, kernel_ms: float = 0.0) -> CertificationResult:
    """CertificationResult from the per-candidate arrays of a certification launch."""
    dups = duplicate_groups(out.certified.astype(bool), out.box1) if check_distinct else []
    return CertificationResult(S, out, dups, kernel_ms)


class DistinctCertifiedSolutions:
    """A growing set of distinct certified solutions of F(x; p) (DistinctCertifiedSolutions,
    src/certification.jl:1561-1750).  add_solution certifies one candidate; add_solutions
    certifies a batch in one launch and then inserts the certificates in order, which gives
    each candidate the status sequential add_solution calls would give it."""

    def __init__(self, F: System, params: Optional[Sequence] = None, *, ctx=None, certifier=None):
        if params is None and F.nparams:
            raise ValueError("The given system expects parameters but none are given.")
        self.F = F
        self.params = None if params is None else list(params)
        self._pts, self._boxes, self._balls = parameter_boxes(self.params or [])
        self._real = is_real_system(F, self._pts)
        if certifier is None:
            from .gpu import default_context
            certifier = (ctx or default_context()).certifier(F)
        self._certifier = certifier
        self._certs: List[SolutionCertificate] = []
        self._box = np.zeros((0, F.nvars, 4))
        self.processed = self.duplicates = self.not_certified = 0

    def __len__(self):
        return len(self._certs)

    def ncertified_distinct(self) -> int:
        return len(self._certs)

    def nprocessed(self) -> int:
        return self.processed

    def nduplicates(self) -> int:
        return self.duplicates

    def nnotcertified(self) -> int:
        return self.not_certified

    def stats(self):
        return dict(processed=self.processed, certified_distinct=len(self._certs), duplicates=self.duplicates,
                    not_certified=self.not_certified)

    def certificates(self) -> List[SolutionCertificate]:
        return list(self._certs)

    def solutions(self) -> List[np.ndarray]:
        return [c.solution for c in self._certs]

    def empty(self):
        self._certs, self._box = [], np.zeros((0, self.F.nvars, 4))
        self.processed = self.duplicates = self.not_certified = 0
        return self

    def _containing(self, x: np.ndarray) -> Optional[SolutionCertificate]:
        """A distinct certificate whose box contains the point x (guaranteed_duplicate_certificate)."""
        if not len(self._certs):
            return None
        B = self._box
        inside = ((B[:, :, 0] <= x.real) & (x.real <= B[:, :, 1]) & (B[:, :, 2] <= x.imag) & (x.imag <= B[:, :, 3])).all(axis=1)
        k = np.flatnonzero(inside)
        return self._certs[int(k[0])] if len(k) else None

    def _overlapping(self, b: np.ndarray) -> Optional[SolutionCertificate]:
        if not len(self._certs):
            return None
        B = self._box
        ov = ((B[:, :, 0] <= b[:, 1]) & (b[:, 0] <= B[:, :, 1]) & (B[:, :, 2] <= b[:, 3]) & (b[:, 2] <= B[:, :, 3])).all(axis=1)
        k = np.flatnonzero(ov)
        return self._certs[int(k[0])] if len(k) else None

    def _insert(self, c: SolutionCertificate):
        """add_certificate! (src/certification.jl:351-375): (added, representative)."""
        other = self._overlapping(c.I)
        if other is not None:
            return False, other
        self._certs.append(c)
        self._box = np.concatenate([self._box, c.I[None]], axis=0)
        return True, c

    def add_solutions(self, S, indices: Optional[Sequence[int]] = None):
        """Add candidates S (distinct_certified_solutions!, src/certification.jl:1768-1860);
        returns (added, status, representative index, certified solution) per candidate,
        status one of "certified_distinct", "duplicate", "not_certified"."""
        S = np.asarray(S, dtype=np.complex128).reshape(-1, self.F.nvars)
        idx = list(range(1, len(S) + 1)) if indices is None else [int(i) for i in indices]
        out: List = [None] * len(S)
        todo = []
        for k, x in enumerate(S):
            c = self._containing(x)
            if c is not None:
                self.processed += 1
                self.duplicates += 1
                out[k] = (False, "duplicate", c.index, None)
            else:
                todo.append(k)
        if todo:
            has = self.F.nparams > 0
            o = self._certifier.certify(S[todo], self._pts if has else None, self._boxes if has else None,
                                        self._balls if has else None, real_system=self._real)
            r = assemble(S[todo], o, check_distinct=False)
            for j, k in enumerate(todo):
                c = r.certificates[j]
                c.index = idx[k]
                self.processed += 1
                if not c.certified:
                    self.not_certified += 1
                    out[k] = (False, "not_certified", 0, None)
                    continue
                added, rep = self._insert(c)
                if added:
                    out[k] = (True, "certified_distinct", c.index, c.solution)
                else:
                    self.duplicates += 1
                    out[k] = (False, "duplicate", rep.index, None)
        return out

    def add_solution(self, sol, i: int = 0):
        """add_solution! (src/certification.jl:1673-1719)."""
        return self.add_solutions([sol], [i])[0]

    def merge(self, other: "DistinctCertifiedSolutions") -> "DistinctCertifiedSolutions":
        """merge! (src/certification.jl:1735-1745): other's certificates added to this set."""
        for c in other.certificates():
            added, _ = self._insert(c)
            self.processed += 1
            if not added:
                self.duplicates += 1
        return self


def distinct_certified_solutions(F: System, S, p: Optional[Sequence] = None, **kw) -> DistinctCertifiedSolutions:
    """DistinctCertifiedSolutions of the candidates S (distinct_certified_solutions,
    src/certification.jl:1752-1766)."""
    d = DistinctCertifiedSolutions(F, p, **kw)
    d.add_solutions(S)
    return d
