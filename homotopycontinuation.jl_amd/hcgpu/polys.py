"""Sparse polynomial systems: the symbolic input the tape producer lowers.

A small stand-in for ModelKit's `@var` / `Expression` / `System`
(src/model_kit/symbolic.jl:1012-1038).  Polynomials are dicts mapping exponent
tuples over (variables..., parameters...) to complex coefficients.  Only what
square polynomial systems need: +, -, *, integer powers, differentiation.
"""
from __future__ import annotations

from typing import Dict, Iterable, List, Sequence, Tuple

Exp = Tuple[int, ...]


class Poly:
    """Polynomial over a fixed ring of `nsym` symbols (variables then parameters)."""

    __slots__ = ("terms", "nsym")

    def __init__(self, terms: Dict[Exp, complex], nsym: int):
        self.terms = {e: complex(c) for e, c in terms.items() if c != 0}
        self.nsym = nsym

    @staticmethod
    def const(c, nsym):
        return Poly({(0,) * nsym: complex(c)}, nsym)

    def _lift(self, other):
        if isinstance(other, Poly):
            if other.nsym != self.nsym:
                raise ValueError("polynomials over different symbol sets")
            return other
        return Poly.const(other, self.nsym)

    def __add__(self, other):
        other = self._lift(other)
        t = dict(self.terms)
        for e, c in other.terms.items():
            t[e] = t.get(e, 0) + c
        return Poly(t, self.nsym)

    __radd__ = __add__

    def __neg__(self):
        return Poly({e: -c for e, c in self.terms.items()}, self.nsym)

    def __sub__(self, other):
        return self + (-self._lift(other))

    def __rsub__(self, other):
        return self._lift(other) - self

    def __mul__(self, other):
        other = self._lift(other)
        t: Dict[Exp, complex] = {}
        for e1, c1 in self.terms.items():
            for e2, c2 in other.terms.items():
                e = tuple(a + b for a, b in zip(e1, e2))
                t[e] = t.get(e, 0) + c1 * c2
        return Poly(t, self.nsym)

    __rmul__ = __mul__

    def __pow__(self, k: int):
        if not isinstance(k, int) or k < 0:
            raise ValueError("only non-negative integer powers")
        r = Poly.const(1, self.nsym)
        b = self
        while k:
            if k & 1:
                r = r * b
            b = b * b
            k >>= 1
        return r

    def diff(self, i: int) -> "Poly":
        t: Dict[Exp, complex] = {}
        for e, c in self.terms.items():
            if e[i] > 0:
                e2 = list(e)
                e2[i] -= 1
                t[tuple(e2)] = t.get(tuple(e2), 0) + c * e[i]
        return Poly(t, self.nsym)

    def degree(self, upto: int | None = None) -> int:
        """Total degree in the first `upto` symbols (all symbols when None)."""
        if not self.terms:
            return 0
        k = self.nsym if upto is None else upto
        return max(sum(e[:k]) for e in self.terms)

    def __call__(self, vals: Sequence[complex]) -> complex:
        s = 0j
        for e, c in self.terms.items():
            m = c
            for v, k in zip(vals, e):
                if k:
                    m *= v ** k
            s += m
        return s

    def __repr__(self):
        return f"Poly({len(self.terms)} terms)"


class System:
    """A square polynomial system F(x; p) with n variables and `nparams` parameters."""

    def __init__(self, polys: Sequence[Poly], nvars: int, nparams: int = 0, names=None):
        self.polys = list(polys)
        self.nvars = nvars
        self.nparams = nparams
        self.names = names
        for f in self.polys:
            if f.nsym != nvars + nparams:
                raise ValueError("symbol count mismatch")

    def __len__(self):
        return len(self.polys)

    @property
    def n(self):
        return self.nvars

    def degrees(self) -> List[int]:
        return [f.degree(self.nvars) for f in self.polys]

    def jacobian(self) -> List[List[Poly]]:
        return [[f.diff(j) for j in range(self.nvars)] for f in self.polys]

    def fix_parameters(self, p: Sequence[complex]) -> "System":
        """Substitute parameter values (fix_parameters, src/systems.jl:17-27)."""
        nv = self.nvars
        out = []
        for f in self.polys:
            t: Dict[Exp, complex] = {}
            for e, c in f.terms.items():
                m = c
                for k, v in zip(e[nv:], p):
                    if k:
                        m *= v ** k
                ee = e[:nv]
                t[ee] = t.get(ee, 0) + m
            out.append(Poly(t, nv))
        return System(out, nv, 0)

    def evaluate(self, x, p=()):
        vals = list(x) + list(p)
        return [f(vals) for f in self.polys]


def variables(nvars: int, nparams: int = 0):
    """Return (x_1..x_n, p_1..p_m) as Poly generators of one symbol ring."""
    nsym = nvars + nparams
    gens = []
    for i in range(nsym):
        e = [0] * nsym
        e[i] = 1
        gens.append(Poly({tuple(e): 1.0}, nsym))
    return gens[:nvars], gens[nvars:]


def support_coefficients(f: Poly, nvars: int):
    """Support and coefficients of f in the variables (support_coefficients, symbolic.jl:1345)."""
    return [e[:nvars] for e in f.terms], list(f.terms.values())
