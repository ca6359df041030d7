"""Seeded synthetic workloads (the benchmark configurations).

dense_system(n, d, seed): n dense polynomials with all monomials of degree <= d,
    iid complex Gaussian coefficients (re, im ~ N(0, 1/2), Julia's randn(ComplexF64)).
sparse_param_system(n, nterms, nparams, seed): n polynomials of degree <= 2 with
    `nterms` monomials each (always the constant and x_i^2 in equation i), whose
    coefficients are affine in `nparams` complex parameters.
shared_factor_system(n, seed): f_i = l(x) q_i(x) + r_i(x) with one shared random
    linear form l, random quadratic forms q_i and dense r_i of degree <= 2, so the
    cubic parts share a factor and many total-degree paths diverge.
"""
from __future__ import annotations

import itertools
from typing import List, Tuple

import numpy as np

from .polys import Poly, System, variables


def cgauss(rng, *shape):
    return (rng.standard_normal(shape) + 1j * rng.standard_normal(shape)) / np.sqrt(2)


def monomials(n: int, d: int, exact: bool = False) -> List[Tuple[int, ...]]:
    out = []
    for deg in range(d + 1):
        if exact and deg != d:
            continue
        for c in itertools.combinations_with_replacement(range(n), deg):
            e = [0] * n
            for i in c:
                e[i] += 1
            out.append(tuple(e))
    return out


def dense_system(n: int, d: int = 3, seed: int = 0) -> System:
    rng = np.random.default_rng(seed)
    mons = monomials(n, d)
    polys = []
    for i in range(n):
        c = cgauss(rng, len(mons))
        polys.append(Poly(dict(zip(mons, c)), n))
    return System(polys, n)


def sparse_param_system(n: int = 10, nterms: int = 20, nparams: int = 16, seed: int = 0) -> System:
    rng = np.random.default_rng(seed)
    mons = monomials(n, 2)
    nsym = n + nparams
    polys = []
    for i in range(n):
        zero = (0,) * n
        sq = tuple(2 if k == i else 0 for k in range(n))
        others = [m for m in mons if m != zero and m != sq]
        pick = rng.choice(len(others), size=nterms - 2, replace=False)
        support = [zero, sq] + [others[k] for k in sorted(pick)]
        terms = {}
        for m in support:
            a = cgauss(rng, nparams + 1)
            terms[m + (0,) * nparams] = a[0]
            for k in range(nparams):
                e = list(m) + [0] * nparams
                e[n + k] = 1
                terms[tuple(e)] = a[k + 1]
        polys.append(Poly(terms, nsym))
    return System(polys, n, nparams)


def shared_factor_system(n: int, seed: int = 0) -> System:
    rng = np.random.default_rng(seed)
    xs, _ = variables(n)
    lin = sum((complex(c) * x for c, x in zip(cgauss(rng, n), xs)), Poly.const(0, n))
    quad = monomials(n, 2, exact=True)
    low = monomials(n, 2)
    polys = []
    for i in range(n):
        q = Poly(dict(zip(quad, cgauss(rng, len(quad)))), n)
        r = Poly(dict(zip(low, cgauss(rng, len(low)))), n)
        polys.append(lin * q + r)
    return System(polys, n)


def parameter_draws(nparams: int, count: int, seed: int) -> np.ndarray:
    rng = np.random.default_rng(seed)
    return cgauss(rng, count, nparams)


# The batched parameter-homotopy workload: n = 10 sparse quadrics (20 monomials each) with
# coefficients affine in 16 parameters; 1,024 start solutions at p0 tracked to 256 targets.
CFG2_SEEDS = dict(system=20, p0=21, targets=22, gamma=23)


def cfg2_problem(ntargets: int = 256):
    """(F, p0, Q): the system, its start parameters and `ntargets` target parameter vectors."""
    F = sparse_param_system(10, 20, 16, seed=CFG2_SEEDS["system"])
    p0 = parameter_draws(16, 1, seed=CFG2_SEEDS["p0"])[0]
    Q = parameter_draws(16, 256, seed=CFG2_SEEDS["targets"])[:ntargets]
    return F, p0, Q


def shared_factor_exprs(n: int, seed: int = 0):
    """The same system as shared_factor_system(n, seed), kept factored for evaluation:
    f_i = l(x) q_i(x) + r_i(x) as an expression DAG (l shared by all equations)."""
    from .expr import ExprSystem, expr_variables
    rng = np.random.default_rng(seed)
    xs, _ = expr_variables(n)
    lc = cgauss(rng, n)
    lin = xs[0] * complex(lc[0])
    for c, x in zip(lc[1:], xs[1:]):
        lin = lin + complex(c) * x
    quad = monomials(n, 2, exact=True)
    low = monomials(n, 2)

    def mono(e):
        m = None
        for i, k in enumerate(e):
            for _ in range(k):
                m = xs[i] if m is None else m * xs[i]
        return m

    exprs = []
    for i in range(n):
        qc = cgauss(rng, len(quad))
        rc = cgauss(rng, len(low))
        q = None
        for c, e in zip(qc, quad):
            t = complex(c) * mono(e)
            q = t if q is None else q + t
        r = None
        for c, e in zip(rc, low):
            t = complex(c) if sum(e) == 0 else complex(c) * mono(e)
            r = t if r is None else r + t
        exprs.append(lin * q + r)
    return ExprSystem(exprs, n)


# The mixed-size, endgame-heavy workload: one system per size, all tracked in one launch.
CFG4_SIZES = (4, 6, 8, 12)
CFG4_SEEDS = dict(system=40, gamma=41)


def cfg4_problem(sizes=CFG4_SIZES):
    """[(H, starts)] per size: total-degree homotopies of shared-factor systems."""
    from .homotopies import total_degree
    out = []
    for k, n in enumerate(sizes):
        F = shared_factor_exprs(n, seed=CFG4_SEEDS["system"] + n)
        out.append(total_degree(F, seed=CFG4_SEEDS["gamma"] + n))
    return out
