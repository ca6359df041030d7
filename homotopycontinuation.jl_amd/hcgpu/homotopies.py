"""Homotopies and start systems (host side).

StraightLineHomotopy   src/homotopies/straight_line_homotopy.jl:10-68
ParameterHomotopy      src/homotopies/parameter_homotopy.jl:10-64
total_degree           src/total_degree.jl:32-117 (non-homogeneous, square case)
start solutions        src/total_degree.jl:235-298
"""
from __future__ import annotations

import ctypes
import itertools
import math
from typing import List, Optional, Sequence

import numpy as np

from .abi import HCG_PARAMETER, HCG_STRAIGHT_LINE, hcg_homotopy_desc, hcg_system
from .polys import Poly, System, variables
from .tape import Tape, tapes


class TapeSystem:
    """A square system given directly by its two straight-line programs, the form in which
    ModelKit hands a compiled system over: the `evaluate_and_jacobian!` tape (u and U
    assignments) and the `evaluate!` tape (u only; also used for `taylor!` and the
    double-double residual), as `jacobian_instruction_sequence` / `instruction_sequence`
    produce them (src/model_kit/instruction_sequence.jl:1-27, 256-272).  Any OpType the C ABI
    accepts may appear, including the fused ADD3/ADD4/MUL3/MUL4/MULADD/MULSUB/SUBMUL/
    MULMULADD/MULMULSUB and the inverse ops (src/model_kit/operations.jl:183-248)."""

    def __init__(self, eval_jac: Tape, eval_: Tape):
        if eval_jac.nvars != eval_.nvars or eval_jac.nparams != eval_.nparams:
            raise ValueError("the two tapes disagree on variables / parameters")
        if eval_.output_dim != eval_.nvars or eval_jac.output_dim != eval_.nvars:
            raise ValueError("only square systems")
        self.tapes = (eval_jac, eval_)
        self.nvars = eval_.nvars
        self.nparams = eval_.nparams

    def __len__(self):
        return self.nvars

    @property
    def n(self):
        return self.nvars


class _SysTapes:
    def __init__(self, system: System, fixed_params: Optional[Sequence[complex]] = None):
        self.system = system
        if isinstance(system, TapeSystem):
            self.eval_jac, self.eval = system.tapes
        elif hasattr(system, "exprs"):
            from .expr import expr_tapes
            self.eval_jac, self.eval = expr_tapes(system)
        else:
            self.eval_jac, self.eval = tapes(system)
        self.fixed = None
        if fixed_params is not None:
            self.fixed = np.ascontiguousarray(np.asarray(fixed_params, dtype=np.complex128).view(np.float64))

    def struct(self) -> hcg_system:
        s = hcg_system()
        s.eval_jac = ctypes.pointer(self.eval_jac.struct())
        s.eval = ctypes.pointer(self.eval.struct())
        s.fixed_params = (self.fixed.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
                          if self.fixed is not None else ctypes.POINTER(ctypes.c_double)())
        return s


class AbstractHomotopy:
    n: int

    def desc(self) -> hcg_homotopy_desc:
        raise NotImplementedError


class StraightLineHomotopy(AbstractHomotopy):
    """H(x, t) = γ t G(x) + (1 - t) F(x)."""

    def __init__(self, G: System, F: System, gamma: complex = 1.0,
                 G_params: Optional[Sequence[complex]] = None, F_params: Optional[Sequence[complex]] = None):
        if len(G) != len(F) or G.nvars != F.nvars:
            raise ValueError("start and target systems differ in size")
        if G.nparams and G_params is None or F.nparams and F_params is None:
            raise ValueError("parameters of a straight-line homotopy's systems must be fixed")
        self.G = _SysTapes(G, G_params if G.nparams else None)
        self.F = _SysTapes(F, F_params if F.nparams else None)
        self.gamma = complex(gamma)
        self.n = F.nvars
        self._desc = None

    def desc(self) -> hcg_homotopy_desc:
        if self._desc is None:
            d = hcg_homotopy_desc()
            d.kind = HCG_STRAIGHT_LINE
            d.n = self.n
            d.start = self.G.struct()
            d.target = self.F.struct()
            d.gamma_re, d.gamma_im = self.gamma.real, self.gamma.imag
            self._desc = d
        return self._desc


def affine_parameter_lift(F: System):
    """Lift a system whose coefficients are affine in its parameters to the system in its
    coefficients: F(x; p) = F̃(x; c(p)) with F̃_i = Σ_α c_{iα} x^α and c(p) affine.

    Because c is affine, c(t p + (1 - t) q) = t c(p) + (1 - t) c(q): the parameter homotopy
    of F from p to q and the one of F̃ from c(p) to c(q) follow the same path x(t).  F̃ is
    much cheaper to evaluate (one interpolation per coefficient instead of re-forming every
    coefficient from all parameters).  Returns (F̃, c) with c(p) evaluated in the same order
    of operations as System.fix_parameters, so c(p) reproduces F(x; p)'s coefficients bit
    for bit."""
    nv = F.nvars
    if hasattr(F, "exprs"):
        raise NotImplementedError("lift the expanded polynomial system")
    index, slots, maps = {}, [], []
    for i, f in enumerate(F.polys):
        for e, c in f.terms.items():
            pe = e[nv:]
            if sum(pe) > 1:
                raise ValueError("coefficients are not affine in the parameters")
            key = (i, e[:nv])
            if key not in index:
                index[key] = len(slots)
                slots.append(key)
                maps.append([])
            maps[index[key]].append((pe.index(1) if sum(pe) == 1 else -1, c))
    m = len(slots)
    terms = [dict() for _ in F.polys]
    for s, (i, alpha) in enumerate(slots):
        e = list(alpha) + [0] * m
        e[nv + s] = 1
        terms[i][tuple(e)] = 1.0
    Ft = System([Poly(t, nv + m) for t in terms], nv, m)

    def cmap(p):
        p = np.asarray(p, dtype=np.complex128)
        out = np.empty(m, dtype=np.complex128)
        for s, lst in enumerate(maps):
            acc = 0j
            for k, c in lst:
                acc = acc + (complex(c) if k < 0 else complex(c) * complex(p[k]))
            out[s] = acc
        return out

    return Ft, cmap


class ParameterHomotopy(AbstractHomotopy):
    """H(x, t) = F(x; t p + (1 - t) q).  `q` may be one vector or a list of target
    vectors (one group per vector, as many_solve's target_parameters).

    lift_affine=True tracks the same homotopy through affine_parameter_lift (for systems
    whose coefficients are affine in the parameters); p and q are then mapped to
    coefficient vectors and `self.F` holds the lifted system."""

    def __init__(self, F: System, start_parameters: Sequence[complex], target_parameters, lift_affine: bool = False):
        if F.nparams == 0:
            raise ValueError("parameter homotopy needs a system with parameters")
        q = np.asarray(target_parameters, dtype=np.complex128).reshape(-1, F.nparams)
        p = np.asarray(start_parameters, dtype=np.complex128)
        if p.shape != (F.nparams,):
            raise ValueError("start parameters have the wrong length")
        self.original = F
        self.lifted = bool(lift_affine)
        self._cmap = None
        if lift_affine:
            F, cmap = affine_parameter_lift(F)
            self._cmap = cmap
            p = cmap(p)
            q = np.stack([cmap(qq) for qq in q])
        self.F = _SysTapes(F)
        self.n = F.nvars
        self.p = p
        self.q = q
        self._desc = None

    def target_matrix(self, target_parameters) -> np.ndarray:
        """Target parameter vectors (one per row) in the parameter space the tapes use."""
        Q = np.asarray(target_parameters, dtype=np.complex128).reshape(-1, self.original.nparams)
        if self._cmap is None:
            return Q
        return np.stack([self._cmap(q) for q in Q])

    def desc(self) -> hcg_homotopy_desc:
        if self._desc is None:
            d = hcg_homotopy_desc()
            d.kind = HCG_PARAMETER
            d.n = self.n
            d.target = self.F.struct()
            d.gamma_re, d.gamma_im = 1.0, 0.0
            self._desc = d
        return self._desc


def total_degree_start_solutions(degrees: Sequence[int]) -> np.ndarray:
    """x = (cis(2πk₁/d₁), …, cis(2πk_n/d_n)); the first coordinate varies fastest
    (Iterators.product order, src/total_degree.jl:235-262)."""
    ranges = [range(d) for d in degrees]
    out = []
    for ks in itertools.product(*reversed(ranges)):
        ks = ks[::-1]
        out.append([complex(math.cos(2 * math.pi * k / d), math.sin(2 * math.pi * k / d)) for k, d in zip(ks, degrees)])
    return np.asarray(out, dtype=np.complex128).reshape(-1, len(degrees))


def total_degree(F: System, gamma: Optional[complex] = None, seed: Optional[int] = None,
                 target_parameters: Optional[Sequence[complex]] = None):
    """Total-degree homotopy H = γ t G + (1-t) F with G_i = s_i (x_i^{d_i} - 1),
    s_i = max_α |c_{iα}| (src/total_degree.jl:49-62, 102-106).  Returns (H, starts)."""
    if target_parameters is not None:
        F = F.fix_parameters(target_parameters)
    if F.nparams:
        raise ValueError("fix the parameters of F first")
    n = F.nvars
    if len(F) != n:
        raise NotImplementedError("only square systems (overdetermined systems are out of scope)")
    D = []
    homogeneous = True
    for f in F.polys:
        degs = [sum(e) for e in f.terms]
        D.append(max(degs))
        homogeneous = homogeneous and len(set(degs)) == 1
    if homogeneous:
        raise NotImplementedError("homogeneous systems (projective start systems) are out of scope")
    scaling = [max(abs(c) for c in f.terms.values()) for f in F.polys]
    xs, ss = variables(n, n)
    G = System([ss[i] * (xs[i] ** D[i] - 1) for i in range(n)], n, n)
    if gamma is None:
        rng = np.random.default_rng(seed)
        gamma = complex(np.exp(2j * np.pi * rng.random()))
    H = StraightLineHomotopy(G, F, gamma=gamma, G_params=scaling)
    return H, total_degree_start_solutions(D)
