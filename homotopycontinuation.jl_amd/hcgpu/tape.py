"""Tape producer: polynomial systems -> straight-line programs in ModelKit's format.

The output is a field-for-field mirror of ModelKit's `InstructionSequence`
(src/model_kit/instruction_sequence.jl:1-27): instructions are 6 x int32
(input1..input4, op, output) with 1-based tape slots, constants occupy slots
1..nconst, then parameters, then variables, then instruction outputs; the
op codes are the `OpType` ordinals (src/model_kit/operations.jl:5-49).  A real
ModelKit tape (from `interpret(T)` + `instruction_sequence`) plugs into the same
C ABI unchanged; this producer exists so the GPU path is usable without Julia.

Lowering (our own, not SymEngine's CSE): every monomial in the variables is
built once as a product of a smaller monomial and one variable; each output
polynomial is a MULADD chain over its terms.  Coefficients that are
polynomials in the parameters (parameter homotopies) are grouped per
x-monomial and built first, so the x-independent part is a separate sub-chain.
"""
from __future__ import annotations

import ctypes
from typing import Dict, List, Optional, Sequence, Tuple

import numpy as np

from .polys import Poly, System

# OpType ordinals, src/model_kit/operations.jl:5-49
OP = {
    name: i
    for i, name in enumerate(
        """STOP CB ACOS ASIN COS COSH EXP INV INV_NOT_ZERO INVSQR NEG SIN SINH SQR SQRT TAN TANH IDENTITY
        ADD DIV MUL SUB POW_INT POW ADD3 MUL3 MULADD MULSUB SUBMUL ADD4 MUL4 MULMULADD MULMULSUB""".split()
    )
}
ARITY = {
    **{k: 1 for k in "CB ACOS ASIN COS COSH EXP INV INV_NOT_ZERO INVSQR NEG SIN SINH SQR SQRT TAN TANH IDENTITY".split()},
    **{k: 2 for k in "ADD DIV MUL SUB POW_INT POW".split()},
    **{k: 3 for k in "ADD3 MUL3 MULADD MULSUB SUBMUL".split()},
    **{k: 4 for k in "ADD4 MUL4 MULMULADD MULMULSUB".split()},
    "STOP": 0,
}

_ARITY_OF = {OP[k]: v for k, v in ARITY.items()}
# powers of one symbol from this exponent up are one OP_POW_INT instead of a chain of products
POW_INT_MIN = 5


class hcg_tape(ctypes.Structure):
    _fields_ = [
        ("instr", ctypes.POINTER(ctypes.c_int32)),
        ("constants", ctypes.POINTER(ctypes.c_double)),
        ("u_assign", ctypes.POINTER(ctypes.c_int32)),
        ("U_assign", ctypes.POINTER(ctypes.c_int32)),
        ("ninstr", ctypes.c_int32),
        ("nconst", ctypes.c_int32),
        ("params_first", ctypes.c_int32),
        ("nparams", ctypes.c_int32),
        ("vars_first", ctypes.c_int32),
        ("nvars", ctypes.c_int32),
        ("t_index", ctypes.c_int32),
        ("nu", ctypes.c_int32),
        ("nU", ctypes.c_int32),
        ("output_dim", ctypes.c_int32),
        ("tape_space", ctypes.c_int32),
        ("pad_", ctypes.c_int32),
    ]


class Tape:
    """An InstructionSequence: instructions, constants, ranges and assignments."""

    def __init__(self, instr, constants, nparams, nvars, u_assign, U_assign, output_dim):
        self.instr = np.ascontiguousarray(np.asarray(instr, dtype=np.int32).reshape(-1, 6))
        self.constants = np.ascontiguousarray(np.asarray(constants, dtype=np.complex128))
        self.nconst = len(self.constants)
        self.nparams = nparams
        self.nvars = nvars
        self.params_first = self.nconst + 1
        self.vars_first = self.nconst + nparams + 1
        self.u_assign = np.ascontiguousarray(np.asarray(u_assign, dtype=np.int32).reshape(-1, 2))
        self.U_assign = np.ascontiguousarray(np.asarray(U_assign, dtype=np.int32).reshape(-1, 2))
        self.output_dim = output_dim
        outs = self.instr[:, 5]
        base = self.nconst + nparams + nvars
        self.tape_space = int(max([base] + list(outs)))
        self._cdata = np.ascontiguousarray(self.constants.view(np.float64))
        self._struct = None

    @property
    def ninstr(self):
        return len(self.instr)

    def struct(self) -> hcg_tape:
        if self._struct is None:
            s = hcg_tape()
            s.instr = self.instr.ctypes.data_as(ctypes.POINTER(ctypes.c_int32))
            s.constants = self._cdata.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
            s.u_assign = self.u_assign.ctypes.data_as(ctypes.POINTER(ctypes.c_int32))
            s.U_assign = self.U_assign.ctypes.data_as(ctypes.POINTER(ctypes.c_int32))
            s.ninstr = len(self.instr)
            s.nconst = self.nconst
            s.params_first = self.params_first
            s.nparams = self.nparams
            s.vars_first = self.vars_first
            s.nvars = self.nvars
            s.t_index = 0
            s.nu = len(self.u_assign)
            s.nU = len(self.U_assign)
            s.output_dim = self.output_dim
            s.tape_space = self.tape_space
            self._struct = s
        return self._struct

    def op_counts(self) -> Dict[str, int]:
        inv = {v: k for k, v in OP.items()}
        out: Dict[str, int] = {}
        for row in self.instr:
            name = inv[int(row[4])]
            out[name] = out.get(name, 0) + 1
        return out


class _Builder:
    def __init__(self, nvars: int, nparams: int):
        self.nvars = nvars
        self.nparams = nparams
        self.consts: List[complex] = []
        self.const_index: Dict[complex, int] = {}
        self.instrs: List[Tuple[str, Tuple[int, ...]]] = []  # (op, args) ; outputs assigned later
        self.mono: Dict[Tuple[int, ...], object] = {}  # exponent -> ("var", i) | ("par", i) | ("ins", k)

    # references: ("c", idx) constant, ("p", i) parameter, ("x", i) variable, ("i", k) instruction
    def const(self, c: complex):
        c = complex(c)
        if c not in self.const_index:
            self.const_index[c] = len(self.consts)
            self.consts.append(c)
        return ("c", self.const_index[c])

    def emit(self, op: str, *args):
        self.instrs.append((op, args))
        return ("i", len(self.instrs) - 1)

    def monomial(self, e: Tuple[int, ...]):
        """Reference to the value of the monomial with exponent e (over vars+params), e != 0."""
        if e in self.mono:
            return self.mono[e]
        deg = sum(e)
        nz = [i for i, k in enumerate(e) if k]
        if deg == 1:
            i = nz[0]
            ref = ("x", i) if i < self.nvars else ("p", i - self.nvars)
        elif e[nz[-1]] >= POW_INT_MIN:
            # a high power of one symbol is one OP_POW_INT (power by squaring), as ModelKit
            # lowers x^k (pow!, src/model_kit/intermediate_representation.jl:224-241)
            last = nz[-1]
            unit = [0] * len(e)
            unit[last] = 1
            if len(nz) == 1:
                ref = self.emit("POW_INT", self.monomial(tuple(unit)), e[last])
            else:
                rest = list(e)
                rest[last] = 0
                power = list(unit)
                power[last] = e[last]
                ref = self.emit("MUL", self.monomial(tuple(rest)), self.monomial(tuple(power)))
        elif len(nz) == 1 and e[nz[0]] == 2:
            base = list(e)
            base[nz[0]] = 1
            ref = self.emit("SQR", self.monomial(tuple(base)))
        else:
            # split off the last symbol: m(e) = m(e - u_last) * sym_last
            last = nz[-1]
            rest = list(e)
            rest[last] -= 1
            unit = [0] * len(e)
            unit[last] = 1
            ref = self.emit("MUL", self.monomial(tuple(rest)), self.monomial(tuple(unit)))
        self.mono[e] = ref
        return ref

    def poly(self, f: Poly):
        """Emit a MULADD chain for f; returns a reference or None for the zero polynomial."""
        nv = self.nvars
        # group terms by x-monomial; coefficient = polynomial in the parameters
        groups: Dict[Tuple[int, ...], Dict[Tuple[int, ...], complex]] = {}
        for e, c in f.terms.items():
            groups.setdefault(e[:nv], {})[e[nv:]] = c
        keys = sorted(groups, key=lambda a: (sum(a), tuple(-k for k in a)))
        acc = None
        const_ref = None
        for xe in keys:
            cref = self._coeff(groups[xe])
            if cref is None:
                continue
            if sum(xe) == 0:
                const_ref = cref
                continue
            m = self.monomial(xe + (0,) * self.nparams)
            one = cref[0] == "c" and self.consts[cref[1]] == 1
            if acc is None:
                acc = m if one else self.emit("MUL", cref, m)
            else:
                acc = self.emit("ADD", m, acc) if one else self.emit("MULADD", cref, m, acc)
        if const_ref is not None:
            acc = const_ref if acc is None else self.emit("ADD", acc, const_ref)
        return acc

    def polys(self, fs: List[Poly]):
        """MULADD chains of several polynomials built in one sweep over their monomials: for
        every monomial (graded order) each polynomial containing it takes its term next.  Each
        polynomial's chain is exactly the one poly() builds (same terms in the same order), but
        a shared monomial is consumed right after it is formed instead of staying live until
        the last polynomial, which keeps the generated straight-line code's register pressure
        low.  Returns one reference (or None) per polynomial."""
        nv = self.nvars
        groups: List[Dict[Tuple[int, ...], Dict[Tuple[int, ...], complex]]] = []
        for f in fs:
            g: Dict[Tuple[int, ...], Dict[Tuple[int, ...], complex]] = {}
            for e, c in f.terms.items():
                g.setdefault(e[:nv], {})[e[nv:]] = c
            groups.append(g)
        order = lambda a: (sum(a), tuple(-k for k in a))
        keys = sorted({k for g in groups for k in g}, key=order)
        acc: List[object] = [None] * len(fs)
        const_ref: List[object] = [None] * len(fs)
        for xe in keys:
            for r, g in enumerate(groups):
                if xe not in g:
                    continue
                cref = self._coeff(g[xe])
                if cref is None:
                    continue
                if sum(xe) == 0:
                    const_ref[r] = cref
                    continue
                m = self.monomial(xe + (0,) * self.nparams)
                one = cref[0] == "c" and self.consts[cref[1]] == 1
                if acc[r] is None:
                    acc[r] = m if one else self.emit("MUL", cref, m)
                else:
                    acc[r] = self.emit("ADD", m, acc[r]) if one else self.emit("MULADD", cref, m, acc[r])
        for r in range(len(fs)):
            if const_ref[r] is not None:
                acc[r] = const_ref[r] if acc[r] is None else self.emit("ADD", acc[r], const_ref[r])
        return acc

    def _coeff(self, cp: Dict[Tuple[int, ...], complex]):
        """Reference to a coefficient that is a polynomial in the parameters."""
        cp = {e: c for e, c in cp.items() if c != 0}
        if not cp:
            return None
        if all(sum(e) == 0 for e in cp):
            return self.const(sum(cp.values()))
        acc = None
        const = 0j
        for pe in sorted(cp, key=lambda a: (sum(a), tuple(-k for k in a))):
            c = cp[pe]
            if sum(pe) == 0:
                const += c
                continue
            m = self.monomial((0,) * self.nvars + pe)
            if acc is None:
                acc = m if c == 1 else self.emit("MUL", self.const(c), m)
            else:
                acc = self.emit("ADD", m, acc) if c == 1 else self.emit("MULADD", self.const(c), m, acc)
        if const != 0:
            acc = self.emit("ADD", acc, self.const(const))
        return acc

    def finish(self, outputs: List[Optional[object]], jac: Optional[List[List[Optional[object]]]], output_dim: int) -> Tape:
        nc = len(self.consts)
        base = nc + self.nparams + self.nvars

        def slot(ref):
            kind, i = ref
            if kind == "c":
                return i + 1
            if kind == "p":
                return nc + i + 1
            if kind == "x":
                return nc + self.nparams + i + 1
            return base + i + 1

        rows = []
        for k, (op, args) in enumerate(self.instrs):
            a = [r if isinstance(r, int) else slot(r) for r in args]
            while len(a) < 4:
                a.append(a[-1])
            rows.append(a + [OP[op], base + k + 1])
        n = base + len(self.instrs)
        rows.append([n, n, n, n, OP["STOP"], n])
        u_assign = [(i + 1, slot(r)) for i, r in enumerate(outputs) if r is not None]
        U_assign = []
        if jac is not None:
            nrow = len(jac)
            for i in range(nrow):
                for j in range(len(jac[i])):
                    r = jac[i][j]
                    if r is not None:
                        U_assign.append((j * nrow + i + 1, slot(r)))
        return Tape(rows, self.consts, self.nparams, self.nvars, u_assign, U_assign, output_dim)


def max_live(t: Tape) -> int:
    """Peak number of simultaneously live complex values of the straight-line code generated
    from `t` (inputs live from first to last use, every output stored right after the last
    instruction writing it; constants are literals)."""
    rows = [r for r in t.instr if r[4] != OP["STOP"]]
    first_input = t.nconst + 1
    last_use: Dict[int, int] = {}
    first_use: Dict[int, int] = {}
    for i, r in enumerate(rows):
        ar = _ARITY_OF[int(r[4])]
        for q in range(ar):
            if r[4] == OP["POW_INT"] and q == 1:
                continue
            sl = r[q]
            if sl >= first_input:
                last_use[sl] = i
                first_use.setdefault(sl, i)
    for _, sl in list(t.u_assign) + list(t.U_assign):
        last_use.setdefault(sl, len(rows))
    start = {}
    for i, r in enumerate(rows):
        start.setdefault(r[5], i)
    delta = [0] * (len(rows) + 2)
    for sl, lu in last_use.items():
        b0 = start.get(sl, first_use.get(sl, lu))
        delta[b0] += 1
        delta[lu + 1] -= 1
    live = peak = 0
    for d in delta:
        live += d
        peak = max(peak, live)
    return peak


# "gpu": emission order by register pressure (below); "cpu": the fewest instructions, i.e.
# every common subexpression shared (what a CPU evaluator wants; set by `shared_lowering`)
_MODE = ["gpu"]


class shared_lowering:
    """Context manager: systems lowered inside share every common subexpression (the fewest
    instructions) instead of recomputing some to save registers.  Values are identical."""

    def __enter__(self):
        _MODE.append("cpu")

    def __exit__(self, *exc):
        _MODE.pop()


def _best(variants):
    """The tape with the smallest register pressure among equivalent emission orders (peak
    live values, with recomputation charged at one live value per 50 extra instructions)."""
    if _MODE[-1] == "cpu":
        return min((make() for make in variants), key=lambda t: len(t.instr))
    tapes_ = [make() for make in variants]
    nmin = min(len(t.instr) for t in tapes_)
    scored = [(max_live(t) + (len(t.instr) - nmin) / 50.0, k, t) for k, t in enumerate(tapes_)]
    return min(scored, key=lambda x: (x[0], x[1]))[2]


def eval_tape(system: System) -> Tape:
    """The evaluate! tape: u = F(x; p).  The polynomials are emitted either one after the
    other or interleaved over their shared monomials, whichever keeps fewer values live;
    both orders compute every output with exactly the same operations."""
    def interleaved():
        b = _Builder(system.nvars, system.nparams)
        return b.finish(b.polys(list(system.polys)), None, len(system.polys))

    def sequential():
        b = _Builder(system.nvars, system.nparams)
        return b.finish([b.poly(f) for f in system.polys], None, len(system.polys))

    return _best([interleaved, sequential])


def jacobian_tape(system: System) -> Tape:
    """The evaluate_and_jacobian! tape: u = F, U = dF/dx (jacobian_instruction_sequence,
    src/model_kit/instruction_sequence.jl:256-272).  Emission order as in eval_tape: all
    n + n^2 polynomials interleaved, each row (f_i and its gradient) interleaved, or one
    polynomial after the other, whichever keeps fewer values live."""
    m, n = len(system.polys), system.nvars
    J = system.jacobian()

    def build(groups, remat=False):
        b = _Builder(system.nvars, system.nparams)
        refs = {}
        for g in groups:
            if remat:  # recompute (bit-identical) monomials per group instead of keeping them live
                b.mono = {e: r for e, r in b.mono.items() if sum(e) == 1}
            for key, r in zip(g, b.polys([system.polys[i] if j < 0 else J[i][j] for i, j in g])):
                refs[key] = r
        outs = [refs[(i, -1)] for i in range(m)]
        jac = [[refs[(i, j)] for j in range(n)] for i in range(m)]
        return b.finish(outs, jac, m)

    everything = [[(i, -1) for i in range(m)] + [(i, j) for i in range(m) for j in range(n)]]
    by_row = [[(i, -1)] + [(i, j) for j in range(n)] for i in range(m)]
    one_by_one = [[(i, -1)] for i in range(m)] + [[(i, j)] for i in range(m) for j in range(n)]
    return _best([lambda: build(everything), lambda: build(by_row), lambda: build(one_by_one),
                  lambda: build(by_row, remat=True)])


def tapes(system: System) -> Tuple[Tape, Tape]:
    return jacobian_tape(system), eval_tape(system)
