"""Unexpanded expressions -> straight-line programs (ModelKit-style front end).

Expressions are hash-consed DAG nodes (structural CSE, as ModelKit's CSE pass
does for its Expression trees), differentiated symbolically with the product and
power rules (differentiate, src/model_kit/symbolic.jl:332-340) and lowered to the
InstructionSequence format of tape.py, fusing a*b + c into OP_MULADD
(src/model_kit/intermediate_representation.jl:339-447 fuses the same way).
Keeping products and powers unexpanded matters numerically: (x - 10)^6 evaluated
factored is accurate near its 6-fold root, expanded it is rounding noise.
"""
from __future__ import annotations

from typing import Dict, List, Optional, Sequence, Tuple

from .polys import Poly, System
from .tape import _ARITY_OF as _ARITY
from .tape import OP, Tape

_INTERN: Dict[tuple, "Expr"] = {}


class Expr:
    __slots__ = ("op", "args", "val", "_key", "__weakref__")

    def __new__(cls, op, args=(), val=None):
        key = (op, tuple(id(a) for a in args), val)
        e = _INTERN.get(key)
        if e is None:
            e = object.__new__(cls)
            e.op, e.args, e.val, e._key = op, tuple(args), val, key
            _INTERN[key] = e
        return e

    # construction with light constant folding
    @staticmethod
    def const(c):
        return Expr("const", (), complex(c))

    @staticmethod
    def lift(x):
        return x if isinstance(x, Expr) else Expr.const(x)

    def is_const(self, v=None):
        return self.op == "const" and (v is None or self.val == v)

    def __add__(self, o):
        o = Expr.lift(o)
        if self.is_const(0):
            return o
        if o.is_const(0):
            return self
        if self.is_const() and o.is_const():
            return Expr.const(self.val + o.val)
        return Expr("add", (self, o))

    __radd__ = __add__

    def __neg__(self):
        if self.is_const():
            return Expr.const(-self.val)
        if self.op == "neg":
            return self.args[0]
        return Expr("neg", (self,))

    def __sub__(self, o):
        o = Expr.lift(o)
        if o.is_const(0):
            return self
        if self.is_const(0):
            return -o
        if self.is_const() and o.is_const():
            return Expr.const(self.val - o.val)
        return Expr("sub", (self, o))

    def __rsub__(self, o):
        return Expr.lift(o) - self

    def __mul__(self, o):
        o = Expr.lift(o)
        if self.is_const(0) or o.is_const(0):
            return Expr.const(0)
        if self.is_const(1):
            return o
        if o.is_const(1):
            return self
        if self.is_const() and o.is_const():
            return Expr.const(self.val * o.val)
        if o.is_const() and not self.is_const():
            return Expr("mul", (o, self))
        return Expr("mul", (self, o))

    __rmul__ = __mul__

    def __pow__(self, k: int):
        if not isinstance(k, int) or k < 0:
            raise ValueError("only non-negative integer powers")
        if k == 0:
            return Expr.const(1)
        if k == 1:
            return self
        if self.is_const():
            return Expr.const(self.val ** k)
        return Expr("pow", (self,), k)

    def __repr__(self):
        if self.op == "const":
            return repr(self.val)
        if self.op in ("var", "par"):
            return f"{self.op}{self.val}"
        return f"{self.op}({', '.join(map(repr, self.args))}{', ' + str(self.val) if self.op == 'pow' else ''})"


def expr_variables(nvars: int, nparams: int = 0):
    return [Expr("var", (), i) for i in range(nvars)], [Expr("par", (), i) for i in range(nparams)]


def differentiate(e: Expr, i: int, memo: Optional[dict] = None) -> Expr:
    memo = {} if memo is None else memo
    k = (e._key, i)
    if k in memo:
        return memo[k]
    op = e.op
    if op in ("const", "par"):
        d = Expr.const(0)
    elif op == "var":
        d = Expr.const(1 if e.val == i else 0)
    elif op == "add":
        d = differentiate(e.args[0], i, memo) + differentiate(e.args[1], i, memo)
    elif op == "sub":
        d = differentiate(e.args[0], i, memo) - differentiate(e.args[1], i, memo)
    elif op == "neg":
        d = -differentiate(e.args[0], i, memo)
    elif op == "mul":
        a, b = e.args
        d = differentiate(a, i, memo) * b + a * differentiate(b, i, memo)
    elif op == "pow":
        a = e.args[0]
        da = differentiate(a, i, memo)
        d = Expr.const(0) if da.is_const(0) else (e.val * (a ** (e.val - 1))) * da
    else:
        raise ValueError(op)
    memo[k] = d
    return d


def to_poly(e: Expr, nsym: int, nvars: int, memo=None) -> Poly:
    """Expand (for degrees and coefficient scaling of total-degree start systems)."""
    memo = {} if memo is None else memo
    if e._key in memo:
        return memo[e._key]
    op = e.op
    if op == "const":
        p = Poly.const(e.val, nsym)
    elif op in ("var", "par"):
        idx = e.val if op == "var" else nvars + e.val
        ex = [0] * nsym
        ex[idx] = 1
        p = Poly({tuple(ex): 1.0}, nsym)
    elif op == "add":
        p = to_poly(e.args[0], nsym, nvars, memo) + to_poly(e.args[1], nsym, nvars, memo)
    elif op == "sub":
        p = to_poly(e.args[0], nsym, nvars, memo) - to_poly(e.args[1], nsym, nvars, memo)
    elif op == "neg":
        p = -to_poly(e.args[0], nsym, nvars, memo)
    elif op == "mul":
        p = to_poly(e.args[0], nsym, nvars, memo) * to_poly(e.args[1], nsym, nvars, memo)
    else:
        p = to_poly(e.args[0], nsym, nvars, memo) ** e.val
    memo[e._key] = p
    return p


class ExprSystem(System):
    """A square system given by unexpanded expressions; `polys` holds the expansion
    (used only for degrees / coefficient scaling), the tapes come from the DAG."""

    def __init__(self, exprs: Sequence[Expr], nvars: int, nparams: int = 0):
        self.exprs = [Expr.lift(e) for e in exprs]
        memo: dict = {}
        polys = [to_poly(e, nvars + nparams, nvars, memo) for e in self.exprs]
        super().__init__(polys, nvars, nparams)

    def fix_parameters(self, p):
        raise NotImplementedError("substitute parameters in the expressions instead")


def _cheap(e: "Expr") -> bool:
    """A product or power of leaves (variables, parameters, constants): one or two operations
    to recompute, while keeping it live across the rows of a system costs registers."""
    return e.op in ("mul", "pow") and all(a.op in ("var", "par", "const") for a in e.args)


class _Lower:
    def __init__(self, nvars, nparams, remat=False):
        self.nvars, self.nparams = nvars, nparams
        self.consts: List[complex] = []
        self.cidx: Dict[complex, int] = {}
        self.instrs: List[Tuple[str, tuple]] = []
        self.ref: Dict[tuple, tuple] = {}
        self.uses: Dict[tuple, int] = {}
        # remat == "use": cheap nodes are not shared but recomputed, bit for bit the same
        # operations, at every use
        self.remat = remat

    def const(self, c):
        c = complex(c)
        if c not in self.cidx:
            self.cidx[c] = len(self.consts)
            self.consts.append(c)
        return ("c", self.cidx[c])

    def count_uses(self, e: Expr, seen):
        if e._key in seen:
            self.uses[e._key] = self.uses.get(e._key, 0) + 1
            return
        seen.add(e._key)
        self.uses[e._key] = self.uses.get(e._key, 0) + 1
        for a in e.args:
            self.count_uses(a, seen)

    def emit(self, op, *args):
        self.instrs.append((op, args))
        return ("i", len(self.instrs) - 1)

    def lower(self, e: Expr):
        if e._key in self.ref and not (self.remat == "use" and _cheap(e)):
            return self.ref[e._key]
        op = e.op
        if op == "const":
            r = self.const(e.val)
        elif op == "var":
            r = ("x", e.val)
        elif op == "par":
            r = ("p", e.val)
        elif op == "add":
            a, b = e.args
            # fuse a*b + c when the product is used only here; the addend (typically the
            # running sum of a long chain) is lowered first, so that the factors of each term
            # are formed right before their multiply-add instead of all ahead of the chain
            if b.op == "mul" and self.uses.get(b._key, 0) == 1 and b._key not in self.ref:
                acc = self.lower(a)
                r = self.emit("MULADD", self.lower(b.args[0]), self.lower(b.args[1]), acc)
            elif a.op == "mul" and self.uses.get(a._key, 0) == 1 and a._key not in self.ref:
                acc = self.lower(b)
                r = self.emit("MULADD", self.lower(a.args[0]), self.lower(a.args[1]), acc)
            else:
                r = self.emit("ADD", self.lower(a), self.lower(b))
        elif op == "sub":
            r = self.emit("SUB", self.lower(e.args[0]), self.lower(e.args[1]))
        elif op == "neg":
            r = self.emit("NEG", self.lower(e.args[0]))
        elif op == "mul":
            r = self.emit("MUL", self.lower(e.args[0]), self.lower(e.args[1]))
        elif op == "pow":
            base = self.lower(e.args[0])
            if e.val == 2:
                r = self.emit("SQR", base)
            elif e.val == 3:
                r = self.emit("CB", base)
            else:
                r = self.emit("POW_INT", base, ("k", e.val))
        else:
            raise ValueError(op)
        self.ref[e._key] = r
        return r

    def finish(self, outs, jac, output_dim):
        nc = len(self.consts)
        base = nc + self.nparams + self.nvars

        def slot(r):
            kind, i = r
            return {"c": i + 1, "p": nc + i + 1, "x": nc + self.nparams + i + 1, "i": base + i + 1, "k": i}[kind]

        rows = []
        for k, (op, args) in enumerate(self.instrs):
            a = [slot(r) for r in args]
            while len(a) < 4:
                a.append(a[-1])
            rows.append(a + [OP[op], base + k + 1])
        n = base + len(self.instrs)
        rows.append([n, n, n, n, OP["STOP"], n])
        u = [(i + 1, slot(r)) for i, r in enumerate(outs) if r is not None]
        U = []
        if jac is not None:
            m = len(jac)
            for i in range(m):
                for j in range(len(jac[i])):
                    if jac[i][j] is not None:
                        U.append((j * m + i + 1, slot(jac[i][j])))
        return Tape(rows, self.consts, self.nparams, self.nvars, u, U, output_dim)


def _live_peak(t: Tape) -> int:
    """Peak number of live complex values of the code generated from `t`: inputs live from
    their first to their last use, temporaries from their definition to their last use, and
    an output slot until the instruction that writes it last (it is stored right there)."""
    rows = [r for r in t.instr if r[4] != OP["STOP"]]
    first_input = t.nconst + 1
    first_use: Dict[int, int] = {}
    last_use: Dict[int, int] = {}
    start: Dict[int, int] = {}
    for i, r in enumerate(rows):
        for q in range(_ARITY[int(r[4])]):
            if r[4] == OP["POW_INT"] and q == 1:
                continue
            if r[q] >= first_input:
                last_use[r[q]] = i
                first_use.setdefault(r[q], i)
        start.setdefault(r[5], i)
        last_use[r[5]] = max(last_use.get(r[5], i), i)
    delta = [0] * (len(rows) + 2)
    for sl, lu in last_use.items():
        delta[start.get(sl, first_use.get(sl, lu))] += 1
        delta[lu + 1] -= 1
    live = peak = 0
    for d in delta:
        live += d
        peak = max(peak, live)
    return peak


def expr_tapes(system: ExprSystem) -> Tuple[Tape, Tape]:
    """(evaluate_and_jacobian! tape, evaluate! tape).  Emission order: every row at once
    with all common subexpressions shared, or row by row with products of leaves recomputed
    at every use (identical operations, so identical values; the code generator keeps the
    recomputations apart), whichever keeps fewer values live."""
    n, np_ = system.nvars, system.nparams
    memo: dict = {}
    J = [[differentiate(e, j, memo) for j in range(n)] for e in system.exprs]
    m = len(system.exprs)

    def lowering(exprs_for_uses, remat):
        L = _Lower(n, np_, remat)
        seen: set = set()
        for e in exprs_for_uses:
            L.count_uses(e, seen)
        return L

    def ej(remat):
        L = lowering(list(system.exprs) + [d for row in J for d in row], remat)
        if not remat:
            outs = [None if e.is_const(0) else L.lower(e) for e in system.exprs]
            jac = [[None if d.is_const(0) else L.lower(d) for d in row] for row in J]
            return L.finish(outs, jac, m)
        outs, jac = [], []
        for e, row in zip(system.exprs, J):
            outs.append(None if e.is_const(0) else L.lower(e))
            jac.append([None if d.is_const(0) else L.lower(d) for d in row])
        return L.finish(outs, jac, m)

    def ev(remat):
        L = lowering(list(system.exprs), remat)
        return L.finish([None if e.is_const(0) else L.lower(e) for e in system.exprs], None, m)

    def best(a, b):  # recomputation charged at one live value per 100 extra instructions
        return min((a, b), key=lambda t: _live_peak(t) + (len(t.instr) - min(len(a.instr), len(b.instr))) / 100.0)

    from .tape import _MODE
    if _MODE[-1] == "cpu":
        return ej(False), ev(False)
    return best(ej(False), ej("use")), best(ev(False), ev("use"))
