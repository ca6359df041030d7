"""Systems of the reference's own tests (test/endgame_test.jl, test/endgame_tracker_test.jl,
test/tracker_test.jl, test/test_systems.jl), written with the polynomial front end."""
import numpy as np

from hcgpu.polys import Poly, System, variables


def beyond_polyhedral():
    (x, y), _ = variables(2)
    return System([2.3 * x**2 + 1.2 * y**2 + 3 * x - 2 * y + 3, 2.3 * x**2 + 1.2 * y**2 + 5 * x + 2 * y - 5], 2)


def cyclic(n):
    z, _ = variables(n)
    eqs = []
    for m in range(n - 1):
        s = Poly.const(0, n)
        for j in range(n):
            p = Poly.const(1, n)
            for k in range(j, j + m + 1):
                p = p * z[k % n]
            s = s + p
        eqs.append(s)
    p = Poly.const(1, n)
    for v in z:
        p = p * v
    eqs.append(p - 1)
    return System(eqs, n)


def hyperbolic():
    (x, z), _ = variables(2)
    y = 1
    return System([0.75 * x**4 + 1.5 * x**2 * y**2 - 2.5 * x**2 * z**2 + 0.75 * y**4 - 2.5 * y**2 * z**2 + 0.75 * z**4,
                   10 * x**2 * z + 10 * y**2 * z - 6 * z**3], 2)


def singular1():
    (x, y), _ = variables(2)
    z = 1
    return System([x**2 + 2 * y**2 + 2j * y * z,
                   (18 + 3j) * x * y + 7j * y**2 - (3 - 18j) * x * z - 14 * y * z - 7j * z**2], 2)


def wilkinson(d):
    (x,), _ = variables(1)
    f = Poly.const(1, 1)
    for i in range(1, d + 1):
        f = f * (x - i)
    return System([f], 1)


def winding_family(d):
    a = [0.257, -0.139, -1.73, -0.199, 1.79, -1.32]
    (x, y), _ = variables(2)
    f1 = (a[0] * x**d + a[1] * y) * (a[2] * x + a[3] * y) + 1
    f2 = (a[0] * x**d + a[1] * y) * (a[4] * x + a[5] * y) + 1
    return System([f1, f2], 2)


def mohab():
    (x, z, y), _ = variables(3)  # variable order [x, z, y] as in the reference test
    return System([
        -9091098778555951517 * x**3 * y**4 * z**2 + 5958442613080401626 * y**2 * z**7
        + 17596733865548170996 * x**2 * z**6 - 17979170986378486474 * x * y * z**6
        - 2382961149475678300 * x**4 * y**3 - 15412758154771986214 * x * y**3 * z**3 + 133,
        -10798198881812549632 * x**6 * y**3 * z - 11318272225454111450 * x * y**9
        - 14291416869306766841 * y**9 * z - 5851790090514210599 * y**2 * z**8
        + 15067068695242799727 * x**2 * y**3 * z**4 + 7716112995720175148 * x**3 * y * z**3 + 171,
        13005416239846485183 * x**7 * y**3 + 4144861898662531651 * x**5 * z**4 - 8026818640767362673 * x**6
        - 6882178109031199747 * x**2 * y**4 + 7240929562177127812 * x**2 * y**3 * z
        + 5384944853425480296 * x * y * z**4 + 88,
    ], 3)


MOHAB_GAMMA = -0.9132549847010242 + 0.4073884300256109j


def tracker_param_system():
    """F = [x^2 - a, x*y - a + b] with parameters (a, b) (test/tracker_test.jl:5-27)."""
    (x, y), (a, b) = variables(2, 2)
    return System([x**2 - a, x * y - a + b], 2, 2)


def _golden_vectors():
    import json
    from pathlib import Path
    d = json.loads((Path(__file__).resolve().parent / "golden" / "param_cases.json").read_text())
    return {case: {k: np.array([complex(a, b) for a, b in v]) for k, v in d[case].items()} for case in ("steiner", "four_bar")}


def steiner():
    """Conics tangent to five given conics (test/test_cases/steiner_higher_prec.jl:35-64):
    unknowns a[1:5] (the tangent conic f = a1 x1^2 + a2 x1 x2 + a3 x2^2 + a4 x1 + a5 x2 + 1)
    and the tangency points y[:, i]; parameters v[1:6, i], the coefficients of the i-th
    given conic g_i.  Per conic: f(y_i) = 0, g_i(y_i) = 0, det[∇f(y_i) ∇g_i(y_i)] = 0.
    Unknowns ordered [a; vec(y)], parameters vec(v) (column-major, as in the reference).
    Returns (F, p, q, s_p, s_q) with the reference's vectors."""
    xs, ps = variables(15, 30)
    a = xs[:5]
    y = [(xs[5 + 2 * i], xs[6 + 2 * i]) for i in range(5)]
    eqs = []
    for i in range(5):
        c = ps[6 * i: 6 * i + 6]
        y1, y2 = y[i]
        f = a[0] * y1**2 + a[1] * y1 * y2 + a[2] * y2**2 + a[3] * y1 + a[4] * y2 + 1
        g = c[0] * y1**2 + c[1] * y1 * y2 + c[2] * y2**2 + c[3] * y1 + c[4] * y2 + c[5]
        f1, f2 = 2 * a[0] * y1 + a[1] * y2 + a[3], a[1] * y1 + 2 * a[2] * y2 + a[4]
        g1, g2 = 2 * c[0] * y1 + c[1] * y2 + c[3], c[1] * y1 + 2 * c[2] * y2 + c[4]
        eqs += [f, g, f1 * g2 - f2 * g1]
    v = _golden_vectors()["steiner"]
    return System(eqs, 15, 30), v["p"], v["q"], v["s_p"], v["s_q"]


def four_bar():
    """Four-bar linkage synthesis (test/test_cases/four_bar.jl:2-25): 24 unknowns
    (x, a, y, b, x̂, â, ŷ, b̂, γ[1:8], γ̂[1:8]), 16 parameters (δ[1:8], δ̂[1:8]).
    Returns (F, p, q, s) with the reference's vectors (s is given to 8 digits)."""
    xs, ps = variables(24, 16)
    x, a, y, b, xh, ah, yh, bh = xs[:8]
    g, gh = xs[8:16], xs[16:24]
    d, dh = ps[:8], ps[8:16]
    D1 = [(ah * x - dh[i] * x) * g[i] + (a * xh - d[i] * xh) * gh[i] + (ah - xh) * d[i] + (a - x) * dh[i] - d[i] * dh[i]
          for i in range(8)]
    D2 = [(bh * y - dh[i] * y) * g[i] + (b * yh - d[i] * yh) * gh[i] + (bh - yh) * d[i] + (b - y) * dh[i] - d[i] * dh[i]
          for i in range(8)]
    D3 = [g[i] * gh[i] + g[i] + gh[i] for i in range(8)]
    v = _golden_vectors()["four_bar"]
    return System(D1 + D2 + D3, 24, 16), v["p"], v["q"], v["s"]


def straight_line_case():
    """test/tracker_test.jl:93-105: G = (0.2 + 0.4i) [x^2 - 1, y - 1], F = [x^2 + y^2 - 2.3,
    2x + 3y - 4]; returns (G, F, starts)."""
    (x, y), _ = variables(2)
    F = System([x**2 + y**2 - 2.3, 2 * x + 3 * y - 4], 2)
    G = System([(0.2 + 0.4j) * (x**2 - 1), (0.2 + 0.4j) * (y - 1)], 2)
    return G, F, np.array([[1, 1], [-1, 1]], dtype=complex)


def framework_454(seed: int = 454):
    """test/tracker_test.jl:137-219 (issue 454): a bar framework with 11 vertices, 15 edges,
    vertices 1, 6, 11 pinned, plus one linear slice a·x = b through the initial configuration.
    The start configuration has a singular Jacobian.  Returns (F, b0, b1, start)."""
    p0 = np.array([[0, 0], [1, 0], [1, 1], [2, 1], [1.5, 0.5], [3.0, 1.0], [4, 1], [5, 1], [4.5, 0.5], [5, 0],
                   [6, 0]], dtype=float)
    E = [(1, 2), (2, 3), (2, 5), (3, 4), (3, 5), (4, 5), (4, 6), (5, 9), (6, 7), (7, 8), (7, 9), (8, 9), (8, 10),
         (9, 10), (10, 11)]
    pinned = [1, 6, 11]
    free = [2, 3, 4, 5, 7, 8, 9, 10]
    xs, (b,) = variables(16, 1)
    pos = {}
    for j, i in enumerate(free):
        pos[i] = (xs[2 * j], xs[2 * j + 1])
    for i in pinned:
        pos[i] = (float(p0[i - 1, 0]), float(p0[i - 1, 1]))
    eqs = []
    for i, j in E:
        e = (pos[i][0] - pos[j][0]) ** 2 + (pos[i][1] - pos[j][1]) ** 2
        e = e - float(((p0[i - 1] - p0[j - 1]) ** 2).sum())
        eqs.append(e)
    a = [0, 1.0] * 8
    lin = sum(a[k] * xs[k] for k in range(16))
    eqs.append(lin - b)
    start = np.array([p0[i - 1, k] for i in free for k in range(2)], dtype=complex)
    b0 = sum(a[k] * start[k] for k in range(16))
    rng = np.random.default_rng(seed)
    p1 = p0 + 1e-3 * rng.standard_normal(p0.shape)
    b1 = sum(a[2 * j + k] * p1[i - 1, k] for j, i in enumerate(free) for k in range(2))
    return System(eqs, 16, 1), np.array([b0]), np.array([b1]), start


def certification_simple1():
    """Simple 1 of test/certification_test.jl:3-8: 18 isolated solutions, 4 of them real."""
    (x, y), _ = variables(2, 0)
    f1 = (x**4 + y**4 - 1) * (x**2 + y**2 - 2) + x**5 * y
    f2 = x**2 + 2 * x * y**2 - 2 * y**2 - 0.5
    return System([f1, f2], 2, 0)


def critical_points():
    """Critical points of the distance from u to the sextic curve f = 0
    (test/certification_test.jl:98-110): [x, y] - u - λ ∇f and f, parameters u."""
    xs, us = variables(3, 2)
    x, y, lam = xs
    f = (x**4 + y**4 - 1) * (x**2 + y**2 - 2) + x**5 * y
    return System([x - us[0] - lam * f.diff(0), y - us[1] - lam * f.diff(1), f], 3, 2)


def toric_ed(A):
    """Critical points of the distance to a toric variety (test/monodromy_test.jl:4-12):
    unknowns t[1:d], y[1:n], parameters u[1:n]; equations φ + y − u and Dφᵀ y with
    φ_j = Π_i t_i^{A[i][j]}.  For A = [3 2 1 0; 0 1 2 3] (the twisted cubic) 21 solutions."""
    d, n = len(A), len(A[0])
    xs, us = variables(d + n, n)
    t, y = xs[:d], xs[d:]
    phi = []
    for j in range(n):
        m = t[0] ** A[0][j]
        for i in range(1, d):
            m = m * t[i] ** A[i][j]
        phi.append(m)
    eqs = [phi[j] + y[j] - us[j] for j in range(n)]
    for k in range(d):
        acc = phi[0].diff(k) * y[0]
        for j in range(1, n):
            acc = acc + phi[j].diff(k) * y[j]
        eqs.append(acc)
    return System(eqs, d + n, n)


def circle_line():
    """x² + y² − 1, a x + b y + c with parameters (a, b, c) (test/monodromy_test.jl:359-362)."""
    (x, y), (a, b, c) = variables(2, 3)
    return System([x**2 + y**2 - 1, a * x + b * y + c], 2, 3)


def quotient_tapes():
    """Hand-written tapes of F = [x0 / x1 - 2, x0 + x1 - 3] (zero at (2, 1)): the eval tape
    divides (OP_DIV), the Jacobian tape inverts and negates (∂f1/∂x0 = 1/x1,
    ∂f1/∂x1 = -(x0/x1)/x1).  Slots: constants 2, 3, 1 (1..3), x0 = 4, x1 = 5."""
    from hcgpu.tape import Tape
    OP_INV, OP_NEG, OP_ADD, OP_DIV, OP_MUL, OP_SUB = 7, 10, 18, 19, 20, 21
    consts = [2.0, 3.0, 1.0]
    ev = [(4, 5, 0, 0, OP_DIV, 6), (6, 1, 0, 0, OP_SUB, 7), (4, 5, 0, 0, OP_ADD, 8), (8, 2, 0, 0, OP_SUB, 9)]
    jac = ev + [(5, 0, 0, 0, OP_INV, 10), (6, 10, 0, 0, OP_MUL, 11), (11, 0, 0, 0, OP_NEG, 12)]
    u = [(1, 7), (2, 9)]
    U = [(1, 10), (3, 12), (2, 3), (4, 3)]  # column-major: (1,1), (1,2), (2,1), (2,2)
    return Tape(ev, consts, 0, 2, u, [], 2), Tape(jac, consts, 0, 2, u, U, 2)


def polyhedral_cases():
    """(name, F, only_non_zero, seed) of the polyhedral runs on the GPU: f of
    test/polyhedral_test.jl:3-11 (8 start solutions, 6 finite; 3 in the torus) and
    cyclic-5 (70 solutions, :39-42)."""
    (x, y), _ = variables(2)
    f = System([2 * y + 3 * y ** 2 - x * y ** 3, x + 4 * x ** 2 - 2 * x ** 3 * y], 2)
    return [("f_affine", f, False, 1), ("f_torus", f, True, 1), ("cyclic5", cyclic(5), True, 3)]


# ------------------------------------------------------------------------------------------
# ModelKit-shaped tapes written by hand: every OpType the C ABI accepts.
# ------------------------------------------------------------------------------------------
class HandTape:
    """Builds an InstructionSequence by hand (slots: constants, parameters, variables, then
    instruction outputs, 1-based as ModelKit's) and its Jacobian tape by forward-mode
    differentiation of the same instructions (each derivative written with ADD, SUB, NEG,
    MUL, DIV, INV, INVSQR and constants)."""

    def __init__(self, nvars, nparams):
        from hcgpu.tape import OP
        self.OP = OP
        self.nvars, self.nparams = nvars, nparams
        self.consts = []
        self.ins = []  # (op name, args: refs or int exponent)

    def c(self, v):
        self.consts.append(complex(v))
        return ("c", len(self.consts) - 1)

    def x(self, i):
        return ("x", i)

    def p(self, i):
        return ("p", i)

    def op(self, name, *args):
        self.ins.append((name, args))
        return ("i", len(self.ins) - 1)

    # forward-mode derivative of every instruction output with respect to x_v
    def _jacobian_instructions(self, outputs):
        base = list(self.ins)
        d = {}  # (ref, v) -> derivative ref, absent = structural zero

        def D(ref, v):
            if ref[0] == "x":
                return ONE if ref[1] == v else None
            if ref[0] in ("c", "p"):
                return None
            return d.get((ref, v))

        def mul(a, b):
            if a is None or b is None:
                return None
            if a == ONE:
                return b
            if b == ONE:
                return a
            return self.op("MUL", a, b)

        def add(a, b):
            if a is None:
                return b
            if b is None:
                return a
            return self.op("ADD", a, b)

        def neg(a):
            return None if a is None else self.op("NEG", a)

        ONE = self.c(1.0)
        for k, (name, args) in enumerate(base):
            out = ("i", k)
            refs = [a for a in args if isinstance(a, tuple)]
            for v in range(self.nvars):
                ds = [D(a, v) for a in refs]
                if all(e is None for e in ds):
                    continue
                a = refs
                if name in ("ADD", "ADD3", "ADD4", "IDENTITY"):
                    r = None
                    for e in ds:
                        r = add(r, e)
                elif name == "SUB":
                    r = add(ds[0], neg(ds[1]))
                elif name == "NEG":
                    r = neg(ds[0])
                elif name in ("MUL", "MUL3", "MUL4"):
                    r = None
                    for j in range(len(a)):
                        if ds[j] is None:
                            continue
                        others = [a[m] for m in range(len(a)) if m != j]
                        prod = others[0] if len(others) == 1 else self.op("MUL" if len(others) == 2 else "MUL3", *others)
                        r = add(r, mul(prod, ds[j]))
                elif name in ("MULADD", "MULSUB", "SUBMUL"):
                    s = -1 if name == "SUBMUL" else 1
                    t = add(mul(a[1], ds[0]), mul(a[0], ds[1]))
                    r = add(t if s > 0 else neg(t), ds[2] if name != "MULSUB" else neg(ds[2]))
                elif name in ("MULMULADD", "MULMULSUB"):
                    t1 = add(mul(a[1], ds[0]), mul(a[0], ds[1]))
                    t2 = add(mul(a[3], ds[2]), mul(a[2], ds[3]))
                    r = add(t1, t2 if name == "MULMULADD" else neg(t2))
                elif name == "SQR":
                    r = mul(self.op("ADD", a[0], a[0]), ds[0])
                elif name == "CB":
                    r = mul(self.op("MUL", self.c(3.0), self.op("SQR", a[0])), ds[0])
                elif name in ("INV", "INV_NOT_ZERO"):
                    r = mul(self.op("NEG", self.op("INVSQR", a[0])), ds[0])
                elif name == "INVSQR":
                    r = mul(self.op("MUL", self.c(-2.0), self.op("POW_INT", a[0], -3)), ds[0])
                elif name == "DIV":
                    r = add(self.op("DIV", ds[0], a[1]) if ds[0] is not None else None,
                            mul(self.op("NEG", self.op("DIV", out, a[1])), ds[1]))
                elif name == "POW_INT":
                    k = args[1]
                    r = mul(self.op("MUL", self.c(float(k)), self.op("POW_INT", a[0], k - 1)), ds[0])
                else:
                    raise ValueError(name)
                if r is not None:
                    d[(out, v)] = r
        jac = [[D(o, v) for v in range(self.nvars)] for o in outputs]
        return jac

    def _slot(self, ref):
        nc, npar = len(self.consts), self.nparams
        if ref[0] == "c":
            return ref[1] + 1
        if ref[0] == "p":
            return nc + ref[1] + 1
        if ref[0] == "x":
            return nc + npar + ref[1] + 1
        return nc + npar + self.nvars + ref[1] + 1

    def tapes(self, outputs, extra_jacobian=()):
        """(eval_jac, eval) Tapes of the outputs.  extra_jacobian: (i, j, ref-builder) terms
        added to U[i, j] in the Jacobian tape only."""
        from hcgpu.tape import Tape
        neval = len(self.ins)
        jac = self._jacobian_instructions(outputs)
        for i, j, build in extra_jacobian:
            term = build(self)
            jac[i][j] = term if jac[i][j] is None else self.op("ADD", jac[i][j], term)
        # constants are numbered first: the slots of every instruction are final only now

        def instr(rows):
            out = []
            for k, (name, args) in enumerate(rows):
                ar = [0, 0, 0, 0]
                for q, a in enumerate(args):
                    ar[q] = a if isinstance(a, int) else self._slot(a)
                out.append((*ar, self.OP[name], self._slot(("i", k))))
            return out

        n = self.nvars
        u = [(i + 1, self._slot(o)) for i, o in enumerate(outputs)]
        U = [(j * n + i + 1, self._slot(jac[i][j])) for i in range(n) for j in range(n) if jac[i][j] is not None]
        ev = Tape(instr(self.ins[:neval]), self.consts, self.nparams, n, u, [], n)
        ej = Tape(instr(self.ins), self.consts, self.nparams, n, u, U, n)
        # the eval tape's tape_space must cover the same slot numbering as the Jacobian tape's
        return ej, ev


FUSED_X = np.array([0.5 + 0.2j, -0.3 + 0.6j, 0.4 - 0.1j])


def fused_op_system():
    """F(x; a) = g(x) - a with n = 3, each g_i written with ModelKit's fused and inverse ops:
      g0 = ADD4(c0 x0, SQR(x1), MULMULADD(x0, x2, c1, x1), IDENTITY(x2))
      g1 = SUBMUL(x1, x2, ADD3(MUL3(c2, x0, x1), CB(x2), DIV(x0, x1 + c4)))
      g2 = MUL4(c3, x0, x1, x2) + INV(x0 + c6) + INVSQR(x1 + c7) + POW_INT(x2, 4)
           + MULMULSUB(x0, x1, x2, c8) + INV_NOT_ZERO(x2 + c9) + POW_INT(x0 + c10, -2)
      f0 = SUB(g0, a0), f1 = MULSUB(1, g1, a1), f2 = MULADD(c11, g2, NEG(a2))  (c11 = 1)
    The Jacobian tape also adds INV_NOT_ZERO(x0 - x0) (an exact zero, so the zero branch) to
    U[0, 0].  Returns (TapeSystem, g as an mpmath function, the start point x*)."""
    from hcgpu.homotopies import TapeSystem
    T = HandTape(3, 3)
    x0, x1, x2 = T.x(0), T.x(1), T.x(2)
    C = [T.c(v) for v in (0.6 + 0.3j, -0.4 + 0.8j, 1.2 - 0.5j, 0.9 + 0.1j, 2.1 + 0.3j, 0, 1.7 - 0.4j,
                          -1.9 + 0.2j, 0.35 - 0.7j, 1.4 + 1.1j, -2.2 + 0.5j, 1.0)]
    g0 = T.op("ADD4", T.op("MUL", C[0], x0), T.op("SQR", x1), T.op("MULMULADD", x0, x2, C[1], x1), T.op("IDENTITY", x2))
    g1 = T.op("SUBMUL", x1, x2, T.op("ADD3", T.op("MUL3", C[2], x0, x1), T.op("CB", x2),
                                     T.op("DIV", x0, T.op("ADD", x1, C[4]))))
    s = T.op("ADD4", T.op("MUL4", C[3], x0, x1, x2), T.op("INV", T.op("ADD", x0, C[6])),
             T.op("INVSQR", T.op("ADD", x1, C[7])), T.op("POW_INT", x2, 4))
    g2 = T.op("ADD3", s, T.op("MULMULSUB", x0, x1, x2, C[8]),
              T.op("ADD", T.op("INV_NOT_ZERO", T.op("ADD", x2, C[9])), T.op("POW_INT", T.op("ADD", x0, C[10]), -2)))
    f0 = T.op("SUB", g0, T.p(0))
    f1 = T.op("MULSUB", C[11], g1, T.p(1))
    f2 = T.op("MULADD", C[11], g2, T.op("NEG", T.p(2)))
    zero_branch = (0, 0, lambda H: H.op("INV_NOT_ZERO", H.op("SUB", H.x(0), H.x(0))))
    ej, ev = T.tapes([f0, f1, f2], extra_jacobian=[zero_branch])
    c = [complex(v) for v in T.consts[:12]]

    def g(x, mp):
        a, b, z = x
        k = [mp.mpc(v) for v in c]
        return [k[0] * a + b ** 2 + (a * z + k[1] * b) + z,
                (k[2] * a * b + z ** 3 + a / (b + k[4])) - b * z,
                k[3] * a * b * z + 1 / (a + k[6]) + 1 / (b + k[7]) ** 2 + z ** 4 + (a * b - z * k[8])
                + 1 / (z + k[9]) + (a + k[10]) ** -2]

    return TapeSystem(ej, ev), g, FUSED_X.copy()
