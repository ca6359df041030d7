"""Every OpType the C ABI accepts, pinned in the oracle against closed forms.

The tapes are written by hand in ModelKit's shape (tests/systems.py::fused_op_system): ADD3,
ADD4, MUL3, MUL4, MULADD, MULSUB, SUBMUL, MULMULADD, MULMULSUB, INV, INV_NOT_ZERO (also at an
exact zero), INVSQR, DIV, IDENTITY, CB, SQR, POW_INT (positive and negative), NEG, ADD, SUB,
MUL (src/model_kit/operations.jl:183-248, emitted by src/model_kit/intermediate_representation.jl
:196-418).  Evaluation, Jacobian, Taylor orders 1..3 (src/model_kit/taylor.jl:272-920) and the
double-double evaluation are compared with the same functions written in mpmath at 200+ bits.
The association orders the reference fixes are checked bitwise on inputs where they matter:
left-associated sums and products in `evaluate` (operations.jl:245-246), pairwise in `taylor!`
(taylor.jl:882-900).
"""
import sys
from pathlib import Path

import mpmath as mp
import numpy as np
import pytest

REPO = Path(__file__).resolve().parents[1]
sys.path.insert(0, str(REPO))
sys.path.insert(0, str(REPO / "homotopycontinuation.jl_amd"))

import oracle  # noqa: E402
import systems  # noqa: E402


def _mpc(z):
    return mp.mpc(complex(z).real, complex(z).imag)


def _points(seed, k=6):
    rng = np.random.default_rng(seed)
    X = systems.FUSED_X + 0.3 * (rng.standard_normal((k, 3)) + 1j * rng.standard_normal((k, 3)))
    A = rng.standard_normal((k, 3)) + 1j * rng.standard_normal((k, 3))
    return X, A


def test_the_system_uses_every_accepted_op():
    F, _, _ = systems.fused_op_system()
    ej, ev = F.tapes
    ops = set(ev.op_counts()) | set(ej.op_counts())
    want = {"ADD3", "ADD4", "MUL3", "MUL4", "MULADD", "MULSUB", "SUBMUL", "MULMULADD", "MULMULSUB", "INV",
            "INV_NOT_ZERO", "INVSQR", "DIV", "IDENTITY", "CB", "SQR", "POW_INT", "NEG", "ADD", "SUB", "MUL"}
    assert want <= ops, want - ops


def test_eval_and_jacobian_match_mpmath():
    F, g, _ = systems.fused_op_system()
    ej, ev = F.tapes
    mp.mp.prec = 200
    X, A = _points(1)
    for x, a in zip(X, A):
        u, U = oracle.sys_eval(ej, x, a)
        u2, _ = oracle.sys_eval(ev, x, a)
        assert np.array_equal(u, u2)  # both tapes compute u with the same instructions
        xm = [_mpc(v) for v in x]
        exact = [gi - _mpc(ai) for gi, ai in zip(g(xm, mp), a)]
        for i in range(3):
            assert abs(complex(exact[i]) - u[i]) <= 1e-12 * max(1.0, abs(complex(exact[i])))
            for j in range(3):
                def gij(h, i=i, j=j):
                    xx = list(xm)
                    xx[j] = xx[j] + h
                    return g(xx, mp)[i]
                dij = complex(mp.diff(gij, 0))
                assert abs(dij - U[i, j]) <= 1e-12 * max(1.0, abs(dij)), (i, j, dij, U[i, j])


@pytest.mark.parametrize("K", [1, 2, 3])
def test_taylor_coefficients_match_mpmath(K):
    F, g, _ = systems.fused_op_system()
    _, ev = F.tapes
    mp.mp.prec = 200
    rng = np.random.default_rng(10 + K)
    X, A = _points(20 + K, 3)
    for x, a in zip(X, A):
        tx = np.vstack([x] + [0.2 * (rng.standard_normal(3) + 1j * rng.standard_normal(3)) for _ in range(K - 1)])
        da = rng.standard_normal(3) + 1j * rng.standard_normal(3)
        out = oracle.sys_taylor(ev, K, tx, a, da)

        def f(s, i):
            xs = [sum(_mpc(tx[k][j]) * s ** k for k in range(K)) for j in range(3)]
            return g(xs, mp)[i] - (_mpc(a[i]) + s * _mpc(da[i]))

        for i in range(3):
            coeffs = mp.taylor(lambda s: f(s, i), 0, K)
            for k in range(K + 1):
                c = complex(coeffs[k])
                assert abs(out[k][i] - c) <= 1e-11 * max(1.0, abs(c)), (K, i, k, out[k][i], c)


def test_double_double_evaluation_matches_mpmath():
    F, g, _ = systems.fused_op_system()
    _, ev = F.tapes
    mp.mp.prec = 320
    rng = np.random.default_rng(5)
    X, A = _points(7, 4)
    for x, a in zip(X, A):
        xl = (rng.standard_normal(3) + 1j * rng.standard_normal(3)) * 1e-17 * np.abs(x)
        uh, ul = oracle.sys_eval_dd(ev, x, xl, a)
        z = [mp.mpc(mp.mpf(h.real) + mp.mpf(l.real), mp.mpf(h.imag) + mp.mpf(l.imag)) for h, l in zip(x, xl)]
        exact = [gi - _mpc(ai) for gi, ai in zip(g(z, mp), a)]
        for i in range(3):
            got = mp.mpc(mp.mpf(uh[i].real) + mp.mpf(ul[i].real), mp.mpf(uh[i].imag) + mp.mpf(ul[i].imag))
            # rounded once to ComplexF64 at the end in the tracker; in double-double here
            assert abs(got - exact[i]) <= 1e-27 * max(1.0, abs(exact[i])), (i, complex(got - exact[i]))
    mp.mp.prec = 200


def _tiny(ops_and_args, nvars):
    """Tapes of the outputs ops_and_args(T) builds over nvars variables (HandTape)."""
    T = systems.HandTape(nvars, 0)
    return T.tapes(ops_and_args(T))


def test_inv_not_zero_returns_zero_at_zero():
    # u = [INV_NOT_ZERO(x0), INV(x1)]: at x0 = 0 the first output is 0, not inf
    ej, ev = _tiny(lambda T: [T.op("INV_NOT_ZERO", T.x(0)), T.op("INV", T.x(1))], 2)
    u, _ = oracle.sys_eval(ev, [0.0, 2.0 + 0j])
    assert u[0] == 0 and u[1] == 0.5
    u, _ = oracle.sys_eval(ev, [4.0 + 0j, 2.0 + 0j])
    assert u[0] == 0.25
    uh, ul = oracle.sys_eval_dd(ev, [0.0, 2.0], [0.0, 0.0])
    assert uh[0] == 0 and ul[0] == 0 and uh[1] == 0.5
    # the fused system's Jacobian tape adds INV_NOT_ZERO(x0 - x0) to U[0, 0]: finite there
    F, g, x = systems.fused_op_system()
    _, U = oracle.sys_eval(F.tapes[0], x, np.zeros(3))
    assert np.isfinite(U).all()


def _left(v, op):
    r = v[0]
    for w in v[1:]:
        r = op(r, w)
    return r


def test_add4_mul4_association_eval_left_taylor_pairwise():
    import operator
    # sums: (a + b) + c + d and (a + b) + (c + d) differ in the last bit for these inputs
    s = [1.0, 2.0 ** -53, 2.0 ** -53, 2.0 ** -52]
    left, pair = _left(s, operator.add), (s[0] + s[1]) + (s[2] + s[3])
    assert left != pair
    ej, ev = _tiny(lambda T: [T.op("ADD4", T.x(0), T.x(1), T.x(2), T.x(3)), T.x(1), T.x(2), T.x(3)], 4)
    u, _ = oracle.sys_eval(ev, s)
    assert u[0].real == left and u[0].imag == 0
    out = oracle.sys_taylor(ev, 1, np.array([s], dtype=complex))
    assert out[0][0].real == pair
    # products: real inputs make the complex product the plain FP64 product
    rng = np.random.default_rng(0)
    for _ in range(200):
        m = list(1 + rng.random(4))
        left, pair = _left(m, operator.mul), (m[0] * m[1]) * (m[2] * m[3])
        if left != pair:
            break
    assert left != pair
    ej, ev = _tiny(lambda T: [T.op("MUL4", T.x(0), T.x(1), T.x(2), T.x(3)), T.x(1), T.x(2), T.x(3)], 4)
    u, _ = oracle.sys_eval(ev, m)
    assert u[0].real == left
    out = oracle.sys_taylor(ev, 1, np.array([m], dtype=complex))
    assert out[0][0].real == pair
    # ADD3 / MUL3 are left-associated in both
    ej, ev = _tiny(lambda T: [T.op("ADD3", T.x(0), T.x(1), T.x(2)), T.x(1), T.x(2)], 3)
    s3 = [1.0, 2.0 ** -53, 2.0 ** -53]
    assert (s3[0] + s3[1]) + s3[2] != s3[0] + (s3[1] + s3[2])
    u, _ = oracle.sys_eval(ev, s3)
    out = oracle.sys_taylor(ev, 1, np.array([s3], dtype=complex))
    assert u[0].real == (s3[0] + s3[1]) + s3[2] == out[0][0].real


def test_mulsub_taylor_sign():
    # MULSUB(a, b, c) = a b - c: its series is Σ a_j b_{k-j} - c_k for every order
    # (the reference negates only through the first mulsub of each order, taylor.jl:841-860)
    ej, ev = _tiny(lambda T: [T.op("MULSUB", T.x(0), T.x(1), T.x(2)), T.op("SUBMUL", T.x(0), T.x(1), T.x(2)),
                              T.x(2)], 3)
    rng = np.random.default_rng(3)
    tx = rng.standard_normal((3, 3)) + 1j * rng.standard_normal((3, 3))
    out = oracle.sys_taylor(ev, 3, tx)
    # the inputs' series are x(s) = tx0 + tx1 s + tx2 s²: exact product coefficients
    a, b, c = tx[:, 0], tx[:, 1], tx[:, 2]
    for k in range(4):
        ab = sum(a[j] * b[k - j] for j in range(k + 1) if j < 3 and k - j < 3)
        ck = c[k] if k < 3 else 0
        assert abs(out[k][0] - (ab - ck)) <= 1e-14 * (1 + abs(ab) + abs(ck))
        assert abs(out[k][1] - (ck - ab)) <= 1e-14 * (1 + abs(ab) + abs(ck))
