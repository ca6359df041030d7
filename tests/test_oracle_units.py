"""Pinning the oracle's building blocks against what mathematics fixes.

Each block of the CPU restatement is checked against an independent answer:
mpmath at high precision (double-double, transcendental kernels), closed-form
polynomial evaluation and exact truncated-series composition (SLP evaluation,
Jacobian, Taylor coefficients), numpy/LAPACK (linear solves, exact inverse
norms for the condition estimate), closed-form valuations of pure power series,
and the literal values the reference's own unit tests assert
(test/norm_test.jl, test/utils_test.jl, test/linear_algebra_test.jl,
test/double_double_test.jl, test/valuation_test.jl).
"""
import math

import mpmath
import numpy as np
import pytest

import oracle
import systems
from hcgpu.configs import dense_system
from hcgpu.expr import ExprSystem, expr_variables
from hcgpu.homotopies import total_degree
from hcgpu.polys import System, variables
from hcgpu.solve import make_options
from hcgpu.tape import tapes

mpmath.mp.prec = 240


# ---------------------------------------------------------------- double-double
def _mp(h, l=0.0):
    return mpmath.mpf(h) + mpmath.mpf(l)


def _rand_dd(rng, scale=1.0):
    h = rng.standard_normal() * scale
    return h, h * 2.0 ** -53 * rng.uniform(-0.5, 0.5)


@pytest.mark.parametrize("op", [0, 1, 2, 3])
def test_dd_ops_match_mpmath(op):
    """DoubleDouble.jl ops (src/DoubleDouble.jl:14-331) to ~2^-100 relative, as
    test/double_double_test.jl:67-95 compares against BigFloat."""
    rng = np.random.default_rng(10 + op)
    worst = 0.0
    for _ in range(2000):
        a = _rand_dd(rng, 10.0 ** rng.integers(-5, 6))
        b = _rand_dd(rng, 10.0 ** rng.integers(-5, 6))
        r = oracle.dd_op(op, a, b)
        A, B = _mp(*a), _mp(*b)
        exact = [A + B, A - B, A * B, A / B][op]
        got = _mp(*r)
        # add/sub: error bounded relative to the operands (cancellation), mul/div: relative to the result
        ref = max(abs(A), abs(B)) if op < 2 else abs(exact)
        worst = max(worst, float(abs(got - exact) / ref))
        assert abs(r[1]) <= 0.5 * math.ulp(r[0]) * 1.0000001 or r[0] == 0.0, "result is not normalized"
    assert worst < 2.0 ** -100, worst


def test_dd_mul_is_not_plain_float():
    """A dropped low-order term would show here: (1 + 2^-60)^2 keeps its 2^-59 term."""
    h, l = oracle.dd_op(2, (1.0, 2.0 ** -60), (1.0, 2.0 ** -60))
    assert h == 1.0 and l == 2.0 ** -59


# ---------------------------------------------------------------- transcendental kernels
@pytest.mark.parametrize("which,name", [(0, "log"), (1, "exp"), (2, "pow"), (3, "cbrt")])
def test_math_kernels_faithful(which, name):
    rng = np.random.default_rng(which)
    worst = 0.0
    for _ in range(3000):
        if which == 0:
            x, y = float(10.0 ** rng.uniform(-300, 300)), 0.0
            exact = mpmath.log(x)
        elif which == 1:
            x, y = float(rng.uniform(-700, 700)), 0.0
            exact = mpmath.exp(x)
        elif which == 2:
            x, y = float(10.0 ** rng.uniform(-20, 20)), float(rng.uniform(-6, 6))
            exact = mpmath.power(mpmath.mpf(x), mpmath.mpf(y))
        else:
            x, y = float(rng.standard_normal() * 10.0 ** rng.uniform(-100, 100)), 0.0
            exact = mpmath.cbrt(abs(x)) * (1 if x >= 0 else -1)
        got = oracle.math_fn(which, x, y)
        ulps = float(abs(mpmath.mpf(got) - exact)) / math.ulp(float(exact)) if float(exact) != 0 else 0.0
        if which == 2:
            # pow = exp(y log x) in FP64: the rounding of y·log x is amplified by |y log x|
            ulps /= 1.0 + abs(y * math.log(x))
        worst = max(worst, ulps)
    # the kernels are built from IEEE-exact operations so both sides agree bit for bit; they are
    # faithful (within ~1 ulp) rather than correctly rounded
    assert worst <= 1.5, f"{name}: {worst} ulp"


def test_pow_in_nthroot_range():
    """nthroot(x, N) = x^(1/N) for N >= 5 (src/utils.jl:408-422) is the only use of pow:
    there |y| <= 1/5 and the error stays within 2 (1 + |log x| / N) ulp (the FP64 rounding of
    log(x)/N, amplified by exp).  Julia's x^(1/N) is correctly rounded, so step sizes may differ
    from the reference's in the last bits; the kernel and the oracle share this function."""
    rng = np.random.default_rng(5)
    for _ in range(2000):
        x = float(10.0 ** rng.uniform(-40, 40))
        N = int(rng.integers(5, 13))
        got = oracle.lib().orc_nthroot(x, N)
        exact = mpmath.root(mpmath.mpf(x), N)
        bound = 2.0 * (1 + abs(math.log(x)) / N) * math.ulp(float(exact))
        assert float(abs(mpmath.mpf(got) - exact)) <= bound


def test_math_kernels_special_values():
    assert oracle.math_fn(0, 1.0) == 0.0
    assert oracle.math_fn(1, 0.0) == 1.0
    assert oracle.math_fn(2, 7.3, 0.0) == 1.0
    assert oracle.math_fn(2, 2.0, 10.0) == 1024.0
    assert oracle.math_fn(3, 27.0) == 3.0
    assert oracle.math_fn(3, -8.0) == -2.0
    assert math.isinf(oracle.math_fn(0, 0.0)) and oracle.math_fn(0, 0.0) < 0
    assert math.isnan(oracle.math_fn(0, -1.0))


# ---------------------------------------------------------------- utils / norm
def test_nthroot_values():
    """test/utils_test.jl nthroot asserts (the kernels are correctly rounded here, so the
    reference's `==` against libm holds up to one ulp)."""
    assert oracle.lib().orc_nthroot(42.2, 0) == 1.0
    assert oracle.lib().orc_nthroot(42.2, 1) == 42.2
    assert oracle.lib().orc_nthroot(42.1, 2) == math.sqrt(42.1)
    assert oracle.lib().orc_nthroot(42.1, 4) == math.sqrt(math.sqrt(42.1))
    assert abs(oracle.lib().orc_nthroot(42.1, 3) - float(mpmath.cbrt(42.1))) <= math.ulp(3.4)
    assert abs(oracle.lib().orc_nthroot(42.1, 5) - 42.1 ** (1 / 5)) <= math.ulp(2.1)


def test_eps_is_julia_eps():
    for x in [1.0, 3.7, 1e-300, 1e300, 2.0 ** -1022, 0.5, 123456.789]:
        assert oracle.lib().orc_eps(x) == np.spacing(x)


def test_inf_norm_and_weighted_norm():
    """test/norm_test.jl:1-36 (InfNorm, WeightedNorm, overflow fallback)."""
    x = np.array([2.0j, 3.0 - 1j, 5.0 + 2.0j])
    y = np.array([-2.0j, 3.0 - 1j, 5.0 + 2.0j])
    assert oracle.norm(x) == pytest.approx(abs(5 + 2j), rel=1e-15)
    assert oracle.norm(x - y) == pytest.approx(4.0, rel=1e-15)
    huge = 2.0 ** 700
    assert oracle.norm(huge * x) == pytest.approx(huge * abs(5 + 2j), rel=1e-15)
    assert oracle.norm(huge * x - huge * y) == pytest.approx(huge * 4.0, rel=1e-15)
    w = np.array([4.0, 2.0, 2.0])
    assert oracle.norm(x, w) == pytest.approx(0.5 * abs(5 + 2j), rel=1e-15)
    assert oracle.norm(x - y, w) == pytest.approx(0.25 * 4.0, rel=1e-15)
    assert oracle.norm(huge * x, w) == pytest.approx(huge * 0.5 * abs(5 + 2j), rel=1e-15)


def test_segment_stepper_forward_and_backward():
    """test/utils_test.jl SegmentStepper testset, value for value."""
    e = 2.0 ** -60
    r = oracle.segment_run(0, 1, [e, -1, 2.0 ** -50, 4, -1])
    assert r[0]["abs_delta"] == 1.0 and r[0]["forward"]
    assert r[0]["ds"] == e and r[0]["s1"] == e and r[0]["t1"] == e and r[0]["t"] == 0 and r[0]["dt"] == e
    assert r[1]["s"] == e and r[1]["t"] == e and r[1]["t1"] == e
    assert r[2]["t"] == e and r[2]["s1"] == e + 2.0 ** -50 and r[2]["t1"] == e + 2.0 ** -50
    assert r[3]["s1"] == 1.0 and r[3]["t1"] == 1.0 and not r[3]["done"]
    assert r[4]["done"]

    e = 2.0 ** -20
    r = oracle.segment_run(1j, 0, [e, -1, 2.0 ** -10, 2, -1])
    assert r[0]["abs_delta"] == 1.0 and not r[0]["forward"]
    assert r[0]["ds"] == e and r[0]["s1"] == 1 - e and r[0]["t1"] == 1j * (1 - e)
    assert r[0]["t"] == 1j and r[0]["dt"] == -1j * e
    assert r[1]["s"] == 1 - e and r[1]["t"] == 1j * (1 - e) and r[1]["t1"] == 1j * (1 - e)
    assert r[2]["t"] == 1j * (1 - e) and r[2]["s1"] == 1 - (e + 2.0 ** -10)
    assert r[2]["t1"] == 1j * (1 - (e + 2.0 ** -10))
    assert r[3]["s1"] == 0.0 and r[3]["t1"] == 0.0 and not r[3]["done"]
    assert r[4]["done"]


# ---------------------------------------------------------------- SLP evaluation
def _poly_eval(system: System, x, p=()):
    z = list(x) + list(p)
    return np.array([sum(c * np.prod([zi ** e for zi, e in zip(z, ex)]) for ex, c in f.terms.items())
                     for f in system.polys])


def _poly_jac(system: System, x, p=()):
    z = list(x) + list(p)
    n = system.nvars
    J = np.zeros((len(system.polys), n), dtype=complex)
    for i, f in enumerate(system.polys):
        for ex, c in f.terms.items():
            for j in range(n):
                if ex[j]:
                    d = list(ex)
                    d[j] -= 1
                    J[i, j] += c * ex[j] * np.prod([zi ** e for zi, e in zip(z, d)])
    return J


@pytest.mark.parametrize("n,seed", [(3, 1), (4, 2), (6, 3)])
def test_slp_eval_and_jacobian_closed_form(n, seed):
    F = dense_system(n, 3, seed=seed)
    ej, ev = tapes(F)
    rng = np.random.default_rng(seed)
    for _ in range(3):
        x = (rng.standard_normal(n) + 1j * rng.standard_normal(n)) * 0.7
        u, U = oracle.sys_eval(ej, x)
        ue, _ = oracle.sys_eval(ev, x)
        np.testing.assert_allclose(u, _poly_eval(F, x), rtol=1e-12, atol=1e-12)
        np.testing.assert_allclose(ue, u, rtol=1e-13, atol=1e-13)
        np.testing.assert_allclose(U, _poly_jac(F, x), rtol=1e-12, atol=1e-12)


def test_slp_with_parameters_closed_form():
    F = systems.tracker_param_system()
    ej, _ = tapes(F)
    x, p = np.array([1.3 - 0.2j, -0.7 + 1j]), np.array([0.4 + 0.1j, -2.0 + 0.5j])
    u, U = oracle.sys_eval(ej, x, p)
    np.testing.assert_allclose(u, [x[0] ** 2 - p[0], x[0] * x[1] - p[0] + p[1]], rtol=1e-14)
    np.testing.assert_allclose(U, [[2 * x[0], 0], [x[1], x[0]]], rtol=1e-14)


def _series_compose(system: System, tx, K, p0=None, p1=None):
    """Exact truncated composition F(x(s); p0 + s p1) mod s^(K+1) with numpy polynomial
    arithmetic; coefficient k of the result is row k (the normalized Taylor coefficient)."""
    n = system.nvars
    xs = [np.array([tx[k][i] if k < len(tx) else 0 for k in range(K + 1)], dtype=complex) for i in range(n)]
    ps = []
    for j in range(system.nparams):
        c = np.zeros(K + 1, dtype=complex)
        c[0] = p0[j]
        if p1 is not None and K >= 1:
            c[1] = p1[j]
        ps.append(c)
    z = xs + ps

    def mul(a, b):
        return np.convolve(a, b)[:K + 1]

    out = []
    for f in system.polys:
        acc = np.zeros(K + 1, dtype=complex)
        for ex, c in f.terms.items():
            term = np.zeros(K + 1, dtype=complex)
            term[0] = c
            for zi, e in zip(z, ex):
                for _ in range(e):
                    term = mul(term, zi)
            acc += term
        out.append(acc)
    return np.array(out).T  # (K+1, m)


@pytest.mark.parametrize("K", [1, 2, 3])
def test_taylor_coefficients_closed_form(K):
    """taylor! (src/model_kit/taylor.jl) = the K-th coefficient of the exact composition."""
    F = dense_system(4, 3, seed=7)
    _, ev = tapes(F)
    rng = np.random.default_rng(K)
    tx = (rng.standard_normal((K, 4)) + 1j * rng.standard_normal((K, 4))) * 0.6
    got = oracle.sys_taylor(ev, K, tx)
    want = _series_compose(F, tx, K)
    np.testing.assert_allclose(got[K], want[K], rtol=1e-12, atol=1e-12)
    np.testing.assert_allclose(got[0], want[0], rtol=1e-12, atol=1e-12)


@pytest.mark.parametrize("K", [1, 2, 3])
def test_taylor_with_parameter_direction(K):
    """Parameter homotopy series p(t + s) = p(t) + s (p - q) (src/homotopies/parameter_homotopy.jl:66-102)."""
    F = systems.tracker_param_system()
    _, ev = tapes(F)
    rng = np.random.default_rng(20 + K)
    tx = rng.standard_normal((K, 2)) + 1j * rng.standard_normal((K, 2))
    p0 = rng.standard_normal(2) + 1j * rng.standard_normal(2)
    p1 = rng.standard_normal(2) + 1j * rng.standard_normal(2)
    got = oracle.sys_taylor(ev, K, tx, p0, p1)
    want = _series_compose(F, tx, K, p0, p1)
    np.testing.assert_allclose(got[K], want[K], rtol=1e-12, atol=1e-12)


def test_factored_expression_tapes_closed_form():
    """The expression front end keeps (x - 10)^6 factored; value, derivative and Taylor
    coefficients match the closed forms and stay accurate next to the 6-fold root."""
    (x,), _ = expr_variables(1)
    F = ExprSystem([(x - 10) ** 6 + 3 * x], 1)
    from hcgpu.expr import expr_tapes
    ej, ev = expr_tapes(F)
    for xv in [10.0 + 1e-3j, 3.0 - 2.0j]:
        u, U = oracle.sys_eval(ej, [xv])
        assert abs(u[0] - ((xv - 10) ** 6 + 3 * xv)) <= 1e-13 * abs(u[0])
        assert abs(U[0, 0] - (6 * (xv - 10) ** 5 + 3)) <= 1e-13 * abs(U[0, 0])
    tx = np.array([[10.0 + 1e-2j], [0.5 + 0.0j], [0.25j]])
    got = oracle.sys_taylor(ev, 3, tx)
    s = np.array([tx[0, 0] - 10, tx[1, 0], tx[2, 0], 0])
    p6 = np.array([1, 0, 0, 0], dtype=complex)
    for _ in range(6):
        p6 = np.convolve(p6, s)[:4]
    want = p6 + 3 * np.array([tx[0, 0], tx[1, 0], tx[2, 0], 0])
    np.testing.assert_allclose(got[:, 0], want, rtol=1e-12, atol=1e-15)


def test_dd_evaluation_matches_mpmath():
    """The double-double residual (interpreted_system.jl eval_ComplexDF64) is accurate far
    beyond FP64 at a near-root where the FP64 evaluation cancels."""
    F = dense_system(3, 3, seed=11)
    _, ev = tapes(F)
    rng = np.random.default_rng(3)
    xh = rng.standard_normal(3) + 1j * rng.standard_normal(3)
    xl = (rng.standard_normal(3) + 1j * rng.standard_normal(3)) * 1e-17
    uh, ul = oracle.sys_eval_dd(ev, xh, xl)
    mpmath.mp.prec = 300
    z = [mpmath.mpc(mpmath.mpf(a.real) + mpmath.mpf(b.real), mpmath.mpf(a.imag) + mpmath.mpf(b.imag))
         for a, b in zip(xh, xl)]
    for i, f in enumerate(F.polys):
        exact = mpmath.mpc(0)
        for ex, c in f.terms.items():
            m = mpmath.mpc(c)
            for zi, e in zip(z, ex):
                m *= zi ** e
            exact += m
        got = mpmath.mpc(mpmath.mpf(uh[i].real) + mpmath.mpf(ul[i].real), mpmath.mpf(uh[i].imag) + mpmath.mpf(ul[i].imag))
        scale = sum(abs(c) * max(1.0, float(abs(max(z, key=abs)))) ** sum(ex) for ex, c in f.terms.items())
        assert abs(got - exact) / scale < 1e-28
    mpmath.mp.prec = 240


# ---------------------------------------------------------------- linear algebra
@pytest.mark.parametrize("n", [3, 8, 13])
def test_lu_solve_matches_lapack(n):
    """test/linear_algebra_test.jl ldiv testset: agreement within cond(A) * rounding."""
    rng = np.random.default_rng(n)
    for _ in range(5):
        A = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
        b = rng.standard_normal(n) + 1j * rng.standard_normal(n)
        x, _ = oracle.linsolve(A, b)
        xr = np.linalg.solve(A, b)
        assert np.linalg.norm(x - xr) / np.linalg.norm(xr) <= np.linalg.cond(A) * 1e-15


def test_lu_solve_skeel_weights():
    """ldiv!(x, J, b, WeightedNorm) (src/linear_algebra.jl:846-862) still solves A x = b."""
    rng = np.random.default_rng(5)
    A = (rng.standard_normal((6, 6)) + 1j * rng.standard_normal((6, 6))) * np.logspace(-6, 6, 6)[:, None]
    b = rng.standard_normal(6) + 1j * rng.standard_normal(6)
    x, _ = oracle.linsolve(A, b, weights=np.ones(6))
    xr = np.linalg.solve(A, b)
    assert np.linalg.norm(x - xr) / np.linalg.norm(xr) <= 1e-9


def test_cond_estimate_within_factor_10():
    """test/linear_algebra_test.jl 'Inf-norm estimator / cond': the Higham estimate of
    cond(diag(d_l) A diag(d_r)) is within [0.1, 10] of the exact ∞-norm condition."""
    rng = np.random.default_rng(17)
    for trial in range(20):
        d_r = rng.random() * np.logspace(-6, 6, 6)
        d_l = rng.random() * np.logspace(6, -6, 6)
        A = rng.standard_normal((6, 6)).astype(complex)
        b = np.ones(6, dtype=complex)

        def exact(M):
            return np.linalg.norm(M, np.inf) * np.linalg.norm(np.linalg.inv(M), np.inf)

        B = A @ np.diag(1 / d_r)
        _, c = oracle.linsolve(B, b)
        assert 0.1 <= exact(B) / c <= 10
        _, c = oracle.linsolve(B, b, dr=d_r)
        assert 0.1 <= exact(B @ np.diag(d_r)) / c <= 10
        _, c = oracle.linsolve(B, b, dl=d_r)
        assert 0.1 <= exact(np.diag(d_r) @ B) / c <= 10
        C = np.diag(1 / d_l) @ A
        _, c = oracle.linsolve(C, b, dl=d_l)
        assert 0.1 <= exact(np.diag(d_l) @ C) / c <= 10
        D = np.diag(1 / d_l) @ A @ np.diag(1 / d_r)
        _, c = oracle.linsolve(D, b, dl=d_l, dr=d_r)
        assert 0.1 <= exact(np.diag(d_l) @ D @ np.diag(d_r)) / c <= 10


def test_rank_of_low_rank_products():
    rng = np.random.default_rng(9)
    for n, r in [(4, 2), (6, 5), (5, 5), (7, 1)]:
        A = (rng.standard_normal((n, r)) + 1j * rng.standard_normal((n, r))) @ \
            (rng.standard_normal((r, n)) + 1j * rng.standard_normal((r, n)))
        assert oracle.rank(A, 1e-12) == r


# ---------------------------------------------------------------- valuation
@pytest.mark.parametrize("v", [-1.0, -1 / 3, 0.5, 2.0])
def test_valuation_of_pure_power_series(v):
    """x(t) = c t^v: both valuation estimates are exactly v (val_x) and v (val_tẋ)
    (src/valuation.jl:52-228 on data whose answer is known in closed form)."""
    t = 1e-6
    c = np.array([1.3 - 0.4j, -0.2 + 2.0j])
    tx = np.concatenate([c * t ** v, c * v * t ** (v - 1), c * v * (v - 1) / 2 * t ** (v - 2),
                         c * v * (v - 1) * (v - 2) / 6 * t ** (v - 3)])
    vx, vt = np.zeros(2), np.zeros(2)
    oracle.lib().orc_valuation(2, tx.view(np.float64).ctypes.data_as(oracle.ctypes.POINTER(oracle.ctypes.c_double)), t,
                               vx.ctypes.data_as(oracle.ctypes.POINTER(oracle.ctypes.c_double)),
                               vt.ctypes.data_as(oracle.ctypes.POINTER(oracle.ctypes.c_double)))
    np.testing.assert_allclose(vx, [v, v], atol=1e-9)
    np.testing.assert_allclose(vt, [v, v], atol=1e-9)


def test_valuation_example_1():
    """test/valuation_test.jl Example 1: (x - 10)^5 tracked to t = 1e-13."""
    (x,), _ = expr_variables(1)
    H, S = total_degree(ExprSystem([(x - 10) ** 5], 1), seed=1)
    code, vx, vt = oracle.track_plain_valuation(H, S[0], 1.0, 1e-13)
    assert code == "success"
    assert abs(vx[0]) <= 10 * 1e-13 ** (1 / 5)
    assert abs(vt[0] - 1 / 5) <= 10 * 1e-13 ** (1 / 5)


def test_valuation_example_2():
    """test/valuation_test.jl Example 2: diverging path, valuation ≈ [-1, -1]."""
    H, S = total_degree(systems.beyond_polyhedral(), gamma=1.3j + 0.4)
    tf = 1e-10
    code, vx, vt = oracle.track_plain_valuation(H, S[2], 1.0, tf, make_options({"parameters": "conservative"}))
    assert code == "success"
    np.testing.assert_allclose(vx, [-1, -1], atol=10 * tf ** 0.5)
    np.testing.assert_allclose(vt, [-1, -1], atol=tf ** 0.5)


def test_valuation_example_3():
    """test/valuation_test.jl Example 3: fractional valuations -1/6, -2/6."""
    H, S = total_degree(systems.winding_family(2), gamma=1.3j + 0.4)
    tf = 1e-10
    code, vx, vt = oracle.track_plain_valuation(H, S[2], 1.0, tf, make_options({"parameters": "conservative"}))
    assert code == "success"
    np.testing.assert_allclose(vx, [-1 / 6, -2 / 6], atol=tf ** (1 / 6))
    np.testing.assert_allclose(vt, [-1 / 6, -2 / 6], atol=tf ** (1 / 6))


def test_tape_emission_orders_are_equivalent():
    """The tape producer may emit the polynomials one after the other, interleaved over shared
    monomials, row by row, or row by row with per-row monomials (whichever keeps fewer values
    live); every order computes each output with the same operations, so evaluation results
    are bit for bit the same."""
    from hcgpu import tape
    from hcgpu.configs import cfg2_problem
    from hcgpu.homotopies import affine_parameter_lift
    F2, _, _ = cfg2_problem(1)
    for S in [dense_system(5, 3, seed=4), affine_parameter_lift(F2)[0]]:
        variants = []
        orig = tape._best
        tape._best = lambda makers: (variants.extend(m() for m in makers), variants[0])[1]
        try:
            tape.jacobian_tape(S)
            tape.eval_tape(S)
        finally:
            tape._best = orig
        assert len(variants) == 6
        rng = np.random.default_rng(1)
        x = rng.standard_normal(S.nvars) + 1j * rng.standard_normal(S.nvars)
        p = rng.standard_normal(S.nparams) + 1j * rng.standard_normal(S.nparams)
        ref_u, ref_U = oracle.sys_eval(variants[0], x, p)
        for t in variants[1:4]:
            u, U = oracle.sys_eval(t, x, p)
            assert np.array_equal(u, ref_u) and np.array_equal(U, ref_U)
        ev = [oracle.sys_eval(t, x, p)[0] for t in variants[4:]]
        assert np.array_equal(ev[0], ev[1]) and np.array_equal(ev[0], ref_u)
        tx = rng.standard_normal((3, S.nvars)) + 1j * rng.standard_normal((3, S.nvars))
        dp = rng.standard_normal(S.nparams) + 1j * rng.standard_normal(S.nparams)
        t0 = oracle.sys_taylor(variants[4], 3, tx, p if S.nparams else None, dp if S.nparams else None)
        t1 = oracle.sys_taylor(variants[5], 3, tx, p if S.nparams else None, dp if S.nparams else None)
        assert np.array_equal(t0, t1)


def test_expression_tape_rematerialization_is_equivalent():
    """Expression systems: the tape with shared subexpressions and the tape that recomputes
    products of leaves at every use evaluate, differentiate and expand (Taylor K = 1..3)
    bit for bit identically; the code generator counts the recomputations' flops once."""
    import ctypes
    from hcgpu import abi, expr
    from hcgpu.configs import shared_factor_exprs
    from hcgpu.homotopies import total_degree
    E = shared_factor_exprs(6, seed=3)
    n = E.nvars

    def lowered(remat):
        memo = {}
        J = [[expr.differentiate(e, j, memo) for j in range(n)] for e in E.exprs]
        Lj = expr._Lower(n, 0, remat)
        seen = set()
        for e in list(E.exprs) + [d for r in J for d in r]:
            Lj.count_uses(e, seen)
        ej = Lj.finish([Lj.lower(e) for e in E.exprs], [[None if d.is_const(0) else Lj.lower(d) for d in r] for r in J],
                       n)
        Le = expr._Lower(n, 0, remat)
        seen = set()
        for e in E.exprs:
            Le.count_uses(e, seen)
        return ej, Le.finish([Le.lower(e) for e in E.exprs], None, n)

    (ej0, ev0), (ej1, ev1) = lowered(False), lowered("use")
    assert len(ej1.instr) > len(ej0.instr) and len(ev1.instr) > len(ev0.instr)
    rng = np.random.default_rng(5)
    x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    u0, U0 = oracle.sys_eval(ej0, x)
    u1, U1 = oracle.sys_eval(ej1, x)
    assert np.array_equal(u0, u1) and np.array_equal(U0, U1)
    assert np.array_equal(oracle.sys_eval(ev0, x)[0], oracle.sys_eval(ev1, x)[0])
    for K in (1, 2, 3):
        tx = rng.standard_normal((K, n)) + 1j * rng.standard_normal((K, n))
        assert np.array_equal(oracle.sys_taylor(ev0, K, tx), oracle.sys_taylor(ev1, K, tx))
    # the chosen tapes: counted flops equal those of the shared tape
    assert expr._live_peak(expr.expr_tapes(E)[1]) <= expr._live_peak(ev0)
    H, _ = total_degree(E, seed=1)
    L = abi.lib()
    need, fl = ctypes.c_size_t(), (ctypes.c_long * 6)()
    abi.check(L.hcg_generate(ctypes.byref(H.desc()), None, 0, ctypes.byref(need), fl))
    orig = expr.expr_tapes
    try:
        expr.expr_tapes = lambda system: lowered(False)
        import hcgpu.homotopies as hm
        H0, _ = hm.total_degree(shared_factor_exprs(6, seed=3), seed=1)
        fl0 = (ctypes.c_long * 6)()
        abi.check(L.hcg_generate(ctypes.byref(H0.desc()), None, 0, ctypes.byref(need), fl0))
    finally:
        expr.expr_tapes = orig
    assert list(fl) == list(fl0)
