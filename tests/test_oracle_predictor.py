"""Closed-form pins of the oracle's predictor (src/predictor.jl:158-371).

The oracle's update!/predict! run through `oracle.predictor_probe` on points of paths whose
x(t) is known in closed form, so each check compares against mathematics, not against a
second transcription of the reference's formulas:

  * x(t) = sqrt(y(t)) with y affine in t: x¹, x², x³ are the normalized derivatives
    d/(2x), -d²/(8x³), d³/(16x⁵); the Padé(2,1) error is (x⁴ - x³²/x²) Δt⁴ + O(Δt⁵), so the
    prediction error falls 16-fold when Δt halves and local_error tends to that constant.
  * x(t) = (a + b t)/(c + d t): Padé(2,1) is exact for every Δt, and the trust region is
    |x²|/|x³| = |c + d t|/|d|.
  * x(t) affine in t: x² = x³ = 0, the Taylor branch, exact.
  * winding number m with x(s) a cubic polynomial of s = t^{1/m}: the cubic Hermite
    predictor in s is exact, for real and for complex t.

A planted Taylor-instead-of-Padé, a wrong λ power in τ, a wrong Hermite node or derivative
factor each fail one of these (test_planted_mistakes_fail re-derives the predictions with
such mistakes on the CPU and checks that the assertions would catch them).
"""
import math
import sys
from pathlib import Path

import mpmath as mp
import numpy as np
import pytest

REPO = Path(__file__).resolve().parents[1]
sys.path.insert(0, str(REPO))
sys.path.insert(0, str(REPO / "homotopycontinuation.jl_amd"))

import oracle  # noqa: E402
from hcgpu.homotopies import ParameterHomotopy  # noqa: E402
from hcgpu.polys import System, variables  # noqa: E402

mp.mp.prec = 200


def _squares_system(A):
    """F_i(x; p) = Σ_j A_ij x_j² - p_i: the path is y = x² = A⁻¹ p(t)."""
    n = len(A)
    xs, ps = variables(n, n)
    return System([sum(A[i][j] * xs[j] ** 2 for j in range(n)) - ps[i] for i in range(n)], n, n)


def _sqrt_path(A, p, q, t):
    """Exact x(t) and its normalized derivatives x¹..x⁴ (mpmath) on the principal branch."""
    Am = mp.matrix([[mp.mpc(complex(a)) for a in row] for row in A])
    pt = [mp.mpf(t) * mp.mpc(complex(a)) + (1 - mp.mpf(t)) * mp.mpc(complex(b)) for a, b in zip(p, q)]
    y = mp.lu_solve(Am, mp.matrix(pt))
    d = mp.lu_solve(Am, mp.matrix([mp.mpc(complex(a)) - mp.mpc(complex(b)) for a, b in zip(p, q)]))
    out = []
    for i in range(len(A)):
        x0 = mp.sqrt(y[i])
        # sqrt(y + s d) = x0 Σ_k binom(1/2, k) (s d / y)^k
        r = d[i] / y[i]
        out.append([x0 * mp.binomial(mp.mpf(1) / 2, k) * r ** k for k in range(5)])
    return out  # out[i][k]


def _cplx(v):
    return complex(v)


A2 = [[1.3 + 0.2j, -0.4 + 0.7j], [0.5 - 0.3j, 0.9 + 0.1j]]
P2 = [1.1 + 0.6j, -0.7 + 0.9j]
Q2 = [0.4 - 1.2j, 1.5 + 0.3j]


def _sqrt_probe(ts, dt):
    F = _squares_system(A2)
    H = ParameterHomotopy(F, P2, Q2)
    xs = [[_cplx(c[0]) for c in _sqrt_path(A2, P2, Q2, t)] for t in ts]
    return oracle.predictor_probe(H, xs, ts, dt), F


def test_taylor_coefficients_match_the_closed_form():
    t = 0.63
    r, _ = _sqrt_probe([t], -0.01)
    ref = _sqrt_path(A2, P2, Q2, t)
    for k in range(4):
        for i in range(2):
            exact = _cplx(ref[i][k])
            assert abs(r["tx"][k][i] - exact) <= 1e-12 * abs(exact), (k, i, r["tx"][k][i], exact)
    assert r["hermite"][0] == 0 and r["order"][0] == 4
    # no previous prediction: local error = τ⁻⁴ (src/predictor.jl:279-281)
    assert r["local_error"][0] == pytest.approx(r["trust_region"][0] ** -4, rel=1e-15)
    # trust region: min_i |x²_i| / |x³_i| (the λ-scaling of src/predictor.jl:254-275 cancels
    # whenever |x¹_i| >= 1e-6); for the square-root path that is |2 y_i / d_i|
    tau = min(abs(_cplx(ref[i][2])) / abs(_cplx(ref[i][3])) for i in range(2))
    assert r["trust_region"][0] == pytest.approx(tau, rel=1e-12)


def test_pade_prediction_error_is_fourth_order_with_the_pade_constant():
    t = 0.63
    ref = _sqrt_path(A2, P2, Q2, t)
    # leading error of the (2,1) Padé approximant: (x⁴ - x³²/x²) Δt⁴
    lead = [abs(_cplx(ref[i][4] - ref[i][3] ** 2 / ref[i][2])) for i in range(2)]
    errs = []
    for dt in (-0.02, -0.01, -0.005):
        r, _ = _sqrt_probe([t], dt)
        exact = [_cplx(c[0]) for c in _sqrt_path(A2, P2, Q2, t + dt)]
        e = max(abs(r["xh"][i] - exact[i]) for i in range(2))
        errs.append(e)
        # weights are 1: the weighted norm is the max norm
        assert e == pytest.approx(max(lead) * dt ** 4, rel=0.1), (dt, e, max(lead) * dt ** 4)
    assert errs[0] / errs[1] == pytest.approx(16, rel=0.1)
    assert errs[1] / errs[2] == pytest.approx(16, rel=0.1)


def test_local_error_uses_the_previous_prediction_and_order_four():
    t0, dt = 0.63, -0.01
    r, _ = _sqrt_probe([t0, t0 + dt], -0.01)
    ref = _sqrt_path(A2, P2, Q2, t0)
    lead = max(abs(_cplx(ref[i][4] - ref[i][3] ** 2 / ref[i][2])) for i in range(2))
    x1 = [_cplx(c[0]) for c in _sqrt_path(A2, P2, Q2, t0 + dt)]
    # the prediction made before the second update is predict!(x(t0), t0, dt)
    dist = max(abs(r["xhats"][1][i] - x1[i]) for i in range(2))
    assert r["local_error"][1] == pytest.approx(dist / abs(dt) ** 4, rel=1e-6)
    assert r["local_error"][1] == pytest.approx(lead, rel=0.1)


def _rational_system():
    """F(x; p) = p₁ x - p₂ (n = 1): x(t) = p₂(t)/p₁(t), a (1,1) rational function of t."""
    (x,), (a, b) = variables(1, 2)
    return System([a * x - b], 1, 2)


@pytest.mark.parametrize("dt", [-0.01, -0.3, -0.9, 0.25 + 0.2j])
def test_pade_is_exact_on_a_rational_path(dt):
    F = _rational_system()
    p, q = [1.0 + 0.5j, 2.0 - 1.0j], [0.3 - 0.8j, -0.6 + 0.4j]
    H = ParameterHomotopy(F, p, q)

    def x_of(t):
        t = mp.mpc(t)
        return (t * mp.mpc(p[1]) + (1 - t) * mp.mpc(q[1])) / (t * mp.mpc(p[0]) + (1 - t) * mp.mpc(q[0]))

    t = 0.9
    r = oracle.predictor_probe(H, [[_cplx(x_of(t))]], [t], dt)
    exact = _cplx(x_of(t + dt))
    assert abs(r["xh"][0] - exact) <= 1e-13 * abs(exact)
    # x(t + s) = (α + β s)/(γ + δ s): x^(k+1)/x^(k) = -δ/γ, so τ = |γ/δ|
    g = t * p[0] + (1 - t) * q[0]
    d = p[0] - q[0]
    assert r["trust_region"][0] == pytest.approx(abs(g / d), rel=1e-12)


def test_affine_path_uses_the_taylor_branch_exactly():
    (x,), (a,) = variables(1, 1)
    F = System([2.0 * x - a], 1, 1)  # x(t) = p(t)/2: affine, x² = x³ = 0
    p, q = [1.0 + 1.0j], [-0.5 + 0.25j]
    H = ParameterHomotopy(F, p, q)
    t, dt = 0.7, -0.45
    xt = (t * p[0] + (1 - t) * q[0]) / 2
    r = oracle.predictor_probe(H, [[xt]], [t], dt)
    assert abs(r["tx"][2][0]) == 0 and abs(r["tx"][3][0]) == 0
    exact = ((t + dt) * p[0] + (1 - t - dt) * q[0]) / 2
    assert abs(r["xh"][0] - exact) <= 4e-16 * abs(exact)


def _winding_system(b):
    """F(x, y; p) = [x² - p₁, y - (b₀ + b₁x + b₂x² + b₃x³)] with p₁(t) = t: x = s = t^{1/2},
    y a cubic polynomial of s (winding number 2 at t = 0)."""
    (x, y), (a,) = variables(2, 1)
    return System([x ** 2 - a, y - (b[0] + b[1] * x + b[2] * x ** 2 + b[3] * x ** 3)], 2, 1)


B = [0.3 - 0.2j, 1.1 + 0.4j, -0.7 + 0.9j, 2.0 - 0.5j]


def _winding_point(t):
    s = complex(mp.sqrt(mp.mpc(t))) if complex(t).imag == 0 and complex(t).real > 0 else None
    if s is None:
        tt = complex(t)
        th = math.atan2(tt.imag, tt.real) % (2 * math.pi)
        s = complex(mp.sqrt(abs(tt)) * mp.expj(th / 2))
    return [s, sum(B[k] * s ** k for k in range(4))], s


@pytest.mark.parametrize("ts,dt", [([0.09, 0.04], -0.03), ([0.2 + 0.05j, 0.1 + 0.03j], -0.05 - 0.02j)])
def test_hermite_predictor_is_exact_on_a_cubic_in_s(ts, dt):
    F = _winding_system(B)
    H = ParameterHomotopy(F, [1.0], [0.0])
    xs = [_winding_point(t)[0] for t in ts]
    r = oracle.predictor_probe(H, xs, ts, dt, winding=2)
    assert list(r["hermite"]) == [1, 1]
    exact, _ = _winding_point(complex(ts[-1]) + dt)
    for i in range(2):
        assert abs(r["xh"][i] - exact[i]) <= 1e-12 * max(1.0, abs(exact[i])), (i, r["xh"][i], exact[i])
    # Hermite trust region: ‖x‖ / ‖ẋ‖ (src/predictor.jl:226-236)
    x, xd = np.array(xs[-1]), r["tx"][1]
    assert r["trust_region"][1] == pytest.approx(np.abs(x).max() / np.abs(xd).max(), rel=1e-15)


def test_planted_mistakes_fail():
    """The assertions above would catch the plausible mistakes: each variant below recomputes the
    relevant prediction with one mistake (on the CPU, outside the oracle) and misses the bound
    its test uses."""
    # Taylor instead of Padé on the rational path: error O(Δt³), far above 1e-13
    p, q = [1.0 + 0.5j, 2.0 - 1.0j], [0.3 - 0.8j, -0.6 + 0.4j]
    t, dt = 0.9, -0.3
    g, dg = t * p[0] + (1 - t) * q[0], p[0] - q[0]
    a, da = t * p[1] + (1 - t) * q[1], p[1] - q[1]
    x0 = a / g
    x1 = (da * g - a * dg) / g ** 2
    x2, x3 = x1 * (-dg / g), x1 * (-dg / g) ** 2
    exact = (a + dt * da) / (g + dt * dg)
    taylor = x0 + dt * (x1 + dt * (x2 + dt * x3))
    pade = x0 + dt * (x1 + dt * x2 / (1 - dt * x3 / x2))
    assert abs(pade - exact) <= 1e-13 * abs(exact)
    assert abs(taylor - exact) > 1e-4 * abs(exact)
    # a wrong λ power in τ (c²/c³ · λ instead of c²/c³): misses rel 1e-12 unless |x¹| = 1
    lam = abs(x1)
    assert abs((abs(x2) / abs(x3)) * lam**2 - abs(g / dg)) > 1e-6 * abs(g / dg)
    # Hermite with t-nodes instead of s-nodes, or without the m s^{m-1} factor: not exact
    ts = [0.09, 0.04]
    s = [math.sqrt(v) for v in ts]
    ys = [sum(B[k] * si ** k for k in range(4)) for si in s]
    dys = [sum(k * B[k] * si ** (k - 1) for k in range(1, 4)) for si in s]  # dy/ds
    dyt = [dy / (2 * si) for dy, si in zip(dys, s)]  # dy/dt

    def hermite(a0, a1, t0, b0, b1, t1, tt):
        u = (tt - t0) / (t1 - t0)
        return ((1 + 2 * u) * (1 - u) ** 2 * a0 + (tt - t0) * (1 - u) ** 2 * a1 + u * u * (3 - 2 * u) * b0
                + (tt - t0) * u * (u - 1) * b1)

    s1 = math.sqrt(0.01)
    y_exact = sum(B[k] * s1 ** k for k in range(4))
    good = hermite(ys[0], dys[0], s[0], ys[1], dys[1], s[1], s1)
    assert abs(good - y_exact) <= 1e-12 * abs(y_exact)
    in_t = hermite(ys[0], dyt[0], ts[0], ys[1], dyt[1], ts[1], 0.01)
    no_factor = hermite(ys[0], dyt[0], s[0], ys[1], dyt[1], s[1], s1)
    assert abs(in_t - y_exact) > 1e-6 and abs(no_factor - y_exact) > 1e-6
    # local_error with order 3 (‖x̂ - x‖/Δs³) would be Δs·lead, a hundredth of the bound
    r, _ = _sqrt_probe([0.63, 0.62], -0.01)
    ref = _sqrt_path(A2, P2, Q2, 0.63)
    lead = max(abs(_cplx(ref[i][4] - ref[i][3] ** 2 / ref[i][2])) for i in range(2))
    x1 = [_cplx(c[0]) for c in _sqrt_path(A2, P2, Q2, 0.62)]
    dist = max(abs(r["xhats"][1][i] - x1[i]) for i in range(2))
    assert abs(dist / 0.01 ** 3 - lead) > 0.5 * lead
