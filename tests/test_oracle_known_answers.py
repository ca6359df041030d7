"""The oracle's end-to-end results against the reference's own known answers.

Counts, winding numbers, valuations and endpoints asserted by test/endgame_test.jl,
test/endgame_tracker_test.jl, test/tracker_test.jl and test/result_test.jl, plus an
independent pin of the benchmark-family endpoints: every nonsingular endpoint of a
dense n = 4 total-degree solve is refined by Newton's method in 200-bit arithmetic
(mpmath) and must agree within 1e-10 relative, and the endpoints must be 81 distinct
roots (Bézout's bound, so the solution set is complete).
"""
import mpmath
import numpy as np
import pytest

import oracle
import systems
from hcgpu.configs import dense_system
from hcgpu.expr import ExprSystem, expr_variables
from hcgpu.homotopies import ParameterHomotopy, StraightLineHomotopy, total_degree
from hcgpu.polys import System, variables
from hcgpu.result import Result, path_results_from_batch
from hcgpu.solve import make_options


def solve_oracle(H, S, **opt):
    o = oracle.track_batch(H, S, make_options(**opt))
    return o, Result(path_results_from_batch(o, S))


def test_endgame_tracker_beyond_polyhedral():
    """test/endgame_tracker_test.jl:2-30."""
    H, S = total_degree(systems.beyond_polyhedral(), gamma=1.3j + 0.4)
    o, R = solve_oracle(H, S)
    assert o.codes() == ["success", "success", "at_infinity", "at_infinity"]
    assert np.isfinite(o.solution[0]).all()
    assert o.winding_number[0] == 0  # `nothing`
    assert o.accuracy[0] < 1e-12 and o.residual[0] < 1e-12
    assert o.accepted_steps[0] + o.rejected_steps[0] < 20 and o.rejected_steps[0] == 0
    assert not o.singular[1] and o.cond[1] < 1e3
    np.testing.assert_allclose(o.valuation[2], [-1, -1], rtol=1e-3)
    assert R[0].multiplicity == 1


def test_endgame_tracker_double_root():
    """test/endgame_tracker_test.jl:31-46: (x - 10)^2 has winding number 2, last path point
    at 0 < t < 0.1, endpoints real up to a tiny imaginary part, singular."""
    (x,), _ = expr_variables(1)
    H, S = total_degree(ExprSystem([(x - 10) ** 2], 1), seed=7)
    o, R = solve_oracle(H, S)
    assert (o.winding_number == 2).all()
    assert ((o.last_t > 0) & (o.last_t < 0.1)).all()
    assert all(r.is_real() for r in R)
    assert all(r.is_real(atol=1e-16) for r in R)
    assert not any(r.is_real(atol=0.0) for r in R)
    assert all(r.is_singular() for r in R)


def test_tracker_parameter_homotopy():
    """test/tracker_test.jl:5-27: F = [x^2 - a, xy - a + b] from p = [1, 0] to q = [2, 4]."""
    F = systems.tracker_param_system()
    H = ParameterHomotopy(F, [1, 0], [2, 4])
    r = oracle.track_plain(H, [1, 1])
    assert r["return_code"] == "success"
    assert r["accepted_steps"] + r["rejected_steps"] <= 5
    np.testing.assert_allclose(r["solution"], [np.sqrt(2), -np.sqrt(2)], rtol=1e-12)
    back = oracle.track_plain(H, r["solution"], t1=0.0, t0=1.0)
    assert back["return_code"] == "success"
    np.testing.assert_allclose(back["solution"], [1, 1], rtol=1e-10)


def _isapprox(a, b, rtol=np.sqrt(np.finfo(float).eps)):
    """Julia's default isapprox on vectors: norm(a - b) <= rtol * max(norm(a), norm(b))."""
    a, b = np.asarray(a), np.asarray(b)
    return np.linalg.norm(a - b) <= rtol * max(np.linalg.norm(a), np.linalg.norm(b))


def test_steiner_higher_precision():
    """test/test_cases/steiner_higher_prec.jl: a parameter path close to a singularity
    (extended-precision residuals for a step or two); s_p at p tracks to s_q at q and back."""
    F, p, q, s_p, s_q = systems.steiner()
    H = ParameterHomotopy(F, p, q)
    r = oracle.track_plain(H, s_p, p=p, q=q)
    assert r["return_code"] == "success"
    assert _isapprox(r["solution"], s_q)
    back = oracle.track_plain(H, r["solution"], t1=0.0, t0=1.0, p=p, q=q)
    assert back["return_code"] == "success"
    assert _isapprox(back["solution"], s_p)


def test_four_bar():
    """test/test_cases/four_bar.jl: iterative refinement in the predictor; the start point is
    given to 8 digits only and the path still ends in success."""
    F, p, q, s = systems.four_bar()
    H = ParameterHomotopy(F, p, q)
    r = oracle.track_plain(H, s, p=p, q=q)
    assert r["return_code"] == "success"


def test_tracker_straight_line_and_invalid_start():
    """test/tracker_test.jl:93-105: both starts succeed; [100, -100] is an invalid start."""
    G, F, S = systems.straight_line_case()
    H = StraightLineHomotopy(G, F)
    for s in S:
        assert oracle.track_plain(H, s)["return_code"] == "success"
    assert oracle.track_plain(H, [100, -100])["return_code"] == "terminated_invalid_startvalue"


def test_tracker_invalid_start_singular_jacobian():
    """test/tracker_test.jl:137-219 (issue 454): a framework configuration with a singular
    Jacobian is classified as an invalid start with singular Jacobian."""
    F, b0, b1, start = systems.framework_454()
    assert max(abs(f(list(start) + list(b0))) for f in F.polys) == 0.0
    H = ParameterHomotopy(F, b0, b1)
    r = oracle.track_plain(H, start, p=b0, q=b1)
    assert r["return_code"] == "terminated_invalid_startvalue_singular_jacobian"


def test_cyclic7():
    """test/endgame_test.jl:2-5: cyclic-7 has 924 nonsingular solutions."""
    H, S = total_degree(systems.cyclic(7), seed=4)
    _, R = solve_oracle(H, S)
    assert R.nsolutions() == 924


def test_hyperbolic_6_6():
    """test/endgame_test.jl:7-24: 12 paths with winding number 3, two results of multiplicity 6."""
    H, S = total_degree(systems.hyperbolic(), seed=5)
    _, R = solve_oracle(H, S)
    assert sum(1 for r in R if r.winding_number == 3) == 12
    assert R.nresults() == 2 and R.nsingular() == 2
    assert all(r.multiplicity == 6 for r in R.results())


def test_singular_1():
    """test/endgame_test.jl:26-38."""
    H, S = total_degree(systems.singular1(), seed=6)
    _, R = solve_oracle(H, S)
    assert R.nsingular() == 1
    assert R.nresults(only_nonsingular=True) == 1


def test_wilkinson_12():
    """test/endgame_test.jl:40-47 (expanded polynomial, only_nonsingular endgame)."""
    H, S = total_degree(systems.wilkinson(12), seed=2)
    o, _ = solve_oracle(H, S, endgame_options={"only_nonsingular": True})
    assert (o.code == 1).all()
    roots = o.solution[:, 0]
    assert sorted(np.round(roots.real).astype(int)) == list(range(1, 13))
    assert np.abs(roots.imag).max() < 1e-4


@pytest.mark.parametrize("d", [2, 6])
def test_power_of_linear_winding(d):
    """test/endgame_test.jl:49-54: (x - 10)^d, every path has winding number d."""
    (x,), _ = expr_variables(1)
    H, S = total_degree(ExprSystem([(x - 10) ** d], 1), seed=d)
    o, _ = solve_oracle(H, S)
    assert int((o.winding_number == d).sum()) == d


def test_beyond_polyhedral_counts():
    """test/endgame_test.jl:56-62 (random γ)."""
    H, S = total_degree(systems.beyond_polyhedral(), seed=11)
    o, _ = solve_oracle(H, S)
    assert int((o.code == 1).sum()) == 2
    assert int(((o.code == 2) | (o.code == 3)).sum()) == 2


@pytest.mark.parametrize("d", [2, 4, 6])
def test_winding_number_family(d):
    """test/endgame_test.jl:64-71: d + 1 successes."""
    H, S = total_degree(systems.winding_family(d), seed=3)
    o, _ = solve_oracle(H, S)
    assert int((o.code == 1).sum()) == d + 1


@pytest.mark.parametrize("gamma", [systems.MOHAB_GAMMA, None])
def test_mohab(gamma):
    """test/endgame_test.jl:78-107: 693 nonsingular solutions, with the γ that once caused
    path jumping and with a random γ."""
    H, S = total_degree(systems.mohab(), gamma=gamma, seed=12)
    _, R = solve_oracle(H, S)
    assert R.nnonsingular() == 693


def test_result_basic_functionality():
    """test/result_test.jl:2-41 through the total-degree start system (polyhedral start
    systems are out of scope, so 9 paths are tracked instead of 7)."""
    F = systems.winding_family(2)
    H, S = total_degree(F, seed=21)
    _, R = solve_oracle(H, S)
    assert R.ntracked() == 9
    assert R.nresults() == 3 and len(R.solutions()) == 3
    assert sum(1 for r in R if r.is_success()) == 3
    assert R.nreal() == 1
    assert R.nreal(real_atol=1e-16) == 1
    assert R.nreal(real_atol=0.0) == 0
    assert R.nreal(real_atol=0.0, real_rtol=1e-16) == 1
    assert R.nreal(real_atol=float("inf"), real_rtol=float("inf")) == 3
    assert R.nnonsingular() == 3 and R.nsingular() == 0
    assert R.nexcess_solutions() == 0


def test_result_singular():
    """test/result_test.jl:43-52: [(29/16) x^3 - 2xy, x^2 - y] has one (singular) solution."""
    (x, y), _ = variables(2)
    F = System([(29 / 16) * x ** 3 - 2 * x * y, x ** 2 - y], 2)
    H, S = total_degree(F, seed=22)
    _, R = solve_oracle(H, S)
    assert R.nresults() == 1


# ---------------------------------------------------------------- refined endpoints
def _newton_mp(F: System, x0, iters=8):
    mpmath.mp.prec = 200
    n = F.nvars
    x = [mpmath.mpc(complex(v)) for v in x0]
    for _ in range(iters):
        f = []
        J = mpmath.matrix(n, n)
        for i, p in enumerate(F.polys):
            s = mpmath.mpc(0)
            for ex, c in p.terms.items():
                m = mpmath.mpc(c)
                for v, e in zip(x, ex):
                    m *= v ** e
                s += m
                for j in range(n):
                    if ex[j]:
                        d = mpmath.mpc(c) * ex[j]
                        for k, (v, e) in enumerate(zip(x, ex)):
                            d *= v ** (e - 1 if k == j else e)
                        J[i, j] += d
            f.append(s)
        dx = mpmath.lu_solve(J, mpmath.matrix(f))
        x = [x[i] - dx[i] for i in range(n)]
    return np.array([complex(v) for v in x])


def test_dense_n4_endpoints_refined_and_complete():
    """81 paths of the dense n = 4 cubic solve: every endpoint is nonsingular, agrees with its
    200-bit Newton refinement within 1e-10 relative, and the 81 endpoints are pairwise distinct."""
    F = dense_system(4, 3, seed=100)
    H, S = total_degree(F, seed=101)
    o = oracle.track_batch(H, S, make_options())
    assert (o.code == 1).all() and not o.singular.any()
    for k in range(81):
        x = o.solution[k]
        xr = _newton_mp(F, x)
        assert np.abs(x - xr).max() / max(1.0, np.abs(xr).max()) <= 1e-10
    D = np.abs(o.solution[:, None, :] - o.solution[None, :, :]).max(axis=2) + np.eye(81)
    assert D.min() > 1e-6
