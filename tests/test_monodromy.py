"""Monodromy start-solution generation (src/monodromy.jl), driven by the oracle's tracker on
the CPU and by libhcgpu on the GPU.

Pinned by: the closed-form solutions of test/tracker_test.jl's system (x = ±√a), and the
total-degree solve of the same system at the same parameters (all nonsingular endpoints,
complete by Bézout for generic coefficients): monodromy must find exactly that set.
"""
import numpy as np
import pytest

import oracle
import systems
from hcgpu.configs import sparse_param_system
from hcgpu.homotopies import total_degree
from hcgpu.monodromy import _new_points, find_start_pair, monodromy_solve
from hcgpu.polys import System, variables
from hcgpu.solve import make_options


class _Edge:
    """The template homotopy with the edge's start and end parameters."""

    def __init__(self, H, p, q):
        self._H, self.p, self.q = H, p, q[None, :]

    def desc(self):
        return self._H.desc()


def oracle_tracker(H, X, p, q):
    o = oracle.track_batch(_Edge(H, p, q), X, make_options(), nthreads=4)
    return o.code, o.solution


def same_set(A, B, tol=1e-8):
    if len(A) != len(B):
        return False
    return len(_new_points(A, B, tol)) == 0 and len(_new_points(B, A, tol)) == 0


def test_find_start_pair_is_a_solution():
    """Linear solve for systems linear in their parameters; Newton's method in (x, p) for
    the others (Steiner's tangency equations have no parameters, x² − a² is quadratic in a)."""
    (x,), (a,) = variables(1, 1)
    for F in [sparse_param_system(3, 6, 4, seed=7), systems.steiner()[0], System([x**2 - a * a], 1, 1),
              systems.toric_ed([[3, 2, 1, 0], [0, 1, 2, 3]])]:
        x0, p0 = find_start_pair(F, np.random.default_rng(1))
        scale = max(1.0, np.abs(np.concatenate([x0, p0])).max())
        assert max(abs(f(list(x0) + list(p0))) for f in F.polys) < 1e-10 * scale ** 6
    with pytest.raises(ValueError):
        find_start_pair(System([x**2 - 1], 1, 0), np.random.default_rng(1))


def test_monodromy_closed_form():
    """F = [x^2 - a, xy - a + b]: from one start pair both solutions x = ±√a, y = (a - b)/x."""
    F = systems.tracker_param_system()
    r = monodromy_solve(F, seed=3, tracker=oracle_tracker)
    a, b = r.parameters
    want = np.array([[s * np.sqrt(a), (a - b) / (s * np.sqrt(a))] for s in (1, -1)])
    assert r.nsolutions() == 2
    assert same_set(r.solutions, want, 1e-10)
    assert r.returncode == "heuristic_stop" and r.new_per_loop[-5:] == [0] * 5


def test_monodromy_equals_total_degree_solve():
    """3 quadrics in 3 unknowns, coefficients affine in 4 parameters: monodromy from one start
    pair finds the 8 solutions the total-degree solve finds at the same parameters."""
    F = sparse_param_system(3, 6, 4, seed=7)
    r = monodromy_solve(F, seed=11, tracker=oracle_tracker, target_solutions_count=8)
    assert r.returncode == "success" and r.nsolutions() == 8
    Fp = F.fix_parameters(r.parameters)
    H, S = total_degree(Fp, seed=5)
    o = oracle.track_batch(H, S, make_options())
    ok = (o.code == 1) & (~o.singular)
    assert ok.sum() == 8
    assert same_set(r.solutions, o.solution[ok], 1e-8)
    for x in r.solutions:
        assert max(abs(f(list(x) + list(r.parameters))) for f in F.polys) < 1e-10


def test_monodromy_toric_ed_21():
    """test/monodromy_test.jl:14-31, 347-356: the twisted cubic's ED degree is 21; monodromy
    from one start pair (found by Newton's method in (x, p)) finds all 21, and tracking them
    to new parameters keeps 21 solutions."""
    F = systems.toric_ed([[3, 2, 1, 0], [0, 1, 2, 3]])
    r = monodromy_solve(F, seed=5, tracker=oracle_tracker, target_solutions_count=21, max_loops_no_progress=50)
    assert r.returncode == "success" and r.nsolutions() == 21
    assert len(_new_points(np.zeros((0, 6), dtype=complex), r.solutions, 1e-8)) == 21  # pairwise distinct
    for x in r.solutions:
        assert max(abs(f(list(x) + list(r.parameters))) for f in F.polys) < 1e-9
    q = (np.random.default_rng(9).standard_normal(4) + 1j * np.random.default_rng(10).standard_normal(4)) / np.sqrt(2)
    from hcgpu.homotopies import ParameterHomotopy
    H = ParameterHomotopy(F, r.parameters, q)
    o = oracle.track_batch(H, r.solutions, make_options())
    assert ((o.code == 1) & (~o.singular)).sum() == 21


def test_verify_solution_completeness():
    """test/monodromy_test.jl:358-384: circle ∩ line has 2 solutions; the trace test accepts
    both and rejects one alone."""
    from hcgpu.monodromy import verify_solution_completeness
    F = systems.circle_line()
    q = [1.0, 2.0, 3.0]
    # the two intersection points of x² + y² = 1 with x + 2y + 3 = 0
    ys = np.roots([5, 12, 8])
    S = np.array([[-3 - 2 * yy, yy] for yy in ys], dtype=complex)
    assert max(abs(f(list(s) + q)) for s in S for f in F.polys) < 1e-12
    assert verify_solution_completeness(F, S, q, tracker=oracle_tracker) is True
    assert verify_solution_completeness(F, S[:1], q, tracker=oracle_tracker) is False
