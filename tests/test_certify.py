"""Certification (src/certification.jl) on the CPU: the oracle's restatement of the FP64
interval Krawczyk test and the host logic around it.

Pinned by:
* every interval operation against exact rational arithmetic (fractions.Fraction): the
  exact range over the operand intervals lies in the computed interval, and the interval
  is no wider than that range moved outward by a few ulps; a variant without the outward
  rounding (planted here, on the CPU) fails the same check;
* the double-double balls of the extended retry the same way (their radii cover the exact
  results and stay within ~1e-29 of the magnitudes);
* the reference's known answers (test/certification_test.jl): the 18 solutions of
  "Simple 1" (4 real, 14 complex, all distinct; doubled: 36 certified, 18 distinct), the
  reality check of x - 1 + 1e-16 i, the 36 critical points (8 real), random candidates
  that are not zeros, the positive solution of x² + y² = 1, x = y, and the 3,264 real
  conics tangent to five given conics;
* the Krawczyk property itself: each certified box contains the endpoint that Newton's
  method in exact rational arithmetic converges to.
"""
import json
import math
import random
from fractions import Fraction
from pathlib import Path

import numpy as np
import pytest

import oracle
import systems
from hcgpu.certify import assemble, duplicate_groups, is_real_system, parameter_boxes
from hcgpu.homotopies import _SysTapes, total_degree
from hcgpu.polys import Poly, System, variables
from hcgpu.solve import make_options

GOLDEN = Path(__file__).resolve().parent / "golden"


# ------------------------------------------------------------------ interval operations
def _fr(x):
    return Fraction(x)


def _exact(op, a, b, c):
    """Exact range of the op over the boxes, from the bounds (monotone pieces) and points."""
    A = [_fr(a[0]), _fr(a[1])]
    B = [_fr(b[0]), _fr(b[1])]
    Cc = [_fr(c[0]), _fr(c[1])]
    if op == 0:
        vals = [x + y for x in A for y in B]
    elif op == 1:
        vals = [x - y for x in A for y in B]
    elif op == 2:
        vals = [x * y for x in A for y in B]
    elif op == 3:
        vals = [x / y for x in A for y in B]
    elif op == 4:
        vals = [x * x for x in A] + ([Fraction(0)] if A[0] <= 0 <= A[1] else [])
    elif op == 5:
        vals = [1 / x for x in A]
    elif op == 6:
        vals = [x * y + z for x in A for y in B for z in Cc]
    elif op == 7:
        vals = [A[0] * y for y in B]
    else:
        vals = [A[0] * y + z for y in B for z in Cc]
    return min(vals), max(vals)


def _rand_interval(rng):
    kind = rng.randrange(6)
    scale = 10.0 ** rng.randint(-8, 8)
    u, v = sorted([rng.uniform(-1, 1) * scale, rng.uniform(-1, 1) * scale])
    if kind == 0:
        return (u, u)  # point
    if kind == 1:
        return (abs(u), abs(u) + abs(v))  # positive
    if kind == 2:
        return (-abs(u) - abs(v), -abs(u))  # negative
    if kind == 3:
        return (0.0, abs(v))
    return (u, v)


@pytest.mark.parametrize("op", range(9))
def test_interval_ops_enclose_exact_results(op):
    rng = random.Random(100 + op)
    n_checked = 0
    for _ in range(1500):
        a, b, c = _rand_interval(rng), _rand_interval(rng), _rand_interval(rng)
        if op == 3 and b[0] <= 0 <= b[1]:
            lo, hi = oracle.ival_op(op, a, b, c)
            assert (math.isnan(lo) and math.isnan(hi)) or (a == (0.0, 0.0) and (lo, hi) == (0.0, 0.0))
            continue
        if op == 5 and a[0] <= 0 <= a[1]:
            assert all(math.isnan(v) for v in oracle.ival_op(op, a, b, c))
            continue
        lo, hi = oracle.ival_op(op, a, b, c)
        elo, ehi = _exact(op, a, b, c)
        assert _fr(lo) <= elo and ehi <= _fr(hi), (op, a, b, c, lo, hi)
        # no wider than the exact range rounded outward by two ulps per bound
        w = 2
        assert lo >= _down(float(elo), w + 1) and hi <= _up(float(ehi), w + 1), (op, a, b, c, lo, hi, elo, ehi)
        n_checked += 1
    assert n_checked > 800


def _up(x, k=1):
    for _ in range(k):
        x = math.nextafter(x, math.inf)
    return x


def _down(x, k=1):
    for _ in range(k):
        x = math.nextafter(x, -math.inf)
    return x


def test_enclosure_check_detects_missing_outward_rounding():
    """A product rounded to nearest but not moved outward (the planted mistake) misses the
    exact product for a large share of random inputs, so the enclosure test above would
    catch a restatement that forgot round_down / round_up."""
    rng = random.Random(7)
    misses = 0
    for _ in range(500):
        a, b = sorted([rng.uniform(0.1, 10), rng.uniform(0.1, 10)]), sorted([rng.uniform(0.1, 10), rng.uniform(0.1, 10)])
        lo = a[0] * b[0]  # round-to-nearest only
        if _fr(lo) > _fr(a[0]) * _fr(b[0]):
            misses += 1
    assert misses > 50


def _rand_dd(rng):
    """A normalized double-double (hi, lo) and a radius."""
    hi = rng.uniform(-1, 1) * 10.0 ** rng.randint(-6, 6)
    lo = math.ulp(hi) * rng.uniform(-0.5, 0.5)
    lo = float(Fraction(hi) + Fraction(lo) - Fraction(hi))
    r = 0.0 if rng.random() < 0.5 else abs(hi) * 10.0 ** rng.randint(-32, -20)
    return hi, lo, r


@pytest.mark.parametrize("op", [0, 1, 2])
def test_double_double_balls_enclose_exact_results(op):
    """The ball of a sum, difference or product contains the exact result for every pair of
    points of the operand balls (checked at their extreme points, where the exact range of
    these operations is attained), and its radius stays within ~1e-29 of the magnitudes."""
    rng = random.Random(300 + op)
    for _ in range(1500):
        a, b = _rand_dd(rng), _rand_dd(rng)
        hi, lo, r = oracle.ball_op(op, a, b)
        m = _fr(hi) + _fr(lo)
        ma, mb = _fr(a[0]) + _fr(a[1]), _fr(b[0]) + _fr(b[1])
        for sa in (-1, 1):
            for sb in (-1, 1):
                x, y = ma + sa * _fr(a[2]), mb + sb * _fr(b[2])
                v = x + y if op == 0 else (x - y if op == 1 else x * y)
                assert abs(v - m) <= _fr(r), (op, a, b, hi, lo, r)
        scale = (abs(ma) + abs(mb)) if op < 2 else abs(ma) * abs(mb)
        extra = (_fr(a[2]) + _fr(b[2])) if op < 2 else abs(ma) * _fr(b[2]) + abs(mb) * _fr(a[2]) + _fr(a[2]) * _fr(b[2])
        assert _fr(r) <= extra * (1 + Fraction(1, 10**12)) + scale * Fraction(1, 10**29) + Fraction(1, 10**290)


def test_rational_parameters_are_enclosed():
    p, box, ball = parameter_boxes([Fraction(1, 3), Fraction(-2, 7), 0.5, Fraction(4, 1), 1 + 2j])
    assert box.shape == (5, 4)
    for k, v in enumerate([Fraction(1, 3), Fraction(-2, 7)]):
        assert _fr(box[k, 0]) < v < _fr(box[k, 1]) and box[k, 2] == box[k, 3] == 0.0
        assert p[k] == complex(float(v), 0)
    assert tuple(box[2]) == (0.5, 0.5, 0.0, 0.0) and tuple(box[3]) == (4.0, 4.0, 0.0, 0.0)
    assert tuple(box[4]) == (1.0, 1.0, 2.0, 2.0)
    for k, v in enumerate([Fraction(1, 3), Fraction(-2, 7)]):  # double-double balls hold the exact rational
        hi, lo, rad = ball[k, :3]
        assert abs(v - _fr(hi) - _fr(lo)) <= _fr(rad) and rad < 1e-32 and ball[k, 5] == 0.0
    assert tuple(ball[2]) == (0.5, 0.0, 0.0, 0.0, 0.0, 0.0)
    big = Fraction(3**40 + 1, 7**25)  # numerator and denominator beyond 2^53
    _, bbox, _ = parameter_boxes([big])
    assert _fr(bbox[0, 0]) <= big <= _fr(bbox[0, 1]) and bbox[0, 1] <= math.nextafter(bbox[0, 0], math.inf)


# ------------------------------------------------------------------ known answers
class OracleCertifier:
    """The oracle's certification behind the interface hcgpu.certify uses for the GPU."""

    class _Ctx:
        def last_kernel_ms(self):
            return 0.0

    def __init__(self, F, nthreads=8):
        self.F, self.n, self.nthreads = F, F.nvars, nthreads
        self.tapes = _SysTapes(F)
        self.struct = self.tapes.struct()
        self.ctx = self._Ctx()

    def certify(self, S, params, params_ival, params_ball=None, refine=True, extended=True, real_system=True,
                group_of=None):
        return oracle.certify(self.struct, self.n, S, params, params_ival, params_ball, refine, extended, real_system,
                              self.nthreads, group_of=group_of)


def certify_cpu(F, S, p=None):
    from hcgpu.certify import certify
    return certify(F, S, p, certifier=OracleCertifier(F))


simple1 = systems.certification_simple1
critical_points_system = systems.critical_points


def solve_nonsingular(F, seed=5):
    H, S = total_degree(F, seed=seed)
    o = oracle.track_batch(H, S, make_options(), nthreads=8)
    ok = (o.code == 1) & (~o.singular)
    return o.solution[ok]


def test_simple1_known_answers():
    F = simple1()
    S = solve_nonsingular(F)
    assert len(S) == 18
    r = certify_cpu(F, S)
    assert r.ncertified() == 18 and r.ndistinct_certified() == 18
    assert r.nreal_certified() == 4 and r.ncomplex_certified() == 14
    assert r.ndistinct_real_certified() == 4 and r.ndistinct_complex_certified() == 14
    d = certify_cpu(F, np.vstack([S, S]))
    assert d.ncertified() == 36 and d.ndistinct_certified() == 18
    assert d.nreal_certified() == 8 and d.ndistinct_real_certified() == 4
    assert sorted(len(g) for g in d.duplicates) == [2] * 18
    assert all(g[1] == g[0] + 18 for g in d.duplicates)


def test_certified_box_contains_the_exact_newton_limit():
    """Newton's method in exact rational arithmetic from the certified midpoint (3 steps,
    quadratic convergence far below the box width) stays inside the box x₁."""
    F = simple1()
    S = solve_nonsingular(F)
    r = certify_cpu(F, S[:4])
    J = F.jacobian()
    for c in r.certificates:
        assert c.certified
        z = [(_fr(v.real), _fr(v.imag)) for v in c.solution]
        for _ in range(3):
            z = _newton_step_exact(F, J, z)
        for i, (re, im) in enumerate(z):
            b = c.I[i]
            assert _fr(b[0]) <= re <= _fr(b[1]) and _fr(b[2]) <= im <= _fr(b[3])


def _cmul(a, b):
    return (a[0] * b[0] - a[1] * b[1], a[0] * b[1] + a[1] * b[0])


def _poly_exact(f: Poly, z):
    acc = (Fraction(0), Fraction(0))
    for e, coef in f.terms.items():
        t = (_fr(complex(coef).real), _fr(complex(coef).imag))
        for i, k in enumerate(e):
            for _ in range(k):
                t = _cmul(t, z[i])
        acc = (acc[0] + t[0], acc[1] + t[1])
    return acc


def _newton_step_exact(F, J, z):
    """One Newton step for a 2 x 2 system in exact complex rationals (Cramer's rule)."""
    f = [_poly_exact(p, z) for p in F.polys]
    A = [[_poly_exact(J[i][j], z) for j in range(2)] for i in range(2)]
    det = (_cmul(A[0][0], A[1][1])[0] - _cmul(A[0][1], A[1][0])[0], _cmul(A[0][0], A[1][1])[1] - _cmul(A[0][1], A[1][0])[1])
    dn = det[0] ** 2 + det[1] ** 2
    inv = (det[0] / dn, -det[1] / dn)
    n0 = _cmul(A[1][1], f[0]); n0 = (n0[0] - _cmul(A[0][1], f[1])[0], n0[1] - _cmul(A[0][1], f[1])[1])
    n1 = _cmul(A[0][0], f[1]); n1 = (n1[0] - _cmul(A[1][0], f[0])[0], n1[1] - _cmul(A[1][0], f[0])[1])
    d0, d1 = _cmul(n0, inv), _cmul(n1, inv)
    return [(z[0][0] - d0[0], z[0][1] - d0[1]), (z[1][0] - d1[0], z[1][1] - d1[1])]


def test_reality_check():
    (x,), _ = variables(1, 0)
    F = System([x - 1 + 1e-16j], 1, 0)
    assert not is_real_system(F)
    r = certify_cpu(F, [[1.0]])
    assert r.certificates[0].certified and not r.certificates[0].real
    G = System([x - 1], 1, 0)
    assert is_real_system(G)
    g = certify_cpu(G, [[1.0]])
    assert g.certificates[0].certified and g.certificates[0].real


def test_critical_points_with_parameters():
    F = critical_points_system()
    u0 = [-0.32, -0.1]
    S = solve_nonsingular(F.fix_parameters(u0))
    assert len(S) == 36
    r = certify_cpu(F, S, u0)
    assert r.ncertified() == 36 and r.ndistinct_certified() == 36
    assert r.nreal_certified() == 8 and r.ndistinct_real_certified() == 8
    rng = np.random.default_rng(1)
    junk = 100 * (rng.standard_normal((10, 3)) + 1j * rng.standard_normal((10, 3))) / np.sqrt(2)
    assert certify_cpu(F, junk, u0).ncertified() < 10
    with pytest.raises(ValueError):
        certify_cpu(F, S)  # parameters are required


def test_parameter_groups_match_separate_calls():
    """Candidates of two parameter values certified in one grouped call give the records of
    two separate calls (the kernel indexes the group's parameters per candidate)."""
    from hcgpu.certify import certify_groups
    F = critical_points_system()
    us = [[-0.32, -0.1], [0.25, 0.4]]
    sols = [solve_nonsingular(F.fix_parameters(u))[:12] for u in us]
    S = np.vstack(sols)
    G = np.repeat([0, 1], [len(sols[0]), len(sols[1])])
    cert = OracleCertifier(F)
    r = certify_groups(F, S, us, G, certifier=cert)
    for g, u in enumerate(us):
        one = certify_cpu(F, sols[g], u)
        rows = [c for c, gg in zip(r.certificates, G) if gg == g]
        assert [c.certified for c in rows] == [c.certified for c in one.certificates]
        assert all(np.array_equal(a.I, b.I) for a, b in zip(rows, one.certificates))
    assert r.ncertified() == len(S) and r.nreal_certified() == sum(c.real for c in r.certificates)


def test_certify_monodromy_result():
    """certify(F, monodromy result): the parameters come from the result, as in
    certify(F, ::MonodromyResult) (src/certification.jl:1470-1477)."""
    from hcgpu.configs import sparse_param_system
    from hcgpu.monodromy import monodromy_solve
    from test_monodromy import oracle_tracker
    F = sparse_param_system(3, 6, 4, seed=7)
    m = monodromy_solve(F, seed=11, tracker=oracle_tracker, target_solutions_count=8)
    from hcgpu.certify import certify
    r = certify(F, m, certifier=OracleCertifier(F))
    assert r.ncertified() == 8 and r.ndistinct_certified() == 8


def test_distinct_certified_solutions_incremental_api():
    """test/certification_test.jl:164-247: statuses, representatives and stats of
    add_solution!, batches and merge! (the interval-tree collision case needs no special
    handling here: boxes are compared directly)."""
    from hcgpu.certify import DistinctCertifiedSolutions, distinct_certified_solutions
    (x, y), _ = variables(2, 0)
    F = System([x**2 + y**2 - 1, x - y], 2, 0)
    sols = solve_nonsingular(F)
    assert len(sols) == 2
    d = DistinctCertifiedSolutions(F, None, certifier=OracleCertifier(F))
    added, status, rep, _ = d.add_solution(sols[0], 1)
    assert (added, status, rep) == (True, "certified_distinct", 1)
    added, status, rep, _ = d.add_solution(sols[0], 2)
    assert (added, status, rep) == (False, "duplicate", 1)
    C = critical_points_system()
    nc = DistinctCertifiedSolutions(C, [-0.32, -0.1], certifier=OracleCertifier(C))
    rng = np.random.default_rng(1)
    added, status, rep, _ = nc.add_solution(100 * (rng.standard_normal(3) + 1j * rng.standard_normal(3)) / np.sqrt(2), 3)
    assert (added, status, rep) == (False, "not_certified", 0)
    assert d.stats() == dict(processed=2, certified_distinct=1, duplicates=1, not_certified=0)
    assert nc.stats() == dict(processed=1, certified_distinct=0, duplicates=0, not_certified=1)
    (z,), (u,) = variables(1, 1)
    Z = System([z**2 - u], 1, 1)
    col = DistinctCertifiedSolutions(Z, [1.0], certifier=OracleCertifier(Z))
    assert col.add_solution([1.0 + 0j], 1)[:3] == (True, "certified_distinct", 1)
    assert col.add_solution([-1.0 + 0j], 2)[:3] == (True, "certified_distinct", 2)
    assert np.allclose(sorted(float(s[0].real) for s in col.solutions()), [-1.0, 1.0], rtol=0, atol=1e-15)
    b = DistinctCertifiedSolutions(F, None, certifier=OracleCertifier(F))
    b.add_solutions(sols[:1])
    b.add_solutions([sols[1], sols[0]])
    assert len(b.solutions()) == 2 and b.ncertified_distinct() == 2 and b.nduplicates() == 1
    left = distinct_certified_solutions(F, sols[:1], certifier=OracleCertifier(F))
    right = distinct_certified_solutions(F, [sols[1], sols[0]], certifier=OracleCertifier(F))
    left.merge(right)
    assert len(left.solutions()) == 2 and left.nduplicates() == 1


def test_certify_accepts_path_results():
    """certify(F, ::Vector{PathResult}) and certify(F, ::PathResult) (src/certification.jl:1479-1506)."""
    from hcgpu.certify import certify
    from hcgpu.result import PathResult
    F = simple1()
    S = solve_nonsingular(F)[:3]
    prs = [PathResult(solution=x, return_code="success") for x in S]
    r = certify(F, prs, certifier=OracleCertifier(F))
    assert r.ncertified() == 3
    assert certify(F, prs[0], certifier=OracleCertifier(F)).ncertified() == 1


def _quotient_system():
    from hcgpu.abi import hcg_system
    import ctypes
    ev, ej = systems.quotient_tapes()
    st = hcg_system()
    st.eval_jac = ctypes.pointer(ej.struct())
    st.eval = ctypes.pointer(ev.struct())
    return st, (ev, ej)


def test_certify_with_quotients():
    """Interval division and inversion (src/interval_arithmetic.jl:196-227, 356-366) inside the
    SLP: the zero (2, 1) of [x0/x1 - 2, x0 + x1 - 3] certifies (real), a perturbed start is
    refined onto it, and a point with x1 near 0 does not certify."""
    st, keep = _quotient_system()
    S = np.array([[2.0, 1.0], [2.0 + 1e-9, 1.0 - 1e-9j], [1.0, 1e-300]], dtype=complex)
    o = oracle.certify(st, 2, S, real_system=True)
    assert list(o.certified) == [1, 1, 0] and list(o.real[:2]) == [1, 1]
    for k in range(2):
        b = o.box1[k]
        assert b[0, 0] <= 2.0 <= b[0, 1] and b[1, 0] <= 1.0 <= b[1, 1]


def test_positive():
    (x, y), _ = variables(2, 0)
    F = System([x**2 + y**2 - 1, x - y], 2, 0)
    S = solve_nonsingular(F)
    r = certify_cpu(F, S)
    assert sum(c.is_positive() for c in r.certificates) == 1
    assert sum(c.is_positive(1) for c in r.certificates) == 1
    assert sum(c.real for c in r.certificates) == 2


def test_3264_real_conics():
    """All 3,264 real conics tangent to five given conics certify as real and distinct
    (test/certification_test.jl:290-295); the ill-conditioned ones (cond(J) up to ~1e18)
    through the double-double retry, three of them only after C is recomputed."""
    F = systems.steiner()[0]
    conics = [Fraction(s) for s in json.loads((GOLDEN / "param_cases.json").read_text())["conics"]["real_conics"]]
    S = np.load(GOLDEN / "conics_3264_real.npy")
    r = certify_cpu(F, S, conics)
    assert r.ndistinct_real_certified() == 3264 and r.ncertified() == 3264
    assert sum(c.prec == 106 for c in r.certificates) > 3000
    d = certify_cpu(F, np.vstack([S[:40], S[:5]]), conics)
    assert d.nreal_certified() == 45 and d.ndistinct_real_certified() == 40 and len(d.duplicates) == 5


def test_duplicate_groups_follow_candidate_order():
    box = np.zeros((4, 1, 4))
    box[:, 0] = [[0, 1, 0, 1], [2, 3, 0, 1], [0.5, 1.5, 0.5, 2], [2.5, 4, 0.5, 0.7]]
    assert duplicate_groups(np.array([1, 1, 1, 1], bool), box) == [[1, 3], [2, 4]]
    assert duplicate_groups(np.array([0, 1, 1, 1], bool), box) == [[2, 4]]
    # a box overlapping only a duplicate is distinct (add_certificate! compares with the
    # distinct certificates only)
    box2 = np.zeros((3, 1, 4))
    box2[:, 0] = [[0, 1, 0, 1], [0.9, 2, 0, 1], [1.5, 3, 0, 1]]
    assert duplicate_groups(np.array([1, 1, 1], bool), box2) == [[1, 2]]


def test_assemble_reports_midpoints_and_counts():
    F = simple1()
    S = solve_nonsingular(F)[:3]
    tapes = _SysTapes(F)  # the struct points into these tapes
    o = oracle.certify(tapes.struct(), 2, S)
    r = assemble(S, o)
    for c in r.certificates:
        assert np.allclose(c.solution, c.solution_candidate, atol=1e-12)
        assert np.all(c.I[:, 0] <= c.I[:, 1]) and np.all(c.I0[:, 0] < c.I[:, 0])
