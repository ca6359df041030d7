"""Polyhedral start systems on the CPU (the oracle tracks both stages).

Known answers of the reference's tests: the start-solution counts and mixed volumes of
test/polyhedral_test.jl:3-35 and test/solve_test.jl:160-176 (8 / 3 / mixed volume 3 /
Bezout 16; 92 / 54 for the system with s^15), the 6 finite and 3 torus solutions of f,
and cyclic-5's 70 solutions (:39-42).  The binomial start solutions are checked by
residual, the cell exponents by exact integer arithmetic in the module itself.
"""
import numpy as np
import pytest

import systems
from hcgpu.monodromy import _new_points
from hcgpu.polyhedral import binomial_solutions, mixed_volume, paths_to_track, polyhedral
from hcgpu.polys import System, variables
from test_monodromy import oracle_tracker


def _f():
    return systems.polyhedral_cases()[0][1]


def test_start_counts_and_mixed_volumes():
    f = _f()
    assert paths_to_track(f) == 8
    assert paths_to_track(f, only_non_zero=True) == 3
    assert mixed_volume(f) == 3
    assert paths_to_track(f, start_system="total_degree") == 16
    (a, b, s), _ = variables(3)
    G = System([a ** 3 * s ** 15 + a * b * s + b ** 3 + s ** 12, a ** 2 * s ** 9 + a * b ** 2 + b * s ** 3,
                a ** 2 * b * s ** 5 + a * s ** 8 + b ** 2], 3)
    assert paths_to_track(G, only_torus=False) == 92
    assert paths_to_track(G, only_torus=True) == 54
    assert mixed_volume(systems.cyclic(5)) == 70


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_mixed_volume_independent_of_lifting(seed):
    """The mixed volume is a property of the supports: every generic lifting sums to it."""
    assert polyhedral(systems.cyclic(5), only_non_zero=True, seed=seed).mixed_volume == 70
    assert polyhedral(_f(), seed=seed).mixed_volume == 8


def test_binomial_start_solutions():
    P = polyhedral(systems.cyclic(5), only_non_zero=True, seed=3)
    for c in P.cells:
        X = binomial_solutions(c, P.A, P.start_coefficients)
        assert len(X) == c.volume and len(_new_points(np.zeros((0, 5), complex), X, 1e-10)) == c.volume
        for i in range(5):
            j0, j1 = c.pairs[i]
            r = (P.start_coefficients[i][j0] * np.prod(X ** P.A[i][j0], axis=1)
                 + P.start_coefficients[i][j1] * np.prod(X ** P.A[i][j1], axis=1))
            assert np.abs(r).max() < 1e-12


def _run(F, nz, seed):
    P = polyhedral(F, only_non_zero=nz, seed=seed)
    S, failed = P.start_solutions(tracker=oracle_tracker)
    H = P.homotopy()
    codes, X = oracle_tracker(H, S, H.p, H.q[0])
    return P, S, failed, np.asarray(codes), X


@pytest.mark.parametrize("name,nz,seed,nstart,nsuccess", [("f_affine", False, 1, 8, 6), ("f_torus", True, 1, 3, 3),
                                                          ("cyclic5", True, 3, 70, 70)])
def test_polyhedral_solve_oracle(name, nz, seed, nstart, nsuccess):
    F = next(c[1] for c in systems.polyhedral_cases() if c[0] == name)
    P, S, failed, codes, X = _run(F, nz, seed)
    assert P.mixed_volume == nstart and len(S) == nstart and failed == 0
    ok = codes == 1
    assert ok.sum() == nsuccess
    assert len(_new_points(np.zeros((0, F.nvars), complex), X[ok], 1e-8)) == nsuccess
    assert max(abs(v) for x in X[ok] for v in F.evaluate(x)) < 1e-9


@pytest.mark.parametrize("seed", range(20))
def test_cyclic5_toric_tapes_build_for_every_seed(seed):
    """Every toric homotopy of every lifting the seed draws lowers to a tape (high powers of s
    are one OP_POW_INT, not a chain of products) and the cells sum to the mixed volume 70."""
    from hcgpu.homotopies import ParameterHomotopy
    from hcgpu.tape import OP
    P = polyhedral(systems.cyclic(5), only_non_zero=True, seed=seed)
    assert P.mixed_volume == 70
    for c in P.cells:
        H = ParameterHomotopy(P.toric_system(c), [0.0], [1.0])
        H.desc()
        ev = H.F.eval
        assert ev.ninstr < 400
        top = int(max(e.max() for e in c.s_exponents))
        if top >= 5:
            assert OP["POW_INT"] in set(ev.instr[:, 4].tolist())
    assert P.max_power == max(int(max(e.max() for e in c.s_exponents)) for c in P.cells)


def test_lifting_keeps_the_lowest_largest_power():
    """Of the liftings drawn, the one with the smallest largest power of s is kept."""
    from hcgpu.polyhedral import PolyhedralStart
    one = PolyhedralStart(systems.cyclic(5), only_non_zero=True, seed=0, liftings=1).max_power
    many = PolyhedralStart(systems.cyclic(5), only_non_zero=True, seed=0, liftings=8).max_power
    assert many <= one


def test_large_supports_raise_instead_of_returning_no_cells():
    """5 dense quartics: 7,875^5 pair tuples (beyond int64): a clear error, never 0 cells."""
    from hcgpu.configs import dense_system
    from hcgpu.polyhedral import EnumerationTooLarge
    F = dense_system(5, 4, seed=0)
    with pytest.raises(EnumerationTooLarge):
        paths_to_track(F)
    with pytest.raises(EnumerationTooLarge):
        polyhedral(F)
    assert paths_to_track(F, start_system="total_degree") == 1024


def test_only_non_zero_given_explicitly_wins_over_only_torus():
    (a, b, s), _ = variables(3)
    G = System([a ** 3 * s ** 15 + a * b * s + b ** 3 + s ** 12, a ** 2 * s ** 9 + a * b ** 2 + b * s ** 3,
                a ** 2 * b * s ** 5 + a * s ** 8 + b ** 2], 3)
    assert paths_to_track(G, only_non_zero=True, only_torus=False) == 54
    assert paths_to_track(G, only_non_zero=False, only_torus=True) == 92
    assert paths_to_track(G, only_torus=True) == 54
    assert polyhedral(G, only_non_zero=True, only_torus=False, seed=0).mixed_volume == 54


def test_failed_toric_paths_become_polyhedral_failed_results():
    """A toric path that fails is kept (start, last point, s) and reported as a
    `polyhedral_failed` path result, so successes + failures = mixed volume."""
    from hcgpu.solve import polyhedral_failed_results
    P = polyhedral(systems.cyclic(5), only_non_zero=True, seed=3)
    calls = []

    def tracker(H, X, p, q):
        codes, sols = oracle_tracker(H, X, p, q)
        codes = np.asarray(codes).copy()
        if not calls:
            codes[0] = 11  # as if the first path of the first cell stopped (step size too small)
        calls.append(1)
        return codes, sols, np.full(len(X), 0.75)

    S, failed = P.start_solutions(tracker=tracker)
    assert failed == 1 and len(S) + failed == P.mixed_volume
    prs = polyhedral_failed_results(P.failed, len(S) + 1)
    assert [r.return_code for r in prs] == ["polyhedral_failed"]
    assert prs[0].t == 0.75 and prs[0].path_number == len(S) + 1 and prs[0].is_failed()
