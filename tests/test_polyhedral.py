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
