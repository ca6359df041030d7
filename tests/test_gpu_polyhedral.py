"""Polyhedral start systems on the GPU: both stages as batched launches, identical return
codes to the oracle's and endpoints within 1e-10, and the reference's counts
(test/polyhedral_test.jl:3-42): 8 starts / 6 solutions, 3 in the torus, cyclic-5's 70."""
import numpy as np
import pytest

import systems
from hcgpu.monodromy import gpu_edge_tracker
from hcgpu.polyhedral import polyhedral
from test_monodromy import oracle_tracker

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("name,nsuccess", [("f_affine", 6), ("f_torus", 3), ("cyclic5", 70)])
def test_polyhedral_gpu_equals_oracle(gpu_ctx, name, nsuccess):
    _, F, nz, seed = next(c for c in systems.polyhedral_cases() if c[0] == name)
    P = polyhedral(F, only_non_zero=nz, seed=seed)
    gt = gpu_edge_tracker(gpu_ctx)
    S, failed = P.start_solutions(tracker=gt)
    So, failed_o = P.start_solutions(tracker=oracle_tracker)
    assert failed == failed_o == 0 and len(S) == P.mixed_volume
    np.testing.assert_allclose(S, So, rtol=0, atol=1e-10)
    H = P.homotopy()
    codes, X = gt(H, S, H.p, H.q[0])
    co, Xo = oracle_tracker(H, So, H.p, H.q[0])
    np.testing.assert_array_equal(np.asarray(codes), np.asarray(co))
    ok = np.asarray(codes) == 1
    assert ok.sum() == nsuccess
    np.testing.assert_allclose(X[ok], Xo[ok], rtol=0, atol=1e-10)


def test_solve_polyhedral(gpu_ctx):
    """solve(F; start_system = :polyhedral) (test/solve_test.jl:156)."""
    from hcgpu.solve import solve
    _, F, nz, seed = systems.polyhedral_cases()[0]
    r = solve(F, start_system="polyhedral", seed=seed, ctx=gpu_ctx)
    assert len(r.path_results) == 8 and sum(p.is_success() for p in r) == 6
    r2 = solve(F, start_system="polyhedral", only_non_zero=True, seed=seed, ctx=gpu_ctx)
    assert len(r2.path_results) == 3 and sum(p.is_success() for p in r2) == 3
