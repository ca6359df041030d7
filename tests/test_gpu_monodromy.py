"""Monodromy on the GPU: every loop edge is one libhcgpu launch over all known solutions.

* the same monodromy run (same seed) on the GPU and on the oracle's tracker gives the same
  solutions in the same order (the trackers agree bit for bit);
* from ONE of the parameter workload's start solutions, monodromy at p0 recovers all 1,024
  committed start solutions (tests/golden/cfg2_starts.npy, produced by a total-degree solve).
"""
import time
from pathlib import Path

import numpy as np
import pytest

import oracle
from hcgpu.configs import cfg2_problem, sparse_param_system
from hcgpu.monodromy import _new_points, gpu_edge_tracker, monodromy_solve
from hcgpu.solve import make_options

pytestmark = pytest.mark.gpu

GOLDEN = Path(__file__).resolve().parent / "golden"


class _Edge:
    def __init__(self, H, p, q):
        self._H, self.p, self.q = H, p, q[None, :]

    def desc(self):
        return self._H.desc()


def _oracle_tracker(H, X, p, q):
    o = oracle.track_batch(_Edge(H, p, q), X, make_options())
    return o.code, o.solution


def test_monodromy_gpu_equals_oracle(gpu_ctx):
    F = sparse_param_system(3, 6, 4, seed=7)
    g = monodromy_solve(F, seed=11, tracker=gpu_edge_tracker(gpu_ctx), target_solutions_count=8)
    o = monodromy_solve(F, seed=11, tracker=_oracle_tracker, target_solutions_count=8)
    assert g.nsolutions() == o.nsolutions() == 8
    assert g.new_per_loop == o.new_per_loop
    np.testing.assert_array_equal(g.solutions, o.solutions)


def test_monodromy_recovers_parameter_workload_starts(gpu_ctx):
    F, p0, _ = cfg2_problem(1)
    S = np.load(GOLDEN / "cfg2_starts.npy")
    t0 = time.perf_counter()
    r = monodromy_solve(F, S[:1], p0, target_solutions_count=len(S), lift_affine=True, seed=0,
                        tracker=gpu_edge_tracker(gpu_ctx), max_loops=80)
    dt = time.perf_counter() - t0
    print(f"monodromy: {r.nsolutions()} solutions in {r.loops} loops, {r.tracked_paths} paths, {dt:.2f} s; "
          f"new per loop {r.new_per_loop}")
    assert r.returncode == "success" and r.nsolutions() == len(S)
    assert len(_new_points(S, r.solutions, 1e-8)) == 0 and len(_new_points(r.solutions, S, 1e-8)) == 0


def test_monodromy_toric_ed_gpu(gpu_ctx):
    """test/monodromy_test.jl:14-31: the twisted cubic's 21 ED critical points, on the GPU,
    identical to the same run on the oracle's tracker."""
    import systems
    F = systems.toric_ed([[3, 2, 1, 0], [0, 1, 2, 3]])
    g = monodromy_solve(F, seed=5, tracker=gpu_edge_tracker(gpu_ctx), target_solutions_count=21,
                        max_loops_no_progress=50)
    o = monodromy_solve(F, seed=5, tracker=_oracle_tracker, target_solutions_count=21, max_loops_no_progress=50)
    assert g.returncode == "success" and g.nsolutions() == 21
    assert g.new_per_loop == o.new_per_loop
    np.testing.assert_array_equal(g.solutions, o.solutions)
    # parameter homotopy with the monodromy result (test/monodromy_test.jl:347-356)
    from hcgpu.solve import solve
    q = (np.random.default_rng(3).standard_normal(4) + 1j * np.random.default_rng(4).standard_normal(4)) / np.sqrt(2)
    res = solve(F, g, target_parameters=q)
    assert res.nsolutions() == 21


def test_verify_solution_completeness_gpu(gpu_ctx):
    """The trace test (test/monodromy_test.jl:358-384) with every homotopy on the GPU."""
    import systems
    from hcgpu.monodromy import verify_solution_completeness
    F = systems.circle_line()
    q = [1.0, 2.0, 3.0]
    ys = np.roots([5, 12, 8])
    S = np.array([[-3 - 2 * yy, yy] for yy in ys], dtype=complex)
    tr = gpu_edge_tracker(gpu_ctx)
    assert verify_solution_completeness(F, S, q, tracker=tr) is True
    assert verify_solution_completeness(F, S[:1], q, tracker=tr) is False
    G = sparse_param_system(3, 6, 4, seed=7)
    r = monodromy_solve(G, seed=11, tracker=tr, target_solutions_count=8)
    assert verify_solution_completeness(G, r.solutions, r.parameters, tracker=tr) is True
    assert verify_solution_completeness(G, r.solutions[:7], r.parameters, tracker=tr) is False
