"""GPU parity: libhcgpu.so (HIP, gfx950) against the CPU oracle on the same inputs.

The bar: identical return codes and counts, identical singular / winding
classification, endpoints equal within 1e-10 relative error for nonsingular
solutions (the kernel follows the oracle's arithmetic contract, so trajectories
normally agree bit for bit; the tolerance absorbs libm differences).
"""
import numpy as np
import pytest

import oracle
import systems
from hcgpu import abi
from hcgpu.configs import dense_system
from hcgpu.homotopies import ParameterHomotopy, StraightLineHomotopy, total_degree
from hcgpu.solve import EndgameTracker, make_options

pytestmark = pytest.mark.gpu

RTOL = 1e-10


def run_both(H, S, gpu_ctx, layout=None, **opt):
    opts = make_options(**opt)
    T = EndgameTracker(H, ctx=gpu_ctx)
    if layout is not None:
        T._program = gpu_ctx.program(H, layout=layout)
    T.options = opts
    g = T.track_batch(S)
    o = oracle.track_batch(H, S, opts)
    return g, o


def assert_parity(g, o, endpoint_rtol=RTOL):
    assert list(g.code) == list(o.code), "return codes differ"
    assert list(g.singular) == list(o.singular)
    assert list(g.winding_number) == list(o.winding_number)
    ok = (o.code == 1) & (~o.singular)
    if ok.any():
        d = np.abs(g.solution[ok] - o.solution[ok]).max(axis=1)
        scale = np.maximum(1.0, np.abs(o.solution[ok]).max(axis=1))
        assert (d / scale).max() <= endpoint_rtol, f"endpoint mismatch {(d / scale).max():.3e}"


def test_beyond_polyhedral(gpu_ctx):
    H, S = total_degree(systems.beyond_polyhedral(), gamma=1.3j + 0.4)
    g, o = run_both(H, S, gpu_ctx)
    print("gpu", g.codes(), g.accepted_steps, "oracle", o.codes(), o.accepted_steps)
    assert_parity(g, o)
    # test/endgame_tracker_test.jl:2-46 on the GPU result
    assert g.codes() == ["success", "success", "at_infinity", "at_infinity"]
    assert g.accuracy[0] < 1e-12 and g.residual[0] < 1e-12
    assert g.accepted_steps[0] < 20 and g.rejected_steps[0] == 0
    assert g.cond[1] < 1e3
    np.testing.assert_allclose(g.valuation[2], [-1, -1], rtol=1e-3)
    assert (g.accepted_steps == o.accepted_steps).all()


def test_parameter_homotopy_tracker_case(gpu_ctx):
    F = systems.tracker_param_system()
    H = ParameterHomotopy(F, [1, 0], [2, 4])
    S = np.array([[1, 1]], dtype=complex)
    g, o = run_both(H, S, gpu_ctx)
    assert_parity(g, o)
    np.testing.assert_allclose(g.solution[0], [np.sqrt(2), -np.sqrt(2)], rtol=1e-12)


def test_dense_n4_config0(gpu_ctx):
    F = dense_system(4, 3, seed=100)
    H, S = total_degree(F, seed=101)
    g, o = run_both(H, S, gpu_ctx)
    assert_parity(g, o)
    assert (g.code == 1).sum() == 81


@pytest.mark.parametrize("name", ["cyclic7", "mohab", "hyperbolic", "winding4", "wilkinson12"])
def test_known_answers_gpu(gpu_ctx, name):
    from hcgpu.result import Result, path_results_from_batch
    if name == "cyclic7":
        H, S = total_degree(systems.cyclic(7), seed=4)
    elif name == "mohab":
        H, S = total_degree(systems.mohab(), gamma=systems.MOHAB_GAMMA)
    elif name == "hyperbolic":
        H, S = total_degree(systems.hyperbolic(), seed=5)
    elif name == "winding4":
        H, S = total_degree(systems.winding_family(4), seed=3)
    else:
        H, S = total_degree(systems.wilkinson(12), seed=2)
    g, o = run_both(H, S, gpu_ctx)
    assert_parity(g, o, endpoint_rtol=1e-8)
    R = Result(path_results_from_batch(g, S))
    if name == "cyclic7":
        assert R.nsolutions() == 924
    elif name == "mohab":
        assert R.nnonsingular() == 693
    elif name == "hyperbolic":
        assert sum(1 for r in R if r.winding_number == 3) == 12
        assert R.nresults() == 2 and R.nsingular() == 2
    elif name == "winding4":
        assert sum(1 for r in R if r.is_success()) == 5
    else:
        assert sorted(np.round(np.real([r.solution[0] for r in R])).astype(int)) == list(range(1, 13))


@pytest.mark.parametrize("refill,bpc", [(1, 0), (5, 2), (64, 0), (1000, 1)])
def test_lane_schedule_does_not_change_results(gpu_ctx, refill, bpc):
    """Which lane tracks which path, and when a wave's idle lanes start their next paths
    (refill_idle: at once, in groups, only when the whole wave is idle), changes only the
    schedule: every path's record and endpoint are bit for bit those of the default launch,
    and agree with the oracle."""
    H, S = total_degree(dense_system(4, 3, seed=100), seed=101)
    S = np.concatenate([S] * 5)  # 405 paths: several waves, each starting many paths
    T = EndgameTracker(H, ctx=gpu_ctx)
    T.options = make_options()
    base = T.track_batch(S)
    T.options = make_options(refill_idle=refill, blocks_per_cu=bpc)
    g = T.track_batch(S)
    assert list(g.code) == list(base.code)
    np.testing.assert_array_equal(g.solution, base.solution)
    assert (g.accepted_steps == base.accepted_steps).all() and (g.rejected_steps == base.rejected_steps).all()
    sub = type(g)(g.records[:81], g.solution[:81], g.last_point[:81], g.valuation[:81])
    assert_parity(sub, oracle.track_batch(H, S[:81], make_options()))


def _isapprox(a, b, rtol=np.sqrt(np.finfo(float).eps)):
    a, b = np.asarray(a), np.asarray(b)
    return np.linalg.norm(a - b) <= rtol * max(np.linalg.norm(a), np.linalg.norm(b))


@pytest.mark.parametrize("layout", ["lane", "group"])
def test_steiner_parameter_case(gpu_ctx, layout):
    """test/test_cases/steiner_higher_prec.jl on the GPU (n = 15: the generic-size linear
    algebra): s_p at p reaches s_q at q, and s_q reaches s_p on the reversed homotopy, with
    extended-precision residuals on the way; results equal the oracle's."""
    F, p, q, s_p, s_q = systems.steiner()
    for H, start, want in ((ParameterHomotopy(F, p, q), s_p, s_q), (ParameterHomotopy(F, q, p), s_q, s_p)):
        g, o = run_both(H, start[None, :], gpu_ctx, layout=layout)
        assert_parity(g, o)
        assert g.codes() == ["success"]
        assert _isapprox(g.solution[0], want)
        assert g.extended_precision_used[0] == 1


@pytest.mark.parametrize("layout", ["lane", "group"])
def test_four_bar_parameter_case(gpu_ctx, layout):
    """test/test_cases/four_bar.jl on the GPU (n = 24, the largest supported size): the
    8-digit start point ends in success, as on the oracle."""
    F, p, q, s = systems.four_bar()
    g, o = run_both(ParameterHomotopy(F, p, q), s[None, :], gpu_ctx, layout=layout)
    assert_parity(g, o)
    assert g.codes() == ["success"]


def test_straight_line_and_invalid_starts(gpu_ctx):
    """test/tracker_test.jl:93-105 and :137-219 on the GPU: two successful starts, an invalid
    start, and a start with a singular Jacobian, classified as the oracle classifies them."""
    G, F, S = systems.straight_line_case()
    S = np.vstack([S, [[100, -100]]]).astype(complex)
    g, o = run_both(StraightLineHomotopy(G, F), S, gpu_ctx)
    assert_parity(g, o)
    assert g.codes() == ["success", "success", "terminated_invalid_startvalue"]
    F, b0, b1, start = systems.framework_454()
    g, o = run_both(ParameterHomotopy(F, b0, b1), start[None, :], gpu_ctx)
    assert_parity(g, o)
    assert g.codes() == ["terminated_invalid_startvalue_singular_jacobian"]
