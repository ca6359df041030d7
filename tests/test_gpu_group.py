"""The lane-group layout (one path per 16 lanes, 32 for n > 16; layout = group) against the
oracle on the GPU parity cases: identical return codes, singular flags, winding numbers,
accepted and rejected step counts, and endpoints bit for bit (the layout performs the
oracle's operations in the oracle's order), plus the mixed-size multi-system launch and the
learned proposer."""
import numpy as np
import pytest

import oracle
import systems
from hcgpu.configs import cfg2_problem, cfg4_problem, dense_system
from hcgpu.gpu import track_batch, track_multi
from hcgpu.homotopies import ParameterHomotopy, StraightLineHomotopy, total_degree
from hcgpu.solve import make_options

pytestmark = pytest.mark.gpu


def run_group(gpu_ctx, H, S, **kw):
    opts = make_options(**kw)
    g = track_batch(gpu_ctx.program(H, layout="group"), S, opts)
    o = oracle.track_batch(H, S, make_options(**{k: v for k, v in kw.items() if k != "learned_step"}))
    return g, o


def assert_same(g, o, label=""):
    assert list(g.code) == list(o.code), (label, g.codes()[:8], o.codes()[:8])
    assert list(g.singular) == list(o.singular), label
    assert list(g.winding_number) == list(o.winding_number), label
    assert (g.accepted_steps == o.accepted_steps).all(), label
    assert (g.rejected_steps == o.rejected_steps).all(), label
    fin = np.isfinite(o.solution).all(axis=1)
    np.testing.assert_array_equal(g.solution[fin], o.solution[fin], err_msg=label)


def test_group_smoke_case(gpu_ctx):
    H, S = total_degree(systems.beyond_polyhedral(), gamma=1.3j + 0.4)
    g, o = run_group(gpu_ctx, H, S)
    print("group", g.codes(), g.accepted_steps, "oracle", o.codes(), o.accepted_steps)
    assert_same(g, o)


def test_group_parameter_tracker_case(gpu_ctx):
    H = ParameterHomotopy(systems.tracker_param_system(), [1, 0], [2, 4])
    g, o = run_group(gpu_ctx, H, np.array([[1, 1]], dtype=complex))
    assert_same(g, o)


@pytest.mark.parametrize("name", ["dense4", "cyclic7", "mohab", "hyperbolic", "winding4", "wilkinson12"])
def test_group_total_degree_cases(gpu_ctx, name):
    if name == "dense4":
        H, S = total_degree(dense_system(4, 3, seed=100), seed=101)
    elif name == "cyclic7":
        H, S = total_degree(systems.cyclic(7), seed=4)
    elif name == "mohab":
        H, S = total_degree(systems.mohab(), gamma=systems.MOHAB_GAMMA)
    elif name == "hyperbolic":
        H, S = total_degree(systems.hyperbolic(), seed=5)
    elif name == "winding4":
        H, S = total_degree(systems.winding_family(4), seed=3)
    else:
        H, S = total_degree(systems.wilkinson(12), seed=2)
    g, o = run_group(gpu_ctx, H, S)
    assert_same(g, o, name)


def test_group_parameter_workload_sample(gpu_ctx):
    F, p0, Q = cfg2_problem(2)
    from pathlib import Path
    S = np.load(Path(__file__).resolve().parent / "golden" / "cfg2_starts.npy")
    H = ParameterHomotopy(F, p0, Q, lift_affine=True)
    starts = np.tile(S, (2, 1))
    groups = np.repeat(np.arange(2, dtype=np.int32), len(S))
    g = track_batch(gpu_ctx.program(H, layout="group"), starts, make_options(), group_of_path=groups)
    print("group kernel ms", gpu_ctx.last_kernel_ms(), gpu_ctx.last_counters())
    o = oracle.track_batch(H, starts, make_options(), group_of_path=groups)
    assert_same(g, o)


def test_group_steiner_and_invalid_starts(gpu_ctx):
    F, p, q, s_p, s_q = systems.steiner()
    g, o = run_group(gpu_ctx, ParameterHomotopy(F, p, q), s_p[None, :])
    assert_same(g, o, "steiner")
    assert g.extended_precision_used[0] == 1
    G, F2, S = systems.straight_line_case()
    S = np.vstack([S, [[100, -100]]]).astype(complex)
    g, o = run_group(gpu_ctx, StraightLineHomotopy(G, F2), S)
    assert_same(g, o, "straight line")
    F3, b0, b1, start = systems.framework_454()
    g, o = run_group(gpu_ctx, ParameterHomotopy(F3, b0, b1), start[None, :])
    assert_same(g, o, "framework 454")
    assert g.codes() == ["terminated_invalid_startvalue_singular_jacobian"]


def test_group_steiner_batch_equals_lane_and_oracle(gpu_ctx):
    """n = 15, past the per-lane kernel's register-resident LU: the Steiner start tracked to 512
    nearby targets (the bench.py steiner workload's first 512) in both layouts.  Codes, step
    counts and endpoints are equal bit for bit between the layouts and to the oracle on a
    sample."""
    rng = np.random.default_rng(5)
    F, p, q, s_p, _ = systems.steiner()
    m = 512
    Q = np.asarray(q)[None, :] * (1 + 1e-3 * (rng.standard_normal((m, len(q))) + 1j * rng.standard_normal((m, len(q)))))
    H = ParameterHomotopy(F, p, Q)
    S = np.tile(np.asarray(s_p)[None, :], (m, 1))
    groups = np.arange(m, dtype=np.int32)
    g = track_batch(gpu_ctx.program(H, layout="group"), S, make_options(), group_of_path=groups)
    ln = track_batch(gpu_ctx.program(H, layout="lane"), S, make_options(), group_of_path=groups)
    assert (g.code == 1).all()
    assert_same(g, ln, "steiner batch, group vs lane")
    k = 32
    o = oracle.track_batch(H, S[:k], make_options(), group_of_path=groups[:k])
    sub = track_batch(gpu_ctx.program(H, layout="group"), S[:k], make_options(), group_of_path=groups[:k])
    assert_same(sub, o, "steiner batch sample, group vs oracle")


def test_group_mixed_sizes_one_launch(gpu_ctx):
    probs = cfg4_problem()
    Hs = [H for H, _ in probs]
    starts = [S if H.n < 12 else S[:2048] for H, S in probs]
    outs = track_multi(gpu_ctx.program_multi(Hs, layout="group"), starts, make_options())
    print("group multi launch ms", gpu_ctx.last_kernel_ms(), gpu_ctx.last_counters())
    for H, S, g in zip(Hs, starts, outs):
        assert_same(g, oracle.track_batch(H, S, make_options()), f"n={H.n}")


def test_group_learned_equals_lane_learned(gpu_ctx):
    """The learned proposer runs in the group layout (network evaluated by the 16 lanes of a
    path together, sums in the per-lane order): results equal the per-lane kernel's."""
    from hcgpu import learned
    F, p0, Q = cfg2_problem(1)
    from pathlib import Path
    S = np.load(Path(__file__).resolve().parent / "golden" / "cfg2_starts.npy")
    H = ParameterHomotopy(F, p0, Q, lift_affine=True)
    learned.load(gpu_ctx, learned.saved_weights("param10_step_mlp"))
    lane = track_batch(gpu_ctx.program(H), S, make_options(learned_step=1))
    grp = track_batch(gpu_ctx.program(H, layout="group"), S, make_options(learned_step=1))
    assert list(grp.code) == list(lane.code)
    assert (grp.accepted_steps == lane.accepted_steps).all()
    assert (grp.records["learned_accepted"] == lane.records["learned_accepted"]).all()
    np.testing.assert_array_equal(grp.solution, lane.solution)


def test_default_layout_by_size(gpu_ctx):
    """Without a tuned choice, GpuContext.program() takes one path per lane up to 12 variables
    and the lane-group layout past that (where the per-lane LU leaves the registers)."""
    F, p, q, _ = systems.four_bar()
    assert gpu_ctx.program(ParameterHomotopy(F, p, q)).layout == "group"
    F, p, q, _, _ = systems.steiner()
    assert gpu_ctx.program(ParameterHomotopy(F, p, q)).layout == "group"
    H, _ = total_degree(dense_system(4, 3, seed=100), seed=101)
    assert gpu_ctx.program(H).layout == "lane"


def _four_bar_batch(m, seed=6):
    F, p, q, s = systems.four_bar()
    rng = np.random.default_rng(seed)
    Q = np.asarray(q)[None, :] * (1 + 1e-3 * (rng.standard_normal((m, len(q))) + 1j * rng.standard_normal((m, len(q)))))
    return ParameterHomotopy(F, p, Q), np.tile(np.asarray(s)[None, :], (m, 1)), np.arange(m, dtype=np.int32)


def test_group32_four_bar_case(gpu_ctx):
    """n = 24 on the 32-lane kernel: test/test_cases/four_bar.jl's 8-digit start ends in
    success, equal to the oracle bit for bit."""
    F, p, q, s = systems.four_bar()
    H = ParameterHomotopy(F, p, q)
    prog = gpu_ctx.program(H, layout="group")
    assert prog.kernel_info()["workgroup"] == 256
    print("group32: four-bar single path", flush=True)
    g = track_batch(prog, np.asarray(s)[None, :], make_options())
    assert g.codes() == ["success"]
    assert_same(g, oracle.track_batch(H, np.asarray(s)[None, :], make_options()), "four-bar, group vs oracle")


def test_group32_four_bar_batch_equals_oracle(gpu_ctx):
    """256 nearby targets (the bench.py fourbar workload's first 256): every path equal to the
    oracle's (codes, step counts, endpoints bit for bit).  With the learned proposer (which the
    oracle does not run), the 32-lane kernel equals the per-lane kernel on the four shortest
    paths."""
    from hcgpu import learned
    H, S, groups = _four_bar_batch(256)
    prog = gpu_ctx.program(H, layout="group")
    print("group32: four-bar batch, plain", flush=True)
    g = track_batch(prog, S, make_options(), group_of_path=groups)
    assert (g.code == 1).mean() > 0.9
    o = oracle.track_batch(H, S, make_options(), group_of_path=groups)
    assert_same(g, o, "four-bar batch, plain")
    learned.load(gpu_ctx, learned.saved_weights("param10_step_mlp"))
    idx = np.argsort(o.accepted_steps + o.rejected_steps, kind="stable")[:4]
    print("group32: four-bar, learned, paths", idx.tolist(), flush=True)
    gl = track_batch(prog, S[idx], make_options(learned_step=1), group_of_path=groups[idx])
    ll = track_batch(gpu_ctx.program(H, layout="lane"), S[idx], make_options(learned_step=1), group_of_path=groups[idx])
    assert gl.records["learned_proposed"].sum() > 0
    assert_same(gl, ll, "four-bar, learned, group vs lane")
    assert (gl.records["learned_accepted"] == ll.records["learned_accepted"]).all()


def _chain20(seed=3):
    """n = 20: 18 linear equations chaining the variables and two quadratics (4 total-degree
    paths); a straight-line homotopy on the 32-lane kernel."""
    from hcgpu.polys import System, variables
    n = 20
    xs, _ = variables(n)
    rng = np.random.default_rng(seed)
    c = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    b = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    eqs = [xs[i] - c[i] * xs[i + 1] - b[i] for i in range(n - 2)]
    eqs += [xs[n - 2] ** 2 + c[n - 2] * xs[n - 1] - b[n - 2], xs[n - 1] ** 2 + c[n - 1] * xs[0] * xs[n - 2] - b[n - 1]]
    return System(eqs, n)


def test_group32_straight_line_case(gpu_ctx):
    H, S = total_degree(_chain20(), seed=4)
    print("group32: straight-line n = 20", flush=True)
    g, o = run_group(gpu_ctx, H, S)
    assert g.codes() == ["success"] * 4
    assert_same(g, o, "chain n=20")


def test_group32_multi_launch_with_a_small_system(gpu_ctx):
    """A launch over n = 4 and n = 20 runs on the 32-lane kernel (the largest system decides),
    where the n = 4 paths leave 28 lanes of their group idle; both equal the oracle."""
    H4, S4 = total_degree(dense_system(4, 3, seed=100), seed=101)
    H20, S20 = total_degree(_chain20(), seed=4)
    print("group32: multi-system launch n = 4, 20", flush=True)
    outs = track_multi(gpu_ctx.program_multi([H4, H20], layout="group"), [S4, S20], make_options())
    assert_same(outs[0], oracle.track_batch(H4, S4, make_options()), "n=4 in the 32-lane kernel")
    assert_same(outs[1], oracle.track_batch(H20, S20, make_options()), "n=20 beside n=4")


def test_group_step_log_equals_oracle_log(gpu_ctx):
    """The step records (HCG_LOG_* layout) written by the group layout equal the oracle's record
    for record, on paths that reach the singular endgame (winding number 3)."""
    H, S = total_degree(systems.hyperbolic(), seed=5)
    S = S[:12]
    g = track_batch(gpu_ctx.program(H, layout="group"), S, make_options(log_steps=1), log_capacity=50000)
    o = oracle.track_batch(H, S, make_options(log_steps=1), log_capacity=50000)
    assert_same(g, o)
    gl, ol = g.log, o.log
    gl = gl[np.lexsort((gl[:, 1], gl[:, 0]))]
    ol = ol[np.lexsort((ol[:, 1], ol[:, 0]))]
    assert gl.shape == ol.shape and len(gl) > 0
    np.testing.assert_array_equal(gl, ol)
