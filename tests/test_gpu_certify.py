"""Certification on the GPU (hcg_certify): one launch over all candidates.

* bit-identical to the oracle: every certificate field, both boxes and the refined point,
  for the FP64 test (Simple 1, the critical points with parameters, random non-solutions)
  and for the double-double retry (the 3,264 real conics' ill-conditioned candidates);
* the reference's known answers (test/certification_test.jl): 18 / 4 real / 14 complex for
  Simple 1 (doubled: 36 certified, 18 distinct), and all 3,264 real conics tangent to five
  given conics certified real and distinct, also with 100 of them given twice.
"""
import json
import time
from fractions import Fraction
from pathlib import Path

import numpy as np
import pytest

import oracle
import systems
from hcgpu.certify import certify, is_real_system, parameter_boxes
from hcgpu.homotopies import _SysTapes
from test_certify import critical_points_system, simple1, solve_nonsingular

pytestmark = pytest.mark.gpu

GOLDEN = Path(__file__).resolve().parent / "golden"
FIELDS = ["certified", "real", "complex_", "newton_code", "newton_iters", "prec", "contraction", "inflation"]


def _conics():
    return [Fraction(s) for s in json.loads((GOLDEN / "param_cases.json").read_text())["conics"]["real_conics"]]


def _compare(gpu_ctx, F, S, p=None):
    cert = gpu_ctx.certifier(F)
    pts, boxes, balls = parameter_boxes(list(p) if p is not None else [])
    has = F.nparams > 0
    real = is_real_system(F, pts)
    g = cert.certify(S, pts if has else None, boxes if has else None, balls if has else None, real_system=real)
    tapes = _SysTapes(F)  # the struct points into these tapes
    o = oracle.certify(tapes.struct(), F.nvars, S, pts if has else None, boxes if has else None,
                       balls if has else None, True, True, real)
    for f in FIELDS:
        np.testing.assert_array_equal(getattr(g, f), getattr(o, f), err_msg=f)
    np.testing.assert_array_equal(g.box1, o.box1)
    np.testing.assert_array_equal(g.box0, o.box0)
    np.testing.assert_array_equal(g.refined, o.refined)
    return g


def test_certify_bitwise_equals_oracle(gpu_ctx):
    F = simple1()
    S = solve_nonsingular(F)
    g = _compare(gpu_ctx, F, np.vstack([S, S]))
    assert g.certified.sum() == 36
    G = critical_points_system()
    u0 = [-0.32, -0.1]
    S2 = solve_nonsingular(G.fix_parameters(u0))
    rng = np.random.default_rng(1)
    junk = 100 * (rng.standard_normal((10, 3)) + 1j * rng.standard_normal((10, 3))) / np.sqrt(2)
    g2 = _compare(gpu_ctx, G, np.vstack([S2, junk]), u0)
    assert g2.certified[:36].all() and g2.certified[36:].sum() < 10


def test_certify_parameter_groups_bitwise_equals_oracle(gpu_ctx):
    G = critical_points_system()
    us = [[-0.32, -0.1], [0.25, 0.4], [1.5, -0.7]]
    sols = [solve_nonsingular(G.fix_parameters(u)) for u in us]
    S = np.vstack(sols)
    grp = np.repeat(np.arange(3), [len(x) for x in sols]).astype(np.int32)
    enc = [parameter_boxes(u) for u in us]
    pts, boxes, balls = (np.stack([e[i] for e in enc]) for i in range(3))
    g = gpu_ctx.certifier(G).certify(S, pts, boxes, balls, real_system=True, group_of=grp)
    tapes = _SysTapes(G)
    o = oracle.certify(tapes.struct(), 3, S, pts, boxes, balls, True, True, True, group_of=grp)
    for f in FIELDS:
        np.testing.assert_array_equal(getattr(g, f), getattr(o, f), err_msg=f)
    np.testing.assert_array_equal(g.box1, o.box1)
    assert g.certified.all()


def test_certify_extended_retry_bitwise_equals_oracle(gpu_ctx):
    F = systems.steiner()[0]
    S = np.load(GOLDEN / "conics_3264_real.npy")[::97]
    g = _compare(gpu_ctx, F, S, _conics())
    assert g.certified.all() and g.real.all() and (g.prec == 106).any()


def test_certify_simple1_known_answers(gpu_ctx):
    F = simple1()
    S = solve_nonsingular(F)
    r = certify(F, S, ctx=gpu_ctx)
    assert (r.ncertified(), r.nreal_certified(), r.ncomplex_certified(), r.ndistinct_certified()) == (18, 4, 14, 18)
    d = certify(F, np.vstack([S, S]), ctx=gpu_ctx)
    assert (d.ncertified(), d.ndistinct_certified(), d.nreal_certified(), d.ndistinct_real_certified()) == (36, 18, 8, 4)


def test_certify_3264_real_conics(gpu_ctx):
    F = systems.steiner()[0]
    S = np.load(GOLDEN / "conics_3264_real.npy")
    conics = _conics()
    t0 = time.perf_counter()
    r = certify(F, S, conics, ctx=gpu_ctx)
    dt = time.perf_counter() - t0
    prec = np.array([c.prec for c in r.certificates])
    print(f"3264 conics: {r.ndistinct_real_certified()} distinct real certified, {int((prec == 106).sum())} by the "
          f"double-double retry; kernel {r.kernel_ms:.2f} ms, call {dt * 1e3:.1f} ms")
    assert r.ndistinct_real_certified() == 3264 and r.ncertified() == 3264
    r2 = certify(F, np.vstack([S, S[:100]]), conics, ctx=gpu_ctx)
    assert r2.ndistinct_certified() == 3264 and len(r2.duplicates) == 100
    # DistinctCertifiedSolutions (test/certification_test.jl:300-314): one at a time, and batched
    from hcgpu.certify import DistinctCertifiedSolutions, distinct_certified_solutions
    d = DistinctCertifiedSolutions(F, conics, ctx=gpu_ctx)
    for x in S[:40]:
        d.add_solution(x, 1)
    assert len(d.solutions()) == 40
    d2 = distinct_certified_solutions(F, np.vstack([S, S[:100]]), conics, ctx=gpu_ctx)
    assert len(d2.solutions()) == 3264 and d2.nduplicates() == 100


def test_certify_solve_result(gpu_ctx):
    """The reference's flow: result = solve(F); certify(F, result) (test/certification_test.jl:3-24)."""
    from hcgpu.solve import solve
    F = simple1()
    res = solve(F, seed=5)
    assert res.nnonsingular() == 18
    r = certify(F, res, ctx=gpu_ctx)
    assert (r.ncertified(), r.nreal_certified(), r.ndistinct_certified()) == (18, 4, 18)


def test_certify_argument_errors(gpu_ctx):
    """Bad arguments are rejected before any launch, with a message."""
    from hcgpu.abi import HcgpuError
    G = critical_points_system()
    cert = gpu_ctx.certifier(G)
    S = np.zeros((3, 3), dtype=complex)
    with pytest.raises(ValueError):
        cert.certify(S)  # parameters missing
    with pytest.raises(HcgpuError, match="out of range"):
        cert.certify(S, [[1.0, 2.0]], group_of=[0, 1, 0])
    with pytest.raises(HcgpuError, match="not in its interval"):
        cert.certify(S[:1], [1.0, 2.0], params_ival=[[2.0, 3.0, 0.0, 0.0], [2.0, 2.0, 0.0, 0.0]])
    assert len(cert.certify(S[:0], [1.0, 2.0]).certified) == 0


def test_certify_with_quotients_bitwise_equals_oracle(gpu_ctx):
    """The interval quotient path (ci_div, ci_inv) on the GPU, on hand-written tapes."""
    import ctypes
    from hcgpu import abi
    from hcgpu.gpu import _CERT_DTYPE, _dp
    from test_certify import _quotient_system
    st, keep = _quotient_system()
    h = ctypes.c_void_p()
    abi.check(abi.lib().hcg_certifier_create(gpu_ctx.handle, ctypes.byref(st), 2, ctypes.byref(h)))
    try:
        S = np.ascontiguousarray(np.array([[2.0, 1.0], [2.0 + 1e-9, 1.0 - 1e-9j], [1.0, 1e-300]], dtype=complex))
        rec = np.zeros(3, dtype=_CERT_DTYPE)
        b1, b0, ref = np.zeros((3, 2, 4)), np.zeros((3, 2, 4)), np.zeros((3, 2), dtype=complex)
        abi.check(abi.lib().hcg_certify(gpu_ctx.handle, h, 3, _dp(S), 0, None, None, None, None, 1, 1, 1,
                                        rec.ctypes.data_as(ctypes.POINTER(abi.hcg_certificate)), _dp(b1), _dp(b0),
                                        _dp(ref)))
    finally:
        abi.lib().hcg_program_destroy(h)
    o = oracle.certify(st, 2, S, real_system=True)
    for f in FIELDS:
        np.testing.assert_array_equal(rec[f], getattr(o, f), err_msg=f)
    np.testing.assert_array_equal(b1, o.box1)
    assert list(rec["certified"]) == [1, 1, 0]
