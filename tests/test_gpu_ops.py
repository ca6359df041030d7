"""GPU parity on a ModelKit-shaped system that uses every OpType the C ABI accepts
(tests/systems.py::fused_op_system: the fused ADD3/ADD4/MUL3/MUL4/MULADD/MULSUB/SUBMUL/
MULMULADD/MULMULSUB, INV, INV_NOT_ZERO at an exact zero in the Jacobian tape, INVSQR, DIV,
IDENTITY, CB, SQR, POW_INT with positive and negative exponents).  The generated HIP evaluates
them in the tracker kernel's eval, Jacobian and Taylor functions; the trajectories must equal
the oracle's, whose op semantics are pinned against mpmath in test_oracle_ops.py."""
import mpmath as mp
import numpy as np
import pytest

import oracle
import systems
from hcgpu.gpu import track_batch
from hcgpu.homotopies import ParameterHomotopy
from hcgpu.solve import make_options

pytestmark = pytest.mark.gpu


def fused_problem(ntargets=64, seed=4):
    F, g, x = systems.fused_op_system()
    mp.mp.prec = 200
    p = np.array([complex(v) for v in g([mp.mpc(complex(v).real, complex(v).imag) for v in x], mp)])
    rng = np.random.default_rng(seed)
    Q = p + 0.4 * (rng.standard_normal((ntargets, 3)) + 1j * rng.standard_normal((ntargets, 3)))
    return F, g, x, p, Q


def test_fused_ops_parameter_homotopy_matches_oracle(gpu_ctx):
    F, g, x, p, Q = fused_problem()
    H = ParameterHomotopy(F, p, Q)
    S = np.tile(x, (len(Q), 1))
    groups = np.arange(len(Q), dtype=np.int32)
    opts = make_options()
    prog = gpu_ctx.program(H)
    r = track_batch(prog, S, opts, group_of_path=groups)
    o = oracle.track_batch(H, S, opts, group_of_path=groups)
    print("codes", np.bincount(r.code), "steps", r.accepted_steps.sum(), "ext", r.extended_precision_used.sum())
    assert list(r.code) == list(o.code)
    assert list(r.accepted_steps) == list(o.accepted_steps)
    assert list(r.rejected_steps) == list(o.rejected_steps)
    assert np.array_equal(r.solution, o.solution)  # bit for bit
    ok = r.code == 1
    assert ok.sum() >= 0.9 * len(Q)
    # every successful endpoint solves g(x) = q (200-bit evaluation of the closed form)
    mp.mp.prec = 200
    for k in np.flatnonzero(ok):
        z = [mp.mpc(v.real, v.imag) for v in r.solution[k]]
        res = max(abs(complex(gi) - qi) for gi, qi in zip(g(z, mp), Q[k]))
        assert res <= 1e-12 * max(1.0, np.abs(Q[k]).max()), (k, res)
