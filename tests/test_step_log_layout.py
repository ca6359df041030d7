"""The step-record layout documented at the C ABI (include/hcgpu_types.h, HCG_LOG_* indices)
is the layout the tracker writes: checked on the oracle's step log (the GPU log equals it
record for record, tests/test_gpu_learned.py::test_step_log_matches_oracle) by consistency
relations that a misplaced field breaks."""
import re
from pathlib import Path

import numpy as np

import oracle
import systems
from hcgpu.homotopies import total_degree
from hcgpu.solve import make_options

REPO = Path(__file__).resolve().parents[1]


def _layout():
    text = (REPO / "include" / "hcgpu_types.h").read_text()
    d = {m.group(1): int(m.group(2)) for m in re.finditer(r"#define HCG_LOG_([A-Z0-9_]+) (\d+)", text)}
    floats = int(re.search(r"#define HCG_STEP_RECORD_FLOATS (\d+)", text).group(1))
    return d, floats


def test_documented_step_record_layout_matches_the_log():
    L, floats = _layout()
    assert floats == 16 and set(L) == {"PATH", "STEP", "FEATURES", "ACCEPTED", "NEWTON_ITERS", "NEXT_LOG_DS",
                                       "FLAGS", "LOG10_ACCURACY", "PLAIN_LOG_DS"}
    H, S = total_degree(systems.beyond_polyhedral(), gamma=1.3j + 0.4)
    o = oracle.track_batch(H, S, make_options(), log_capacity=4096)
    log = o.log
    assert log.shape[1] == floats and len(log) > 0
    F = L["FEATURES"]
    for p in range(len(S)):
        R = log[log[:, L["PATH"]] == p]
        R = R[np.argsort(R[:, L["STEP"]])]
        # one record per step, numbered 0, 1, 2, ...
        assert np.array_equal(R[:, L["STEP"]], np.arange(len(R)))
        acc = R[:, L["ACCEPTED"]]
        assert set(np.unique(acc)) <= {0.0, 1.0}
        # accepted + rejected steps of the path result (endgame steps included)
        assert int(acc.sum()) == o.accepted_steps[p] and int((1 - acc).sum()) == o.rejected_steps[p]
        assert (R[:, L["NEWTON_ITERS"]] >= 1).all() and (R[:, L["NEWTON_ITERS"]] <= 11).all()
        # the step size a record announces is the first feature (log Δs) of the next record
        np.testing.assert_allclose(R[1:, F + 0], R[:-1, L["NEXT_LOG_DS"]], rtol=0, atol=1e-6)
        # feature 3 is t after the step: below 1, never increasing, unchanged by a rejected step
        t = R[:, F + 3]
        assert 0 <= t[0] < 1.0 and (np.diff(t) <= 0).all()
        assert (t[1:][acc[1:] == 0] == t[:-1][acc[1:] == 0]).all()
        # flags are sums of 1, 2 and 4; accuracy is a log10
        assert set(np.unique(R[:, L["FLAGS"]])) <= set(range(8))
        assert (R[acc == 1, L["LOG10_ACCURACY"]] < 0).all()
        # without the learned proposer the plain proposal is the next step unless the segment
        # end clamps it (it is then larger)
        assert (R[:, L["PLAIN_LOG_DS"]] >= R[:, L["NEXT_LOG_DS"]] - 1e-6).all()
