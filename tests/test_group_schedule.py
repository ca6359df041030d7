"""Slot schedules of the lane-group layout, built on the CPU for every system family the GPU
tests and the benchmark track.  hcg_group_schedule_info schedules each homotopy's tapes for
16-lane groups and checks the schedule against the tapes (every instruction placed once,
each operand produced at least two slots before its use, no LDS register overwritten while a
value in it is still to be read, every output register holding its tape value), so a
malformed schedule fails here with the checker's message."""
import numpy as np
import pytest

from hcgpu.configs import cfg2_problem, cfg4_problem, dense_system
from hcgpu.gpu import group_schedule_info
from hcgpu.homotopies import ParameterHomotopy, total_degree
import systems


def _info(hs):
    info = group_schedule_info(hs)
    for s in info["systems"]:
        assert s["ej_slots"] * 16 >= s["ej_instr"] > 0
        assert s["e_slots"] * 16 >= s["e_instr"] > 0
    return info


def test_parameter_workload_schedule():
    F, p0, Q = cfg2_problem(4)
    H = ParameterHomotopy(F, p0, Q, lift_affine=True)
    info = _info([H])
    s = info["systems"][0]
    assert info["nn"] == 12
    # the rows advance side by side: most slots are full
    assert s["ej_instr"] / (16 * s["ej_slots"]) > 0.6, s
    assert info["lds_per_path"] <= 8192, info


@pytest.mark.parametrize("n", [4, 8])
def test_dense_total_degree_schedule(n):
    H, _ = total_degree(dense_system(n, 3, seed=1), seed=2)
    info = _info([H])
    assert info["nn"] == (4 if n <= 4 else 8)


def test_mixed_workload_schedule_in_one_program():
    Hs = [H for H, _ in cfg4_problem()]
    info = _info(Hs)
    assert info["nn"] == 12 and len(info["systems"]) == 4


def test_every_op_type_and_quotients_schedule():
    from hcgpu.homotopies import ParameterHomotopy as PH
    F, _, _ = systems.fused_op_system()
    _info([PH(F, np.zeros(3), [np.zeros(3)])])
    for name in ("cyclic7", "mohab", "hyperbolic", "wilkinson12"):
        F = {"cyclic7": lambda: systems.cyclic(7), "mohab": systems.mohab, "hyperbolic": systems.hyperbolic,
             "wilkinson12": lambda: systems.wilkinson(12)}[name]()
        _info([total_degree(F, seed=3)[0]])
    F, p, q, _, _ = systems.steiner()
    info = _info([PH(F, p, q)])
    assert info["nn"] == 16


def test_four_bar_takes_the_32_lane_kernel():
    F, p, q, _ = systems.four_bar()  # n = 24
    assert group_schedule_info([ParameterHomotopy(F, p, q)])["nn"] == 24


def test_systems_beyond_24_variables_are_refused():
    from hcgpu.abi import HcgpuError
    with pytest.raises(HcgpuError, match="out of range"):
        group_schedule_info([total_degree(dense_system(25, 1, seed=1), seed=2)[0]])
