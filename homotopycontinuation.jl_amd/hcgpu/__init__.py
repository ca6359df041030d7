"""hcgpu: MI355X batched path tracker behind HomotopyContinuation.jl's solve / track API.

The compute path is libhcgpu.so (HIP for gfx950, hipRTC-compiled per system);
this package is the host-side mirror of the reference's interface.
"""
from .abi import HcgpuError, RETURN_CODES, default_options
from .homotopies import ParameterHomotopy, StraightLineHomotopy, total_degree, total_degree_start_solutions
from .polys import Poly, System, variables
from .result import PathResult, Result
from .solve import EndgameTracker, Solver, solve, solver_startsolutions, track
from .tape import Tape, eval_tape, jacobian_tape

__all__ = [
    "HcgpuError", "RETURN_CODES", "default_options", "ParameterHomotopy", "StraightLineHomotopy", "total_degree",
    "total_degree_start_solutions", "Poly", "System", "variables", "PathResult", "Result", "EndgameTracker",
    "Solver", "solve", "solver_startsolutions", "track", "Tape", "eval_tape", "jacobian_tape",
]
