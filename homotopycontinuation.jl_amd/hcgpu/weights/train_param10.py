"""Re-creates param10_step_mlp.npy on the GPU: the plain tracker logs its steps on 16
training targets of the n = 10 parameter workload (seeded draws disjoint from the
benchmark's 256 targets), and the HIP training kernels fit the step-size network.

Run from the repository root on a GPU host:
    python homotopycontinuation.jl_amd/hcgpu/weights/train_param10.py
"""
import sys
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
REPO = HERE.parents[2]
sys.path.insert(0, str(REPO / "homotopycontinuation.jl_amd"))

from hcgpu import learned  # noqa: E402
from hcgpu.configs import cfg2_problem, parameter_draws  # noqa: E402
from hcgpu.gpu import default_context, track_batch  # noqa: E402
from hcgpu.homotopies import ParameterHomotopy  # noqa: E402
from hcgpu.solve import make_options  # noqa: E402


def main():
    F, p0, _ = cfg2_problem(1)
    S = np.load(REPO / "tests" / "golden" / "cfg2_starts.npy")
    Qt = parameter_draws(16, 16, seed=30)
    H = ParameterHomotopy(F, p0, Qt, lift_affine=True)
    ctx = default_context()
    starts = np.tile(S, (len(Qt), 1))
    groups = np.repeat(np.arange(len(Qt), dtype=np.int32), len(S))
    out = track_batch(ctx.program(H), starts, make_options(), group_of_path=groups, log_capacity=2_000_000)
    w, loss, info = learned.fit_from_log(ctx, out.log, out.code, epochs=8)
    np.save(HERE / "param10_step_mlp.npy", w)
    print(f"trained on {info['records']} records ({info['weighted']} with a step target), final loss {loss:.4f}")


if __name__ == "__main__":
    main()
