"""Re-creates param10_step_mlp.npy on the GPU: the plain tracker logs its steps on 16
training targets of the n = 10 parameter workload (seeded draws disjoint from the
benchmark's 256 targets), and the HIP training kernels fit the step-size network.  The offset
of the log-step output is then chosen on the same training targets
(learned.select_output_offset: fastest among the offsets that keep the plain tracker's return
codes and endpoints on every training path).

Run from the repository root on a GPU host:
    python homotopycontinuation.jl_amd/hcgpu/weights/train_param10.py [--offset-only]
--offset-only keeps the saved network and chooses its output offset again.
"""
import json
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
    prog = ctx.program(H)
    if "--offset-only" in sys.argv:
        w = learned.saved_weights("param10_step_mlp")
        w = learned.with_output_offset(w, -float(json.loads((HERE / "param10_step_mlp.json").read_text())["kappa"])) \
            if (HERE / "param10_step_mlp.json").exists() else w
    else:
        out = track_batch(prog, starts, make_options(), group_of_path=groups, log_capacity=2_000_000)
        w, loss, info = learned.fit_from_log(ctx, out.log, out.code, epochs=8)
        print(f"trained on {info['records']} records ({info['weighted']} with a step target), final loss {loss:.4f}")

    def run(opts):
        o = track_batch(prog, starts, opts, group_of_path=groups)
        return o.code, o.solution, ctx.last_kernel_ms()
    w, kappa, table = learned.select_output_offset(ctx, w, [run])
    np.save(HERE / "param10_step_mlp.npy", w)
    (HERE / "param10_step_mlp.json").write_text(json.dumps({"kappa": kappa, "table": table}, indent=1) + "\n")
    print(f"output offset {kappa}: " + json.dumps(table))


if __name__ == "__main__":
    main()
