"""Re-creates mixed_step_mlp.npy on the GPU: the plain tracker logs its steps on
shared-factor training systems (n = 4, 6, 8; seeds disjoint from the mixed benchmark's),
and the HIP training kernels fit the step-size / at-infinity network on the log.  The offset
of the log-step output is then chosen on the same training systems
(learned.select_output_offset: fastest among the offsets that keep the plain tracker's return
codes and endpoints on every training path).

Run from the repository root on a GPU host:
    python homotopycontinuation.jl_amd/hcgpu/weights/train_mixed.py [--offset-only]
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
from hcgpu.configs import shared_factor_exprs  # noqa: E402
from hcgpu.gpu import default_context, track_batch  # noqa: E402
from hcgpu.homotopies import total_degree  # noqa: E402
from hcgpu.solve import make_options  # noqa: E402

TRAINING = [(8, 200 + k) for k in range(3)] + [(6, 210 + k) for k in range(6)] + [(4, 220 + k) for k in range(10)]


def training_homotopies():
    return [total_degree(shared_factor_exprs(n, seed=s), seed=s + 1000) for n, s in TRAINING]


def offset_runs(ctx):
    """One callable per training system: track its paths with the given options."""
    runs = []
    for H, S in training_homotopies():
        prog = ctx.program(H)

        def run(opts, prog=prog, S=S):
            out = track_batch(prog, S, opts)
            return out.code, out.solution, ctx.last_kernel_ms()
        runs.append(run)
    return runs


def main():
    ctx = default_context()
    if "--offset-only" in sys.argv:
        w = learned.saved_weights("mixed_step_mlp")
        w = learned.with_output_offset(w, -float(json.loads((HERE / "mixed_step_mlp.json").read_text())["kappa"])) \
            if (HERE / "mixed_step_mlp.json").exists() else w
    else:
        parts = []
        for H, S in training_homotopies():
            out = track_batch(ctx.program(H), S, make_options(), log_capacity=800_000)
            parts.append((out.log, out.code))
        log, codes = learned.combine_logs(parts)
        w, loss, info = learned.fit_from_log(ctx, log, codes, epochs=8)
        print(f"trained on {info['records']} records ({info['weighted']} with a step target), "
              f"{int(np.isin(codes, (2, 3)).sum())}/{len(codes)} paths at infinity, final loss {loss:.4f}")
    w, kappa, table = learned.select_output_offset(ctx, w, offset_runs(ctx))
    np.save(HERE / "mixed_step_mlp.npy", w)
    (HERE / "mixed_step_mlp.json").write_text(json.dumps({"kappa": kappa, "table": table}, indent=1) + "\n")
    print(f"output offset {kappa}: " + json.dumps(table))


if __name__ == "__main__":
    main()
