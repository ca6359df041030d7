#!/usr/bin/env python3
"""Benchmark: paths/s of the MI355X batched path tracker on one GPU.

Workload (one "step"): the dense total-degree solve -- n = 8 variables, 8 dense cubic
polynomials with all 165 monomials of degree <= 3, iid complex Gaussian coefficients
(seed 1), total-degree start system x_i^3 - 1, 6,561 paths, straight-line homotopy with
γ from seed 2, ComplexF64 -- tracked to t = 0 with the endgame in ONE persistent
kernel launch.  Inputs are resident in HBM before the timed region (torch allocates
the buffers, libhcgpu's hcg_track_batch_device launches on them).

The JSON line carries the counted-FLOP roofline of the tracker kernel (HIP events on
the launch stream) and the CPU baseline: the oracle (a C restatement of the
reference tracker) on 16 host threads over the same 6,561 paths.
"""
import argparse
import ctypes
import json
import os
import sys
import time
from pathlib import Path

REPO = Path(__file__).resolve().parent
sys.path.insert(0, str(REPO / "homotopycontinuation.jl_amd"))

FP64_PEAK_TFLOPS = 78.6  # MI355X FP64 (vector and matrix)
HBM_PEAK_GBS = 8000.0


def measured_traffic(workload: str):
    """HBM bytes per launch from the committed FETCH_SIZE / WRITE_SIZE passes of this workload,
    each scaled by what the counter reads for a plain copy of known size; None if not profiled."""
    try:
        t = json.loads((REPO / "profiles" / "traffic.json").read_text())[workload]
    except (OSError, KeyError, ValueError):
        return None
    return round((t["fetch_kib"] * t["fetch_scale"] + t["write_kib"] * t["write_scale"]) * 1024.0)


def build_workload(learned: bool):
    from hcgpu.configs import dense_system
    from hcgpu.homotopies import total_degree
    F = dense_system(8, 3, seed=1)
    H, S = total_degree(F, seed=2)
    return H, S


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--cpu-threads", type=int, default=16)
    ap.add_argument("--copy-calibration-mib", type=int, default=0,
                    help="only run one plain device-to-device copy of this many MiB (the reference kernel for "
                         "scaling FETCH_SIZE / WRITE_SIZE counter readings) and exit")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU only: paths are independent and this build has no multi-GPU path "
              "(see DESIGN.md for how the batch would shard)", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch

    if args.copy_calibration_mib:
        nbytes = args.copy_calibration_mib << 20
        src = torch.ones(nbytes // 4, dtype=torch.float32, device="cuda:0")
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_calibration_bytes": nbytes, "read_bytes": nbytes, "written_bytes": nbytes}))
        return

    from hcgpu import abi
    from hcgpu.abi import check, hcg_batch, hcg_path_result, lib
    from hcgpu.gpu import GpuContext
    from hcgpu.solve import make_options

    H, S = build_workload(False)
    npaths, n = S.shape
    ctx = GpuContext(0)
    prog = ctx.program(H)  # hipRTC (code-object cache)
    opts = make_options()

    dev = torch.device("cuda:0")
    starts = torch.from_numpy(np.ascontiguousarray(S).view(np.float64).copy()).to(dev)
    res = torch.empty(npaths * ctypes.sizeof(hcg_path_result), dtype=torch.uint8, device=dev)
    sol = torch.empty(npaths * n * 2, dtype=torch.float64, device=dev)
    last = torch.empty_like(sol)
    val = torch.empty(npaths * n, dtype=torch.float64, device=dev)
    torch.cuda.synchronize()

    b = hcg_batch()
    b.npaths = npaths
    b.starts = ctypes.cast(starts.data_ptr(), ctypes.POINTER(ctypes.c_double))
    b.t1, b.t0 = 1.0, 0.0
    L = lib()
    L.hcg_track_batch_device.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.POINTER(hcg_batch),
                                         ctypes.POINTER(abi.hcg_options), ctypes.c_void_p, ctypes.c_void_p,
                                         ctypes.c_void_p, ctypes.c_void_p]

    def launch():
        check(L.hcg_track_batch_device(ctx.handle, prog._h, ctypes.byref(b), ctypes.byref(opts), res.data_ptr(),
                                       sol.data_ptr(), last.data_ptr(), val.data_ptr()))
        return ctx.last_kernel_ms(), ctx.last_counters()

    for _ in range(args.warmup):
        launch()
    torch.cuda.synchronize()
    kms, flops, steps, newton = [], [], [], []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        ms, c = launch()
        kms.append(ms)
        flops.append(c["flops"])
        steps.append(c["steps"])
        newton.append(c["newton_iters"])
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0

    # results of the last launch
    rec = np.frombuffer(res.cpu().numpy().tobytes(), dtype=np.dtype(
        [(f, np.int32 if t is ctypes.c_int32 else np.float64) for f, t in hcg_path_result._fields_], align=True))
    gsol = sol.cpu().numpy().view(np.complex128).reshape(npaths, n)
    nonsing = int(((rec["code"] == 1) & (rec["singular"] == 0)).sum())
    counts = dict(success=int((rec["code"] == 1).sum()), nonsingular=nonsing,
                  singular=int(((rec["code"] == 1) & (rec["singular"] != 0)).sum()),
                  at_infinity=int(((rec["code"] == 2) | (rec["code"] == 3)).sum()),
                  failed=int(((rec["code"] != 1) & (rec["code"] != 2) & (rec["code"] != 3)).sum()))

    value = npaths * args.steps / wall
    mean_ms = float(np.mean(kms))
    mean_flops = float(np.mean(flops))
    achieved = mean_flops / (mean_ms * 1e-3) / 1e12
    out = {
        "metric": "paths/sec: learned-step GPU vs plain GPU vs host CPU, same solution sets",
        "value": round(value, 1),
        "unit": "paths/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(wall / args.steps * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded complex Gaussian coefficients)",
        "config": {"workload": "dense total-degree solve: n=8, 8 dense cubics (165 monomials each), "
                               "6561 paths, one persistent launch (configs[1])",
                   "n": n, "paths": npaths, "system_seed": 1, "gamma_seed": 2, "tracker": "plain (adaptive rule)"},
        "roofline": {"bound": "mfma", "bound_detail": "FP64 compute roof (MI355X FP64 vector = matrix peak)",
                     "achieved": round(achieved, 4), "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                     "frac": round(achieved / FP64_PEAK_TFLOPS, 5), "traffic": measured_traffic("dense_n8_6561paths"),
                     "kernel": "hcg_track_kernel", "kernel_ms": round(mean_ms, 3),
                     "flops_per_launch": mean_flops, "flops_per_step": mean_flops / max(1, np.mean(steps))},
        "steps_per_path": round(float(np.mean(steps)) / npaths, 3),
        "newton_iters_per_step": round(float(np.mean(newton)) / max(1.0, float(np.mean(steps))), 3),
        "solution_counts": counts,
    }

    if not args.no_cpu:
        sys.path.insert(0, str(REPO))
        import oracle  # CPU baseline only (checker / baseline, never the measured GPU path)
        threads = args.cpu_threads
        t0 = time.perf_counter()
        o = oracle.track_batch(H, S, opts, nthreads=threads)
        dt = time.perf_counter() - t0
        same_codes = bool((o.code == rec["code"]).all())
        ok = (o.code == 1) & (~o.singular)
        rel = np.abs(gsol[ok] - o.solution[ok]).max(axis=1) / np.maximum(1.0, np.abs(o.solution[ok]).max(axis=1))
        out["cpu_baseline"] = {"value": round(npaths / dt, 1), "unit": "paths/s", "cores": threads, "kind": "port",
                               "sample": f"all {npaths} paths of the workload, oracle (C restatement of the "
                                         f"reference tracker) on {threads} threads"}
        out["gpu_vs_cpu"] = round(value / (npaths / dt), 2)
        out["parity"] = {"identical_return_codes": same_codes,
                         "max_rel_endpoint_diff": float(rel.max()) if ok.any() else 0.0}
    print(json.dumps(out))


if __name__ == "__main__":
    main()
