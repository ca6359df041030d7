#!/usr/bin/env python3
"""Benchmark: paths/s of the MI355X batched path tracker on one GPU.

Default workload ("param10", one step = one launch over the whole batch): n = 10 sparse
quadrics (20 monomials each) whose coefficients are affine in 16 complex parameters;
1,024 start solutions at p0 (tests/golden/cfg2_starts.npy) tracked to 256 target
parameter vectors = 262,144 paths of uneven length in ONE persistent launch, ComplexF64,
with the endgame.  Timed: the learned-step tracker (`value`) and the plain tracker, both
with inputs resident in HBM (hcg_track_batch_device).  Beside them: the CPU oracle (a C
restatement of the reference tracker) on a bounded sample of the same batch.

Other workloads (--workload): "dense4" = 4 dense cubics in 4 unknowns, 81 total-degree paths;
"dense8" = 8 dense cubics in 8 unknowns, 6,561 total-degree paths; "mixed" = shared-factor
systems of n = 4, 6, 8, 12 (538,812 paths) in one launch; "steiner" = the Steiner conics
(n = 15) tracked to 8,192 nearby targets; "fourbar" = the four-bar linkage (n = 24) likewise.  The default line also carries the
mixed workload's learned and plain launches (`mixed_workload`).

The JSON line carries the counted-FLOP roofline of the timed kernel (HIP events on the
launch stream) and, when profiled, its HBM traffic from profiles/traffic.json.
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
METRIC = "paths/sec: learned-step GPU vs plain GPU vs host CPU, same solution sets"


HBM_PEAK_GBS = 8000.0  # MI355X HBM3E


def _profile_entry(workload: str):
    try:
        return json.loads((REPO / "profiles" / "traffic.json").read_text())[workload]
    except (OSError, KeyError, ValueError):
        return None


def measured_traffic(workload: str):
    """HBM bytes per launch from the committed FETCH_SIZE / WRITE_SIZE passes of this workload,
    each scaled by what the counter reads for a plain copy of known size; None if not profiled."""
    t = _profile_entry(workload)
    if t is None or "fetch_kib" not in t:
        return None
    return round((t["fetch_kib"] * t["fetch_scale"] + t["write_kib"] * t["write_scale"]) * 1024.0)


def roofline(flops, kernel_ms, workload, kernel="hcg_track_kernel", info=None, grid_waves=None):
    """The tracker kernel against its algorithmic roof, MI355X's FP64 vector peak (no MFMA is
    used: the linear algebra is a few percent of a step's flops at these sizes), with what the
    profiles measured beside it: HBM traffic per launch (the bound the per-lane kernel actually
    hits) and, when profiled, the FP64 instruction counters and resident waves."""
    achieved = flops / (kernel_ms * 1e-3) / 1e12
    traffic = measured_traffic(workload)
    r = {"bound": "fp64_vector", "bound_detail": "FP64 vector peak of MI355X, 78.6 TFLOP/s (counted FP64 flops "
                                                  "of the launch / kernel time)",
         "achieved": round(achieved, 4), "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
         "frac": round(achieved / FP64_PEAK_TFLOPS, 5), "traffic": traffic,
         "kernel": kernel, "kernel_ms": round(kernel_ms, 3), "flops_per_launch": flops}
    prof = _profile_entry(workload) or {}
    if traffic is not None:
        gbs = traffic / (kernel_ms * 1e-3) / 1e9
        r["measured_bound"] = {"what": "HBM bandwidth (FETCH_SIZE + WRITE_SIZE of a profiled launch of this "
                                       "workload / this launch's kernel time)",
                               "achieved_gbs": round(gbs, 1), "peak_gbs": HBM_PEAK_GBS,
                               "frac": round(gbs / HBM_PEAK_GBS, 4), "profile": prof.get("source")}
    if "fp64_wave_flops" in prof:
        # SQ_INSTS_VALU_FLOPS_FP64 adds one lane's flops per wave instruction (ADD + MUL + 2 FMA +
        # TRANS); times the 64 lanes of a wave it bounds the FP64 lane flops executed
        bound = prof["fp64_wave_flops"] * 64
        r["fp64_counters"] = {
            "counter": prof["fp64_counter"], "per_launch": prof["fp64_wave_flops"],
            "executed_lane_flops_upper_bound": bound,
            "upper_bound_tflops": round(bound / (kernel_ms * 1e-3) / 1e12, 3),
            "upper_bound_frac": round(bound / (kernel_ms * 1e-3) / 1e12 / FP64_PEAK_TFLOPS, 4),
            "executed_bound_over_counted_flops": round(bound / flops, 3),
            "profile": prof.get("counters_source")}
    if info is not None:
        occ = dict(info)
        if grid_waves is not None:
            occ["grid_waves"] = int(grid_waves)
            occ["resident_waves_per_simd"] = round(min(grid_waves, 4 * info["cus"] * info["waves_per_simd_built"])
                                                   / (4 * info["cus"]), 3)
        if "mean_occupancy_per_cu" in prof:  # rocprofv3's MeanOccupancyPerCU of a profiled launch
            occ["measured_mean_waves_per_simd"] = round(prof["mean_occupancy_per_cu"] / 4.0, 3)
        r["occupancy"] = occ
    return r


def grid_waves(prog, npaths):
    """Waves the launch starts.  Per-lane kernel: one per 64 paths, at most one per wave slot
    (4 SIMDs x the waves per SIMD it was built for, per CU).  Lane-group kernel: four per
    workgroup of 16 paths (8 paths with 32-lane groups, n > 16), at most two workgroups per CU."""
    info = prog.kernel_info()
    if getattr(prog, "layout", "lane") == "group":
        per_wg = 16 if prog.n <= 16 else 8
        return 4 * min(-(-npaths // per_wg), 2 * info["cus"])
    return min(-(-npaths // 64), 4 * info["cus"] * info["waves_per_simd_built"])


def host_info():
    model = ""
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return {"nproc": os.cpu_count(), "usable_cpus": len(os.sched_getaffinity(0)), "cpu_model": model}


def native_oracle():
    """The oracle for the CPU-baseline timing: rebuilt here with -O3 -march=native."""
    sys.path.insert(0, str(REPO))
    import oracle  # CPU baseline only (never the measured GPU path)
    try:
        build = oracle.use_native_build()
    except RuntimeError:
        build = "gcc -O2 -mavx2 (already loaded)"
    return oracle, build


def DeviceBatch(*args, **kw):
    """A batch already resident in HBM (torch allocations), launched with hcg_track_batch_device."""
    from hcgpu.device import DeviceBatch as _DeviceBatch
    return _DeviceBatch(*args, **kw)


def timed(batch, opts, steps, warmup):
    import numpy as np
    import torch
    for _ in range(warmup):
        batch.launch(opts)
    torch.cuda.synchronize()
    kms, cs = [], []
    t0 = time.perf_counter()
    for _ in range(steps):
        ms, c = batch.launch(opts)
        kms.append(ms)
        cs.append(c)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    return dict(wall=wall, kernel_ms=float(np.mean(kms)), flops=float(np.mean([c["flops"] for c in cs])),
                steps=float(np.mean([c["steps"] for c in cs])), newton=float(np.mean([c["newton_iters"] for c in cs])))


def counts(rec):
    import numpy as np
    code, sing = rec["code"], rec["singular"] != 0
    return dict(success=int((code == 1).sum()), nonsingular=int(((code == 1) & ~sing).sum()),
                singular=int(((code == 1) & sing).sum()), at_infinity=int(np.isin(code, (2, 3)).sum()),
                failed=int((~np.isin(code, (1, 2, 3))).sum()))


def learned_weights(ctx, F, p0, S):
    """The committed network for this workload, or one trained now on the GPU from the plain
    tracker's step log over 16 training targets (seeded draws disjoint from the benchmark's)."""
    import numpy as np
    from hcgpu import learned
    from hcgpu.configs import parameter_draws
    from hcgpu.gpu import track_batch
    from hcgpu.homotopies import ParameterHomotopy
    from hcgpu.solve import make_options
    w = learned.saved_weights("param10_step_mlp")
    if w is not None:
        return w, "committed (hcgpu/weights/param10_step_mlp.npy)"
    Qt = parameter_draws(16, 16, seed=30)
    Ht = ParameterHomotopy(F, p0, Qt, lift_affine=True)
    st = np.tile(S, (len(Qt), 1))
    gt = np.repeat(np.arange(len(Qt), dtype=np.int32), len(S))
    out = track_batch(ctx.program(Ht), st, make_options(), group_of_path=gt, log_capacity=2_000_000)
    w, loss, info = learned.fit_from_log(ctx, out.log, out.code)
    return w, f"trained now on {info['records']} step records (loss {loss:.4f})"


def certify_leg(ctx, F, Q, sol, rec, groups, args):
    """Interval Krawczyk certification of every successful endpoint of the learned run, each
    against its own target parameters, in one grouped launch (outside the timed region)."""
    import numpy as np
    from hcgpu.certify import certify_groups
    ok = rec["code"] == 1
    S, G = sol[ok], groups[ok]
    cert = ctx.certifier(F)
    t0 = time.perf_counter()
    r = certify_groups(F, S, Q, G, certifier=cert)
    wall = time.perf_counter() - t0
    kms = ctx.last_kernel_ms()
    leg = {"candidates": int(len(S)), "certified": r.ncertified(), "distinct_certified": r.ndistinct_certified(),
           "by_double_double_retry": int(((r.prec == 106) & r.certified).sum()),
           "kernel": "hcg_certify_kernel", "kernel_ms": round(kms, 3),
           "value": round(len(S) / (kms * 1e-3), 1), "unit": "certificates/s",
           "host_call_s": round(wall, 3)}
    if not args.no_cpu:
        sys.path.insert(0, str(REPO))
        import oracle  # CPU baseline only (never the measured GPU path)
        from hcgpu.certify import parameter_boxes
        from hcgpu.homotopies import _SysTapes
        m = min(len(S), 4096)
        enc = [parameter_boxes(list(q)) for q in Q]
        pts = np.stack([e[0] for e in enc])
        tapes = _SysTapes(F)
        t0 = time.perf_counter()
        o = oracle.certify(tapes.struct(), F.nvars, S[:m], pts, None, None, True, True, False, args.cpu_threads,
                           group_of=G[:m])
        dt = time.perf_counter() - t0
        g = cert.certify(S[:m], pts, real_system=False, group_of=G[:m])
        same = all(np.array_equal(getattr(g, f), getattr(o, f)) for f in ("certified", "real", "complex_", "prec",
                                                                           "contraction"))
        leg["cpu_baseline"] = {"value": round(m / dt, 1), "unit": "certificates/s", "cores": args.cpu_threads,
                               "kind": "port", "sample": f"first {m} candidates, oracle, {dt:.2f} s"}
        leg["parity_vs_oracle_sample"] = {"identical_records_and_boxes": bool(same and np.array_equal(g.box1, o.box1))}
    return leg


def training_leg(ctx, F, p0, S, args):
    """The learned proposer's training workload: step records logged by the plain tracker on
    16 training targets (seeded draws disjoint from the benchmark's), then Adam on the GPU
    (hcg_mlp_train) for 2 epochs; throughput in records per second (outside the timed region)."""
    import numpy as np
    from hcgpu import learned
    from hcgpu.configs import parameter_draws
    from hcgpu.gpu import track_batch
    from hcgpu.homotopies import ParameterHomotopy
    from hcgpu.solve import make_options
    Qt = parameter_draws(16, 16, seed=30)
    Ht = ParameterHomotopy(F, p0, Qt, lift_affine=True)
    st = np.tile(S, (len(Qt), 1))
    gt = np.repeat(np.arange(len(Qt), dtype=np.int32), len(S))
    o = track_batch(ctx.program(Ht), st, make_options(), group_of_path=gt, log_capacity=2_000_000)
    X, Y, W = learned.training_set(o.log, o.code)
    mu, sd = learned.feature_scaling(X)
    Xn = ((X - mu) / sd).astype(np.float32)
    epochs = 2
    t0 = time.perf_counter()
    _, loss = learned.train(ctx, learned.init_params(0), Xn, Y, W, epochs=epochs, batch=4096, lr=2e-3, seed=0)
    wall = time.perf_counter() - t0
    leg = {"records": int(len(X)), "epochs": epochs, "batch": 4096, "value": round(len(X) * epochs / wall, 1),
           "unit": "records/s (forward + backward + Adam)", "wall_s": round(wall, 3), "final_loss": round(float(loss), 5)}
    if not args.no_cpu:
        sys.path.insert(0, str(REPO))
        from oracle import mlp_ref  # CPU baseline only
        m = 4096
        t0 = time.perf_counter()
        mlp_ref.loss_and_grad(learned.init_params(0), Xn[:m].astype(np.float64), Y[:m].astype(np.float64),
                              W[:m].astype(np.float64))
        dt = time.perf_counter() - t0
        leg["cpu_baseline"] = {"value": round(m / dt, 1), "unit": "records/s (forward + backward)", "cores": 1,
                               "kind": "port", "sample": f"one batch of {m} records, numpy float64 reference, {dt:.2f} s"}
    return leg


# per-lane (waves per SIMD, workgroups per CU) candidates the benchmark's autotuning pass times
BENCH_LANE_CANDIDATES = ((1, 8), (2, 8), (2, 16), (4, 16))


def graph_leg(ctx, prog, batch, opts, rl, rec_l, sol_l, H, S, args):
    """The same launch captured once as a HIP graph (hcg_graph_create) and replayed: the whole
    batch, and a one-target batch (1,024 paths) where launch overhead is a larger share."""
    import numpy as np
    import torch

    def same(a, b):
        return bool(np.array_equal(a[0].view(np.uint8), b[0].view(np.uint8)) and
                    np.array_equal(a[1].view(np.uint8), b[1].view(np.uint8)))

    def replay(g, k):
        kms = []
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(k):
            kms.append(g.launch()[0])
        return (time.perf_counter() - t0) / k * 1e3, float(np.mean(kms))

    leg = {}
    g = batch.graph(opts)
    try:
        g.launch()
        wall_ms, kms = replay(g, args.steps)
        rec_g, sol_g = batch.results()[:2]
        leg["full_batch"] = {"paths": batch.npaths, "direct_ms_per_batch": round(rl["wall"] / args.steps * 1e3, 3),
                             "graph_ms_per_batch": round(wall_ms, 3), "graph_kernel_ms": round(kms, 3),
                             "identical_to_direct": same((rec_g, sol_g), (rec_l, sol_l))}
    finally:
        g.close()
    small = DeviceBatch(ctx, prog, S, H.p, H.q[:1], np.zeros(len(S), dtype=np.int32))
    k = max(20, args.steps)
    small.launch(opts)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(k):
        small.launch(opts)
    direct_ms = (time.perf_counter() - t0) / k * 1e3
    ref = small.results()
    g = small.graph(opts)
    try:
        g.launch()
        wall_ms, kms = replay(g, k)
        leg["one_target_batch"] = {"paths": small.npaths, "direct_ms_per_batch": round(direct_ms, 3),
                                   "graph_ms_per_batch": round(wall_ms, 3), "graph_kernel_ms": round(kms, 3),
                                   "identical_to_direct": same(small.results(), ref)}
    finally:
        g.close()
    return leg


def bench_param10(args, out):
    import numpy as np
    from hcgpu import autotune, learned
    from hcgpu.configs import cfg2_problem
    from hcgpu.gpu import GpuContext
    from hcgpu.homotopies import ParameterHomotopy
    from hcgpu.solve import make_options
    F, p0, Q = cfg2_problem(256)
    S = np.load(REPO / "tests" / "golden" / "cfg2_starts.npy")
    H = ParameterHomotopy(F, p0, Q, lift_affine=True)
    starts = np.tile(S, (len(Q), 1))
    groups = np.repeat(np.arange(len(Q), dtype=np.int32), len(S))
    ctx = GpuContext(0)
    w, wsrc = learned_weights(ctx, F, p0, S)
    learned.load(ctx, w)
    plain_o, learned_o = make_options(), make_options(learned_step=1)
    if args.layout == "auto":
        # the autotuner on the measured configuration (learned step, whole batch), before the
        # timed region; its choice is cached per device and GpuContext.program() builds it
        cands = autotune.candidates_for(H.n, lane=BENCH_LANE_CANDIDATES, group=(2,))
        t0 = time.perf_counter()
        tuned = autotune.tune(ctx, H, starts, learned_o, groups, candidates=cands, force=args.retune)
        tune_s = time.perf_counter() - t0
        prog = ctx.program(H)
        launch = {"chosen_by": "autotuner", "tuning_s": round(tune_s, 1), "table": tuned["table"]}
    else:
        prog = ctx.program(H, layout=args.layout)
        launch = {"chosen_by": "--layout"}
    plain_o.blocks_per_cu = learned_o.blocks_per_cu = prog.blocks_per_cu
    info = prog.kernel_info()
    launch.update({"layout": prog.layout, "waves_per_simd_built": info["waves_per_simd_built"],
                   "blocks_per_cu": prog.blocks_per_cu or None})
    batch = DeviceBatch(ctx, prog, starts, H.p, H.q, groups)
    npaths = len(starts)
    rp = timed(batch, plain_o, args.steps, args.warmup)
    rec_p, sol_p = batch.results()[:2]
    rl = timed(batch, learned_o, args.steps, args.warmup)
    rec_l, sol_l = batch.results()[:2]

    value = npaths * args.steps / rl["wall"]
    plain_value = npaths * args.steps / rp["wall"]
    ok = (rec_p["code"] == 1) & (rec_l["code"] == 1) & (rec_p["singular"] == 0)
    rel = (np.abs(sol_l[ok] - sol_p[ok]).max(axis=1) / np.maximum(1.0, np.abs(sol_p[ok]).max(axis=1))) if ok.any() \
        else np.zeros(1)
    lp = int(rec_l["learned_proposed"].sum())
    out.update({
        "value": round(value, 1), "ms_per_step": round(rl["wall"] / args.steps * 1e3, 3),
        "config": {"workload": "param10: n=10 sparse quadrics (20 monomials, coefficients affine in 16 params), "
                               "1024 starts x 256 targets = 262144 paths, one persistent launch, learned-step tracker",
                   "n": 10, "paths": npaths, "starts": len(S), "targets": len(Q), "params": 16,
                   "affine_parameter_lift": True, "weights": wsrc,
                   "layout": "one path per lane" if prog.layout == "lane" else "one path per 16-lane group",
                   "launch": launch},
        "roofline": roofline(rl["flops"], rl["kernel_ms"], "param10_learned" + ("" if prog.layout == "lane" else "_group"),
                             kernel="hcg_track_kernel" if prog.layout == "lane" else "hcg_group_kernel",
                             info=prog.kernel_info(), grid_waves=grid_waves(prog, npaths)),
        "plain_gpu": {"value": round(plain_value, 1), "unit": "paths/s", "kernel_ms": round(rp["kernel_ms"], 3),
                      "steps_per_path": round(rp["steps"] / npaths, 3),
                      "roofline_frac": round(rp["flops"] / (rp["kernel_ms"] * 1e-3) / 1e12 / FP64_PEAK_TFLOPS, 5)},
        "learned": {"steps_per_path": round(rl["steps"] / npaths, 3), "proposals": lp,
                    "acceptance_rate": round(int(rec_l["learned_accepted"].sum()) / max(1, lp), 4),
                    "speedup_vs_plain": round(value / plain_value, 3)},
        "steps_per_path": round(rl["steps"] / npaths, 3),
        "newton_iters_per_step": round(rl["newton"] / max(1.0, rl["steps"]), 3),
        "solution_counts": counts(rec_l),
        "parity_learned_vs_plain": {"identical_return_codes": bool((rec_l["code"] == rec_p["code"]).all()),
                                    "max_rel_endpoint_diff": float(rel.max())},
    })
    if not args.no_graph:
        out["graph_replay"] = graph_leg(ctx, prog, batch, learned_o, rl, rec_l, sol_l, H, S, args)
    if not args.no_certify:
        out["certification"] = certify_leg(ctx, F, Q, sol_l, rec_l, groups, args)
    if not args.no_training:
        out["learned_training"] = training_leg(ctx, F, p0, S, args)
    if not args.no_cpu:
        sys.path.insert(0, str(REPO))
        import oracle  # CPU baseline only (never the measured GPU path)
        from hcgpu.tape import shared_lowering
        with shared_lowering():  # the CPU evaluates tapes with every subexpression shared
            Hc = ParameterHomotopy(F, p0, Q, lift_affine=True)
        m = args.cpu_targets * len(S)
        t0 = time.perf_counter()
        o = oracle.track_batch(Hc, starts[:m], plain_o, nthreads=args.cpu_threads, group_of_path=groups[:m])
        dt = time.perf_counter() - t0
        okc = (o.code == 1) & (~o.singular)
        relc = np.abs(sol_l[:m][okc] - o.solution[okc]).max(axis=1) / np.maximum(1.0, np.abs(o.solution[okc]).max(axis=1))
        out["cpu_baseline"] = {"value": round(m / dt, 1), "unit": "paths/s", "cores": args.cpu_threads, "kind": "port",
                               "build": args.oracle_build, "host": host_info(),
                               "sample": f"first {args.cpu_targets} of the 256 targets ({m} paths) of the same batch, "
                                         f"oracle (C restatement of the reference tracker, plain rule), {dt:.1f} s"}
        out["gpu_vs_cpu"] = round(value / (m / dt), 2)
        out["plain_gpu_vs_cpu"] = round(plain_value / (m / dt), 2)
        out["parity_vs_oracle_sample"] = {
            "identical_return_codes": bool((o.code == rec_l["code"][:m]).all()),
            "max_rel_endpoint_diff_learned": float(relc.max()) if okc.any() else 0.0}


# dense total-degree workloads: (n, system seed, gamma seed), the sizes and seeds of the GPU tests
DENSE = {"dense4": (4, 100, 101), "dense8": (8, 1, 2)}


def bench_dense(args, out):
    from hcgpu.configs import dense_system
    from hcgpu.gpu import GpuContext
    from hcgpu.homotopies import total_degree
    from hcgpu.solve import make_options
    from hcgpu import autotune
    n, sseed, gseed = DENSE[args.workload]
    H, S = total_degree(dense_system(n, 3, seed=sseed), seed=gseed)
    ctx = GpuContext(0)
    opts = make_options()
    if args.layout == "auto":
        # 3^n paths (81 or 6,561) fill 2 or 103 waves of the per-lane kernel, far fewer than the
        # GPU's 1,024 SIMDs, or 21 or 1,641 waves of the lane-group kernel (4 paths per wave):
        # the tuner times both
        cands = autotune.candidates_for(H.n, lane=((1, 4), (1, 8), (2, 8)), group=(1, 2))
        t0 = time.perf_counter()
        tuned = autotune.tune(ctx, H, S, opts, candidates=cands, force=args.retune)
        launch = {"chosen_by": "autotuner", "tuning_s": round(time.perf_counter() - t0, 1), "table": tuned["table"]}
        prog = ctx.program(H)
    else:
        prog = ctx.program(H, waves_per_eu=1) if args.layout == "lane" else ctx.program(H, layout="group")
        launch = {"chosen_by": "--layout"}
    opts.blocks_per_cu = prog.blocks_per_cu
    launch.update({"layout": prog.layout, "waves_per_simd_built": prog.kernel_info()["waves_per_simd_built"],
                   "blocks_per_cu": prog.blocks_per_cu or None})
    batch = DeviceBatch(ctx, prog, S)
    r = timed(batch, opts, args.steps, args.warmup)
    rec = batch.results()[0]
    npaths = len(S)
    out.update({"value": round(npaths * args.steps / r["wall"], 1), "ms_per_step": round(r["wall"] / args.steps * 1e3, 3),
                "config": {"workload": f"{args.workload}: n={n}, {n} dense cubics ({len(S)} total-degree paths), "
                                       "one persistent launch, plain tracker", "n": n, "paths": npaths,
                           "launch": launch},
                "roofline": roofline(r["flops"], r["kernel_ms"], f"dense_n{n}_{npaths}paths" + ("" if prog.layout == "lane" else "_group"),
                                     kernel="hcg_track_kernel" if prog.layout == "lane" else "hcg_group_kernel",
                                     info=prog.kernel_info(), grid_waves=grid_waves(prog, npaths)),
                "steps_per_path": round(r["steps"] / npaths, 3), "solution_counts": counts(rec)})
    if not args.no_cpu:
        sys.path.insert(0, str(REPO))
        import oracle  # CPU baseline only
        from hcgpu.tape import shared_lowering
        with shared_lowering():  # the CPU evaluates tapes with every subexpression shared
            Hc, _ = total_degree(dense_system(n, 3, seed=sseed), seed=gseed)
        t0 = time.perf_counter()
        o = oracle.track_batch(Hc, S, make_options(), nthreads=args.cpu_threads)
        dt = time.perf_counter() - t0
        out["cpu_baseline"] = {"value": round(npaths / dt, 1), "unit": "paths/s", "cores": args.cpu_threads,
                               "kind": "port", "build": args.oracle_build, "host": host_info(), "sample": f"all {npaths} paths, oracle, {dt:.1f} s"}
        out["parity_vs_oracle"] = {"identical_return_codes": bool((o.code == rec["code"]).all())}


def nearby_targets(q, m, seed):
    """m targets near q: every coefficient perturbed by 1e-3 relative (seeded)."""
    import numpy as np
    rng = np.random.default_rng(seed)
    return np.asarray(q)[None, :] * (1 + 1e-3 * (rng.standard_normal((m, len(q))) + 1j * rng.standard_normal((m, len(q)))))


def steiner_problem(m, seed=5):
    """Steiner conics (n = 15, 30 parameters; test/test_cases/steiner_higher_prec.jl): the
    reference's start solution s_p at p tracked to m targets near its q."""
    import numpy as np
    sys.path.insert(0, str(REPO / "tests"))
    import systems
    F, p, q, s_p, _ = systems.steiner()
    return F, np.asarray(p), nearby_targets(q, m, seed), np.asarray(s_p)


def fourbar_problem(m, seed=6):
    """Four-bar linkage synthesis (n = 24, 16 parameters; test/test_cases/four_bar.jl): the
    reference's start point s (8 digits) at p tracked to m targets near its q."""
    import numpy as np
    sys.path.insert(0, str(REPO / "tests"))
    import systems
    F, p, q, s = systems.four_bar()
    return F, np.asarray(p), nearby_targets(q, m, seed), np.asarray(s)


def bench_steiner(args, out):
    import numpy as np
    from hcgpu import autotune
    from hcgpu.gpu import GpuContext
    from hcgpu.homotopies import ParameterHomotopy
    from hcgpu.solve import make_options
    m = 8192
    fourbar = args.workload == "fourbar"
    F, p, Q, s_p = fourbar_problem(m) if fourbar else steiner_problem(m)
    H = ParameterHomotopy(F, p, Q)
    S = np.tile(s_p[None, :], (m, 1))
    groups = np.arange(m, dtype=np.int32)
    ctx = GpuContext(0)
    opts = make_options()
    if fourbar:
        # n = 24: the 32-lane group kernel unless --layout lane.  The tuner is not run: it times
        # every candidate twice on the whole batch, and the per-lane kernel takes ~39 s here
        prog = ctx.program(H, layout="lane" if args.layout == "lane" else "group")
        launch = {"chosen_by": "--layout" if args.layout != "auto" else "default for n = 24 (lane-group, 32 lanes)"}
    elif args.layout == "auto":
        # n = 15 is past the per-lane kernel's register-resident LU (n <= 12): the tuner times the
        # lane-group kernel beside it
        cands = autotune.candidates_for(H.n, lane=((1, 4), (1, 8), (2, 8)), group=(1, 2))
        t0 = time.perf_counter()
        tuned = autotune.tune(ctx, H, S, opts, groups, candidates=cands, force=args.retune)
        launch = {"chosen_by": "autotuner", "tuning_s": round(time.perf_counter() - t0, 1), "table": tuned["table"]}
        prog = ctx.program(H)
    else:
        prog = ctx.program(H, layout=args.layout)
        launch = {"chosen_by": "--layout"}
    opts.blocks_per_cu = prog.blocks_per_cu
    launch.update({"layout": prog.layout, "waves_per_simd_built": prog.kernel_info()["waves_per_simd_built"],
                   "blocks_per_cu": prog.blocks_per_cu or None})
    batch = DeviceBatch(ctx, prog, S, H.p, H.q, groups)
    r = timed(batch, opts, args.steps, args.warmup)
    rec, sol = batch.results()[:2]
    out.update({"value": round(m * args.steps / r["wall"], 1), "ms_per_step": round(r["wall"] / args.steps * 1e3, 3),
                "config": {"workload": (f"fourbar: n=24, 16 parameters, the reference's four-bar start tracked to {m} "
                                        "targets near its q in one persistent launch, plain tracker") if fourbar else
                                       (f"steiner: n=15, 30 parameters, the reference's Steiner start tracked to {m} "
                                        "targets near its q in one persistent launch, plain tracker"), "n": H.n,
                           "paths": m, "launch": launch},
                "roofline": roofline(r["flops"], r["kernel_ms"], ("fourbar" if fourbar else "steiner") + f"_{m}paths" + ("" if prog.layout == "lane" else "_group"),
                                     kernel="hcg_track_kernel" if prog.layout == "lane" else "hcg_group_kernel",
                                     info=prog.kernel_info(), grid_waves=grid_waves(prog, m)),
                "steps_per_path": round(r["steps"] / m, 3), "solution_counts": counts(rec)})
    if not args.no_cpu:
        sys.path.insert(0, str(REPO))
        import oracle  # CPU baseline only
        k = 512
        t0 = time.perf_counter()
        o = oracle.track_batch(H, S[:k], opts, nthreads=args.cpu_threads, group_of_path=groups[:k])
        dt = time.perf_counter() - t0
        ok = (o.code == 1) & (rec["code"][:k] == 1)
        rel = np.abs(sol[:k][ok] - o.solution[ok]).max(axis=1) / np.maximum(1.0, np.abs(o.solution[ok]).max(axis=1))
        out["cpu_baseline"] = {"value": round(k / dt, 1), "unit": "paths/s", "cores": args.cpu_threads, "kind": "port",
                               "build": args.oracle_build, "host": host_info(),
                               "sample": f"first {k} of the {m} paths, oracle, {dt:.1f} s"}
        out["gpu_vs_cpu"] = round(out["value"] / (k / dt), 2)
        out["parity_vs_oracle_sample"] = {"identical_return_codes": bool((o.code == rec["code"][:k]).all()),
                                          "max_rel_endpoint_diff": float(rel.max()) if ok.any() else 0.0}


def bench_mixed(args, out):
    import numpy as np
    from hcgpu import learned
    from hcgpu.configs import cfg4_problem
    from hcgpu.gpu import GpuContext, track_multi
    from hcgpu.solve import make_options
    probs = cfg4_problem()
    ctx = GpuContext(0)
    prog = ctx.program_multi([H for H, _ in probs], layout="group" if args.layout == "group" else "lane")
    starts = [S for _, S in probs]
    total = sum(len(S) for S in starts)
    learned.load(ctx, learned.saved_weights("mixed_step_mlp"))

    def run(opts):
        for _ in range(args.warmup):
            track_multi(prog, starts, opts)
        kms, fl, outs = [], [], None
        for _ in range(args.steps):
            outs = track_multi(prog, starts, opts)
            kms.append(ctx.last_kernel_ms())
            fl.append(ctx.last_counters()["flops"])
        return float(np.mean(kms)), float(np.mean(fl)), outs

    pms, pfl, pouts = run(make_options())
    lms, lfl, louts = run(make_options(learned_step=1))
    same = all(bool((a.code == b.code).all()) for a, b in zip(pouts, louts))
    jump = 0.0  # largest relative endpoint difference learned vs plain over paths both end in success
    for a, b in zip(pouts, louts):
        ok = (a.code == 1) & (b.code == 1)
        if ok.any():
            r = np.abs(b.solution[ok] - a.solution[ok]).max(axis=1) / np.maximum(1.0, np.abs(a.solution[ok]).max(axis=1))
            jump = max(jump, float(r.max()))
    out.update({"value": round(total / lms * 1e3, 1), "ms_per_step": round(lms, 3),
                "config": {"workload": "mixed: shared-factor systems n=4,6,8,12 (538812 total-degree paths) in one "
                                       "launch, learned-step tracker; value from the kernel time (inputs are copied "
                                       "from the host each launch)", "paths": total,
                           "weights": "committed (hcgpu/weights/mixed_step_mlp.npy)"},
                "roofline": roofline(lfl, lms, "mixed_learned", kernel="hcg_track_multi_kernel",
                                     info=prog.kernel_info(), grid_waves=grid_waves(prog, total)),
                "plain_gpu": {"value": round(total / pms * 1e3, 1), "unit": "paths/s", "kernel_ms": round(pms, 3)},
                "learned": {"speedup_vs_plain": round(pms / lms, 3),
                            "steps_per_path": round(sum(int((o.accepted_steps + o.rejected_steps).sum()) for o in louts)
                                                    / total, 3),
                            "acceptance_rate": round(sum(int(o.learned_accepted.sum()) for o in louts)
                                                     / max(1, sum(int(o.learned_proposed.sum()) for o in louts)), 4)},
                "plain_steps_per_path": round(sum(int((o.accepted_steps + o.rejected_steps).sum()) for o in pouts)
                                              / total, 3),
                "identical_return_codes_learned_vs_plain": same,
                "max_rel_endpoint_diff_learned_vs_plain": jump,
                "solution_counts": {f"n{H.n}": counts(o.records) for (H, _), o in zip(probs, louts)}})
    if not args.no_cpu:
        sys.path.insert(0, str(REPO))
        import oracle  # CPU baseline only
        # the oracle evaluates the same systems with their common subexpressions shared (the
        # GPU tapes recompute some products to save registers; both give identical values)
        from hcgpu.tape import shared_lowering
        with shared_lowering():
            cpu_probs = cfg4_problem()
        cpu_s, sample = 0.0, []
        for (H, S), g in zip(cpu_probs, louts):
            m = min(len(S), 2048)
            t0 = time.perf_counter()
            o = oracle.track_batch(H, S[:m], make_options(), nthreads=args.cpu_threads)
            dt = time.perf_counter() - t0
            cpu_s += dt * len(S) / m
            sample.append(f"n={H.n}: {m} of {len(S)} paths in {dt:.1f} s")
            if not (o.code == g.code[:m]).all():
                out.setdefault("parity_vs_oracle_sample", {})[f"n{H.n}"] = "return codes differ"
        out["cpu_baseline"] = {"value": round(total / cpu_s, 1), "unit": "paths/s", "cores": args.cpu_threads,
                               "kind": "port", "build": args.oracle_build, "host": host_info(), "sample": "oracle, per size, extrapolated to the full batch: "
                                                         + "; ".join(sample)}
        out["gpu_vs_cpu"] = round(out["value"] / (total / cpu_s), 2)


def mixed_leg(args):
    """The mixed workload (configs n = 4, 6, 8, 12 in one launch) beside the main line: one
    learned and one plain launch after a warm-up each, kernel time, no CPU baseline."""
    sub = {}
    a = argparse.Namespace(**vars(args))
    a.steps, a.warmup, a.no_cpu, a.layout = 1, 1, True, "lane"
    bench_mixed(a, sub)
    return {"workload": sub["config"]["workload"], "paths": sub["config"]["paths"],
            "learned_paths_per_s": sub["value"], "learned_kernel_ms": sub["ms_per_step"],
            "plain_paths_per_s": sub["plain_gpu"]["value"], "plain_kernel_ms": sub["plain_gpu"]["kernel_ms"],
            "learned_over_plain": sub["learned"]["speedup_vs_plain"],
            "learned_steps_per_path": sub["learned"]["steps_per_path"],
            "plain_steps_per_path": sub["plain_steps_per_path"],
            "learned_acceptance_rate": sub["learned"]["acceptance_rate"],
            "identical_return_codes_learned_vs_plain": sub["identical_return_codes_learned_vs_plain"],
            "max_rel_endpoint_diff_learned_vs_plain": sub["max_rel_endpoint_diff_learned_vs_plain"],
            "roofline_frac_learned": sub["roofline"]["frac"]}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", choices=["param10", "dense4", "dense8", "mixed", "steiner", "fourbar"], default="param10")
    ap.add_argument("--retune", action="store_true", help="tune again even when the autotuner has a cached choice")
    ap.add_argument("--layout", choices=["auto", "lane", "group"], default="auto",
                    help="auto: the autotuner's choice (param10, dense4, dense8, steiner; the mixed workload takes lane); lane: one path "
                         "per lane (hipRTC-compiled tapes); group: one path per 16 lanes")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--no-certify", action="store_true", help="skip the certification leg (param10)")
    ap.add_argument("--no-training", action="store_true", help="skip the proposer-training leg (param10)")
    ap.add_argument("--no-graph", action="store_true", help="skip the HIP-graph replay leg (param10)")
    ap.add_argument("--no-mixed", action="store_true", help="skip the mixed-workload leg (param10)")
    ap.add_argument("--cpu-threads", type=int, default=len(os.sched_getaffinity(0)),
                    help="threads of the CPU baseline (default: every CPU this process may use)")
    ap.add_argument("--cpu-targets", type=int, default=192, help="targets of the CPU baseline sample (param10)")
    ap.add_argument("--copy-calibration-mib", type=int, default=0,
                    help="only run one plain device-to-device copy of this many MiB (the reference kernel for "
                         "scaling FETCH_SIZE / WRITE_SIZE counter readings) and exit")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU only: paths are independent, so a batch would shard by path index "
              "(one persistent launch per GPU), and this build has no multi-GPU path", file=sys.stderr)
        sys.exit(2)
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
    args.oracle_build = None
    if not args.no_cpu:
        _, args.oracle_build = native_oracle()
    out = {"metric": METRIC, "value": None, "unit": "paths/s", "n_gpus": 1, "steps": args.steps,
           "warmup": args.warmup, "ms_per_step": None, "higher_is_better": True, "scaling": "weak",
           "vs_baseline": None, "dtype": "f64", "data": "synthetic (seeded complex Gaussian coefficients)"}
    {"param10": bench_param10, "dense4": bench_dense, "dense8": bench_dense, "mixed": bench_mixed,
     "steiner": bench_steiner, "fourbar": bench_steiner}[args.workload](args, out)
    if args.workload == "param10" and not args.no_mixed:
        out["mixed_workload"] = mixed_leg(args)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
