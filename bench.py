"""Benchmark: batched primal-dual SQP iterations on one MI355X.

Workload (the north-star configuration): B = 4096 instances, n = 36, m = 12,
N = 50, 8 equality rows per stage, tanh dynamics with analytic Jacobians, FP32.
A "step" is one SQP iteration over the whole batch; steps run as MPC ticks of
3 iterations, each tick starting with the warm-start shift of the previous
solution.  Also reported: the plain path (one instance per launch sequence,
timed on a sample of instances), per-phase kernel times from HIP events on
the launch stream, the LQR kernel's roofline and the CPU baseline (the FP64
C restatement of the same iteration, on a bounded sample of instances).

usage: python bench.py [--gpus 1] [--steps 30] [--warmup 6]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)
# The serving leg runs one HIP stream per (shape bucket, staging area): one
# each for the two fast shapes, two for the slow one.  With a hardware queue
# for each of them a small shape's launch never queues behind a large shape's
# kernels (HIP's default is 4; read when HIP initialises, so set before
# anything touches the device).
if int(os.environ.get("GPU_MAX_HW_QUEUES", "4")) < 8:
    os.environ["GPU_MAX_HW_QUEUES"] = "8"

PEAK_FP32 = 157.3e12     # MI355X dense FP32 (matrix) FLOP/s
PEAK_HBM = 8.0e12        # B/s


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def make_batch(B, seed, dtype):
    from mpx_amd import synthetic as syn
    from mpx_amd.problem import DeviceBatch
    insts = [syn.make_instance(3, i, seed) for i in range(B)]
    starts = [syn.initial_iterate(d, "rollout") for d in insts]
    batch = DeviceBatch.from_instances(insts, dtype=dtype)
    it = batch.iterate_from(*[np.stack([s[k] for s in starts]) for k in range(5)])
    return insts, starts, batch, it


def traffic_per_launch(kernel, batch):
    """HBM bytes per launch of the kernel from the committed counter summary
    (profiles/traffic.json: FETCH_SIZE + WRITE_SIZE per dispatch, scaled by the
    same counters' reading for a copy of known size), or None."""
    path = os.path.join(ROOT, "profiles", "traffic.json")
    try:
        with open(path) as f:
            rec = json.load(f).get(kernel)
    except (OSError, ValueError):
        return None
    if not rec or rec.get("batch") != batch:
        return None
    return rec["bytes_per_launch"]


def cpu_baseline(insts, starts, budget_s):
    """FP64 CPU restatement timed on a bounded sample (instance-iterations/s)."""
    try:
        from oracle import cport
        if cport.available():
            return cport.bench(insts, starts, iters=3, budget_s=budget_s)
    except Exception as exc:  # the C port is optional; fall back to the numpy oracle
        log("C port unavailable:", exc)
    from oracle import pdilqr as orc
    t0, done = time.perf_counter(), 0
    for d, s in zip(insts, starts):
        orc.solve(d, *s, 3)
        done += 3
        if time.perf_counter() - t0 > budget_s:
            break
    dt = time.perf_counter() - t0
    return dict(inst_iters_per_s=done / dt, cores=1, kind="port",
                sample=f"{done} instance-iterations of the numpy FP64 oracle, 1 thread")


def config2_leg(plain_sample=16, iters=5, seed=0):
    """Config 2: B=1024, n=24, m=12, N=100, |u| <= 1, tanh dynamics, FP32, 5 SQP
    iterations from the rollout warm start; batched (one launch sequence for
    all instances) against plain (one instance per launch sequence, timed on
    a sample).  Returns SQP iterations per second over the whole batch."""
    import torch
    from mpx_amd import synthetic as syn
    from mpx_amd.problem import DeviceBatch, SQPSolver
    B = syn.WORKLOADS[2]["batch"]
    insts = [syn.make_instance(2, i, seed) for i in range(B)]
    starts = [syn.initial_iterate(d, "rollout") for d in insts]
    batch = DeviceBatch.from_instances(insts, dtype=torch.float32)
    host = [np.stack([s[k] for s in starts]) for k in range(5)]
    solver = SQPSolver(batch)
    it = batch.iterate_from(*host)
    solver.iterate(it)                          # one-time set-up outside the timing
    reps = 3
    ms = []
    for _ in range(reps):
        it = batch.iterate_from(*host)
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(iters):
            solver.iterate(it)
        e1.record()
        torch.cuda.synchronize()
        ms.append(e0.elapsed_time(e1) / iters)
    ms_iter = float(np.median(ms))
    S = min(plain_sample, B)
    singles = []
    for i in range(S):
        b1 = DeviceBatch.from_instances([insts[i]], dtype=torch.float32)
        singles.append((SQPSolver(b1), b1.iterate_from(*[s[None] for s in starts[i]])))
    for s1, it1 in singles:
        s1.iterate(it1)
    torch.cuda.synchronize()
    p0 = time.perf_counter()
    for s1, it1 in singles:
        for _ in range(iters):
            s1.iterate(it1)
    torch.cuda.synchronize()
    plain_per_inst_iter = (time.perf_counter() - p0) / (iters * S)
    batched = 1e3 / ms_iter
    plain = 1.0 / (plain_per_inst_iter * B)
    return {"workload": "B=1024 n=24 m=12 N=100 |u|<=1, tanh dynamics, 5 iterations from the rollout start",
            "ms_per_iteration": ms_iter, "batched_iters_per_s": batched, "plain_iters_per_s": plain,
            "batched_vs_plain": batched / plain,
            "plain_sample": f"{S} instances x {iters} iterations, one launch sequence each"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=30)
    ap.add_argument("--warmup", type=int, default=6)
    ap.add_argument("--batch", type=int, default=4096)
    ap.add_argument("--plain-sample", type=int, default=64)
    ap.add_argument("--cpu-budget", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--tune", default="", help="segments,block_threads,trial_group: skip the autotuner")
    ap.add_argument("--serve", type=float, default=3.0,
                    help="seconds of the mixed serving stream after the batched run (0: skip)")
    ap.add_argument("--serve-rate", type=float, default=2000.0, help="serving arrivals per second")
    ap.add_argument("--serve-sweep", default="10000,40000",
                    help="further arrival rates (comma list, '' for none), each served for --serve seconds "
                         "with a wall-clock cap; reported in serving_sweep")
    ap.add_argument("--no-config2", action="store_true", help="skip the config-2 batched-vs-plain leg")
    ap.add_argument("--calibrate-copy", type=int, default=0,
                    help="MiB: only copy a buffer of this size twice (reference for the HBM byte counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus must be 1 "
                                   "(instances are independent: shard the batch, one process per GPU)"}))
        sys.exit(2)

    import torch
    if args.calibrate_copy:
        nbytes = args.calibrate_copy << 20
        src = torch.ones(nbytes // 4, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(2):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"calibrate_copy_bytes_read": nbytes, "calibrate_copy_bytes_written": nbytes}))
        return
    from mpx_amd import _lib as L
    from mpx_amd import autotune, flops
    from mpx_amd.problem import DeviceBatch, SQPSolver

    torch.cuda.set_device(0)
    B, n, m, N, p = args.batch, 36, 12, 50, 8
    t0 = time.perf_counter()
    insts, starts, batch, it = make_batch(B, args.seed, torch.float32)
    log(f"generated B={B} instances in {time.perf_counter() - t0:.1f} s")

    # previous tick: 3 iterations from the rollout start, then shift (the warm start)
    if args.tune:
        s_, t_, g_ = (int(x) for x in args.tune.split(","))
        tune = dict(segments=s_, block_threads=t_, trial_group=g_)
    else:
        tune = autotune.tune(batch, it, iters=3, verbose=True)
    log(f"launch shape: {tune}")
    solver = SQPSolver(batch, segments=tune["segments"], block_threads=tune["block_threads"],
                       trial_group=tune["trial_group"])
    for _ in range(3):
        solver.iterate(it)

    def run(steps):
        for k in range(steps):
            if k % 3 == 0:
                solver.shift(it, 1)
            solver.iterate(it)

    run(args.warmup)
    torch.cuda.synchronize()
    L.profile_enable(True)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    w0 = time.perf_counter()
    e0.record()
    run(args.steps)
    e1.record()
    torch.cuda.synchronize()
    wall = time.perf_counter() - w0
    L.profile_enable(False)
    phase_ms, calls = L.profile_collect()
    ms_total = e0.elapsed_time(e1)
    ms_step = ms_total / args.steps
    trials = float(solver.stats_t["trials"].float().mean().item())
    status_bad = int((solver.stats_t["status"] != 0).sum().item())
    per_phase = [x / max(calls, 1) for x in phase_ms]
    log(f"timed {args.steps} steps: {ms_step:.3f} ms/step (wall {wall:.3f} s); phases ms "
        f"lin={per_phase[0]:.3f} lqr={per_phase[1]:.3f} costate={per_phase[2]:.3f} ls={per_phase[3]:.3f}; "
        f"trials={trials:.2f}")

    # roofline of the dominant kernel: the LQR recursion (phase 1 events)
    wave = tune["segments"] == 1 and n + 1 <= 48 and m <= 16 and p <= 16
    if wave:
        kname, lqr_f = "k_lqr_wave", flops.riccati_bwd_flops(n, m, N, p) * B
    elif tune["segments"] == 1:
        kname, lqr_f = "k_lqr_seq", flops.lqr_flops(n, m, N, p, 1) * B
    else:
        kname, lqr_f = "scan kernels", flops.lqr_flops(n, m, N, p, tune["segments"]) * B
    lqr_s = per_phase[1] / 1e3
    achieved = lqr_f / lqr_s
    roofline = {"bound": "mfma", "kernel": kname,
                "achieved": achieved / 1e12, "peak": PEAK_FP32 / 1e12, "unit": "TFLOP/s",
                "frac": achieved / PEAK_FP32, "traffic": traffic_per_launch(kname, B),
                "flops_per_launch": lqr_f, "ms_per_launch": per_phase[1]}

    # plain path: one instance per launch sequence, sample of instances
    S = min(args.plain_sample, B)
    singles = []
    for i in range(S):
        b1 = DeviceBatch.from_instances([insts[i]], dtype=torch.float32)
        it1 = b1.iterate_from(*[s[None] for s in starts[i]])
        singles.append((b1, it1))
    tune1 = tune if args.tune else autotune.tune(singles[0][0], singles[0][1], iters=3, verbose=True)
    solvers1 = [SQPSolver(b1, segments=tune1["segments"], block_threads=tune1["block_threads"],
                          trial_group=tune1["trial_group"]) for b1, _ in singles]
    for (b1, it1), s1 in zip(singles, solvers1):
        s1.iterate(it1)
    torch.cuda.synchronize()
    p0 = time.perf_counter()
    for (b1, it1), s1 in zip(singles, solvers1):
        for _ in range(3):
            s1.iterate(it1)
    torch.cuda.synchronize()
    plain_per_inst_iter = (time.perf_counter() - p0) / (3 * S)
    plain_iters_per_s = 1.0 / (plain_per_inst_iter * B)
    value = 1e3 / ms_step
    log(f"plain: {plain_per_inst_iter * 1e6:.1f} us per instance-iteration (tune {tune1}); "
        f"batched/plain = {value / plain_iters_per_s:.1f}x")

    config2 = None
    if not args.no_config2:
        config2 = config2_leg(seed=args.seed)
        log(f"config 2: {config2['ms_per_iteration']:.3f} ms per batched iteration, "
            f"batched/plain = {config2['batched_vs_plain']:.1f}x")

    cpu = None
    if not args.no_cpu:
        c = cpu_baseline(insts, starts, args.cpu_budget)
        cpu = {"value": c["inst_iters_per_s"] / B, "unit": "SQP iters/s at B=4096 (extrapolated)",
               "inst_iters_per_s": c["inst_iters_per_s"], "cores": c["cores"], "kind": c["kind"],
               "sample": c["sample"]}

    serving = None
    serving_sweep = []
    if args.serve > 0:
        from mpx_amd import serve_bench
        sv = serve_bench.run(args.serve_rate, args.serve, seed=args.seed, max_wall=args.serve + 20.0)
        serving = {k: sv[k] for k in ("rate", "duration", "arrivals", "served", "solves_per_s", "p50_ms",
                                      "p99_ms", "p99_ms_all_arrivals", "backlog", "turned_away",
                                      "device_bytes_held", "by_shape")}
        log(f"serving at {args.serve_rate:.0f}/s: p50 {sv['p50_ms']:.2f} ms, p99 {sv['p99_ms']:.2f} ms, "
            f"{sv['solves_per_s']:.0f} solves/s")
        for rate in [float(r) for r in args.serve_sweep.split(",") if r.strip()]:
            sw = serve_bench.run(rate, args.serve, seed=args.seed, max_wall=2 * args.serve)
            serving_sweep.append({k: sw[k] for k in ("rate", "duration", "arrivals", "served", "solves_per_s",
                                                     "p50_ms", "p99_ms", "p99_ms_all_arrivals", "backlog",
                                                     "turned_away", "device_bytes_held")})
            log(f"serving at {rate:.0f}/s: {sw['solves_per_s']:.0f} solves/s, p50 {sw['p50_ms']:.2f} ms, "
                f"p99 {sw['p99_ms']:.2f} ms, backlog {sw['backlog']}, turned away {sw['turned_away']}")

    out = {
        "metric": "SQP iters/sec at B=4096 n=36 N=50 vs plain; p50/p99 MPC tick latency",
        "value": value,
        "unit": "SQP iterations/s (each over all 4096 instances)",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32", "data": "synthetic (seeded instances of the workload, rollout warm start)",
        "config": {"workload": "batched primal-dual SQP, B=4096 n=36 m=12 N=50 p=8, tanh dynamics, "
                               "3 iterations per tick from the shifted previous solution",
                   "batch": B, "n": n, "m": m, "N": N, "p": p,
                   "segments": tune["segments"], "block_threads": tune["block_threads"],
                   "trial_group": tune["trial_group"]},
        "instance_iters_per_s": value * B,
        "plain_iters_per_s": plain_iters_per_s,
        "batched_vs_plain": value / plain_iters_per_s,
        "phase_ms": {"linearize": per_phase[0], "lqr": per_phase[1], "costates": per_phase[2],
                     "linesearch": per_phase[3]},
        "mean_trials": trials, "nonzero_status": status_bad,
        "roofline": roofline,
        "cpu_baseline": cpu,
        "config2": config2,
        "serving": serving,
        "serving_sweep": serving_sweep,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
