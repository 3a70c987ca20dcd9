"""FP32 device path against FP64, iteration by iteration, at full batch size.

The FP64 iterates come from the C restatement of the oracle (checked against
the numpy oracle in test_cport.py), fed its own output each iteration.  The
error model is the same C restatement compiled in FP32 arithmetic (inputs and
iterates rounded to FP32, every product and sum in float): the error that this
algorithm has in FP32 on each instance, measured, not guessed.  The device's
FP32 iterates must stay within ERR_FACTOR times that error (plus a floor of a
few FP32 ulps): two FP32 implementations that differ only in summation order
have errors of the same size, and the maximum over hundreds of instances of
their ratio stays within a small factor.

The KKT residual (2-norm of stationarity, defects, equality rows and perturbed
complementarity) of every iterate is evaluated in FP64 on the device (the FP64
residual kernel is checked against the oracle in test_gpu_parity.py).
Relative KKT eps = kkt / max(1, kkt at the start).

* Quadruped-sized workload (n=36, m=12, N=50, 8 equality rows, B=4096, three
  iterations from a rollout start): |eps_fp32 - eps_fp64| <= 1e-5 for every
  instance at every iteration (the target), and within the error model.
* Config 2 at full shape (n=24, m=12, N=100, |u| <= 1, B=1024, five
  iterations from a rollout): within the error model; |eps_fp32 - eps_fp64|
  <= 1e-5 wherever the FP32 restatement itself meets 1e-5 (at iteration 5 it
  does not: instance 983's |d eps| is about 2e-5 in FP32 arithmetic on the CPU
  too -- the problem's FP32 sensitivity, not the device's).
* Humanoid-sized workload (n=76, m=32, N=50, 12 equality rows and |u| <= 1,
  B=256, ten iterations from a zero trajectory): within the error model at
  every iteration.  Per-iteration tables go to gpurun_out/ as JSON when that
  folder exists.
"""
import json
import os

import numpy as np
import pytest

torch = pytest.importorskip("torch")

from mpx_amd import synthetic as syn  # noqa: E402
from oracle import cport  # noqa: E402

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not cport.available(), reason="C restatement of the oracle unavailable")]

KEYS = ("X", "U", "V", "Yeq", "Ybox")
NAMES = ("x", "u", "v", "y_eq", "y_box")


def _rel(a, b):
    """max over instances of |a - b|_inf / max(1, |b|_inf)."""
    B = a.shape[0]
    if a.size == 0:
        return 0.0
    d = np.abs(a - b).reshape(B, -1).max(axis=1)
    s = np.maximum(1.0, np.abs(b).reshape(B, -1).max(axis=1))
    return float(np.max(d / s))


ERR_FACTOR = 4.0
FLOOR = 1e-6


def _compare(wl, B, iters, start):
    from mpx_amd.problem import DeviceBatch, SQPSolver
    insts = [syn.make_instance(wl, i) for i in range(B)]
    starts = [syn.initial_iterate(d, start) for d in insts]
    b32 = DeviceBatch.from_instances(insts, dtype=torch.float32)
    b64 = DeviceBatch.from_instances(insts, dtype=torch.float64)
    s32 = SQPSolver(b32, segments=1)
    s64 = SQPSolver(b64, segments=1)
    it32 = b32.iterate_from(*[np.stack([s[k] for s in starts]) for k in range(5)])

    def kkt64(host):
        return s64.kkt(b64.iterate_from(*[host[k] for k in KEYS])).cpu().numpy()

    ref = {k: np.stack([s[j] for s in starts]) for j, k in enumerate(KEYS)}
    c32 = dict(ref)
    k0 = np.maximum(1.0, kkt64(ref))
    threads = min(16, os.cpu_count() or 1)
    rows = []
    for k in range(iters):
        out = cport.solve_batch(insts, [tuple(ref[key][i] for key in KEYS) for i in range(B)], 1,
                                threads=threads)
        ref = dict(zip(KEYS, out[:5]))
        out = cport.solve_batch(insts, [tuple(c32[key][i] for key in KEYS) for i in range(B)], 1,
                                threads=threads, f32=True)
        c32 = dict(zip(KEYS, out[:5]))
        s32.iterate(it32)
        torch.cuda.synchronize()
        got = {key: it32[key].double().cpu().numpy() for key in KEYS}
        e64, e32, ec = kkt64(ref) / k0, kkt64(got) / k0, kkt64(c32) / k0
        row = dict(iter=k + 1, eps64_max=float(e64.max()), eps32_max=float(e32.max()),
                   d_eps=float(np.abs(e32 - e64).max()), d_eps_cpu32=float(np.abs(ec - e64).max()))
        for key, name in zip(KEYS, NAMES):
            row[name] = _rel(got[key], ref[key])
            row[name + "_cpu32"] = _rel(c32[key], ref[key])
        rows.append(row)
    hdr = (f"{'it':>3} " + " ".join(f"{nm:>9} {'(cpu32)':>9}" for nm in NAMES) +
           f" {'eps64':>9} {'|d eps|':>9} {'(cpu32)':>9}")
    print(f"\nworkload {wl}, B={B}, FP32 device (and FP32 C restatement) vs FP64, max over instances\n" + hdr)
    for r in rows:
        print(f"{r['iter']:>3} " + " ".join(f"{r[nm]:9.2e} {r[nm + '_cpu32']:9.2e}" for nm in NAMES) +
              f" {r['eps64_max']:9.2e} {r['d_eps']:9.2e} {r['d_eps_cpu32']:9.2e}")
    if os.path.isdir("gpurun_out"):
        with open(os.path.join("gpurun_out", f"accuracy_workload{wl}.json"), "w") as f:
            json.dump(dict(workload=wl, batch=B, start=start, rows=rows), f, indent=1)
    return rows


def _within_model(rows):
    """Device FP32 error <= ERR_FACTOR x the FP32 restatement's error + FLOOR,
    for every iterate component and for eps, at every iteration."""
    for r in rows:
        for name in NAMES + ("d_eps",):
            bound = ERR_FACTOR * r[name + "_cpu32"] + FLOOR
            assert r[name] <= bound, (r["iter"], name, r[name], bound)


def test_workload3_fp32_kkt_within_1e5():
    rows = _compare(3, 4096, syn.WORKLOADS[3]["iters"], "rollout")
    for r in rows:
        assert r["d_eps"] <= 1e-5, r
    _within_model(rows)


def test_workload2_fp32_against_fp32_arithmetic():
    """Config 2 at its full shape: B=1024, n=24, m=12, N=100, |u| <= 1, five
    iterations from the rollout warm start."""
    rows = _compare(2, 1024, syn.WORKLOADS[2]["iters"], "rollout")
    _within_model(rows)
    for r in rows:
        if r["d_eps_cpu32"] <= 1e-5:
            assert r["d_eps"] <= 1e-5, r


def test_workload4_fp32_vs_fp64_per_iteration():
    """Config 4 (B=256, n=76, m=32, N=50, 12 rows, |u| <= 1, ten iterations
    from zero), iteration by iteration against the error model."""
    rows = _compare(4, 256, syn.WORKLOADS[4]["iters"], "zero")
    _within_model(rows)
