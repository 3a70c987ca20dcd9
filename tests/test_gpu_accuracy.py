"""FP32 device path against FP64, iteration by iteration, at full batch size.

The FP64 iterates come from the C restatement of the oracle (checked against
the numpy oracle in test_cport.py), fed its own output each iteration.  The
KKT residual (2-norm of stationarity, defects, equality rows and perturbed
complementarity) of both iterates is evaluated in FP64 on the device (the
FP64 residual kernel is checked against the oracle in test_gpu_parity.py).
Relative KKT eps = kkt / max(1, kkt at the start).

* Quadruped-sized workload (n=36, m=12, N=50, 8 equality rows, B=4096, three
  iterations from a rollout start): |eps_fp32 - eps_fp64| <= 1e-5 for every
  instance at every iteration.
* Humanoid-sized workload (n=76, m=32, N=50, 12 equality rows and |u| <= 1,
  B=256, ten iterations from a zero trajectory): the per-iteration relative
  errors of solution, multipliers and eps are printed and, when a
  ``gpurun_out`` folder exists, written there as JSON.  Those ten iterations
  are the pre-asymptotic phase (the box's fraction-to-boundary rule holds the
  step near 0.1-0.4 and eps is O(1-10); most instances converge between
  iterations 20 and 30), where FP32 rounding moves the iterate by about 1e-4
  relative; the bounds asserted are about three times the measured errors.
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


def _rel(a, b):
    """max over instances of |a - b|_inf / max(1, |b|_inf)."""
    B = a.shape[0]
    if a.size == 0:
        return 0.0
    d = np.abs(a - b).reshape(B, -1).max(axis=1)
    s = np.maximum(1.0, np.abs(b).reshape(B, -1).max(axis=1))
    return float(np.max(d / s))


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
    k0 = np.maximum(1.0, kkt64(ref))
    threads = min(16, os.cpu_count() or 1)
    rows = []
    for k in range(iters):
        out = cport.solve_batch(insts, [tuple(ref[key][i] for key in KEYS) for i in range(B)], 1,
                                threads=threads)
        ref = dict(zip(KEYS, out[:5]))
        s32.iterate(it32)
        torch.cuda.synchronize()
        got = {key: it32[key].double().cpu().numpy() for key in KEYS}
        e64, e32 = kkt64(ref) / k0, kkt64(got) / k0
        rows.append(dict(iter=k + 1, x=_rel(got["X"], ref["X"]), u=_rel(got["U"], ref["U"]),
                         v=_rel(got["V"], ref["V"]), y_eq=_rel(got["Yeq"], ref["Yeq"]),
                         y_box=_rel(got["Ybox"], ref["Ybox"]), eps64_max=float(e64.max()),
                         eps32_max=float(e32.max()), d_eps=float(np.abs(e32 - e64).max()),
                         d_eps_rel=float(np.max(np.abs(e32 - e64) / np.maximum(1.0, e64)))))
    hdr = (f"{'it':>3} {'X':>9} {'U':>9} {'V':>9} {'Y_eq':>9} {'Y_box':>9} {'eps64':>9} {'eps32':>9} "
           f"{'|d eps|':>9}")
    print(f"\nworkload {wl}, B={B}, FP32 vs FP64, max over instances\n" + hdr)
    for r in rows:
        print(f"{r['iter']:>3} {r['x']:9.2e} {r['u']:9.2e} {r['v']:9.2e} {r['y_eq']:9.2e} {r['y_box']:9.2e} "
              f"{r['eps64_max']:9.2e} {r['eps32_max']:9.2e} {r['d_eps']:9.2e}")
    if os.path.isdir("gpurun_out"):
        with open(os.path.join("gpurun_out", f"accuracy_workload{wl}.json"), "w") as f:
            json.dump(dict(workload=wl, batch=B, start=start, rows=rows), f, indent=1)
    return rows


def test_workload3_fp32_kkt_within_1e5():
    rows = _compare(3, 4096, syn.WORKLOADS[3]["iters"], "rollout")
    for r in rows:
        assert r["d_eps"] <= 1e-5, r


def test_workload4_fp32_vs_fp64_per_iteration():
    rows = _compare(4, 256, syn.WORKLOADS[4]["iters"], "zero")
    for r in rows:
        assert max(r["x"], r["u"], r["v"]) <= 2e-3, r
        assert max(r["y_eq"], r["y_box"]) <= 5e-3, r
        assert r["d_eps_rel"] <= 1e-3, r
