"""Multi-iteration solve on the device with runOffline's stop rule
(mpx/utils/mpc_wrapper.py:425-467) against the oracle's ``solve_offline``.

FP64 on the device must stop every instance after the same number of
iterations as the oracle and land on the same iterate (1e-8 relative: the
iterates of many iterations accumulate rounding differences of order 1e-13
per iteration).  The oracle steps with the C restatement (equal to the numpy
oracle to 1e-11, tests/test_cport.py) so the humanoid-sized case finishes in
seconds.
"""
import numpy as np
import pytest

torch = pytest.importorskip("torch")

from mpx_amd import synthetic as syn  # noqa: E402
from oracle import cport  # noqa: E402
from oracle import pdilqr as orc  # noqa: E402

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not cport.available(), reason="C restatement of the oracle unavailable")]

KEYS = ("X", "U", "V", "Yeq", "Ybox")


def _cstep(inst, X, U, V, Yeq, Ybox):
    out = cport.solve_batch([inst], [(X, U, V, Yeq, Ybox)], 1, threads=1)
    return dict(zip(KEYS, (o[0] for o in out[:5])))


def _rel(a, b):
    if np.size(b) == 0:
        return 0.0
    return float(np.max(np.abs(a - b)) / max(1.0, np.max(np.abs(b))))


def _device_solve(insts, starts, dtype, **kw):
    from mpx_amd.problem import DeviceBatch, SQPSolver
    batch = DeviceBatch.from_instances(insts, dtype=dtype)
    solver = SQPSolver(batch)
    it = batch.iterate_from(*[np.stack([s[k] for s in starts]) for k in range(5)])
    iters, cost, cc, launched = solver.solve(it, **kw)
    torch.cuda.synchronize()
    got = {k: it[k].double().cpu().numpy() for k in KEYS}
    return iters.cpu().numpy(), cost.double().cpu().numpy(), cc.double().cpu().numpy(), launched, got


@pytest.mark.parametrize("wl,shape,start,B", [
    (3, {}, "rollout", 4),                                  # n=36, m=12, N=50, p=8 (full shape)
    (4, dict(n=24, m=8, N=20, p=6), "zero", 3),             # config 4's structure: rows and |u| <= 1
    (2, dict(n=8, m=4, N=12), "rollout", 6),
])
@pytest.mark.parametrize("check_every", [0, 3])
def test_fp64_stops_with_oracle(wl, shape, start, B, check_every):
    print(f"solve: workload {wl} {shape} B={B} check_every={check_every}", flush=True)
    insts = [syn.make_instance(wl, i, **shape) for i in range(B)]
    starts = [syn.initial_iterate(d, start) for d in insts]
    refs = [orc.solve_offline(d, *s, step=_cstep) for d, s in zip(insts, starts)]
    iters, cost, cc, launched, got = _device_solve(insts, starts, torch.float64, check_every=check_every)
    want = np.array([r["iters"] for r in refs])
    print("iterations: oracle", want.tolist(), "device", iters.tolist(), "enqueued", launched)
    assert iters.tolist() == want.tolist()
    if check_every:
        assert launched <= want.max() + check_every
    else:
        assert launched == 101
    for i, r in enumerate(refs):
        for k in KEYS:
            assert _rel(got[k][i], r[k]) < 1e-8, (i, k, _rel(got[k][i], r[k]))
        assert abs(cost[i] - r["l2"]) <= 1e-9 * max(1.0, r["l2"])
        assert abs(cc[i] - r["cc"]) <= 1e-12 + 1e-6 * r["cc"]


def test_stop_rule_max_iter_and_fp32():
    """max_iter bounds the loop at max_iter + 1 iterations (runOffline's
    i > max_iter); FP32 stops within one iteration of FP64 on the quadruped
    shape."""
    insts = [syn.make_instance(4, i, n=12, m=6, N=12, p=4) for i in range(4)]
    starts = [syn.initial_iterate(d, "zero") for d in insts]
    iters, _, _, launched, _ = _device_solve(insts, starts, torch.float64, max_iter=2)
    assert iters.tolist() == [3, 3, 3, 3] and launched == 3
    insts = [syn.make_instance(3, i) for i in range(8)]
    starts = [syn.initial_iterate(d, "rollout") for d in insts]
    i64, *_ = _device_solve(insts, starts, torch.float64, check_every=2)
    i32, *_ = _device_solve(insts, starts, torch.float32, check_every=2)
    assert np.all(np.abs(i32 - i64) <= 1), (i32, i64)


def test_fp32_humanoid_shape_stops_near_fp64_oracle():
    """Config 4 at full shape (n=76, m=32, N=50, 12 rows, |u| <= 1, from zero;
    the FP64 device path does not fit this shape's LDS plan, so FP32): the
    device stops within two iterations of the FP64 oracle.  (In FP32 the sum
    of squared stage costs, O(1e4-1e5) here, carries an absolute rounding of
    about 1e-3, the size of the stop rule's own threshold, so the last
    iterations' decisions may differ.)"""
    insts = [syn.make_instance(4, i) for i in range(3)]
    starts = [syn.initial_iterate(d, "zero") for d in insts]
    want = np.array([orc.solve_offline(d, *s, step=_cstep)["iters"] for d, s in zip(insts, starts)])
    iters, _, _, launched, _ = _device_solve(insts, starts, torch.float32, check_every=4)
    print("iterations: oracle", want.tolist(), "device FP32", iters.tolist(), "enqueued", launched)
    assert np.all(np.abs(iters - want) <= 2), (iters, want)


def test_solve_graph_replays_equal_eager():
    """The whole solve (101 gated iterations, no host synchronisation) captured
    as one HIP graph gives the eager result on every replay."""
    from mpx_amd.problem import DeviceBatch, SQPSolver
    insts = [syn.make_instance(2, i, n=8, m=4, N=12) for i in range(5)]
    starts = [syn.initial_iterate(d, "rollout") for d in insts]
    batch = DeviceBatch.from_instances(insts, dtype=torch.float64)
    solver = SQPSolver(batch)
    host = [np.stack([s[k] for s in starts]) for k in range(5)]
    it = batch.iterate_from(*host)
    iters_e, *_ = solver.solve(it)
    eager = {k: it[k].clone() for k in KEYS}
    it2 = batch.iterate_from(*host)
    init = {k: it2[k].clone() for k in KEYS}
    g = solver.capture_solve(it2)
    for rep in range(3):
        for k in KEYS:
            it2[k].copy_(init[k])
        g.replay()
        torch.cuda.synchronize()
        assert torch.equal(g.iters, iters_e), rep
        for k in KEYS:
            assert torch.equal(it2[k], eager[k]), (rep, k)
