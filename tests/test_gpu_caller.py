"""The caller-dynamics SQP iteration (MPXKKT_MODEL_CALLER) through the C ABI
against the oracle's sqp_step on the same caller functions.

The problem (tests/caller_problem.py) has nonlinear dynamics with a bilinear
state-control coupling evaluated on the host, a non-quadratic cost with a
cross term M != 0, and a Hessian triple that leaves out part of the cost's
curvature, as the reference's Gauss-Newton hessian_approx does.  Each
iteration: host linearisation -> mpxkkt_sqp_direction -> host evaluation of
the trial points -> mpxkkt_sqp_accept.  FP64 must equal the oracle to 1e-9
(relative) at every iteration, with the same accepted step.
"""
import numpy as np
import pytest

torch = pytest.importorskip("torch")

from caller_problem import linearization, make_caller_instance, start, trials  # noqa: E402
from oracle import pdilqr as orc  # noqa: E402

pytestmark = pytest.mark.gpu

KEYS = ("X", "U", "V", "Yeq", "Ybox")


def _rel(a, b):
    return float(np.max(np.abs(a - b)) / max(1.0, np.max(np.abs(b)))) if np.size(b) else 0.0


def _run(insts, iters, dtype, tol):
    from mpx_amd.caller import CallerBatch, CallerSolver
    i0 = insts[0]
    n, m, N, p = i0["n"], i0["m"], i0["N"], i0["p"]
    B = len(insts)
    box = i0["ulo"] is not None
    st = lambda k: np.stack([d[k] for d in insts])
    batch = CallerBatch(B, n, m, N, st("x0"), D=st("D") if p else None, E=st("E") if p else None,
                        e=st("e") if p else None, ulo=st("ulo") if box else None,
                        uhi=st("uhi") if box else None, dtype=dtype)
    solver = CallerSolver(batch)
    starts = [start(d) for d in insts]
    cur = [list(s) for s in starts]
    it = batch.iterate_from(*[np.stack([s[k] for s in starts]) for k in range(5)])
    Ltr = int(solver.params.max_trials)
    beta = float(solver.params.beta)
    for k in range(iters):
        refs = [orc.sqp_step(d, *c) for d, c in zip(insts, cur)]
        # the caller linearises at U moved inside the box, as the library does
        Uin = [r["Uin"] for r in refs]
        lins = [linearization(d, c[0], u) for d, c, u in zip(insts, cur, Uin)]
        batch.set_linearization(**{key: np.stack([l[key] for l in lins]) for key in lins[0]})
        dX, dU, amax = solver.direction(it)
        torch.cuda.synchronize()
        dXh, dUh, ah = dX.double().cpu().numpy(), dU.double().cpu().numpy(), amax.double().cpu().numpy()
        for i, r in enumerate(refs):
            assert _rel(dXh[i], r["dX"]) < tol and _rel(dUh[i], r["dU"]) < tol, (k, i)
            assert abs(ah[i] - r["amax"]) <= tol * max(1.0, r["amax"]), (k, i, ah[i], r["amax"])
        tg = np.zeros((B, Ltr, N + 1))
        tf = np.zeros((B, Ltr, N, n))
        for i, d in enumerate(insts):
            tg[i], tf[i] = trials(d, cur[i][0], Uin[i], dXh[i], dUh[i], ah[i], beta, Ltr)
        stats = solver.accept(it, tg, tf)
        torch.cuda.synchronize()
        got = {key: it[key].double().cpu().numpy() for key in KEYS}
        alpha = stats["alpha"].double().cpu().numpy()
        ntr = stats["trials"].cpu().numpy()
        for i, r in enumerate(refs):
            assert abs(alpha[i] - r["alpha"]) <= tol * max(1.0, r["alpha"]), (k, i, alpha[i], r["alpha"])
            # every candidate was evaluated by the caller: all count as examined
            assert ntr[i] == Ltr, (k, i)
            for key in KEYS:
                assert _rel(got[key][i], r[key]) < tol, (k, i, key, _rel(got[key][i], r[key]))
        cur = [[r[key] for key in KEYS] for r in refs]
    return cur


@pytest.mark.parametrize("shape", [dict(n=8, m=4, N=10), dict(n=8, m=4, N=10, p=2, box=True),
                                   dict(n=20, m=6, N=16, p=3, box=True)])
def test_caller_fp64_matches_oracle(shape):
    print(f"caller form FP64 {shape}", flush=True)
    insts = [make_caller_instance(i, **shape) for i in range(3)]
    _run(insts, 4, torch.float64, 1e-9)


def test_caller_fp32_close_to_fp64_oracle():
    print("caller form FP32", flush=True)
    insts = [make_caller_instance(i, n=12, m=4, N=12, p=2, box=True) for i in range(4)]
    _run(insts, 3, torch.float32, 2e-3)


def test_caller_entry_points_refuse_other_models():
    """sqp_iterate refuses a caller-model problem and direction refuses a device
    model family (argument errors, no device work)."""
    import ctypes as C
    from mpx_amd import _lib as L
    from mpx_amd.caller import CallerBatch, CallerSolver
    d = make_caller_instance(0, n=4, m=2, N=3)
    batch = CallerBatch(1, 4, 2, 3, d["x0"][None], dtype=torch.float64)
    solver = CallerSolver(batch)
    it = batch.iterate_from(*[s[None] for s in start(d)])
    ci = batch.c_iterate(it)
    st = L.load().mpxkkt_sqp_iterate(None, C.byref(batch.problem), C.byref(solver.params), C.byref(ci), None,
                                     C.c_void_p(solver.ws.data_ptr()), solver.ws_bytes)
    assert st == L.ERR_ARG and b"direction" in L.load().mpxkkt_last_error()
