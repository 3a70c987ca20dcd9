"""The SRBD quadruped SQP iteration on the device (mpxkkt_srbd_iterate)
against the oracle: the linearisation (analytic Jacobians, gradient,
Gauss-Newton Hessian) against oracle/srbd.py, pinned by finite differences
in tests/test_srbd.py, and each iterate against the oracle's sqp_step on the
same functions.  FP64 to 1e-9; FP32 within the stated tolerance."""
import numpy as np
import pytest

torch = pytest.importorskip("torch")

import srbd_problem as sp  # noqa: E402
from oracle import pdilqr as orc  # noqa: E402
from oracle import srbd  # noqa: E402

pytestmark = pytest.mark.gpu

KEYS = ("X", "U", "V")


def _rel(a, b):
    return float(np.max(np.abs(a - b)) / max(1.0, np.max(np.abs(b))))


def _setup(B, N, perturb, dtype):
    from mpx_amd.srbd import SRBDSolver
    probs = [sp.make(i, N=N, perturb=perturb) for i in range(B)]
    W = sp.weight()
    solver = SRBDSolver(sp.MASS, sp.INERTIA, sp.DT, W, np.stack([p[1] for p in probs]),
                        np.stack([p[2] for p in probs]), np.stack([p[0] for p in probs]), dtype=dtype)
    insts = [srbd.caller_instance(sp.consts(), W, ref, par, x0, N) for x0, ref, par in probs]
    starts = [sp.start(p[0], N) for p in probs]
    it = solver.iterate_from(*[np.stack([s[k] for s in starts]) for k in range(3)])
    cur = [(s[0], s[1], s[2], np.zeros((N, 0)), np.zeros((N, 24))) for s in starts]
    return solver, insts, it, cur


@pytest.mark.parametrize("perturb", [False, True])
def test_srbd_fp64_matches_oracle(perturb):
    """perturb=False is the controller's own reset start (X = x0 tiled, U = 0,
    zero angular rate and orientation error: the mjx zero-vector conventions)."""
    print(f"srbd FP64 perturb={perturb}", flush=True)
    B, N = 3, 12
    solver, insts, it, cur = _setup(B, N, perturb, torch.float64)
    for k in range(3):
        refs = [orc.sqp_step(d, *c) for d, c in zip(insts, cur)]
        solver.iterate(it)
        torch.cuda.synchronize()
        lin = {key: v.double().cpu().numpy() for key, v in solver.lin.items() if v is not None}
        for i, (d, c, r) in enumerate(zip(insts, cur, refs)):
            L = r["lqr"]
            for t in (0, N // 2, N - 1):
                fx, fu = d["fjac"](c[0][t], c[1][t], t)
                np.testing.assert_allclose(lin["A"][i, t], fx, atol=1e-11, err_msg=f"A {k} {i} {t}")
                np.testing.assert_allclose(lin["B"][i, t], fu, atol=1e-11, err_msg=f"B {k} {i} {t}")
                Qh, Rh, Mh = d["hess"](c[0][t], c[1][t], t)
                np.testing.assert_allclose(lin["Q"][i, t], Qh, rtol=1e-12, atol=1e-9)
                np.testing.assert_allclose(lin["R"][i, t], Rh, rtol=1e-12, atol=1e-9)
                np.testing.assert_allclose(lin["M"][i, t], Mh, atol=1e-9)
                np.testing.assert_allclose(lin["q"][i, t], L["q"][t], rtol=1e-11, atol=1e-8)
                np.testing.assert_allclose(lin["r"][i, t], L["r"][t], rtol=1e-11, atol=1e-8)
                np.testing.assert_allclose(lin["fbar"][i, t], d["f"](c[0][t], c[1][t], t), atol=1e-12)
            np.testing.assert_allclose(lin["QN"][i], L["QN"], rtol=1e-12, atol=1e-9)
            got = {key: it[key][i].double().cpu().numpy() for key in KEYS}
            for key in KEYS:
                assert _rel(got[key], r[key]) < 1e-9, (k, i, key, _rel(got[key], r[key]))
        cur = [(r["X"], r["U"], r["V"], r["Yeq"], r["Ybox"]) for r in refs]


def test_srbd_fp32_close_to_fp64_oracle():
    print("srbd FP32", flush=True)
    B, N = 4, 25
    solver, insts, it, cur = _setup(B, N, True, torch.float32)
    for k in range(3):
        refs = [orc.sqp_step(d, *c) for d, c in zip(insts, cur)]
        solver.iterate(it)
        torch.cuda.synchronize()
        for i, r in enumerate(refs):
            got = {key: it[key][i].double().cpu().numpy() for key in KEYS}
            errs = {key: _rel(got[key], r[key]) for key in KEYS}
            print(k + 1, i, {key: f"{v:.2e}" for key, v in errs.items()})
            assert errs["X"] < 1e-3 and errs["U"] < 1e-2, (k, i, errs)
        cur = [(r["X"], r["U"], r["V"], r["Yeq"], r["Ybox"]) for r in refs]


def _gait_inputs(B, rng):
    x = np.zeros((B, 13))
    x[:, 0:3] = [0.0, 0.0, 0.33] + 0.05 * rng.standard_normal((B, 3))
    yaw = rng.uniform(-0.5, 0.5, B)
    x[:, 3] = np.cos(yaw / 2)
    x[:, 6] = np.sin(yaw / 2)
    x[:, 7:10] = 0.2 * rng.standard_normal((B, 3))
    foot = np.tile(sp.LEGS0, (B, 1)) + np.repeat(x[:, 0:3] * [1, 1, 0], 4, axis=0).reshape(B, 12)
    cmd = np.concatenate([rng.uniform(-0.5, 0.5, (B, 2)), np.zeros((B, 3)), rng.uniform(-0.3, 0.3, (B, 1)),
                          np.full((B, 1), 0.33)], axis=1)
    contact = rng.integers(0, 2, (B, 4)).astype(float)
    return x, foot, cmd, contact


@pytest.mark.parametrize("dtype,tol", [(torch.float64, 1e-12), (torch.float32, 2e-5)])
def test_reference_generator_matches_oracle(dtype, tol):
    """Three consecutive ticks of the gait timer and reference generator for
    B controllers (phases and liftoff points carried on the device) against
    the oracle's restatement of reference_generator_srbd."""
    from mpx_amd.srbd import SRBDGait
    rng = np.random.default_rng(3)
    B, N, duty, freq, sh, clr = 6, 25, 0.65, 1.35, 0.065, 0.4
    t0 = rng.uniform(0, 1, (B, 4))
    lo0 = np.tile(sp.LEGS0, (B, 1))
    gait = SRBDGait(B, N, sp.DT, sp.MASS, duty, freq, sh, clr, sp.LEGS0, t0, lo0, dtype=dtype)
    timer, lo = t0.copy(), lo0.copy()
    for tick in range(3):
        x, foot, cmd, contact = _gait_inputs(B, rng)
        ref_d, par_d = gait.generate(x, foot, cmd, contact, advance_dt=0.02)
        torch.cuda.synchronize()
        for b in range(B):
            _, timer[b] = srbd.timer_run(duty, freq, timer[b], 0.02)
            ref, par, lo[b], fd = srbd.reference_generator(N, sp.DT, sp.MASS, sp.LEGS0, timer[b], x[b], foot[b],
                                                           cmd[b], duty, freq, sh, lo[b], contact[b], clr)
            np.testing.assert_allclose(ref_d[b].double().cpu().numpy(), ref, atol=tol, rtol=0, err_msg=f"{tick} {b}")
            np.testing.assert_allclose(par_d[b].double().cpu().numpy(), par, atol=tol, rtol=0, err_msg=f"{tick} {b}")
            np.testing.assert_allclose(gait.foot_ref_dot[b].double().cpu().numpy(), fd, atol=10 * tol, rtol=0)
        np.testing.assert_allclose(gait.t_timer.double().cpu().numpy(), timer, atol=tol, rtol=0)
        np.testing.assert_allclose(gait.liftoff.double().cpu().numpy(), lo, atol=tol, rtol=0)


def test_srbd_controller_tick_on_device_fp64():
    """A whole SRBD controller tick with no host step: gait timer and reference
    generator, two SQP iterations, warm-start shift -- against the oracle's
    same sequence."""
    from mpx_amd.problem import shift_warmstart
    from mpx_amd.srbd import SRBDGait, SRBDSolver
    rng = np.random.default_rng(5)
    B, N = 3, 12
    x, foot, cmd, contact = _gait_inputs(B, rng)
    t0 = rng.uniform(0, 1, (B, 4))
    lo0 = np.tile(sp.LEGS0, (B, 1))
    gait = SRBDGait(B, N, sp.DT, sp.MASS, 0.65, 1.35, 0.065, 0.4, sp.LEGS0, t0, lo0, dtype=torch.float64)
    ref_d, par_d = gait.generate(x, foot, cmd, contact, advance_dt=0.02)
    W = sp.weight()
    solver = SRBDSolver(sp.MASS, sp.INERTIA, sp.DT, W, ref_d.cpu().numpy(), par_d.cpu().numpy(), x,
                        dtype=torch.float64)
    # the solver reads the generator's device arrays in place
    solver.t["reference"].copy_(ref_d)
    solver.t["parameter"].copy_(par_d)
    starts = [sp.start(x[b], N) for b in range(B)]
    it = solver.iterate_from(*[np.stack([s[k] for s in starts]) for k in range(3)])
    for _ in range(2):
        solver.iterate(it)
    shift_warmstart(solver.batch, it, 1)
    torch.cuda.synchronize()
    for b in range(B):
        timer = srbd.timer_run(0.65, 1.35, t0[b], 0.02)[1]
        ref, par, _, _ = srbd.reference_generator(N, sp.DT, sp.MASS, sp.LEGS0, timer, x[b], foot[b], cmd[b], 0.65,
                                                  1.35, 0.065, lo0[b], contact[b], 0.4)
        d = srbd.caller_instance(sp.consts(), W, ref, par, x[b], N)
        cur = (*starts[b], np.zeros((N, 0)), np.zeros((N, 24)))
        for _ in range(2):
            r = orc.sqp_step(d, *cur)
            cur = (r["X"], r["U"], r["V"], r["Yeq"], r["Ybox"])
        want = [orc.shift_warmstart(a, 1) for a in cur[:3]]
        for key, w in zip(KEYS, want):
            assert _rel(it[key][b].double().cpu().numpy(), w) < 1e-9, (b, key)
