"""GPU parity of the HIP SQP iteration against the FP64 oracle.

FP64 on the device must reproduce the oracle's iterates to 1e-9 (same
algorithm, only rounding differs); FP32 must stay within the stated
tolerances of the FP64 oracle, iteration by iteration.
"""
import numpy as np
import pytest

torch = pytest.importorskip("torch")

from mpx_amd import synthetic as syn  # noqa: E402
from oracle import pdilqr as orc  # noqa: E402

pytestmark = pytest.mark.gpu

# FP32 tolerances (relative to max(1, |oracle|)) per SQP iteration
TOL32_X = 2e-4
TOL32_KKT = 1e-5      # |eps_gpu - eps_fp64|, eps = kkt / max(1, kkt at the start)


def _solver(insts, dtype, **tune):
    from mpx_amd.problem import DeviceBatch, SQPSolver
    batch = DeviceBatch.from_instances(insts, dtype=dtype)
    return batch, SQPSolver(batch, **tune)


def _rel(a, b):
    return float(np.max(np.abs(a - b)) / max(1.0, np.max(np.abs(b))))


def _run(wl, B, iters, dtype, start=None, shape=None, **tune):
    shape = shape or {}
    insts = [syn.make_instance(wl, i, **shape) for i in range(B)]
    start = start or syn.WORKLOADS.get(wl, syn.WORKLOADS[2])["start"]
    starts = [syn.initial_iterate(d, start) for d in insts]
    batch, solver = _solver(insts, dtype, **tune)
    it = batch.iterate_from(*[np.stack([s[k] for s in starts]) for k in range(5)])
    refs = [orc.solve(d, *s, iters) for d, s in zip(insts, starts)]
    out = []
    for k in range(iters):
        st = solver.iterate(it)
        torch.cuda.synchronize()
        g = {key: it[key].double().cpu().numpy() for key in ("X", "U", "V", "Yeq", "Ybox")}
        out.append((g, {key: v.cpu().numpy() for key, v in st.items()}))
    return insts, starts, refs, out, solver, it


@pytest.mark.parametrize("segments", [1, 3])
@pytest.mark.parametrize("wl,shape,iters", [
    (1, dict(n=6, m=3, N=9), 1),
    (2, dict(n=8, m=4, N=12), 4),
    (3, dict(n=8, m=4, N=10, p=2), 3),
    (4, dict(n=8, m=6, N=8, p=2), 6),
])
def test_fp64_matches_oracle(wl, shape, iters, segments):
    insts, starts, refs, out, _, _ = _run(wl, 3, iters, torch.float64, shape=shape, segments=segments)
    for k in range(iters):
        g, st = out[k]
        for i in range(3):
            r = refs[i][k]
            assert abs(st["alpha"][i] - r["alpha"]) < 1e-12, (k, i, st["alpha"][i], r["alpha"])
            for key in ("X", "U", "V"):
                assert _rel(g[key][i], r[key]) < 1e-9, (k, i, key, _rel(g[key][i], r[key]))
            if insts[i]["p"] > 0:
                assert _rel(g["Yeq"][i], r["Yeq"]) < 1e-9
            if insts[i].get("ulo") is not None:
                assert _rel(g["Ybox"][i], r["Ybox"]) < 1e-9


@pytest.mark.parametrize("segments,flags", [(1, 0), (1, 1), (4, 0)])
def test_fp32_workload3_shape(segments, flags):
    """n=36, m=12, N=50, p=8 (the benchmark shape), 3 iterations, FP32; flags=1
    forms the products on VALU instead of the matrix cores."""
    insts, starts, refs, out, solver, it = _run(3, 4, 3, torch.float32, segments=segments, flags=flags)
    k0 = [max(1.0, orc.kkt_residual(d, *s)) for d, s in zip(insts, starts)]
    for k in range(3):
        g, st = out[k]
        for i in range(4):
            r = refs[i][k]
            assert st["status"][i] == 0
            assert _rel(g["X"][i], r["X"]) < TOL32_X, (k, i, _rel(g["X"][i], r["X"]))
            for key in ("U", "V", "Yeq"):
                assert _rel(g[key][i], r[key]) < TOL32_X * 10, (k, i, key, _rel(g[key][i], r[key]))
            # relative KKT residual of the FP32 iterate against the oracle's (both in FP64)
            e32 = orc.kkt_residual(insts[i], *[g[key][i] for key in ("X", "U", "V", "Yeq", "Ybox")]) / k0[i]
            e64 = orc.kkt_residual(insts[i], *[r[key] for key in ("X", "U", "V", "Yeq", "Ybox")]) / k0[i]
            assert abs(e32 - e64) <= TOL32_KKT, (k, i, e32, e64)


def test_lq_exact_one_step_fp32():
    """Workload 1 (LTV, n=m=12, N=50): one step reaches the optimum."""
    insts, starts, refs, out, solver, it = _run(1, 2, 1, torch.float32)
    g, st = out[0]
    for i in range(2):
        assert st["alpha"][i] == 1.0
        assert _rel(g["X"][i], refs[i][0]["X"]) < TOL32_X
        k0 = orc.kkt_residual(insts[i], *starts[i])
        k_ref = orc.kkt_residual(insts[i], *[refs[i][0][key] for key in ("X", "U", "V", "Yeq", "Ybox")])
        k_gpu = float(solver.kkt(it)[i].cpu())
        assert abs(k_gpu - k_ref) / max(1.0, k0) < TOL32_KKT


def test_kkt_and_model_eval_fp64():
    insts = [syn.make_instance(4, i, n=8, m=6, N=8, p=2) for i in range(2)]
    starts = [syn.initial_iterate(d, "zero") for d in insts]
    batch, solver = _solver(insts, torch.float64)
    it = batch.iterate_from(*[np.stack([s[k] for s in starts]) for k in range(5)])
    for _ in range(3):
        solver.iterate(it)
    kk = solver.kkt(it).cpu().numpy()
    g, c = solver.model_eval(it["X"], it["U"])
    for i in range(2):
        host = {k: it[k][i].cpu().numpy() for k in ("X", "U", "V", "Yeq", "Ybox")}
        ref = orc.kkt_residual(insts[i], host["X"], host["U"], host["V"], host["Yeq"], host["Ybox"])
        assert abs(kk[i] - ref) < 1e-9 * max(1.0, ref)
        gr, cr = orc.model_evaluator_helper(insts[i], host["X"], host["U"])
        assert _rel(g[i].cpu().numpy(), gr) < 1e-12 and _rel(c[i].cpu().numpy(), cr) < 1e-12


def test_shift_and_nan_fallback():
    insts = [syn.make_instance(2, i, n=5, m=2, N=6) for i in range(3)]
    starts = [syn.initial_iterate(d, "rollout") for d in insts]
    batch, solver = _solver(insts, torch.float64)
    it = batch.iterate_from(*[np.stack([s[k] for s in starts]) for k in range(5)])
    solver.iterate(it)
    before = {k: it[k].cpu().numpy().copy() for k in ("X", "U", "V", "Ybox")}
    it["U"][1, 0, 0] = float("nan")
    reset = solver.shift(it, 2).cpu().numpy()
    assert list(reset) == [0, 1, 0]
    for i in (0, 2):
        for k in ("X", "U", "V", "Ybox"):
            assert np.array_equal(it[k][i].cpu().numpy(), orc.shift_warmstart(before[k][i], 2))
    assert np.all(it["X"][1].cpu().numpy() == insts[1]["x0"][None])
    assert np.all(it["U"][1].cpu().numpy() == 0) and np.all(it["V"][1].cpu().numpy() == 0)


@pytest.mark.parametrize("wl,shape,iters", [
    (1, dict(n=12, m=12, N=20), 1),            # LTV, one tile (n + 1 <= 16)
    (1, dict(n=40, m=10, N=8), 1),             # LTV, three tiles: stage-varying Q, B, R
    (2, dict(n=15, m=16, N=14), 3),            # box, n + 1 = 16 exactly, m = 16
    (3, dict(n=16, m=3, N=11, p=16), 2),       # two tiles, p = 16
    (3, dict(n=31, m=7, N=9, p=5), 2),         # n + 1 = 32
    (4, dict(n=47, m=16, N=6, p=16), 4),       # three tiles full, box and rows
])
def test_fp32_register_wave_kernel(wl, shape, iters):
    """FP32 with the register-resident wave LQR kernel (default for these
    shapes) against the FP64 oracle, and against the LDS block kernel
    (flags bit 1) on the same inputs."""
    insts, starts, refs, out, _, _ = _run(wl, 3, iters, torch.float32, shape=shape)
    _, _, _, out_blk, _, _ = _run(wl, 3, iters, torch.float32, shape=shape, flags=2)
    for k in range(iters):
        g, st = out[k]
        gb, _ = out_blk[k]
        for i in range(3):
            r = refs[i][k]
            assert st["status"][i] == 0
            for key, tol in (("X", TOL32_X), ("U", 10 * TOL32_X), ("V", 10 * TOL32_X)):
                assert _rel(g[key][i], r[key]) < tol, (k, i, key, _rel(g[key][i], r[key]))
                assert _rel(g[key][i], gb[key][i]) < tol, (k, i, key, _rel(g[key][i], gb[key][i]))


@pytest.mark.parametrize("shape", [
    dict(n=3, m=1, N=7, p=0),      # one tile: 1 matrix step in the last state tile, 1 over m, none over p
    dict(n=6, m=5, N=7, p=3),      # 2 / 2 / 1 steps
    dict(n=10, m=9, N=7, p=6),     # 3 / 3 / 2
    dict(n=20, m=13, N=6, p=11),   # two tiles: 2 / 4 / 3
    dict(n=26, m=16, N=6, p=16),   # 3 / 4 / 4
    dict(n=32, m=2, N=5, p=1),     # three tiles: 1 / 1 / 1
    dict(n=41, m=11, N=5, p=9),    # 3 / 3 / 3
])
def test_fp32_wave_kernel_step_counts(shape):
    """The wave LQR kernel skips the matrix steps past n + 1 in the last state
    tile (one instantiation per step count) and past m and p (a switch per
    section): shapes that reach every step count, against the FP64 oracle and
    the LDS block kernel."""
    insts, starts, refs, out, _, _ = _run(3, 3, 2, torch.float32, shape=shape)
    _, _, _, out_blk, _, _ = _run(3, 3, 2, torch.float32, shape=shape, flags=2)
    for k in range(2):
        g, st = out[k]
        gb, _ = out_blk[k]
        for i in range(3):
            r = refs[i][k]
            assert st["status"][i] == 0
            for key, tol in (("X", TOL32_X), ("U", 10 * TOL32_X), ("V", 10 * TOL32_X)):
                assert _rel(g[key][i], r[key]) < tol, (k, i, key, _rel(g[key][i], r[key]))
                assert _rel(g[key][i], gb[key][i]) < tol, (k, i, key, _rel(g[key][i], gb[key][i]))


@pytest.mark.parametrize("threads", [64, 256, 512])
@pytest.mark.parametrize("wl,shape", [
    (4, dict(n=40, m=20, N=6, p=6)),   # 16 < m <= 32, boxes and equality rows
    (2, dict(n=24, m=32, N=5)),        # m = 32, boxes, no equality rows
])
def test_block_kernel_block_sizes(wl, shape, threads):
    """The LDS block Riccati kernel (flags bit 1) for 16 < m <= 32, where one
    wave factors H in registers: with one wave per block (it then forms S_t in
    one pass) and with several (the others form Q + A'T1 + rho_c D'D while it
    factors H, and S_t = that - W'W follows), in FP32 and FP64 against the
    FP64 oracle."""
    for dtype, tol in ((torch.float32, TOL32_X), (torch.float64, 1e-9)):
        insts, starts, refs, out, _, _ = _run(wl, 3, 2, dtype, shape=shape, flags=2, block_threads=threads)
        for k in range(2):
            g, st = out[k]
            for i in range(3):
                r = refs[i][k]
                assert st["status"][i] == 0, (dtype, k, i)
                for key, t in (("X", tol), ("U", 10 * tol), ("V", 10 * tol)):
                    assert _rel(g[key][i], r[key]) < t, (dtype, k, i, key, _rel(g[key][i], r[key]))


@pytest.mark.parametrize("shape", [
    dict(n=12, m=5, N=9, p=3),     # one tile: offset 12 alone in step 4 of 4
    dict(n=20, m=13, N=8, p=11),   # two tiles: step 2 of 2
    dict(n=24, m=7, N=8, p=6),     # two tiles: step 3 of 3
    dict(n=32, m=9, N=7, p=4),     # three tiles: the last state tile is offset 32 alone (no step left)
    dict(n=36, m=12, N=10, p=8),   # the benchmark's tiles: step 2 of 2
])
def test_fp32_wave_kernel_homogeneous_step(shape):
    """n % 4 == 0: the wave LQR kernel leaves the homogeneous offset n out of
    its contractions over the state and adds its part on the VALU, with A~
    formed in the kernel (flags 0) and read from f_x (flags bit 3); flags bit 4
    keeps the offset on the matrix cores.  All three against the FP64 oracle
    and against each other."""
    runs = {f: _run(3, 3, 2, torch.float32, shape=shape, flags=f) for f in (0, 8, 16)}
    refs = runs[0][2]
    for k in range(2):
        for i in range(3):
            r = refs[i][k]
            for f, res in runs.items():
                g, st = res[3][k]
                assert st["status"][i] == 0, (f, k, i)
                for key, tol in (("X", TOL32_X), ("U", 10 * TOL32_X), ("V", 10 * TOL32_X)):
                    assert _rel(g[key][i], r[key]) < tol, (f, k, i, key, _rel(g[key][i], r[key]))
                    g16 = runs[16][3][k][0]
                    assert _rel(g[key][i], g16[key][i]) < tol, (f, k, i, key, _rel(g[key][i], g16[key][i]))


@pytest.mark.parametrize("dtype", ["f64", "f32"])
@pytest.mark.parametrize("group", [1, 3, 0])
@pytest.mark.parametrize("wl,shape,iters,start", [
    (2, dict(n=8, m=4, N=12), 4, "zero"),            # box, cold start: several trials
    (3, dict(n=36, m=12, N=10, p=8), 2, None),       # rows 16-byte aligned: float4 products
    (4, dict(n=60, m=9, N=6, p=66), 3, None),        # two equality-row slots per lane
    (4, dict(n=76, m=32, N=5, p=12), 3, "zero"),     # two state rows per lane (FP32: FP64 LQR LDS too large)
])
def test_stage_parallel_line_search(wl, shape, iters, start, group, dtype):
    """The stage-parallel line search (k_ls_setup / k_trials / k_ls_decide / k_ls_update,
    the default) against the single-kernel line search (flags bit 2) on the
    same inputs, and both against the oracle: the same accepted step and trial
    count, iterates equal to rounding."""
    if dtype == "f64" and shape["n"] > 64:
        pytest.skip("the FP64 sequential LQR kernel needs more LDS than a block has at n > 64")
    dt = torch.float64 if dtype == "f64" else torch.float32
    kw = dict(shape=shape, start=start, trial_group=group)
    insts, starts, refs, out, _, _ = _run(wl, 3, iters, dt, **kw)
    _, _, _, old, _, _ = _run(wl, 3, iters, dt, flags=4, **kw)
    tol = 1e-10 if dtype == "f64" else TOL32_X
    for k in range(iters):
        (g, st), (go, sto) = out[k], old[k]
        for i in range(3):
            r = refs[i][k]
            assert st["status"][i] == sto["status"][i], (k, i)
            assert st["trials"][i] == sto["trials"][i], (k, i, st["trials"][i], sto["trials"][i])
            # FP32: the two paths sum in different orders, so the fraction-to-boundary step
            # can differ in its last bits after a few cold-start iterations
            assert st["alpha"][i] == pytest.approx(sto["alpha"][i], rel=1e-6 if dtype == "f64" else 1e-4), (k, i)
            if dtype == "f64":
                assert abs(st["alpha"][i] - r["alpha"]) < 1e-12, (k, i)
            for key in ("X", "U", "V", "Yeq", "Ybox"):
                if g[key].size == 0:
                    continue
                assert _rel(g[key][i], go[key][i]) < 10 * tol, (k, i, key, _rel(g[key][i], go[key][i]))
                if dtype == "f64" or key == "X":
                    assert _rel(g[key][i], r[key]) < 10 * tol, (k, i, key, _rel(g[key][i], r[key]))


@pytest.mark.parametrize("group", [1, 0])
@pytest.mark.parametrize("wl,shape,iters,start", [
    (2, dict(n=8, m=4, N=12), 4, "zero"),              # box, one tile, cold start: several trials
    (2, dict(n=20, m=13, N=37), 3, "zero"),            # box, two tiles, 3 waves of 16 stages (the last partial)
    (3, dict(n=36, m=12, N=50, p=8), 2, None),         # the benchmark shape, equality rows
    (3, dict(n=47, m=16, N=17, p=16), 2, "zero"),      # three full tiles, 16 rows, N = 16 + 1
    (4, dict(n=12, m=8, N=21, p=4), 3, None),          # boxes and equality rows with 128-bit loads
])
def test_trials_matrix_core_form(wl, shape, iters, start, group):
    """FP32 linearisation and trials on the matrix cores (k_linearize_mma and
    k_trials_mma, the default for these shapes: 16 stages per wave, products
    of the stage-invariant matrices with the stages' vectors as columns)
    against the row-per-lane k_linearize_inst and k_trials (flags bit 4) and
    the oracle: the same accepted step and trial count, iterates (with the box
    projection and multiplier resets of the linearisation) equal to rounding."""
    kw = dict(shape=shape, start=start, trial_group=group)
    insts, starts, refs, out, _, _ = _run(wl, 3, iters, torch.float32, **kw)
    _, _, _, rows, _, _ = _run(wl, 3, iters, torch.float32, flags=16, **kw)
    for k in range(iters):
        (g, st), (gr, sr) = out[k], rows[k]
        for i in range(3):
            assert st["status"][i] == sr["status"][i], (k, i)
            assert st["trials"][i] == sr["trials"][i], (k, i, st["trials"][i], sr["trials"][i])
            # two FP32 linearisations: the steps (and with boxes the
            # fraction-to-boundary bound on a) agree to rounding, not bit for bit
            assert st["alpha"][i] == pytest.approx(sr["alpha"][i], rel=TOL32_X), (k, i)
            for key in ("X", "U", "V", "Yeq", "Ybox"):
                if g[key].size == 0:
                    continue
                assert _rel(g[key][i], gr[key][i]) < 10 * TOL32_X, (k, i, key, _rel(g[key][i], gr[key][i]))
            assert _rel(g["X"][i], refs[i][k]["X"]) < 10 * TOL32_X, (k, i)


@pytest.mark.parametrize("wl,shape,iters,start", [
    (2, dict(n=8, m=4, N=12), 4, "zero"),
    (3, dict(n=36, m=12, N=10, p=8), 2, None),
    (4, dict(n=60, m=9, N=6, p=64), 3, None),
])
def test_stage_parallel_linearization(wl, shape, iters, start):
    """flags bit 3 (k_linearize_stage, one wave per stage) against the oracle
    in FP64 and against the instance-per-block linearisation in FP32."""
    insts, starts, refs, out, _, _ = _run(wl, 3, iters, torch.float64, shape=shape, start=start, flags=8)
    for k in range(iters):
        g, st = out[k]
        for i in range(3):
            r = refs[i][k]
            assert abs(st["alpha"][i] - r["alpha"]) < 1e-12, (k, i)
            for key in ("X", "U", "V"):
                assert _rel(g[key][i], r[key]) < 1e-9, (k, i, key)
    _, _, _, a, _, _ = _run(wl, 3, iters, torch.float32, shape=shape, start=start, flags=8)
    _, _, _, b, _, _ = _run(wl, 3, iters, torch.float32, shape=shape, start=start)
    for k in range(iters):
        for key in ("X", "U", "V"):
            assert _rel(a[k][0][key], b[k][0][key]) < TOL32_X, (k, key)


def _cond_instances(B, cond, n=12, m=12, N=20):
    """LTV instances (workload 1's distributions) whose control cost R_t has
    eigenvalues log-spaced over [1, cond] in a random basis, so that
    H_t = R_t + B_t' S_{t+1} B_t (B_t entries ~ 0.02) has about that condition."""
    insts = []
    for i in range(B):
        d = syn.make_instance(1, 60 + i, n=n, m=m, N=N)
        rng = np.random.default_rng([17, i])
        Uq, _ = np.linalg.qr(rng.standard_normal((m, m)))
        R = (Uq * np.logspace(0.0, np.log10(cond), m)) @ Uq.T
        d["R"] = np.broadcast_to(0.5 * (R + R.T), (N, m, m)).copy()
        insts.append(d)
    return insts


def test_fp32_wave_kernel_ill_conditioned_H():
    """The register Cholesky of the wave LQR kernel at cond(H_t) >= 5e5 against
    the FP64 oracle: the FP32 error of the step (X, U after one SQP iteration
    of the LQ problem) stays within 4x the error of the same algorithm in FP32
    arithmetic (the C restatement built with FP32 reals, tests/test_cport.py),
    plus a few FP32 ulps."""
    from oracle import cport
    insts = _cond_instances(4, 1e6)
    starts = [syn.initial_iterate(d, "tile") for d in insts]
    # the condition of H_t the kernel meets
    conds = []
    for d, s in zip(insts, starts):
        st = orc.sqp_step(d, *s)
        L, sol = st["lqr"], orc.lqr_riccati(st["lqr"])
        for t in range(d["N"]):
            H = L["R"][t] + L["B"][t].T @ sol["S"][t + 1] @ L["B"][t]
            conds.append(np.linalg.cond(H))
    assert max(conds) >= 5e5, max(conds)
    batch, solver = _solver(insts, torch.float32)
    it = batch.iterate_from(*[np.stack([s[k] for s in starts]) for k in range(5)])
    solver.iterate(it)
    torch.cuda.synchronize()
    refs = [orc.sqp_step(d, *s) for d, s in zip(insts, starts)]
    c32 = cport.solve_batch(insts, starts, 1, f32=True)
    for j, k in enumerate(("X", "U")):
        want = np.stack([r[k] for r in refs])
        got = it[k].double().cpu().numpy()
        err = max(_rel(got[i], want[i]) for i in range(len(insts)))
        err32 = max(_rel(c32[j][i], want[i]) for i in range(len(insts)))
        print(f"cond(H) max {max(conds):.2e}: {k} device FP32 error {err:.2e}, FP32 restatement {err32:.2e}")
        assert err <= 4.0 * err32 + 1e-6, (k, err, err32)


def _indefinite_instances(B, n=10, m=6, N=12):
    """Tanh-model instances whose control Hessian R_t has a strongly negative
    eigenvalue, so that H_t = R_t + B'S B is indefinite at some stage and the
    modified pivots of the Cholesky factorisation (D2) decide the step."""
    insts = []
    for i in range(B):
        d = syn.make_instance(2, 80 + i, n=n, m=m, N=N, box=False)
        R = d["R"][0].copy()
        R[1, 1] = -5.0
        d["R"] = R[None]
        insts.append(d)
    return insts


@pytest.mark.parametrize("dtype", ["f64", "f32"])
def test_indefinite_stage_hessian(dtype):
    """Indefinite H_t: the device follows the oracle's modified Cholesky
    (hess_floor = 1e-2) -- FP64 to 1e-9, FP32 (register wave kernel) to the
    FP32 tolerance -- and the iterate stays finite."""
    insts = _indefinite_instances(3)
    starts = [syn.initial_iterate(d, "rollout") for d in insts]
    L0 = orc.sqp_step(insts[0], *starts[0], params=dict(hess_floor=1e-2))["lqr"]
    sol = orc.lqr_riccati(L0)
    mins = [np.linalg.eigvalsh(L0["R"][t] + L0["B"][t].T @ sol["S"][t + 1] @ L0["B"][t]).min()
            for t in range(insts[0]["N"])]
    assert min(mins) < 0, "the instance must make H_t indefinite"
    dt = torch.float64 if dtype == "f64" else torch.float32
    batch, solver = _solver(insts, dt, hess_floor=1e-2)
    it = batch.iterate_from(*[np.stack([s[k] for s in starts]) for k in range(5)])
    refs = [orc.solve(d, *s, 2, params=dict(hess_floor=1e-2)) for d, s in zip(insts, starts)]
    tol = 1e-9 if dtype == "f64" else TOL32_X
    for k in range(2):
        solver.iterate(it)
        torch.cuda.synchronize()
        for i in range(3):
            for key in ("X", "U", "V"):
                got = it[key][i].double().cpu().numpy()
                assert np.all(np.isfinite(got))
                assert _rel(got, refs[i][k][key]) < (1 if key == "X" else 10) * tol, (k, i, key)
