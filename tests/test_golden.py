"""Committed FP64 fixtures (tests/golden/*.npz, written by
tests/golden/make_golden.py): the oracle still reproduces them (CPU), every
LQR solver of the oracle reproduces the dense-KKT solution stored with the LQR
fixture (CPU), and the device FP64 path reproduces the stored iterates (GPU)."""
import os

import numpy as np
import pytest

from oracle import pdilqr as orc

HERE = os.path.join(os.path.dirname(__file__), "golden")
INST_KEYS = ("A", "B", "W", "Q", "R", "QN", "xref", "uref", "D", "E", "e", "ulo", "uhi", "x0")
ITER_KEYS = ("X", "U", "V", "Yeq", "Ybox")
WLS = (1, 2, 3, 4)


def _load(wl):
    z = dict(np.load(os.path.join(HERE, f"sqp_w{wl}.npz")))
    _, n, m, N, p, iters, tanh = (int(v) for v in z["meta"])
    inst = dict(n=n, m=m, N=N, p=p, model="tanh" if tanh else "ltv",
                h=float(z["h_kappa"][0]), kappa=float(z["h_kappa"][1]))
    for k in INST_KEYS:
        inst[k] = z.get("in_" + k)
    start = tuple(z["start_" + k] for k in ITER_KEYS)
    return inst, start, iters, z


def _rel(a, b):
    return float(np.max(np.abs(a - b), initial=0.0) / max(1.0, np.max(np.abs(b), initial=0.0)))


@pytest.mark.parametrize("wl", WLS)
def test_oracle_reproduces_fixture(wl):
    inst, start, iters, z = _load(wl)
    for j, st in enumerate(orc.solve(inst, *start, iters)):
        for k in ITER_KEYS:
            assert _rel(st[k], z[f"out{j}_{k}"]) < 1e-12, (j, k)
        assert st["trials"] == int(z[f"out{j}_trials"][()])
        assert abs(st["alpha"] - float(z[f"out{j}_alpha"][()])) < 1e-14


def test_lqr_solvers_match_dense_fixture():
    z = dict(np.load(os.path.join(HERE, "lqr_w1.npz")))
    L = {k[2:]: v for k, v in z.items() if k.startswith("L_")}
    for sol in (orc.lqr_riccati(L), orc.lqr_scan(L, P=1), orc.lqr_scan(L, P=3),
                orc.lqr_scan(L, P=4, tree=True)):
        for k in ("dx", "du", "lam"):
            assert _rel(sol[k], z["sol_" + k]) < 1e-10, k


@pytest.mark.gpu
@pytest.mark.parametrize("segments", [1, 3])
@pytest.mark.parametrize("wl", WLS)
def test_device_fp64_reproduces_fixture(wl, segments):
    torch = pytest.importorskip("torch")
    from mpx_amd.problem import DeviceBatch, SQPSolver
    inst, start, iters, z = _load(wl)
    batch = DeviceBatch.from_instances([inst], dtype=torch.float64)
    solver = SQPSolver(batch, segments=segments, trial_group=1)   # one trial at a time, as the oracle counts
    it = batch.iterate_from(*[s[None] for s in start])
    for j in range(iters):
        st = solver.iterate(it)
        torch.cuda.synchronize()
        assert int(st["trials"][0]) == int(z[f"out{j}_trials"][()])
        for k in ITER_KEYS:
            got = it[k][0].cpu().numpy()
            assert _rel(got, z[f"out{j}_{k}"]) < 1e-9, (j, k)
