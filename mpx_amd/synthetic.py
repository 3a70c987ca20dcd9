"""Seeded synthetic MPC instances for the five benchmark workloads.

Every instance is a dict of float64 numpy arrays plus a few scalars; the same
dict feeds the FP64 oracle (tests) and the device packer
(:mod:`mpx_amd.problem`).  Instance ``i`` of workload ``k`` is drawn from
``numpy.random.default_rng([k, i, seed])``, so it does not depend on the batch
size it is generated in.

Shapes (one instance; N = horizon, p = equality rows per stage):

* ``model == "ltv"``: ``A`` (N, n, n), ``B`` (N, n, m) -- x' = A_t x + B_t u
* ``model == "tanh"``: ``A`` (n, n), ``B`` (n, m), ``W`` (n, n) --
  x' = x + h (A x + B u + kappa tanh(W x)), h = 0.02, kappa = 0.1
* ``Q`` (Tq, n, n) with Tq in {1, N} (1 = time-invariant), ``R`` (Tr, m, m),
  ``QN`` (n, n), ``xref`` (N+1, n), ``uref`` (N, m)
* ``D`` (N, p, n), ``E`` (N, p, m), ``e`` (N, p): rows D_t x_t + E_t u_t + e_t = 0
* ``ulo``/``uhi`` (m,) box on u, or None
* ``x0`` (n,)

Reading of the workload text where it is ambiguous (kept fixed everywhere):

* workload 1 draws Q_t, R_t afresh for every stage; the terminal weight is a
  fresh diag(U[1,10]) draw times 10.
* workloads 2-5 use time-invariant per-instance Q = diag(U[1,10]),
  R = diag(U[0.1,1]), Q_N = 10 Q.  "A, B, W seeded as above" is read so that
  the discrete Jacobians have workload 1's scale: A, W entries N(0,1)/sqrt(n),
  B entries N(0,1) (so h*B has entries N(0, 0.02^2), as B_t in workload 1).
* workload 3 has p = 8 equality rows (entries N(0,1), rhs 0) and no box;
  workload 4 has p = 12 rows and |u| <= 1, with the rows' state part D_t scaled
  by 1/sqrt(n) (with N(0,1) entries, |D_t x| ~ sqrt(n) cannot be met by |u| <= 1
  and the instance is infeasible); workloads 2 and 5 have |u| <= 1.
"""
from __future__ import annotations

import numpy as np

H_STEP = 0.02
KAPPA = 0.1

# workload id -> (n, m, N, p, box, model, batch, sqp iterations, start)
WORKLOADS = {
    1: dict(n=12, m=12, N=50, p=0, box=False, model="ltv", batch=1, iters=1, start="tile"),
    2: dict(n=24, m=12, N=100, p=0, box=True, model="tanh", batch=1024, iters=5, start="rollout"),
    3: dict(n=36, m=12, N=50, p=8, box=False, model="tanh", batch=4096, iters=3, start="rollout"),
    4: dict(n=76, m=32, N=50, p=12, box=True, model="tanh", batch=256, iters=10, start="zero"),
}

# serving mix (workload 5): (n, m, N) and probabilities
SERVING_SHAPES = [(12, 12, 25), (36, 12, 50), (76, 32, 100)]
SERVING_PROBS = [0.5, 0.35, 0.15]


def _diag_uniform(rng, k, lo, hi):
    return np.diag(rng.uniform(lo, hi, size=k))


def _block_spd(rng, n, bs=4, lo=1.0, hi=10.0):
    """Block-diagonal SPD matrix with bs x bs blocks, eigenvalues in [lo, hi]."""
    Q = np.zeros((n, n))
    for s in range(0, n, bs):
        k = min(bs, n - s)
        G = rng.standard_normal((k, k))
        Uq, _ = np.linalg.qr(G)
        lam = rng.uniform(lo, hi, size=k)
        Q[s:s + k, s:s + k] = (Uq * lam) @ Uq.T
    return 0.5 * (Q + Q.T)


def make_instance(workload, i=0, seed=0, n=None, m=None, N=None, p=None, box=None,
                  model=None):
    """One seeded instance of ``workload`` (1-5).  Keyword overrides change the
    shape while keeping the workload's distributions (used for small tests)."""
    if workload == 5:
        base = dict(WORKLOADS[2])
        base.update(p=0, box=True)
    else:
        base = dict(WORKLOADS[workload])
    for key, val in dict(n=n, m=m, N=N, p=p, box=box, model=model).items():
        if val is not None:
            base[key] = val
    n_, m_, N_, p_ = base["n"], base["m"], base["N"], base["p"]
    rng = np.random.default_rng([int(workload), int(i), int(seed)])
    inst = dict(n=n_, m=m_, N=N_, p=p_, model=base["model"], h=H_STEP, kappa=KAPPA)
    if base["model"] == "ltv":
        G = rng.standard_normal((N_, n_, n_)) / np.sqrt(n_)
        inst["A"] = np.eye(n_)[None] + H_STEP * G
        inst["B"] = rng.normal(0.0, H_STEP, size=(N_, n_, m_))
        inst["Q"] = np.stack([_diag_uniform(rng, n_, 1.0, 10.0) for _ in range(N_)])
        inst["R"] = np.stack([_diag_uniform(rng, m_, 0.1, 1.0) for _ in range(N_)])
        inst["QN"] = 10.0 * _diag_uniform(rng, n_, 1.0, 10.0)
    else:
        inst["A"] = rng.standard_normal((n_, n_)) / np.sqrt(n_)
        inst["B"] = rng.standard_normal((n_, m_))
        inst["W"] = rng.standard_normal((n_, n_)) / np.sqrt(n_)
        if workload == 4:
            Q = _block_spd(rng, n_)
        else:
            Q = _diag_uniform(rng, n_, 1.0, 10.0)
        inst["Q"] = Q[None]
        inst["R"] = _diag_uniform(rng, m_, 0.1, 1.0)[None]
        inst["QN"] = 10.0 * Q
    inst["xref"] = np.zeros((N_ + 1, n_))
    inst["uref"] = np.zeros((N_, m_))
    inst["D"] = rng.standard_normal((N_, p_, n_))
    if base["box"]:
        # rows that must hold together with |u| <= 1: state part scaled so that
        # D_t x stays O(1) and the instance stays feasible
        inst["D"] /= np.sqrt(n_)
    inst["E"] = rng.standard_normal((N_, p_, m_))
    inst["e"] = np.zeros((N_, p_))
    if base["box"]:
        inst["ulo"] = -np.ones(m_)
        inst["uhi"] = np.ones(m_)
    else:
        inst["ulo"] = None
        inst["uhi"] = None
    inst["x0"] = rng.standard_normal(n_)
    return inst


def rollout(inst, U):
    """States from x0 under controls U (N, m) with the instance's dynamics."""
    N, n = inst["N"], inst["n"]
    X = np.zeros((N + 1, n))
    X[0] = inst["x0"]
    for t in range(N):
        x, u = X[t], U[t]
        if inst["model"] == "ltv":
            X[t + 1] = inst["A"][t] @ x + inst["B"][t] @ u
        else:
            X[t + 1] = x + inst["h"] * (inst["A"] @ x + inst["B"] @ u
                                        + inst["kappa"] * np.tanh(inst["W"] @ x))
    return X


def initial_iterate(inst, start):
    """Starting (X, U, V, Y_eq, Y_box) for an instance.

    ``tile``: X_t = x0, U = 0, V = 0 (the controllers' reset state);
    ``rollout``: X = dynamics rollout of U = 0 from x0 (warm, feasible start);
    ``zero``: X = 0, U = 0, V = 0.  Multipliers of equality rows and box start at 0.
    """
    N, n, m, p = inst["N"], inst["n"], inst["m"], inst["p"]
    U = np.zeros((N, m))
    if start == "tile":
        X = np.tile(inst["x0"], (N + 1, 1))
    elif start == "rollout":
        X = rollout(inst, U)
    elif start == "zero":
        X = np.zeros((N + 1, n))
    else:
        raise ValueError(f"unknown start {start!r}")
    V = np.zeros((N + 1, n))
    Yeq = np.zeros((N, p))
    Ybox = np.zeros((N, 2 * m))
    return X, U, V, Yeq, Ybox


def make_batch(workload, B, seed=0, **shape):
    """List of ``B`` instances of one workload (same shape)."""
    return [make_instance(workload, i, seed, **shape) for i in range(B)]
