"""A nonlinear problem given as the caller's functions (the MPXKKT_MODEL_CALLER
form): dynamics with a bilinear state-control coupling and a non-quadratic
cost with a state-control cross term, with exact derivatives and a
Gauss-Newton-style Hessian triple (Q, R, M) that leaves out the cost's
non-quadratic curvature, as hessian_approx does in the reference
(mpx/utils/objectives.py:39-75).

    f(x, u)  = x + h (A x + B u + kappa sin(W x) + eps (P x) * (G u))
    l_t      = 1/2 |x - xr_t|_Q^2 + 1/2 |u - ur_t|_R^2 + x' M0 u + w sum log cosh x
    l_N      = 1/2 |x - xr_N|_QN^2 + w sum log cosh x
    hess     = (Q, R, M0) for t < N, (QN, 0, 0) at t = N

Test infrastructure only.
"""
from __future__ import annotations

import numpy as np

H = 0.02
KAPPA = 0.1
EPS = 0.05
WLC = 0.3


def make_caller_instance(i, n=8, m=4, N=10, p=0, box=False, seed=0, cross=0.05):
    rng = np.random.default_rng([77, int(i), int(seed)])
    A = rng.standard_normal((n, n)) / np.sqrt(n)
    B = rng.standard_normal((n, m))
    W = rng.standard_normal((n, n)) / np.sqrt(n)
    P = rng.standard_normal((n, n)) / np.sqrt(n)
    G = rng.standard_normal((n, m)) / np.sqrt(m)
    Q = np.diag(rng.uniform(1.0, 10.0, n))
    R = np.diag(rng.uniform(0.1, 1.0, m))
    QN = 10.0 * Q
    M0 = cross * rng.standard_normal((n, m))
    xr = 0.3 * rng.standard_normal((N + 1, n))
    ur = 0.1 * rng.standard_normal((N, m))
    zm = np.zeros(m)

    def f(x, u, t):
        return x + H * (A @ x + B @ u + KAPPA * np.sin(W @ x) + EPS * (P @ x) * (G @ u))

    def fjac(x, u, t):
        fx = np.eye(n) + H * (A + KAPPA * np.cos(W @ x)[:, None] * W + EPS * (G @ u)[:, None] * P)
        fu = H * (B + EPS * (P @ x)[:, None] * G)
        return fx, fu

    def cost(x, u, t):
        lc = WLC * np.sum(np.log(np.cosh(x)))
        if t == N:
            d = x - xr[N]
            return 0.5 * d @ QN @ d + lc
        d, e = x - xr[t], u - ur[t]
        return 0.5 * d @ Q @ d + 0.5 * e @ R @ e + x @ M0 @ u + lc

    def grad(x, u, t):
        if t == N:
            return QN @ (x - xr[N]) + WLC * np.tanh(x), zm.copy()
        return (Q @ (x - xr[t]) + M0 @ u + WLC * np.tanh(x), R @ (u - ur[t]) + M0.T @ x)

    def hess(x, u, t):
        if t == N:
            return QN.copy(), np.zeros((m, m)), np.zeros((n, m))
        return Q.copy(), R.copy(), M0.copy()

    inst = dict(n=n, m=m, N=N, p=p, model="caller", f=f, fjac=fjac, cost=cost, grad=grad, hess=hess,
                x0=rng.standard_normal(n))
    D = rng.standard_normal((N, p, n)) / np.sqrt(n)
    E = rng.standard_normal((N, p, m))
    inst["D"], inst["E"], inst["e"] = D, E, 0.1 * rng.standard_normal((N, p))
    inst["ulo"] = -np.ones(m) if box else None
    inst["uhi"] = np.ones(m) if box else None
    return inst


def start(inst):
    """Rollout of U = 0 from x0, zero multipliers."""
    N, n, m, p = inst["N"], inst["n"], inst["m"], inst["p"]
    X = np.zeros((N + 1, n))
    X[0] = inst["x0"]
    U = np.zeros((N, m))
    for t in range(N):
        X[t + 1] = inst["f"](X[t], U[t], t)
    return X, U, np.zeros((N + 1, n)), np.zeros((N, p)), np.zeros((N, 2 * m))


def linearization(inst, X, U):
    """The caller's arrays at (X, U): f_x, f_u, f(x_t, u_t), Q, R, M, q, r, g, QN."""
    N, n, m = inst["N"], inst["n"], inst["m"]
    out = dict(A=np.zeros((N, n, n)), B=np.zeros((N, n, m)), fbar=np.zeros((N, n)), Q=np.zeros((N, n, n)),
               R=np.zeros((N, m, m)), M=np.zeros((N, n, m)), q=np.zeros((N + 1, n)), r=np.zeros((N, m)),
               g=np.zeros((N + 1, 1)))
    for t in range(N):
        x, u = X[t], U[t]
        out["A"][t], out["B"][t] = inst["fjac"](x, u, t)
        out["fbar"][t] = inst["f"](x, u, t)
        out["Q"][t], out["R"][t], out["M"][t] = inst["hess"](x, u, t)
        out["q"][t], out["r"][t] = inst["grad"](x, u, t)
        out["g"][t, 0] = inst["cost"](x, u, t)
    zm = np.zeros(m)
    out["QN"] = inst["hess"](X[N], zm, N)[0]
    out["q"][N] = inst["grad"](X[N], zm, N)[0]
    out["g"][N, 0] = inst["cost"](X[N], zm, N)
    return out


def trials(inst, X, U, dX, dU, amax, beta, L):
    """trial_g (L, N+1) and trial_f (L, N, n) at a_j = amax beta^j."""
    N, n, m = inst["N"], inst["n"], inst["m"]
    tg = np.zeros((L, N + 1))
    tf = np.zeros((L, N, n))
    a = amax
    for j in range(L):
        x, u = X + a * dX, U + a * dU
        for t in range(N):
            tg[j, t] = inst["cost"](x[t], u[t], t)
            tf[j, t] = inst["f"](x[t], u[t], t)
        tg[j, N] = inst["cost"](x[N], np.zeros(m), N)
        a *= beta
    return tg, tf
