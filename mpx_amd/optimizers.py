"""Drop-in mirror of the reference solver's Python interface.

The reference imports ``mpx.primal_dual_ilqr.primal_dual_ilqr.optimizers``
(mpx/utils/mpc_wrapper.py:10) and calls

* ``optimizers.mpc(cost, dynamics, hessian_approx, limited_memory, reference,
  parameter, W, x0, X_in, U_in, V_in) -> (X, U, V)`` -- one SQP iteration
  (mpx/utils/mpc_wrapper.py:74, :150-158, :244; mpx/examples/acrobat.py:101-113),
* ``optimizers.model_evaluator_helper(cost, dynamics, x0, X, U) -> (g, c)``
  (mpx/utils/mpc_wrapper.py:422).

Here the same calls run one batched primal-dual SQP iteration of the HIP
library, in one of two forms:

* ``dynamics`` a device model family (:mod:`mpx_amd.models`): dynamics and
  cost are evaluated inside the kernels.  ``cost`` is a
  :class:`~mpx_amd.models.TrackingCost` or None, in which case ``W`` supplies
  Q and R as its diagonal blocks (the reference's weight matrix);
  ``reference`` is the state reference (N+1, n) (or a (x_ref, u_ref) pair).
  The tracking cost's Gauss-Newton Hessian is exact, so ``hessian_approx``
  must be None or the cost's own; stage parameters are carried by the model
  family, so ``parameter`` must be None.
* ``dynamics`` a plain callable ``dynamics(x, u, t, parameter)`` with
  ``cost(W, reference, x, u, t)`` and ``hessian_approx(W, reference, x, u,
  t) -> (Q, R, M)`` -- the reference's callables, unchanged
  (mpx/utils/models.py:6, mpx/utils/objectives.py:15, :75).  They are
  evaluated on the host (numpy); the iteration runs on the device in the
  caller-dynamics form (:mod:`mpx_amd.caller`).  The reference takes f_x, f_u
  and the cost gradient from JAX; here pass ``dynamics_jac(x, u, t,
  parameter) -> (f_x, f_u)`` and ``cost_grad(W, reference, x, u, t) -> (q,
  r)``, or leave them out for central differences (step 1e-6 (1 + |z_i|)).

A leading batch axis on x0 / X_in / U_in / V_in batches the call, as
``jax.vmap`` does at mpx/utils/mpc_wrapper_srbd.py:77 (in the callable form
``W``, ``reference`` and ``parameter`` then carry the batch axis too).  As in
the reference, outputs are fresh arrays and numerical failure shows as NaN in
U[..., 0, 0], never as an exception.

Optional keyword ``constraints`` adds equality rows and boxes, whose
multipliers ``V_eq`` (N, p) and ``V_ineq`` (N, 2m) are then returned as well
(the arrays mpx/examples/acrobat.py:96-97 prepares).
"""
from __future__ import annotations

import numpy as np
import torch

from .models import TrackingCost, instance_dict
from .problem import DeviceBatch, SQPSolver

def _as_np(a):
    if isinstance(a, torch.Tensor):
        return a.detach().cpu().double().numpy()
    return np.asarray(a, dtype=np.float64)


def _pick(a, b, base_ndim, batched):
    """Instance b of a model array: arrays with one axis more than their
    per-instance rank carry a leading batch axis; others are shared."""
    a = np.asarray(a, dtype=np.float64)
    return a[b] if (batched and a.ndim == base_ndim + 1) else a


def _instances(cost, dynamics, W, reference, x0, constraints):
    x0 = _as_np(x0)
    batched = x0.ndim == 2
    B = x0.shape[0] if batched else 1
    n = x0.shape[-1]
    if isinstance(reference, (tuple, list)):
        x_ref, u_ref = _as_np(reference[0]), _as_np(reference[1])
    else:
        x_ref, u_ref = _as_np(reference), None
    if cost is None:
        cost = TrackingCost.from_weights(_as_np(W), n)
    ltv = dynamics.model == "ltv"
    keys = ("A", "B") + (() if ltv else ("W",))
    insts = []
    for b in range(B):
        dyn = type(dynamics)(**{k: _pick(getattr(dynamics, k), b, 3 if ltv else 2, batched) for k in keys})
        dyn.h, dyn.kappa = dynamics.h, dynamics.kappa
        # Q, R: (n, n) shared, (B, n, n) per instance, (B, N, n, n) per instance and stage;
        # unbatched (N, n, n) is per stage
        Qb = np.asarray(cost.Q, np.float64)
        Rb = np.asarray(cost.R, np.float64)
        Q = Qb[b] if (batched and Qb.ndim >= 3) else Qb
        R = Rb[b] if (batched and Rb.ndim >= 3) else Rb
        cst = TrackingCost(Q=Q, R=R, QN=_pick(cost.QN, b, 2, batched))
        xr = _pick(x_ref, b, 2, batched)
        ur = None if u_ref is None else _pick(u_ref, b, 2, batched)
        insts.append(instance_dict(dyn, cst, x0[b] if batched else x0, xr, ur, constraints))
    return insts, batched


def _central(fn, z, h=1e-6):
    """Central-difference Jacobian of fn at z (columns: d fn / d z_j)."""
    z = np.asarray(z, dtype=np.float64)
    cols = []
    for j in range(z.size):
        e = np.zeros_like(z)
        e[j] = h * (1.0 + abs(z[j]))
        cols.append((np.asarray(fn(z + e), np.float64) - np.asarray(fn(z - e), np.float64)) / (2.0 * e[j]))
    return np.stack(cols, axis=-1)


def _callable_iteration(cost, dynamics, hessian_approx, reference, parameter, W, x0, X_in, U_in, V_in,
                        constraints, V_eq, V_ineq, dtype, dynamics_jac, cost_grad, tune):
    from .caller import CallerBatch, CallerSolver
    x0 = _as_np(x0)
    batched = x0.ndim == 2
    X, U, V = _as_np(X_in), _as_np(U_in), _as_np(V_in)
    if not batched:
        x0, X, U, V = x0[None], X[None], U[None], V[None]
    B, N1, n = X.shape
    N, m = U.shape[1], U.shape[2]
    per = lambda a, b: (a[b] if (batched and a is not None and np.ndim(a) > 0 and np.shape(a)[0] == B
                                 and not np.isscalar(a)) else a)
    c = constraints
    D = None if c is None or c.D is None else np.broadcast_to(np.asarray(c.D, np.float64), (B,) + np.shape(c.D)[-3:])
    E = None if D is None else np.broadcast_to(np.asarray(c.E, np.float64), (B,) + np.shape(c.E)[-3:])
    e = None if D is None else (np.zeros(D.shape[:3]) if c.e is None else
                                np.broadcast_to(np.asarray(c.e, np.float64), D.shape[:3]))
    box = c is not None and c.u_lo is not None
    ulo = np.broadcast_to(np.asarray(c.u_lo, np.float64), (B, m)) if box else None
    uhi = np.broadcast_to(np.asarray(c.u_hi, np.float64), (B, m)) if box else None
    if box:
        # linearise where the library iterates: U moved inside the box (include/mpx_kkt.h)
        w = 1e-3 * (uhi - ulo)
        U = np.where(U >= uhi[:, None], (uhi - w)[:, None], U)
        U = np.where(U <= ulo[:, None], (ulo + w)[:, None], U)
    batch = CallerBatch(B, n, m, N, x0, D=D, E=E, e=e, ulo=ulo, uhi=uhi, dtype=dtype)
    tune = dict(tune)
    solver = CallerSolver(batch, **tune)
    lin = dict(A=np.zeros((B, N, n, n)), B=np.zeros((B, N, n, m)), fbar=np.zeros((B, N, n)),
               Q=np.zeros((B, N, n, n)), R=np.zeros((B, N, m, m)), M=np.zeros((B, N, n, m)),
               q=np.zeros((B, N + 1, n)), r=np.zeros((B, N, m)), g=np.zeros((B, N + 1, 1)),
               QN=np.zeros((B, n, n)))
    zm = np.zeros(m)
    fns = []
    for b in range(B):
        Wb, refb, parb = per(W, b), per(reference, b), per(parameter, b)
        f = (lambda Wb_, parb_: lambda x, u, t: np.asarray(dynamics(x, u, t, parb_), np.float64))(Wb, parb)
        l = (lambda Wb_, refb_: lambda x, u, t: float(cost(Wb_, refb_, x, u, t)))(Wb, refb)
        fns.append((f, l))
        for t in range(N1):
            x = X[b, t]
            u = U[b, t] if t < N else zm
            l_t = lambda x_, u_, t_=t: l(x_, u_, t_)
            if cost_grad is not None:
                q, r = cost_grad(Wb, refb, x, u, t)
            else:
                q = _central(lambda z: l_t(z, u), x)
                r = _central(lambda z: l_t(x, z), u)
            Qh, Rh, Mh = (np.asarray(a, np.float64) for a in hessian_approx(Wb, refb, x, u, t))
            lin["q"][b, t], lin["g"][b, t, 0] = q, l_t(x, u)
            if t == N:
                lin["QN"][b] = Qh
                continue
            lin["r"][b, t], lin["Q"][b, t], lin["R"][b, t], lin["M"][b, t] = r, Qh, Rh, Mh
            lin["fbar"][b, t] = f(x, u, t)
            if dynamics_jac is not None:
                fx, fu = dynamics_jac(x, u, t, parb)
            else:
                fx = _central(lambda z: f(z, u, t), x)
                fu = _central(lambda z: f(x, z, t), u)
            lin["A"][b, t], lin["B"][b, t] = fx, fu
    batch.set_linearization(**lin)
    Ye = None if V_eq is None else (_as_np(V_eq) if batched else _as_np(V_eq)[None])
    Yb = None if V_ineq is None else (_as_np(V_ineq) if batched else _as_np(V_ineq)[None])
    it = batch.iterate_from(X, U, V, Ye, Yb)
    dX, dU, _ = solver.direction(it)
    steps = solver.trial_steps().double().cpu().numpy()
    dX, dU = dX.double().cpu().numpy(), dU.double().cpu().numpy()
    Ltr = steps.shape[1]
    tg = np.zeros((B, Ltr, N1))
    tf = np.zeros((B, Ltr, N, n))
    for b, (f, l) in enumerate(fns):
        for j in range(Ltr):
            xs, us = X[b] + steps[b, j] * dX[b], U[b] + steps[b, j] * dU[b]
            for t in range(N):
                tg[b, j, t] = l(xs[t], us[t], t)
                tf[b, j, t] = f(xs[t], us[t], t)
            tg[b, j, N] = l(xs[N], zm, N)
    solver.accept(it, tg, tf)
    out = [it[k].double().cpu().numpy() for k in ("X", "U", "V", "Yeq", "Ybox")]
    if not batched:
        out = [o[0] for o in out]
    if constraints is not None:
        return out[0], out[1], out[2], out[3], out[4]
    return out[0], out[1], out[2]


def mpc(cost, dynamics, hessian_approx, limited_memory, reference, parameter, W, x0, X_in, U_in, V_in,
        constraints=None, V_eq=None, V_ineq=None, dtype=torch.float32, dynamics_jac=None, cost_grad=None,
        **tune):
    """One primal-dual SQP iteration; returns (X, U, V) (plus (V_eq, V_ineq)
    when ``constraints`` is given), as numpy arrays."""
    tune = dict(tune)
    tune.setdefault("limited_memory", 1 if limited_memory else 0)
    if callable(dynamics) and not hasattr(dynamics, "model"):
        if hessian_approx is None or not callable(cost):
            raise ValueError("callable dynamics need a callable cost and hessian_approx "
                             "(the reference's objectives.py:15 / :75 contracts)")
        return _callable_iteration(cost, dynamics, hessian_approx, reference, parameter, W, x0, X_in, U_in,
                                   V_in, constraints, V_eq, V_ineq, dtype, dynamics_jac, cost_grad, tune)
    if parameter is not None:
        raise ValueError("stage parameters belong to the model family here; pass parameter=None "
                         "or use callable dynamics(x, u, t, parameter)")
    if hessian_approx is not None and not (cost is not None and getattr(hessian_approx, "__self__", None) is cost):
        raise ValueError("the device model families use the tracking cost's exact Gauss-Newton Hessian; "
                         "pass hessian_approx=None (or cost.hessian_approx), or use callable dynamics and cost "
                         "to supply your own Hessian")
    insts, batched = _instances(cost, dynamics, W, reference, x0, constraints)
    batch = DeviceBatch.from_instances(insts, dtype=dtype)
    X, U, V = _as_np(X_in), _as_np(U_in), _as_np(V_in)
    if not batched:
        X, U, V = X[None], U[None], V[None]
    Ye = None if V_eq is None else (_as_np(V_eq) if batched else _as_np(V_eq)[None])
    Yb = None if V_ineq is None else (_as_np(V_ineq) if batched else _as_np(V_ineq)[None])
    it = batch.iterate_from(X, U, V, Ye, Yb)
    solver = SQPSolver(batch, **tune)
    solver.iterate(it)
    out = [it[k].double().cpu().numpy() for k in ("X", "U", "V", "Yeq", "Ybox")]
    if not batched:
        out = [o[0] for o in out]
    if constraints is not None:
        return out[0], out[1], out[2], out[3], out[4]
    return out[0], out[1], out[2]


def model_evaluator_helper(cost, dynamics, x0, X, U, W=None, reference=None, dtype=torch.float64):
    """(g, c): per-stage cost (N+1,) and constraint rows (N+1, n) =
    [x0 - X_0, f(x_t, u_t) - x_{t+1}] on the device."""
    Xn = _as_np(X)
    batched = Xn.ndim == 3
    if reference is None:
        N1, n = Xn.shape[-2], Xn.shape[-1]
        reference = np.zeros((Xn.shape[0], N1, n) if batched else (N1, n))
    insts, _ = _instances(cost, dynamics, W, reference, x0, None)
    batch = DeviceBatch.from_instances(insts, dtype=dtype)
    solver = SQPSolver(batch)
    Xd = torch.as_tensor(Xn if batched else Xn[None], dtype=dtype, device=batch.device).contiguous()
    Un = _as_np(U)
    Ud = torch.as_tensor(Un if batched else Un[None], dtype=dtype, device=batch.device).contiguous()
    g, c = solver.model_eval(Xd, Ud)
    g, c = g.cpu().numpy(), c.cpu().numpy()
    return (g, c) if batched else (g[0], c[0])
