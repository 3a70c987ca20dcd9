"""Drop-in mirror of the reference solver's Python interface.

The reference imports ``mpx.primal_dual_ilqr.primal_dual_ilqr.optimizers``
(mpx/utils/mpc_wrapper.py:10) and calls

* ``optimizers.mpc(cost, dynamics, hessian_approx, limited_memory, reference,
  parameter, W, x0, X_in, U_in, V_in) -> (X, U, V)`` -- one SQP iteration
  (mpx/utils/mpc_wrapper.py:74, :150-158, :244; mpx/examples/acrobat.py:101-113),
* ``optimizers.model_evaluator_helper(cost, dynamics, x0, X, U) -> (g, c)``
  (mpx/utils/mpc_wrapper.py:422).

Here the same calls run one batched primal-dual SQP iteration of the HIP
library.  ``dynamics`` is a device model family (:mod:`mpx_amd.models`);
``cost`` is a :class:`~mpx_amd.models.TrackingCost` or None, in which case
``W`` supplies Q and R as its diagonal blocks (the reference's weight matrix);
``reference`` is the state reference (N+1, n) (or a (x_ref, u_ref) pair);
``hessian_approx`` and ``parameter`` are accepted for signature
compatibility (the Gauss-Newton Hessian of the tracking cost is exact; stage
parameters are carried by the model family).  A leading batch axis on x0 /
X_in / U_in / V_in batches the call, as ``jax.vmap`` does at
mpx/utils/mpc_wrapper_srbd.py:77.  As in the reference, outputs are fresh
arrays and numerical failure shows as NaN in U[..., 0, 0], never as an
exception.

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


def mpc(cost, dynamics, hessian_approx, limited_memory, reference, parameter, W, x0, X_in, U_in, V_in,
        constraints=None, V_eq=None, V_ineq=None, dtype=torch.float32, **tune):
    """One primal-dual SQP iteration; returns (X, U, V) (plus (V_eq, V_ineq)
    when ``constraints`` is given), as numpy arrays."""
    del hessian_approx, parameter
    insts, batched = _instances(cost, dynamics, W, reference, x0, constraints)
    batch = DeviceBatch.from_instances(insts, dtype=dtype)
    X, U, V = _as_np(X_in), _as_np(U_in), _as_np(V_in)
    if not batched:
        X, U, V = X[None], U[None], V[None]
    Ye = None if V_eq is None else (_as_np(V_eq) if batched else _as_np(V_eq)[None])
    Yb = None if V_ineq is None else (_as_np(V_ineq) if batched else _as_np(V_ineq)[None])
    it = batch.iterate_from(X, U, V, Ye, Yb)
    tune = dict(tune)
    tune.setdefault("limited_memory", 1 if limited_memory else 0)
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
