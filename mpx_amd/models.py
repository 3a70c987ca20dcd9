"""Device-evaluated model families that stand in for the reference's callables.

The reference passes JAX callables to the solver: ``dynamics(x, u, t,
parameter)`` (mpx/utils/models.py:6), ``cost(W, reference, x, u, t)``
(mpx/utils/objectives.py:15) and ``hessian_approx(W, reference, x, u, t) ->
(Q, R, M)`` (mpx/utils/objectives.py:75).  On the device the dynamics and the
tracking cost are evaluated inside the HIP kernels, so here they are described
by their data:

* :class:`TanhDynamics` -- x' = x + h (A x + B u + kappa tanh(W x)), per-instance
  A, B, W (the synthetic whole-body-sized family of the benchmarks);
* :class:`LTVDynamics` -- x' = A_t x + B_t u with caller-supplied per-stage
  Jacobians (the "caller supplies the Jacobians" case);
* :class:`TrackingCost` -- 1/2 |x - x_ref|_Q^2 + 1/2 |u - u_ref|_R^2 with
  terminal Q_N (its Gauss-Newton Hessian (Q, R, 0) is exact, which is what
  ``hessian_approx`` returns for it);
* :class:`Constraints` -- rows D_t x + E_t u + e_t = 0 and boxes on u.

Arrays may carry a leading batch axis (the ``jax.vmap`` of
mpx/utils/mpc_wrapper_srbd.py:77); shapes follow :mod:`mpx_amd.synthetic`.
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

H_STEP = 0.02
KAPPA = 0.1


@dataclass
class TanhDynamics:
    A: np.ndarray
    B: np.ndarray
    W: np.ndarray
    h: float = H_STEP
    kappa: float = KAPPA
    model: str = field(default="tanh", init=False)


@dataclass
class LTVDynamics:
    A: np.ndarray   # (N, n, n) per stage
    B: np.ndarray   # (N, n, m)
    model: str = field(default="ltv", init=False)
    h: float = H_STEP
    kappa: float = KAPPA


@dataclass
class TrackingCost:
    Q: np.ndarray           # (n, n) or (N, n, n)
    R: np.ndarray           # (m, m) or (N, m, m)
    QN: np.ndarray          # (n, n)

    @classmethod
    def from_weights(cls, W, n, terminal_scale=1.0):
        """Q, R from the diagonal blocks of a weight matrix W ((n+m) x (n+m)),
        the role W plays in mpx/utils/objectives.py:27-31."""
        W = np.asarray(W, dtype=np.float64)
        return cls(Q=W[..., :n, :n], R=W[..., n:, n:], QN=terminal_scale * W[..., :n, :n])

    def hessian_approx(self):
        """Gauss-Newton (Q, R, M) triple (exact for this cost)."""
        m = self.R.shape[-1]
        n = self.Q.shape[-1]
        return self.Q, self.R, np.zeros(self.Q.shape[:-2] + (n, m))


@dataclass
class Constraints:
    D: np.ndarray | None = None     # (N, p, n)
    E: np.ndarray | None = None     # (N, p, m)
    e: np.ndarray | None = None     # (N, p)
    u_lo: np.ndarray | None = None  # (m,)
    u_hi: np.ndarray | None = None  # (m,)


def instance_dict(dynamics, cost, x0, x_ref, u_ref, constraints=None):
    """One instance in the layout of :mod:`mpx_amd.synthetic` (numpy, float64)."""
    x0 = np.asarray(x0, dtype=np.float64)
    n = x0.shape[-1]
    x_ref = np.asarray(x_ref, dtype=np.float64)
    N = x_ref.shape[0] - 1
    R = np.asarray(cost.R, dtype=np.float64)
    m = R.shape[-1]
    d = dict(n=n, m=m, N=N, model=dynamics.model, h=float(dynamics.h), kappa=float(dynamics.kappa))
    d["A"] = np.asarray(dynamics.A, dtype=np.float64)
    d["B"] = np.asarray(dynamics.B, dtype=np.float64)
    if dynamics.model == "tanh":
        d["W"] = np.asarray(dynamics.W, dtype=np.float64)
    Q = np.asarray(cost.Q, dtype=np.float64)
    d["Q"] = Q if Q.ndim == 3 else Q[None]
    d["R"] = R if R.ndim == 3 else R[None]
    d["QN"] = np.asarray(cost.QN, dtype=np.float64)
    d["xref"] = x_ref
    d["uref"] = np.zeros((N, m)) if u_ref is None else np.asarray(u_ref, dtype=np.float64)
    c = constraints or Constraints()
    if c.D is not None:
        d["D"], d["E"] = np.asarray(c.D, np.float64), np.asarray(c.E, np.float64)
        d["e"] = np.zeros(d["D"].shape[:2]) if c.e is None else np.asarray(c.e, np.float64)
        d["p"] = d["D"].shape[1]
    else:
        d["D"], d["E"], d["e"], d["p"] = np.zeros((N, 0, n)), np.zeros((N, 0, m)), np.zeros((N, 0)), 0
    d["ulo"] = None if c.u_lo is None else np.asarray(c.u_lo, np.float64)
    d["uhi"] = None if c.u_hi is None else np.asarray(c.u_hi, np.float64)
    d["x0"] = x0
    return d
