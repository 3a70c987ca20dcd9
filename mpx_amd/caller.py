"""SQP iterations on the caller's own dynamics and cost (MPXKKT_MODEL_CALLER).

The reference solver takes ``dynamics(x, u, t, parameter)`` and
``cost(W, reference, x, u, t)`` as callables (mpx/utils/models.py:6,
mpx/utils/objectives.py:15, bound at mpx/utils/mpc_wrapper.py:66-74),
differentiates them with JAX and calls ``hessian_approx(W, reference, x, u,
t) -> (Q, R, M)`` (mpx/utils/objectives.py:75, :202).  Here the caller
evaluates the same quantities and the library does the rest of the iteration
on the device:

1. the caller fills the linearisation tensors of :class:`CallerBatch` at the
   current iterate (f_x, f_u, f(x_t, u_t), Q, R, M, q, r, l_t, QN);
2. :meth:`CallerSolver.direction` runs the augmentation, the LQR and the
   line-search set-up and returns the step dX, dU and the first trial step
   a_max (trial j is at a_max beta^j);
3. the caller evaluates its cost and dynamics at the trial points and
   :meth:`CallerSolver.accept` picks the step and updates the iterate.

PyTorch provides device memory and the stream; the arithmetic is in the HIP
library.
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import torch

from . import _lib as L
from .problem import _ptr, _stream, _view


class CallerBatch:
    """B same-shaped instances whose dynamics and cost are the caller's.

    Holds the device tensors the caller refills every iteration (``lin``) and
    the library-side data: equality rows D_t x + E_t u + e_t = 0 (``D`` (B, N,
    p, n), ``E`` (B, N, p, m), ``e`` (B, N, p)), boxes ``ulo``/``uhi`` (B, m)
    and the initial states ``x0`` (B, n).
    """

    def __init__(self, B, n, m, N, x0, D=None, E=None, e=None, ulo=None, uhi=None, cross=True,
                 dtype=torch.float32, device="cuda"):
        self.B, self.n, self.m, self.N = B, n, m, N
        self.p = 0 if D is None else int(np.shape(D)[2])
        self.box = ulo is not None
        self.dtype = dtype
        self.device = torch.device(device)
        z = lambda *shape: torch.zeros(shape, dtype=dtype, device=self.device)
        dev = lambda a: torch.as_tensor(np.ascontiguousarray(a), dtype=dtype, device=self.device)
        self.lin = dict(A=z(B, N, n, n), B=z(B, N, n, m), fbar=z(B, N, n), Q=z(B, N, n, n), R=z(B, N, m, m),
                        M=z(B, N, n, m) if cross else None, q=z(B, N + 1, n), r=z(B, N, m), g=z(B, N + 1, 1),
                        QN=z(B, n, n))
        self.t = dict(x0=dev(x0), D=None if D is None else dev(D), E=None if E is None else dev(E),
                      e=None if e is None else dev(e), ulo=None if ulo is None else dev(ulo),
                      uhi=None if uhi is None else dev(uhi))
        self._build()

    def _build(self):
        n, m, N, p = self.n, self.m, self.N, self.p
        lin, t = self.lin, self.t
        pr = L.Problem()
        pr.dtype = L.F64 if self.dtype == torch.float64 else L.F32
        pr.model = L.MODEL_CALLER
        pr.batch, pr.n, pr.m, pr.horizon, pr.n_eq = self.B, n, m, N, p
        pr.has_box = 1 if self.box else 0
        pr.A = _view(lin["A"], N * n * n, n * n)
        pr.B = _view(lin["B"], N * n * m, n * m)
        pr.W = pr.xref = pr.uref = _view(None, 0, 0)
        pr.Q = _view(lin["Q"], N * n * n, n * n)
        pr.R = _view(lin["R"], N * m * m, m * m)
        pr.QN = _view(lin["QN"], n * n, 0)
        pr.fbar = _view(lin["fbar"], N * n, n)
        pr.M = _view(lin["M"], N * n * m, n * m) if lin["M"] is not None else _view(None, 0, 0)
        pr.q = _view(lin["q"], (N + 1) * n, n)
        pr.r = _view(lin["r"], N * m, m)
        pr.g = _view(lin["g"], N + 1, 1)
        if p > 0:
            pr.D = _view(t["D"], N * p * n, p * n)
            pr.E = _view(t["E"], N * p * m, p * m)
            pr.e = _view(t["e"], N * p, p)
        else:
            pr.D = pr.E = pr.e = _view(None, 0, 0)
        if self.box:
            pr.ulo = _view(t["ulo"], m, 0)
            pr.uhi = _view(t["uhi"], m, 0)
        else:
            pr.ulo = pr.uhi = _view(None, 0, 0)
        pr.x0 = _view(t["x0"], n, 0)
        self.problem = pr

    def set_linearization(self, **arrays):
        """Copy the caller's arrays (host or device, with a leading batch axis)
        into the linearisation tensors: A, B, fbar, Q, R, M, q, r, g, QN."""
        for k, v in arrays.items():
            dst = self.lin[k]
            if dst is None:
                if v is not None and np.any(np.asarray(v) != 0):
                    raise ValueError("this batch was built without a cross term M (cross=False)")
                continue
            src = v if isinstance(v, torch.Tensor) else torch.as_tensor(np.asarray(v))
            dst.copy_(src.reshape(dst.shape).to(dtype=self.dtype))

    def iterate_from(self, X, U, V, Yeq=None, Ybox=None):
        def dev(a, shape):
            if a is None:
                return torch.zeros(shape, dtype=self.dtype, device=self.device)
            return torch.as_tensor(np.ascontiguousarray(a), dtype=self.dtype, device=self.device).clone()
        B, n, m, N, p = self.B, self.n, self.m, self.N, self.p
        return dict(X=dev(X, (B, N + 1, n)), U=dev(U, (B, N, m)), V=dev(V, (B, N + 1, n)),
                    Yeq=dev(Yeq, (B, N, p)), Ybox=dev(Ybox, (B, N, 2 * m)))

    def c_iterate(self, it):
        ci = L.Iterate()
        ci.X, ci.U, ci.V = it["X"].data_ptr(), it["U"].data_ptr(), it["V"].data_ptr()
        ci.Y_eq = it["Yeq"].data_ptr() if self.p > 0 else None
        ci.Y_box = it["Ybox"].data_ptr() if self.box else None
        return ci


class CallerSolver:
    """direction / accept halves of the caller-dynamics SQP iteration."""

    def __init__(self, batch: CallerBatch, params=None, **tune):
        self.batch = batch
        self.lib = L.load()
        self.params = params if params is not None else L.default_params()
        for k, v in tune.items():
            setattr(self.params, k, v)
        nbytes = C.c_size_t(0)
        L.check(self.lib.mpxkkt_workspace_size(C.byref(batch.problem), C.byref(self.params), C.byref(nbytes)))
        self.ws_bytes = int(nbytes.value)
        self.ws = torch.empty(self.ws_bytes, dtype=torch.uint8, device=batch.device)
        b = batch
        self.dX = torch.zeros((b.B, b.N + 1, b.n), dtype=b.dtype, device=b.device)
        self.dU = torch.zeros((b.B, b.N, b.m), dtype=b.dtype, device=b.device)
        self.amax = torch.zeros(b.B, dtype=b.dtype, device=b.device)
        self.stats_t = dict(
            alpha=torch.zeros(b.B, dtype=b.dtype, device=b.device),
            rho=torch.zeros(b.B, dtype=b.dtype, device=b.device),
            merit0=torch.zeros(b.B, dtype=b.dtype, device=b.device),
            dmerit=torch.zeros(b.B, dtype=b.dtype, device=b.device),
            status=torch.zeros(b.B, dtype=torch.int32, device=b.device),
            trials=torch.zeros(b.B, dtype=torch.int32, device=b.device))
        st = L.Stats()
        for k, v in self.stats_t.items():
            setattr(st, k, v.data_ptr())
        self._stats = st

    def direction(self, it):
        """(dX, dU, a_max) of the step from the linearisation now in the batch."""
        d = L.Direction()
        d.dX, d.dU, d.alpha_max = self.dX.data_ptr(), self.dU.data_ptr(), self.amax.data_ptr()
        ci = self.batch.c_iterate(it)
        L.check(self.lib.mpxkkt_sqp_direction(_stream(), C.byref(self.batch.problem), C.byref(self.params),
                                              C.byref(ci), C.byref(d), C.c_void_p(self.ws.data_ptr()),
                                              self.ws_bytes))
        return self.dX, self.dU, self.amax

    def trial_steps(self, n_trials=None):
        """a_j = a_max beta^j, j < n_trials, per instance (device, (B, L))."""
        L_ = int(n_trials or self.params.max_trials)
        powers = torch.tensor([self.params.beta ** j for j in range(L_)], dtype=self.batch.dtype,
                              device=self.batch.device)
        return self.amax[:, None] * powers[None, :]

    def accept(self, it, trial_g, trial_f):
        """Line-search choice and update in place on ``it`` from the caller's
        trial evaluations trial_g (B, L, N+1) and trial_f (B, L, N, n)."""
        b = self.batch
        dev = lambda a: (a if isinstance(a, torch.Tensor) else torch.as_tensor(np.asarray(a))).to(
            device=b.device, dtype=b.dtype).contiguous()
        tg, tf = dev(trial_g), dev(trial_f)
        Ltr = int(tg.shape[1])
        if tuple(tg.shape) != (b.B, Ltr, b.N + 1) or tuple(tf.shape) != (b.B, Ltr, b.N, b.n):
            raise ValueError(f"trial arrays must be (B, L, N+1) and (B, L, N, n), got {tuple(tg.shape)}, "
                             f"{tuple(tf.shape)}")
        ci = b.c_iterate(it)
        L.check(self.lib.mpxkkt_sqp_accept(_stream(), C.byref(b.problem), C.byref(self.params), C.byref(ci),
                                           Ltr, _ptr(tg), _ptr(tf), C.byref(self._stats),
                                           C.c_void_p(self.ws.data_ptr()), self.ws_bytes))
        return self.stats_t
