"""The reference's single-rigid-body quadruped controller problem on the device.

``BatchedMPCControllerWrapper`` of mpx/utils/mpc_wrapper_srbd.py:58-77 binds
``quadruped_srbd_dynamics`` (mpx/utils/models.py:6-37), ``quadruped_srbd_obj``
and ``quadruped_srbd_hessian_gn`` (mpx/utils/objectives.py:15-75) into
``optimizers.mpc`` and runs it under ``jit(vmap)``.  :class:`SRBDSolver` runs
that SQP iteration for a whole batch with one stream-ordered library call,
``mpxkkt_srbd_iterate``: the model's linearisation (analytic Jacobians,
gradient, Gauss-Newton Hessian triple), the LQR direction, the model at the
line-search trial points and the acceptance all happen on the device.

Per instance: ``W`` (24, 24) residual weight, ``reference`` (N+1, 17) =
[p_ref, quat_ref, dp_ref, omega_ref, contact], ``parameter`` (N+1, 16) =
[contact, p_legs]; the model constants mass, inertia and dt.

:class:`SRBDGait` runs the controller's gait timer and reference generator
(``timer_run`` and ``reference_generator_srbd``, mpx/utils/mpc_utils.py:7-14,
:146-268) on the device, writing reference and parameter in that layout, so a
whole controller tick -- timer, reference, SQP iterations, warm-start shift --
needs no host step.
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import torch

from . import _lib as L
from .caller import CallerBatch, CallerSolver
from .problem import _stream, _view

N_STATE, N_CTRL, N_REF, N_PAR, N_RES = 13, 12, 17, 16, 24


class SRBDSolver:
    """Batched SQP iterations of the SRBD problem; B instances of horizon N."""

    def __init__(self, mass, inertia, dt, W, reference, parameter, x0, dtype=torch.float32, device="cuda",
                 params=None, **tune):
        reference = np.asarray(reference, np.float64)
        parameter = np.asarray(parameter, np.float64)
        B, N1 = reference.shape[0], reference.shape[1]
        if reference.shape != (B, N1, N_REF) or parameter.shape != (B, N1, N_PAR):
            raise ValueError(f"reference must be (B, N+1, {N_REF}) and parameter (B, N+1, {N_PAR})")
        W = np.broadcast_to(np.asarray(W, np.float64), (B, N_RES, N_RES))
        self.N = N1 - 1
        self.batch = CallerBatch(B, N_STATE, N_CTRL, self.N, np.asarray(x0, np.float64), cross=True, dtype=dtype,
                                 device=device)
        self.solver = CallerSolver(self.batch, params=params, **tune)
        dev = lambda a: torch.as_tensor(np.ascontiguousarray(a), dtype=dtype, device=self.batch.device)
        self.t = dict(W=dev(W), reference=dev(reference), parameter=dev(parameter))
        I = np.asarray(inertia, np.float64).reshape(3, 3)
        m = L.Srbd()
        m.mass, m.dt = float(mass), float(dt)
        for k, v in enumerate(I.ravel()):
            m.inertia[k] = float(v)
        for k, v in enumerate(np.linalg.inv(I).ravel()):
            m.inertia_inv[k] = float(v)
        m.W = _view(self.t["W"], N_RES * N_RES, 0)
        m.reference = _view(self.t["reference"], N1 * N_REF, N_REF)
        m.parameter = _view(self.t["parameter"], N1 * N_PAR, N_PAR)
        self.model = m

    @property
    def lin(self):
        """The linearisation at the iterate the last iteration started from."""
        return self.batch.lin

    def iterate_from(self, X, U, V):
        return self.batch.iterate_from(X, U, V)

    def iterate(self, it):
        """One SQP iteration in place on ``it``; returns the stats tensors."""
        s = self.solver
        ci = self.batch.c_iterate(it)
        L.check(s.lib.mpxkkt_srbd_iterate(_stream(), C.byref(self.batch.problem), C.byref(self.model),
                                          C.byref(s.params), C.byref(ci), C.byref(s._stats),
                                          C.c_void_p(s.ws.data_ptr()), s.ws_bytes))
        return s.stats_t


class SRBDGait:
    """Device gait state of B controllers and the reference generator.

    ``t_timer`` (B, 4) gait phases and ``liftoff`` (B, 12) are kept on the
    device between ticks; :meth:`generate` writes ``reference`` (B, N+1, 17),
    ``parameter`` (B, N+1, 16) and ``foot_ref_dot`` (B, N+1, 12).
    """

    def __init__(self, B, N, dt, mass, duty_factor, step_freq, step_height, clearance_speed, foot0, t_timer,
                 liftoff, dtype=torch.float32, device="cuda"):
        self.B, self.N, self.dtype = B, N, dtype
        self.device = torch.device(device)
        dev = lambda a: torch.as_tensor(np.ascontiguousarray(a, np.float64), dtype=dtype,
                                        device=self.device).clone()
        self.t_timer = dev(np.broadcast_to(np.asarray(t_timer, np.float64), (B, 4)))
        self.liftoff = dev(np.broadcast_to(np.asarray(liftoff, np.float64), (B, 12)))
        z = lambda *shape: torch.zeros(shape, dtype=dtype, device=self.device)
        self.reference, self.parameter, self.foot_ref_dot = z(B, N + 1, N_REF), z(B, N + 1, N_PAR), z(B, N + 1, 12)
        g = L.SrbdGait()
        g.dt, g.mass, g.duty_factor, g.step_freq = float(dt), float(mass), float(duty_factor), float(step_freq)
        g.step_height, g.clearance_speed, g.advance_dt = float(step_height), float(clearance_speed), 0.0
        for k, v in enumerate(np.asarray(foot0, np.float64).ravel()):
            g.foot0[k] = float(v)
        self.gait = g

    def generate(self, x, foot, command, contact, advance_dt=0.0):
        """Reference and parameter for states x (B, 13), feet (B, 12), commands
        (B, 7) = [v, omega, base height] and measured contacts (B, 4); with
        advance_dt > 0 the gait phases first advance by that step."""
        dev = lambda a: (a if isinstance(a, torch.Tensor) else torch.as_tensor(np.asarray(a))).to(
            device=self.device, dtype=self.dtype).contiguous()
        x, foot, command, contact = dev(x), dev(foot), dev(command), dev(contact)
        self.gait.advance_dt = float(advance_dt)
        dt = L.F64 if self.dtype == torch.float64 else L.F32
        p = lambda t: C.c_void_p(t.data_ptr())
        L.check(L.load().mpxkkt_srbd_reference(_stream(), dt, self.B, self.N, C.byref(self.gait),
                                               p(self.t_timer), p(x), p(foot), p(command), p(contact),
                                               p(self.liftoff), p(self.reference), p(self.parameter),
                                               p(self.foot_ref_dot)))
        return self.reference, self.parameter
