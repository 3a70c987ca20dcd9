"""Continuous-batching MPC serving on one GPU.

Controllers of differing (n, m, N) open, tick and close over time.  Each one
occupies a slot of the device arrays of its shape bucket for its whole life:
its problem data, warm-start trajectory (X, U, V) and multipliers stay
resident in HBM across ticks (the cross-tick state the reference keeps in
``mpx_data`` / ``BatchedMPCControllerWrapper.batch_X0``,
mpx/utils/mpc_wrapper.py:16-28, mpx/utils/mpc_wrapper_srbd.py:56-58,
:135-137).  Admission is by an HBM byte budget: a controller that does not fit
waits until others close.

Requests (a controller's new measured state x0 and the number of SQP
iterations for this tick) queue as they arrive.  Every round gathers the
pending requests of each bucket, grouped by iteration count, into a staging
prefix of the bucket, runs the batched SQP iterations of each group on its own
HIP stream (groups overlap on the device), then scatters the iterate back into
the slots with the warm-start shift and the per-controller NaN fallback of
update_and_extract (mpx/utils/mpc_wrapper.py:79-96).  A controller's result
does not depend on which other requests share its round: every launch shape is
fixed per bucket, and the kernels never mix instances.
"""
from __future__ import annotations

import collections
import time
from dataclasses import dataclass

import numpy as np
import torch

from . import _lib as L
from .problem import DeviceBatch, SQPSolver

_PROBLEM_KEYS = ("A", "B", "W", "Q", "R", "QN", "xref", "uref", "D", "E", "e", "ulo", "uhi", "x0")
_ITER_KEYS = ("X", "U", "V", "Yeq", "Ybox")


def shape_key(inst):
    return (inst["n"], inst["m"], inst["N"], inst["p"], inst.get("ulo") is not None, inst["model"])


@dataclass
class Request:
    rid: int
    cid: int
    iters: int
    t_arrive: float
    x0: np.ndarray
    t_done: float = 0.0
    u0: np.ndarray | None = None
    reset: bool = False


@dataclass
class _Controller:
    cid: int
    bucket: "_Bucket"
    slot: int
    nbytes: int
    ticks: int = 0


class _Bucket:
    """Resident slot arrays for one shape, plus per-stream staging batches."""

    def __init__(self, key, proto, capacity, max_group, dtype, device, n_streams, tune):
        self.key = key
        self.capacity = capacity
        self.free = list(range(capacity - 1, -1, -1))
        self.slots = DeviceBatch.from_instances([proto] * capacity, dtype=dtype, device=device)
        self.state = self.slots.iterate_from(*[None] * 5)
        self.max_group = max_group
        self.params = L.default_params(**tune)
        # staging batch, iterate and workspace per stream: groups of one round run
        # concurrently, so nothing they write is shared
        self.stage, self.stage_it, self.ws = [], [], []
        for _ in range(n_streams):
            st = DeviceBatch.from_instances([proto] * max_group, dtype=dtype, device=device)
            self.stage.append(st)
            self.stage_it.append(st.iterate_from(*[None] * 5))
            self.ws.append(torch.empty(SQPSolver.workspace_bytes(st, self.params), dtype=torch.uint8,
                                       device=device))
        self.solvers = {}

    def solver(self, k, si):
        if (k, si) not in self.solvers:
            self.solvers[(k, si)] = SQPSolver(self.stage[si].prefix(k), params=self.params, workspace=self.ws[si])
        return self.solvers[(k, si)]

    def load(self, slot, inst):
        """Write an instance's problem data and a cold start into its slot."""
        dt, dev = self.slots.dtype, self.slots.device
        one = DeviceBatch.from_instances([inst], dtype=dt, device=dev)
        for k in _PROBLEM_KEYS:
            if self.slots.t.get(k) is not None:
                self.slots.t[k][slot].copy_(one.t[k][0])
        N, n = inst["N"], inst["n"]
        self.state["X"][slot] = torch.as_tensor(inst["x0"], dtype=dt, device=dev).expand(N + 1, n)
        for k in ("U", "V", "Yeq", "Ybox"):
            self.state[k][slot].zero_()


class MPCServer:
    """Admits controllers by HBM budget and serves their ticks with continuous batching."""

    def __init__(self, hbm_budget_bytes, max_group=256, slots_per_bucket=512, n_streams=4,
                 dtype=torch.float32, device="cuda", shift=1, tune=None):
        self.budget = int(hbm_budget_bytes)
        self.used = 0
        self.max_group = max_group
        self.slots_per_bucket = slots_per_bucket
        self.dtype, self.device = dtype, torch.device(device)
        self.streams = [torch.cuda.Stream(device=self.device) for _ in range(n_streams)]
        self.shift = shift
        self.tune = tune or {}
        self.buckets = {}
        self.controllers = {}
        self.pending = collections.deque()
        self._cid = 0
        self._rid = 0

    # --- admission -------------------------------------------------------
    def controller_bytes(self, inst):
        """Resident HBM bytes of one controller: problem data + warm start + its
        share of the bucket's staging and workspace."""
        n, m, N, p = inst["n"], inst["m"], inst["N"], inst["p"]
        es = 4 if self.dtype == torch.float32 else 8
        prob = 3 * n * n + n * m + m * m + (N + 1) * n + N * m + N * p * (n + m + 1) + 3 * n + 2 * m
        state = 2 * (N + 1) * n + N * m + N * p + 2 * N * m
        work = N * (2 * n * n + 2 * m * n + 6 * n + 4 * m) + (N + 1) * (n * n + 8 * n)
        return es * (2 * (prob + state) + work)

    def open(self, inst):
        """Admit a controller if its bytes fit the budget; returns its id or None."""
        need = self.controller_bytes(inst)
        if self.used + need > self.budget:
            return None
        key = shape_key(inst)
        b = self.buckets.get(key)
        if b is None:
            b = _Bucket(key, inst, self.slots_per_bucket, self.max_group, self.dtype, self.device,
                        len(self.streams), self.tune)
            self.buckets[key] = b
        if not b.free:
            return None
        slot = b.free.pop()
        b.load(slot, inst)
        self._cid += 1
        self.controllers[self._cid] = _Controller(self._cid, b, slot, need)
        self.used += need
        return self._cid

    def close(self, cid):
        c = self.controllers.pop(cid)
        c.bucket.free.append(c.slot)
        self.used -= c.nbytes

    # --- requests ----------------------------------------------------------
    def request(self, cid, x0, iters, t_arrive=None):
        self._rid += 1
        r = Request(self._rid, cid, int(iters), time.perf_counter() if t_arrive is None else t_arrive,
                    np.asarray(x0, dtype=np.float64))
        self.pending.append(r)
        return r

    def run_round(self):
        """Serve every pending request once; returns the completed requests."""
        if not self.pending:
            return []
        reqs = list(self.pending)
        self.pending.clear()
        groups = collections.defaultdict(list)
        seen = set()
        for r in reqs:
            if r.cid in seen:            # one tick per controller per round
                self.pending.append(r)
                continue
            seen.add(r.cid)
            c = self.controllers[r.cid]
            groups[(c.bucket.key, r.iters)].append(r)
        main = torch.cuda.current_stream(self.device)
        for s in self.streams:
            s.wait_stream(main)          # slot loads of newly opened controllers
        launched = []
        si = 0
        for (key, iters), rs in groups.items():
            b = self.buckets[key]
            for g0 in range(0, len(rs), b.max_group):
                chunk = rs[g0:g0 + b.max_group]
                stream = self.streams[si % len(self.streams)]
                launched.append(self._launch(b, chunk, iters, stream, si % len(self.streams)))
                si += 1
        done = []
        for ev, chunk, u0, reset in launched:
            ev.synchronize()
            t = time.perf_counter()
            u0h = u0.numpy()
            rh = reset.numpy()
            for j, r in enumerate(chunk):
                r.u0, r.reset, r.t_done = u0h[j].copy(), bool(rh[j]), t
                self.controllers[r.cid].ticks += 1
                done.append(r)
        return done

    def _launch(self, b, chunk, iters, stream, si):
        """Gather -> iterate -> shift -> scatter for one group on ``stream``.

        Every device tensor of the launch is allocated while ``stream`` is
        current, so the caching allocator orders its reuse after this stream's
        work; host inputs and outputs go through pinned buffers with
        asynchronous copies (the pinned allocator holds a buffer until its copy
        has completed)."""
        k = len(chunk)
        slots_h = torch.tensor([self.controllers[r.cid].slot for r in chunk], dtype=torch.long).pin_memory()
        x0_h = torch.from_numpy(np.stack([r.x0 for r in chunk])).to(self.dtype).pin_memory()
        with torch.cuda.stream(stream):
            slots = slots_h.to(self.device, non_blocking=True)
            st, st_it = b.stage[si], b.stage_it[si]
            for key in _PROBLEM_KEYS:
                src = b.slots.t.get(key)
                if src is not None:
                    torch.index_select(src, 0, slots, out=st.t[key][:k])
            st.t["x0"][:k].copy_(x0_h, non_blocking=True)
            for key in _ITER_KEYS:
                torch.index_select(b.state[key], 0, slots, out=st_it[key][:k])
            it = {key: st_it[key][:k] for key in _ITER_KEYS}
            solver = b.solver(k, si)
            for _ in range(iters):
                solver.iterate(it)
            u0_h = torch.empty((k, st.m), dtype=self.dtype, pin_memory=True)
            u0_h.copy_(it["U"][:, 0, :], non_blocking=True)
            reset = solver.shift(it, self.shift)
            reset_h = torch.empty(k, dtype=torch.int32, pin_memory=True)
            reset_h.copy_(reset, non_blocking=True)
            for key in _ITER_KEYS:
                b.state[key].index_copy_(0, slots, it[key])
            ev = torch.cuda.Event()
            ev.record(stream)
        return ev, chunk, u0_h, reset_h
