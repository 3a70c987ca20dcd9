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
iterations for this tick) queue as they arrive.  ``step()`` never blocks: it
hands back the ticks whose work has finished on the device, and every bucket
with a free staging area takes the requests pending for it (at most
``max_group``, oldest first) as one launch on that area's own HIP stream, so
launches of different buckets, and successive launches of one bucket, overlap
on the device while new requests keep arriving.  A launch gathers its
controllers' slots into the staging prefix sorted by iteration count
(descending), runs SQP iteration j on the prefix of requests that need more
than j iterations, and as soon as a cohort has had all its iterations it
applies the warm-start shift with the per-controller NaN fallback of
update_and_extract (mpx/utils/mpc_wrapper.py:79-96) on that cohort's slice
(a failed controller restarts from x0, u_ref and zero multipliers and its
tick returns the next control of the plan it started from, U0[1]), copies the
first controls to the host, scatters the slice back into the slots and
records an event.

Memory: a bucket's slot arrays start small and double when full, each growth
admitted against the HBM budget; the budget counts every byte the engine
holds on the device -- slot arrays at their capacity, and per staging area
its batch arrays and the library's workspace (``mpxkkt_workspace_size``).
An arrival that would exceed the budget is turned away.  A controller's result does not depend on
which other requests share its launch: every launch shape is fixed per
bucket, and the kernels never mix instances.
"""
from __future__ import annotations

import collections
import time
from dataclasses import dataclass, field

import numpy as np
import torch

from . import _lib as L
from .problem import DeviceBatch, SQPSolver, shift_warmstart

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
    ticks: int = 0
    busy: bool = False


@dataclass
class _Launch:
    bucket: "_Bucket"
    si: int
    # cohorts in completion order: (event, requests, pinned u0, pinned reset flags)
    cohorts: list = field(default_factory=list)


def _tensor_bytes(tensors):
    return sum(t.numel() * t.element_size() for t in tensors if t is not None)


class _Bucket:
    """Resident slot arrays for one shape (growable), plus staging batches,
    each with its own stream and workspace."""

    def __init__(self, key, proto, capacity, max_group, dtype, device, n_streams, params):
        self.key = key
        self.proto = proto
        self.capacity = capacity
        self.free = list(range(capacity - 1, -1, -1))
        self.slots = DeviceBatch.from_instances([proto] * capacity, dtype=dtype, device=device)
        self.state = self.slots.iterate_from(*[None] * 5)
        self.slot_bytes = (_tensor_bytes(self.slots.t.values()) + _tensor_bytes(self.state.values())) // capacity
        self.max_group = max_group
        self.params = params
        self.stage, self.stage_it, self.ws, self.streams, self.io = [], [], [], [], []
        self.idle = []
        self.fixed_bytes = 0
        n, m = proto["n"], proto["m"]
        for si in range(n_streams):
            st = DeviceBatch.from_instances([proto] * max_group, dtype=dtype, device=device)
            self.stage.append(st)
            self.stage_it.append(st.iterate_from(*[None] * 5))
            self.ws.append(torch.empty(SQPSolver.workspace_bytes(st, self.params), dtype=torch.uint8,
                                       device=device))
            self.fixed_bytes += (_tensor_bytes(st.t.values()) + _tensor_bytes(self.stage_it[-1].values())
                                 + self.ws[-1].numel())
            self.streams.append(torch.cuda.Stream(device=device))
            # host <-> device transfer buffers of the area, allocated once: a
            # pinned allocation inside the serving loop can wait for the device
            self.io.append(dict(
                slots_h=torch.empty(max_group, dtype=torch.long, pin_memory=True),
                x0_h=torch.empty((max_group, n), dtype=dtype, pin_memory=True),
                u0_h=torch.empty((max_group, m), dtype=dtype, pin_memory=True),
                reset_h=torch.empty(max_group, dtype=torch.int32, pin_memory=True),
                slots=torch.empty(max_group, dtype=torch.long, device=device),
                reset=torch.empty(max_group, dtype=torch.int32, device=device),
                u0=torch.empty((max_group, m), dtype=dtype, device=device),
                u_next=torch.empty((max_group, m), dtype=dtype, device=device),
                u_reset=torch.empty((max_group, m), dtype=dtype, device=device)))
            self.idle.append(si)
        self.solvers = {}

    def grow(self, capacity):
        """Move the slot arrays to a larger capacity.  The copy runs on the
        current stream after the work already queued on the bucket's staging
        streams (whose launches gather from and scatter to the old arrays);
        later launches wait for the current stream before they gather."""
        cur = torch.cuda.current_stream(self.slots.device)
        for s in self.streams:
            cur.wait_stream(s)
        new = DeviceBatch.from_instances([self.proto] * capacity, dtype=self.slots.dtype,
                                         device=self.slots.device)
        st = new.iterate_from(*[None] * 5)
        old = self.capacity
        for k, v in self.slots.t.items():
            if v is not None:
                new.t[k][:old].copy_(v)
        for k, v in self.state.items():
            st[k][:old].copy_(v)
        self.slots, self.state = new, st
        self.free = list(range(capacity - 1, old - 1, -1)) + self.free
        self.capacity = capacity

    def solver(self, k, si):
        if (k, si) not in self.solvers:
            self.solvers[(k, si)] = SQPSolver(self.stage[si].prefix(k), params=self.params, workspace=self.ws[si])
        return self.solvers[(k, si)]

    def load(self, slot, inst):
        """Write an instance's problem data and a cold start into its slot: one
        host-to-device copy of the packed arrays, then device-side copies into
        the slot (on the current stream; launches wait for it)."""
        dt, dev = self.slots.dtype, self.slots.device
        dst, parts = [], []
        for k in _PROBLEM_KEYS:
            d = self.slots.t.get(k)
            if d is not None:
                dst.append(d[slot])
                parts.append(np.broadcast_to(np.asarray(inst[k]), d.shape[1:]).ravel())
        flat = torch.from_numpy(np.concatenate(parts)).to(device=dev, dtype=dt)
        off = 0
        for d in dst:
            d.view(-1).copy_(flat[off:off + d.numel()])
            off += d.numel()
        N, n = inst["N"], inst["n"]
        self.state["X"][slot].copy_(self.slots.t["x0"][slot].expand(N + 1, n))
        for k in ("U", "V", "Yeq", "Ybox"):
            self.state[k][slot].zero_()


def _workspace_bytes_for(proto, k, dtype, params):
    """mpxkkt_workspace_size for a batch of k instances of proto's shape (a
    host-side query on a problem struct with placeholder pointers)."""
    import ctypes as C
    n, m, N, p = proto["n"], proto["m"], proto["N"], proto["p"]
    pr = L.Problem()
    pr.dtype = L.F64 if dtype == torch.float64 else L.F32
    pr.model = L.MODEL_TANH if proto["model"] == "tanh" else L.MODEL_LTV
    pr.batch, pr.n, pr.m, pr.horizon, pr.n_eq = k, n, m, N, p
    pr.has_box = 1 if proto.get("ulo") is not None else 0
    for name, typ in L.Problem._fields_:
        if typ is L.View and name not in L.CALLER_VIEWS:
            getattr(pr, name).ptr = 256     # never dereferenced by the size query
    nbytes = C.c_size_t(0)
    L.check(L.load().mpxkkt_workspace_size(C.byref(pr), C.byref(params), C.byref(nbytes)))
    return int(nbytes.value)


def _bucket_params(proto, dtype, device, tune):
    """Launch parameters of a bucket and its time per SQP iteration (ms, or
    None): ``tune`` as given (a dict), or, with ``tune='auto'``, the
    autotuner's choice for a group of 16 of its shape."""
    if tune != "auto":
        return L.default_params(**(tune or {})), None
    from . import autotune
    insts = [proto] * 16
    b = DeviceBatch.from_instances(insts, dtype=dtype, device=device)
    it0 = b.iterate_from(*[None] * 5)
    cands = [dict(segments=1, block_threads=t, trial_group=g, flags=f)
             for t in (128, 256, 512) for g in (1, 0) for f in (0, 8)]
    best = autotune.tune(b, it0, iters=2, cands=cands)
    return L.default_params(**{k: v for k, v in best.items() if k != "ms"}), best["ms"]


class MPCServer:
    """Admits controllers by HBM budget and serves their ticks with continuous
    batching.  ``slots_per_bucket``: a new bucket's initial slot capacity
    (doubled whenever it is full, up to ``max_slots_per_bucket``, each growth
    admitted against the budget)."""

    def __init__(self, hbm_budget_bytes, max_group=256, slots_per_bucket=64, n_streams=1,
                 dtype=torch.float32, device="cuda", shift=1, tune=None, n_streams_slow=2, slow_ms=2.0,
                 max_slots_per_bucket=1 << 16, max_pending=2048):
        self.budget = int(hbm_budget_bytes)
        # ticks queued and not yet launched: beyond this a request is refused
        # (an overloaded server sheds load instead of letting every latency grow)
        self.max_pending = int(max_pending)
        self.max_group = max_group
        self.slots_per_bucket = slots_per_bucket
        self.max_slots = max_slots_per_bucket
        self.turned_away = dict(budget=0, slots=0, queue=0)
        # staging areas (each with its stream) per bucket: n_streams, or
        # n_streams_slow for a bucket whose SQP iteration takes slow_ms or more
        # (known with tune='auto'): a fast bucket serves a launch in well under
        # a millisecond and gains nothing from overlapping its own launches,
        # while every extra stream competes with the others for the device
        self.n_streams = n_streams
        self.n_streams_slow = n_streams_slow
        self.slow_ms = slow_ms
        self.dtype, self.device = dtype, torch.device(device)
        self.shift = shift
        self.tune = tune
        self.buckets = {}
        self._probe, self._params, self._bucket_ms = {}, {}, {}
        self.controllers = {}
        self.pending = collections.deque()
        self.in_flight = []
        self._cid = 0
        self._rid = 0

    # --- admission -------------------------------------------------------
    @property
    def used(self):
        """Device bytes the engine holds: slot arrays at their capacity plus
        every staging area's arrays and workspace."""
        return sum(b.fixed_bytes + b.capacity * b.slot_bytes for b in self.buckets.values())

    def _growth(self, b):
        """Capacity a full bucket would grow to, or None at the slot limit."""
        cap = min(2 * b.capacity, self.max_slots)
        return cap if cap > b.capacity else None

    def bytes_to_admit(self, inst):
        """Device bytes that admitting a controller of this shape now would add
        (0 when its bucket has a free slot; None when the bucket is at its slot
        limit).  A bucket not yet built is measured by building its staging
        batch's workspace query on the host (no device memory)."""
        b = self.buckets.get(shape_key(inst))
        if b is None:
            return self._new_bucket_bytes(inst)
        if b.free:
            return 0
        cap = self._growth(b)
        return None if cap is None else (cap - b.capacity) * b.slot_bytes

    def _new_bucket_bytes(self, inst):
        key = shape_key(inst)
        if key not in self._probe:
            n, m, N, p = inst["n"], inst["m"], inst["N"], inst["p"]
            box = inst.get("ulo") is not None
            es = 4 if self.dtype == torch.float32 else 8
            nw = n * n if inst["model"] == "tanh" else 0
            nab = (n * n + n * m) if inst["model"] == "tanh" else N * (n * n + n * m)
            prob = (nab + nw + n * n * (1 if inst["Q"].shape[0] == 1 else N)
                    + m * m * (1 if inst["R"].shape[0] == 1 else N) + n * n + (N + 1) * n + N * m
                    + N * p * (n + m + 1) + (2 * m if box else 0) + n)
            state = 2 * (N + 1) * n + N * m + N * p + 2 * N * m
            ws = _workspace_bytes_for(inst, self.max_group, self.dtype, self.params_for(inst))
            ns = self.n_streams_slow if (self._bucket_ms.get(key) or 0.0) >= self.slow_ms else self.n_streams
            self._probe[key] = ((prob + state) * es, ns * ((prob + state) * es * self.max_group + ws))
        slot, fixed = self._probe[key]
        return fixed + self.slots_per_bucket * slot

    def params_for(self, inst):
        key = shape_key(inst)
        if key not in self._params:
            self._params[key], self._bucket_ms[key] = _bucket_params(inst, self.dtype, self.device, self.tune)
        return self._params[key]

    def room_for(self, inst):
        """Whether ``open`` would admit a controller of this instance's shape now
        (only the shape fields of ``inst`` are read)."""
        need = self.bytes_to_admit(inst)
        return need is not None and self.used + need <= self.budget

    def open(self, inst):
        """Admit a controller if the device bytes it needs fit the budget (and
        its bucket's slot limit); returns its id or None."""
        need = self.bytes_to_admit(inst)
        if need is None:
            self.turned_away["slots"] += 1
            return None
        if self.used + need > self.budget:
            self.turned_away["budget"] += 1
            return None
        key = shape_key(inst)
        b = self.buckets.get(key)
        if b is None:
            params = self.params_for(inst)
            ms = self._bucket_ms.get(key)
            ns = self.n_streams_slow if ms is not None and ms >= self.slow_ms else self.n_streams
            b = _Bucket(key, inst, self.slots_per_bucket, self.max_group, self.dtype, self.device, ns, params)
            self.buckets[key] = b
        elif not b.free:
            b.grow(self._growth(b))
        slot = b.free.pop()
        b.load(slot, inst)
        self._cid += 1
        self.controllers[self._cid] = _Controller(self._cid, b, slot)
        return self._cid

    def close(self, cid):
        """Release a controller's slot; its pending ticks are dropped.  A
        controller whose tick is on the device cannot close until it returns."""
        c = self.controllers[cid]
        if c.busy:
            raise RuntimeError(f"controller {cid} has a tick in flight")
        self.pending = collections.deque(r for r in self.pending if r.cid != cid)
        del self.controllers[cid]
        c.bucket.free.append(c.slot)

    # --- requests ----------------------------------------------------------
    @property
    def accepting(self):
        """Whether a request would be queued now (the pending queue has room)."""
        return len(self.pending) < self.max_pending

    def request(self, cid, x0, iters, t_arrive=None):
        """Queue one tick of controller ``cid`` (open) with measured state x0 and
        ``iters`` >= 1 SQP iterations; returns the request, or None when the
        pending queue is full (the tick is refused: counted in turned_away)."""
        if cid not in self.controllers:
            raise KeyError(f"controller {cid} is not open")
        iters = int(iters)
        if iters < 1:
            raise ValueError(f"a tick needs at least one SQP iteration (got {iters})")
        x0 = np.asarray(x0, dtype=np.float64)
        if x0.shape != (self.controllers[cid].bucket.slots.n,):
            raise ValueError(f"x0 has shape {x0.shape}, controller {cid} expects "
                             f"({self.controllers[cid].bucket.slots.n},)")
        if not self.accepting:
            self.turned_away["queue"] += 1
            return None
        self._rid += 1
        r = Request(self._rid, cid, iters, time.perf_counter() if t_arrive is None else t_arrive, x0)
        self.pending.append(r)
        return r

    @property
    def busy(self):
        return bool(self.pending) or bool(self.in_flight)

    def step(self):
        """Non-blocking: returns the ticks completed since the last call and
        launches pending requests onto idle staging areas."""
        done = self._harvest(block=False)
        self._dispatch()
        return done

    def run_round(self):
        """Serve every request pending now; blocks until they are done and
        returns them (with anything else that finished meanwhile)."""
        want = {r.rid for r in self.pending}
        done = []
        while want:
            self._dispatch()
            if not self.in_flight:
                raise RuntimeError("pending ticks that cannot be launched")
            got = self._harvest(block=True)
            done.extend(got)
            want -= {r.rid for r in got}
        return done

    # --- device work ---------------------------------------------------------
    def _harvest(self, block):
        done = []
        keep = []
        for ln in self.in_flight:
            while ln.cohorts:
                ev, reqs, u0_h, reset_h = ln.cohorts[0]
                if block:
                    ev.synchronize()
                elif not ev.query():
                    break
                t = time.perf_counter()
                u0 = u0_h.numpy()
                rs = reset_h.numpy()
                for j, r in enumerate(reqs):
                    r.u0, r.reset, r.t_done = u0[j].copy(), bool(rs[j]), t
                    c = self.controllers[r.cid]
                    c.ticks += 1
                    c.busy = False
                    done.append(r)
                ln.cohorts.pop(0)
            if ln.cohorts:
                keep.append(ln)
            else:
                ln.bucket.idle.append(ln.si)
        self.in_flight = keep
        return done

    def _dispatch(self):
        if not self.pending:
            return
        by_bucket = collections.OrderedDict()
        rest = collections.deque()
        taken = set()
        for r in self.pending:
            c = self.controllers[r.cid]
            b = c.bucket
            if c.busy or r.cid in taken or not b.idle or len(by_bucket.get(b.key, ())) >= b.max_group:
                rest.append(r)           # one tick per controller in flight; wait for a free area
                continue
            taken.add(r.cid)
            by_bucket.setdefault(b.key, []).append(r)
        # the pass over the queue is complete: now mark the launched controllers
        for cid in taken:
            self.controllers[cid].busy = True
        self.pending = rest
        for key, reqs in by_bucket.items():
            b = self.buckets[key]
            self.in_flight.append(self._launch(b, reqs, b.idle.pop(0)))

    def _launch(self, b, reqs, si):
        """Gather -> iterate (by prefix) -> per cohort: first controls, shift,
        scatter, event; all on the staging area's stream.

        Every device tensor of the launch is allocated while that stream is
        current, so the caching allocator orders its reuse after the stream's
        work; host inputs and outputs go through pinned buffers with
        asynchronous copies (the pinned allocator holds a buffer until its copy
        has completed)."""
        reqs = sorted(reqs, key=lambda r: -r.iters)
        k = len(reqs)
        iters = [r.iters for r in reqs]
        stream = b.streams[si]
        stream.wait_stream(torch.cuda.current_stream(self.device))   # slot loads of new controllers
        io = b.io[si]          # free: the area's previous launch has completed
        io["slots_h"][:k].copy_(torch.tensor([self.controllers[r.cid].slot for r in reqs], dtype=torch.long))
        io["x0_h"][:k].copy_(torch.from_numpy(np.stack([r.x0 for r in reqs])))
        ln = _Launch(b, si)
        with torch.cuda.stream(stream):
            slots = io["slots"][:k]
            slots.copy_(io["slots_h"][:k], non_blocking=True)
            x0_h = io["x0_h"][:k]
            st, st_it = b.stage[si], b.stage_it[si]
            for key in _PROBLEM_KEYS:
                src = b.slots.t.get(key)
                if src is not None:
                    torch.index_select(src, 0, slots, out=st.t[key][:k])
            st.t["x0"][:k].copy_(x0_h, non_blocking=True)
            for key in _ITER_KEYS:
                torch.index_select(b.state[key], 0, slots, out=st_it[key][:k])
            # NaN fallback inputs: the next control of the plan each tick starts
            # from (returned if the tick fails) and the reset control u_ref
            N = st.N
            io["u_next"][:k].copy_(st_it["U"][:k, min(1, N - 1), :])
            io["u_reset"][:k].copy_(st.t["uref"][:k, 0, :])
            for j in range(iters[0]):
                kj = sum(1 for v in iters if v > j)
                b.solver(kj, si).iterate({key: st_it[key][:kj] for key in _ITER_KEYS})
                lo = sum(1 for v in iters if v > j + 1)
                if lo < kj:
                    ln.cohorts.append(self._finish(b, io, st, st_it, slots, reqs, lo, kj))
        return ln

    def _finish(self, b, io, st, st_it, slots, reqs, lo, hi):
        """First controls, shift and scatter of requests lo..hi-1 of a launch
        (on the current stream).  A tick whose first control is NaN returns
        the next control of the plan it started from (U0[1]) and its
        controller restarts from (x0, u_ref, 0), as update_and_extract does
        (mpx/utils/mpc_wrapper.py:87-95)."""
        it = {key: st_it[key][lo:hi] for key in _ITER_KEYS}
        u0 = io["u0"][lo:hi]
        u0.copy_(it["U"][:, 0, :])
        reset = io["reset"][lo:hi]
        reset.zero_()
        shift_warmstart(st.slice(lo, hi), it, self.shift, u_reset=io["u_reset"][lo:hi], reset=reset)
        torch.where(reset[:, None] != 0, io["u_next"][lo:hi], u0, out=u0)
        u0_h = io["u0_h"][lo:hi]
        u0_h.copy_(u0, non_blocking=True)
        reset_h = io["reset_h"][lo:hi]
        reset_h.copy_(reset, non_blocking=True)
        for key in _ITER_KEYS:
            b.state[key].index_copy_(0, slots[lo:hi], it[key])
        ev = torch.cuda.Event()
        ev.record(torch.cuda.current_stream(self.device))
        return ev, reqs[lo:hi], u0_h, reset_h
