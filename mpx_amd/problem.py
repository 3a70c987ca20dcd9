"""Device-resident batches of MPC instances and the batched SQP solver.

``DeviceBatch`` packs same-shaped instances (the dicts of
:mod:`mpx_amd.synthetic`) into struct-of-arrays HBM tensors, instance-major,
and fills the ``mpxkkt_problem`` struct with strided views (stride 0 for
time-invariant Q, R and for per-instance A, B, W).  ``SQPSolver`` owns the
workspace and calls the C ABI on torch's current stream.  PyTorch only
provides device memory and the stream here; all arithmetic is in the HIP
library.
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import torch

from . import _lib as L


def _ptr(t):
    return C.c_void_p(t.data_ptr()) if t is not None else C.c_void_p(0)


def _view(t, sb, st):
    v = L.View()
    v.ptr = t.data_ptr() if t is not None else None
    v.batch_stride = sb
    v.stage_stride = st
    return v


class DeviceBatch:
    """Same-shaped instances packed on the device.

    ``arrays`` maps field -> numpy/torch array with a leading batch axis, using
    the instance shapes of :mod:`mpx_amd.synthetic` (Q/R with a stage axis of
    length 1 or N).
    """

    def __init__(self, dims, arrays, dtype=torch.float32, device="cuda"):
        self.n, self.m, self.N, self.p = dims["n"], dims["m"], dims["N"], dims["p"]
        self.model = dims["model"]
        self.h, self.kappa = dims.get("h", 0.02), dims.get("kappa", 0.1)
        self.box = arrays.get("ulo") is not None
        self.dtype = dtype
        self.device = torch.device(device)
        self.t = {}
        for k, v in arrays.items():
            if v is None:
                self.t[k] = None
                continue
            if isinstance(v, torch.Tensor):
                self.t[k] = v.to(device=self.device, dtype=dtype).contiguous()
            else:
                self.t[k] = torch.as_tensor(np.ascontiguousarray(v), dtype=dtype, device=self.device)
        self.B = int(self.t["x0"].shape[0])
        self._build()

    @classmethod
    def from_instances(cls, insts, dtype=torch.float32, device="cuda"):
        i0 = insts[0]
        n, m, N, p = i0["n"], i0["m"], i0["N"], i0["p"]
        dims = dict(n=n, m=m, N=N, p=p, model=i0["model"], h=i0["h"], kappa=i0["kappa"])
        st = lambda key: np.stack([d[key] for d in insts])
        arr = dict(A=st("A"), B=st("B"), QN=st("QN"), xref=st("xref"), uref=st("uref"), x0=st("x0"))
        if i0["model"] == "tanh":
            arr["W"] = st("W")
        for key in ("Q", "R"):
            T = max(d[key].shape[0] for d in insts)
            arr[key] = np.stack([np.broadcast_to(d[key], (T,) + d[key].shape[1:]) for d in insts])
        if p > 0:
            arr["D"], arr["E"], arr["e"] = st("D"), st("E"), st("e")
        else:
            arr["D"] = arr["E"] = arr["e"] = None
        if i0.get("ulo") is not None:
            arr["ulo"], arr["uhi"] = st("ulo"), st("uhi")
        else:
            arr["ulo"] = arr["uhi"] = None
        return cls(dims, arr, dtype=dtype, device=device)

    def _build(self):
        n, m, N, p = self.n, self.m, self.N, self.p
        t = self.t
        pr = L.Problem()
        pr.dtype = L.F64 if self.dtype == torch.float64 else L.F32
        pr.model = L.MODEL_TANH if self.model == "tanh" else L.MODEL_LTV
        pr.batch, pr.n, pr.m, pr.horizon, pr.n_eq = self.B, n, m, N, p
        pr.has_box = 1 if self.box else 0
        pr.h, pr.kappa = float(self.h), float(self.kappa)
        if self.model == "tanh":
            pr.A = _view(t["A"], n * n, 0)
            pr.B = _view(t["B"], n * m, 0)
            pr.W = _view(t["W"], n * n, 0)
        else:
            pr.A = _view(t["A"], N * n * n, n * n)
            pr.B = _view(t["B"], N * n * m, n * m)
            pr.W = _view(None, 0, 0)
        Tq, Tr = t["Q"].shape[1], t["R"].shape[1]
        pr.Q = _view(t["Q"], Tq * n * n, n * n if Tq > 1 else 0)
        pr.R = _view(t["R"], Tr * m * m, m * m if Tr > 1 else 0)
        pr.QN = _view(t["QN"], n * n, 0)
        pr.xref = _view(t["xref"], (N + 1) * n, n)
        pr.uref = _view(t["uref"], N * m, m)
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

    def prefix(self, k):
        """The first k instances as a batch of their own (same device arrays; the
        instance-major layout makes any prefix contiguous)."""
        if not 1 <= k <= self.B:
            raise ValueError(f"prefix size {k} outside [1, {self.B}]")
        return self.slice(0, k)

    def slice(self, a, b):
        """Instances a..b-1 as a batch of their own: every view's pointer moves by
        a batch strides (same device arrays; ``t`` still names the whole batch)."""
        if not 0 <= a < b <= self.B:
            raise ValueError(f"slice [{a}, {b}) outside [0, {self.B})")
        other = object.__new__(DeviceBatch)
        other.__dict__.update(self.__dict__)
        pr = L.Problem()
        C.pointer(pr)[0] = self.problem
        es = 8 if self.dtype == torch.float64 else 4
        for name, typ in L.Problem._fields_:
            if typ is L.View:
                v = getattr(pr, name)
                if v.ptr:
                    v.ptr = v.ptr + a * v.batch_stride * es
        pr.batch = b - a
        other.problem = pr
        other.B = b - a
        return other

    def iterate_from(self, X, U, V, Yeq=None, Ybox=None):
        """Device iterate dict from host or device arrays (copied, contiguous)."""
        def dev(a, shape):
            if a is None:
                return torch.zeros(shape, dtype=self.dtype, device=self.device)
            if isinstance(a, torch.Tensor):
                return a.to(device=self.device, dtype=self.dtype).clone().contiguous()
            return torch.as_tensor(np.ascontiguousarray(a), dtype=self.dtype, device=self.device)
        B, n, m, N, p = self.B, self.n, self.m, self.N, self.p
        return dict(X=dev(X, (B, N + 1, n)), U=dev(U, (B, N, m)), V=dev(V, (B, N + 1, n)),
                    Yeq=dev(Yeq, (B, N, p)), Ybox=dev(Ybox, (B, N, 2 * m)))

    def c_iterate(self, it):
        ci = L.Iterate()
        ci.X, ci.U, ci.V = it["X"].data_ptr(), it["U"].data_ptr(), it["V"].data_ptr()
        ci.Y_eq = it["Yeq"].data_ptr() if self.p > 0 else None
        ci.Y_box = it["Ybox"].data_ptr() if self.box else None
        return ci


def _stream():
    return C.c_void_p(torch.cuda.current_stream().cuda_stream)


def shift_warmstart(batch, it, shift, u_reset=None, reset=None):
    """Warm-start shift with the per-instance NaN fallback on ``batch`` (any
    slice) and its iterate ``it``; returns the int32 reset flags."""
    if reset is None:
        reset = torch.zeros(batch.B, dtype=torch.int32, device=batch.device)
    ci = batch.c_iterate(it)
    L.check(L.load().mpxkkt_shift_warmstart(_stream(), C.byref(batch.problem), int(shift), C.byref(ci),
                                            _ptr(u_reset), _ptr(reset)))
    return reset


class SQPSolver:
    """Batched primal-dual SQP iterations on one device batch.

    ``workspace``: optional preallocated uint8 device tensor (shared by solvers
    that never run concurrently); sized by :meth:`workspace_bytes`.
    """

    def __init__(self, batch: DeviceBatch, params=None, workspace=None, **tune):
        self.batch = batch
        self.lib = L.load()
        self.params = params if params is not None else L.default_params()
        for k, v in tune.items():
            setattr(self.params, k, v)
        self.ws_bytes = self.workspace_bytes(batch, self.params)
        if workspace is not None:
            if workspace.numel() < self.ws_bytes:
                raise L.KKTError(f"workspace of {workspace.numel()} bytes < {self.ws_bytes} needed")
            self.ws = workspace
        else:
            self.ws = torch.empty(self.ws_bytes, dtype=torch.uint8, device=batch.device)
        B = batch.B
        self.stats_t = dict(
            alpha=torch.zeros(B, dtype=batch.dtype, device=batch.device),
            rho=torch.zeros(B, dtype=batch.dtype, device=batch.device),
            merit0=torch.zeros(B, dtype=batch.dtype, device=batch.device),
            dmerit=torch.zeros(B, dtype=batch.dtype, device=batch.device),
            status=torch.zeros(B, dtype=torch.int32, device=batch.device),
            trials=torch.zeros(B, dtype=torch.int32, device=batch.device))
        st = L.Stats()
        for k, v in self.stats_t.items():
            setattr(st, k, v.data_ptr())
        self._stats = st

    @staticmethod
    def workspace_bytes(batch, params):
        nbytes = C.c_size_t(0)
        L.check(L.load().mpxkkt_workspace_size(C.byref(batch.problem), C.byref(params), C.byref(nbytes)))
        return int(nbytes.value)

    def iterate(self, it):
        """One SQP iteration in place on ``it``; returns the stats tensors."""
        ci = self.batch.c_iterate(it)
        L.check(self.lib.mpxkkt_sqp_iterate(_stream(), C.byref(self.batch.problem),
                                            C.byref(self.params), C.byref(ci), C.byref(self._stats),
                                            C.c_void_p(self.ws.data_ptr()), self.ws_bytes))
        return self.stats_t

    def solve(self, it, max_iter=100, dcost_tol=1e-3, cc_tol=1e-5, check_every=0):
        """SQP iterations in place on ``it`` until runOffline's stop rule holds
        for every instance (mpx/utils/mpc_wrapper.py:425-467).  Returns
        (iters [B] int32, cost [B], cc [B], iterations enqueued): the
        iterations each instance ran, sum g_t^2 and sum |c|^2 after its last
        one.  ``check_every`` = 0 enqueues max_iter + 1 gated iterations with
        no host synchronisation; k > 0 reads the running count every k."""
        b = self.batch
        iters = torch.zeros(b.B, dtype=torch.int32, device=b.device)
        cost = torch.zeros(b.B, dtype=b.dtype, device=b.device)
        cc = torch.zeros(b.B, dtype=b.dtype, device=b.device)
        opts = L.default_solve_opts(max_iter=int(max_iter), dcost_tol=float(dcost_tol),
                                    cc_tol=float(cc_tol), check_every=int(check_every))
        launched = C.c_int32(0)
        ci = b.c_iterate(it)
        L.check(self.lib.mpxkkt_solve(_stream(), C.byref(b.problem), C.byref(self.params), C.byref(opts),
                                      C.byref(ci), C.byref(self._stats), _ptr(iters), _ptr(cost), _ptr(cc),
                                      C.c_void_p(self.ws.data_ptr()), self.ws_bytes, C.byref(launched)))
        return iters, cost, cc, launched.value

    def capture_solve(self, it, max_iter=100, dcost_tol=1e-3, cc_tol=1e-5):
        """Record the whole solve (max_iter + 1 gated iterations and their stop
        tests, no host synchronisation) as one HIP graph on the iterate ``it``;
        ``replay()`` re-runs it from whatever ``it`` then holds."""
        return SolveGraph(self, it, max_iter, dcost_tol, cc_tol)

    def kkt(self, it):
        out = torch.zeros(self.batch.B, dtype=self.batch.dtype, device=self.batch.device)
        ci = self.batch.c_iterate(it)
        L.check(self.lib.mpxkkt_kkt_residual(_stream(), C.byref(self.batch.problem),
                                             C.byref(self.params), C.byref(ci),
                                             C.c_void_p(out.data_ptr()),
                                             C.c_void_p(self.ws.data_ptr()), self.ws_bytes))
        return out

    def model_eval(self, X, U):
        b = self.batch
        g = torch.zeros((b.B, b.N + 1), dtype=b.dtype, device=b.device)
        c = torch.zeros((b.B, b.N + 1, b.n), dtype=b.dtype, device=b.device)
        L.check(self.lib.mpxkkt_model_eval(_stream(), C.byref(b.problem), _ptr(X), _ptr(U), _ptr(g),
                                           _ptr(c)))
        return g, c

    def shift(self, it, shift, u_reset=None, reset=None):
        return shift_warmstart(self.batch, it, shift, u_reset, reset)

    def capture(self, it, iters=1, shift=None):
        """Record ``iters`` SQP iterations on the iterate ``it`` (and, if
        ``shift`` is given, the warm-start shift after them) as one HIP graph.

        The graph reads and writes the device arrays of ``it`` and of the
        batch in place: refresh them with ``copy_`` between replays.  Every
        kernel of the iteration is a plain launch (no memset or host sync), so
        the captured graph replays the eager launch sequence exactly.
        """
        return SQPGraph(self, it, iters, shift)


class SQPGraph:
    """A captured sequence of SQP iterations; ``replay()`` launches it on the
    current stream.  ``stats`` and ``reset`` are the tensors the last
    captured iteration and the shift write."""

    def __init__(self, solver, it, iters, shift):
        dev = solver.batch.device
        self.it = it
        self.reset = (torch.zeros(solver.batch.B, dtype=torch.int32, device=dev)
                      if shift is not None else None)
        # one eager pass on a copy first, so that one-time set-up (code object
        # load, LDS attribute) happens outside the capture
        scratch = {k: v.clone() for k, v in it.items()}
        solver.iterate(scratch)
        side = torch.cuda.Stream(device=dev)
        side.wait_stream(torch.cuda.current_stream(dev))
        self.graph = torch.cuda.CUDAGraph()
        with torch.cuda.graph(self.graph, stream=side):
            for _ in range(iters):
                solver.iterate(it)
            if shift is not None:
                solver.shift(it, shift, reset=self.reset)
        torch.cuda.current_stream(dev).wait_stream(side)
        self.stats = solver.stats_t

    def replay(self):
        self.graph.replay()
        return self.stats


class SolveGraph:
    """A captured multi-iteration solve; ``iters``, ``cost`` and ``cc`` are
    the tensors each replay writes."""

    def __init__(self, solver, it, max_iter, dcost_tol, cc_tol):
        b = solver.batch
        dev = b.device
        self.it = it
        self.iters = torch.zeros(b.B, dtype=torch.int32, device=dev)
        self.cost = torch.zeros(b.B, dtype=b.dtype, device=dev)
        self.cc = torch.zeros(b.B, dtype=b.dtype, device=dev)
        opts = L.default_solve_opts(max_iter=int(max_iter), dcost_tol=float(dcost_tol), cc_tol=float(cc_tol),
                                    check_every=0)
        launched = C.c_int32(0)

        def enqueue(target):
            ci = b.c_iterate(target)
            L.check(solver.lib.mpxkkt_solve(_stream(), C.byref(b.problem), C.byref(solver.params),
                                            C.byref(opts), C.byref(ci), C.byref(solver._stats),
                                            _ptr(self.iters), _ptr(self.cost), _ptr(self.cc),
                                            C.c_void_p(solver.ws.data_ptr()), solver.ws_bytes,
                                            C.byref(launched)))
        # one eager pass on a copy first: one-time set-up stays outside the capture
        enqueue({k: v.clone() for k, v in it.items()})
        side = torch.cuda.Stream(device=dev)
        side.wait_stream(torch.cuda.current_stream(dev))
        self.graph = torch.cuda.CUDAGraph()
        with torch.cuda.graph(self.graph, stream=side):
            enqueue(it)
        torch.cuda.current_stream(dev).wait_stream(side)

    def replay(self):
        self.graph.replay()
        return self.iters
