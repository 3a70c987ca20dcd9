"""mpxkkt_sqp_iterate_noalloc (the library's static device workspace, no
allocation) through the C ABI against mpxkkt_sqp_iterate with a caller
workspace: the same kernels on the same inputs, so every output must be
bit-identical.  Two calls in a row, the second on a smaller problem (the
static array then holds the first call's contents), and a problem too large
for the array, which is refused before any device work."""
import ctypes as C

import numpy as np
import pytest

torch = pytest.importorskip("torch")

from mpx_amd import synthetic as syn  # noqa: E402

pytestmark = pytest.mark.gpu

KEYS = ("X", "U", "V", "Yeq", "Ybox")


def _pair(wl, B, shape, start, dtype):
    from mpx_amd.problem import DeviceBatch, SQPSolver
    insts = [syn.make_instance(wl, i, **shape) for i in range(B)]
    starts = [syn.initial_iterate(d, start) for d in insts]
    batch = DeviceBatch.from_instances(insts, dtype=dtype)
    host = [np.stack([s[k] for s in starts]) for k in range(5)]
    return batch, SQPSolver(batch), batch.iterate_from(*host), batch.iterate_from(*host)


def _noalloc(solver, batch, it):
    from mpx_amd import _lib as L
    b = batch
    stats = {k: torch.zeros_like(v) for k, v in solver.stats_t.items()}
    st = L.Stats()
    for k, v in stats.items():
        setattr(st, k, v.data_ptr())
    ci = b.c_iterate(it)
    stream = C.c_void_p(torch.cuda.current_stream().cuda_stream)
    L.check(L.load().mpxkkt_sqp_iterate_noalloc(stream, C.byref(b.problem), C.byref(solver.params),
                                                C.byref(ci), C.byref(st)))
    return stats


@pytest.mark.parametrize("dtype", [torch.float32, torch.float64])
def test_noalloc_equals_workspace_entry_point(dtype):
    from mpx_amd import _lib as L
    B1 = 48 if dtype == torch.float32 else 24         # within the 64 MiB static workspace
    cases = [(3, B1, dict(n=36, m=12, N=50, p=8), "rollout", 2),     # first call: larger
             (2, 12, dict(n=10, m=4, N=20), "rollout", 3)]            # second call: smaller
    for wl, B, shape, start, iters in cases:
        print(f"noalloc: workload {wl} B={B} {shape} {dtype}", flush=True)
        batch, solver, it_a, it_b = _pair(wl, B, shape, start, dtype)
        assert solver.ws_bytes <= L.NOALLOC_WORKSPACE_BYTES
        for k in range(iters):
            sa = solver.iterate(it_a)
            sb = _noalloc(solver, batch, it_b)
            torch.cuda.synchronize()
            for key in KEYS:
                assert torch.equal(it_a[key], it_b[key]), (wl, k, key)
            for key in sa:
                assert torch.equal(sa[key], sb[key]), (wl, k, key)


def test_noalloc_refuses_problem_beyond_static_workspace():
    from mpx_amd import _lib as L
    batch, solver, it_a, _ = _pair(3, 512, dict(n=36, m=12, N=50, p=8), "rollout", torch.float32)
    assert solver.ws_bytes > L.NOALLOC_WORKSPACE_BYTES
    ci = batch.c_iterate(it_a)
    st = L.load().mpxkkt_sqp_iterate_noalloc(None, C.byref(batch.problem), C.byref(solver.params),
                                             C.byref(ci), None)
    assert st == L.ERR_WORKSPACE
    assert b"static workspace" in L.load().mpxkkt_last_error()
