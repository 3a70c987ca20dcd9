"""A captured SQP iteration (HIP graph) replays the eager launch sequence:
several replays, with eager work and host synchronisation between them, each
equal bit for bit to the eager iterations from the same state."""
import numpy as np
import pytest

torch = pytest.importorskip("torch")

from mpx_amd import synthetic as syn  # noqa: E402

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("wl,shape,segments,dtype", [
    (3, dict(n=36, m=12, N=50, p=8), 1, torch.float32),
    (2, dict(n=8, m=4, N=12), 3, torch.float64),
    (4, dict(n=8, m=6, N=8, p=2), 1, torch.float32),
])
def test_graph_replays_equal_eager(wl, shape, segments, dtype):
    from mpx_amd.problem import DeviceBatch, SQPSolver
    B, iters = 16, 2
    insts = [syn.make_instance(wl, i, **shape) for i in range(B)]
    starts = [syn.initial_iterate(d, syn.WORKLOADS[wl]["start"]) for d in insts]
    batch = DeviceBatch.from_instances(insts, dtype=dtype)
    it0 = batch.iterate_from(*[np.stack([s[k] for s in starts]) for k in range(5)])
    eager = SQPSolver(batch, segments=segments)
    cap = SQPSolver(batch, segments=segments)
    it_g = {k: v.clone() for k, v in it0.items()}
    graph = cap.capture(it_g, iters=iters, shift=1)
    it_e = {k: v.clone() for k, v in it0.items()}
    for rep in range(3):
        for k in it_g:                     # refresh the graph's static inputs
            it_g[k].copy_(it_e[k])
        for _ in range(iters):
            st_e = eager.iterate(it_e)
        reset_e = eager.shift(it_e, 1)
        torch.cuda.synchronize()
        st_g = graph.replay()
        torch.cuda.synchronize()
        for k in it_e:
            assert torch.equal(it_g[k], it_e[k]), (rep, k)
        for k in ("alpha", "rho", "trials", "status"):
            assert torch.equal(st_g[k], st_e[k]), (rep, k)
        assert torch.equal(graph.reset, reset_e)
        # eager work between replays, as a serving loop interleaves it
        _ = (it_e["X"] * 2.0).sum().item()
