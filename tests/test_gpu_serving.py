"""Serving: continuous batching gives every controller the same result alone
and in the mix; admission follows the HBM budget; a NaN in one controller
triggers its own fallback and leaves its batch neighbours untouched."""
import numpy as np
import pytest

torch = pytest.importorskip("torch")

from mpx_amd import synthetic as syn  # noqa: E402

pytestmark = pytest.mark.gpu

SHAPES = [(12, 12, 25), (36, 12, 50), (12, 12, 25), (24, 12, 20), (36, 12, 50)]
ITERS = [[1, 3, 2, 8, 1], [2, 2, 5, 1, 4], [3, 1, 1, 2, 6]]


def _server(**kw):
    from mpx_amd.serving import MPCServer
    return MPCServer(8 * 2**30, max_group=8, slots_per_bucket=8, n_streams=3, **kw)


def _insts():
    return [syn.make_instance(5, 100 + i, n=n, m=m, N=N) for i, (n, m, N) in enumerate(SHAPES)]


def _x(inst, k):
    rng = np.random.default_rng([7, k, inst["n"]])
    return inst["x0"] + 0.05 * k * rng.standard_normal(inst["n"])


def test_alone_equals_mix():
    insts = _insts()
    mix = _server()
    cids = [mix.open(d) for d in insts]
    mix_u = {i: [] for i in range(len(insts))}
    for k, iters in enumerate(ITERS):
        reqs = [mix.request(c, _x(d, k), it) for c, d, it in zip(cids, insts, iters)]
        done = mix.run_round()
        assert len(done) == len(insts)
        for i, r in enumerate(reqs):
            mix_u[i].append(r.u0.copy())
    alone = _server()
    cids_a = [alone.open(d) for d in insts]
    for k, iters in enumerate(ITERS):
        for i, (c, d, it) in enumerate(zip(cids_a, insts, iters)):
            r = alone.request(c, _x(d, k), it)
            assert len(alone.run_round()) == 1
            # same launch shape per bucket and no cross-instance arithmetic: identical bits
            assert np.array_equal(r.u0, mix_u[i][k]), (i, k)
    for c_m, c_a in zip(cids, cids_a):
        cm, ca = mix.controllers[c_m], alone.controllers[c_a]
        for key in ("X", "U", "V", "Ybox"):
            assert torch.equal(cm.bucket.state[key][cm.slot], ca.bucket.state[key][ca.slot])


def test_admission_by_budget():
    from mpx_amd.serving import MPCServer
    d = _insts()[1]
    probe = MPCServer(1 << 40, max_group=2, slots_per_bucket=4, n_streams=1)
    need = probe.controller_bytes(d)
    srv = MPCServer(2 * need + need // 2, max_group=2, slots_per_bucket=4, n_streams=1)
    a, b = srv.open(d), srv.open(d)
    assert a is not None and b is not None
    assert srv.open(d) is None              # over budget
    srv.close(a)
    assert srv.open(d) is not None          # freed bytes are reusable


def test_nan_fallback_isolated():
    insts = _insts()[:3]
    clean, dirty = _server(), _server()
    cc = [clean.open(d) for d in insts]
    cd = [dirty.open(d) for d in insts]
    for srv, cids, bad in ((clean, cc, False), (dirty, cd, True)):
        reqs = []
        for i, (c, d) in enumerate(zip(cids, insts)):
            x = _x(d, 1).copy()
            if bad and i == 1:
                x[0] = float("nan")
            reqs.append(srv.request(c, x, 2))
        srv.run_round()
        if bad:
            assert reqs[1].reset and not reqs[0].reset and not reqs[2].reset
        else:
            assert not any(r.reset for r in reqs)
        srv._last = reqs
    for i in (0, 2):
        assert np.array_equal(clean._last[i].u0, dirty._last[i].u0)


def test_overlapped_launches_equal_alone():
    """Requests launched by the non-blocking step() while earlier launches are
    still on the device (a controller's next tick waits for its previous one)
    get the same bits as the same ticks served one at a time."""
    insts = _insts()
    srv = _server()
    cids = [srv.open(d) for d in insts]
    reqs, done = {}, []
    for k, iters in enumerate(ITERS):
        for i, (c, d, it) in enumerate(zip(cids, insts, iters)):
            reqs[(i, k)] = srv.request(c, _x(d, k), it)
            done += srv.step()
    while srv.busy:
        done += srv.step()
    assert len(done) == len(reqs)
    alone = _server()
    cids_a = [alone.open(d) for d in insts]
    for k, iters in enumerate(ITERS):
        for i, (c, d, it) in enumerate(zip(cids_a, insts, iters)):
            r = alone.request(c, _x(d, k), it)
            assert len(alone.run_round()) == 1
            assert np.array_equal(r.u0, reqs[(i, k)].u0), (i, k)


def test_close_drops_pending_ticks():
    insts = _insts()[:2]
    srv = _server()
    a, b = srv.open(insts[0]), srv.open(insts[1])
    srv.request(a, _x(insts[0], 0), 2)
    rb = srv.request(b, _x(insts[1], 0), 2)
    srv.close(a)                      # its queued tick goes with it
    done = srv.run_round()
    assert [r.rid for r in done] == [rb.rid]
    srv.request(b, _x(insts[1], 1), 3)
    srv.step()                        # launched, not yet harvested
    with pytest.raises(RuntimeError):
        srv.close(b)
    while srv.busy:
        srv.step()
    srv.close(b)
