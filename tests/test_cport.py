"""The FP64 C restatement (oracle/pdilqr_c.c, CPU baseline) reproduces the numpy
oracle's iterates: two independent implementations of the same decisions."""
import numpy as np
import pytest

from mpx_amd import synthetic as syn
from oracle import cport
from oracle import pdilqr as orc

pytestmark = pytest.mark.skipif(not cport.available(), reason="oracle C port not built (make -C oracle)")


@pytest.mark.parametrize("wl,shape,start,iters", [
    (1, dict(n=6, m=3, N=9), "tile", 1),
    (2, dict(n=8, m=4, N=12), "rollout", 4),
    (3, dict(n=8, m=4, N=10, p=2), "rollout", 3),
    (4, dict(n=8, m=6, N=8, p=2), "zero", 6),
    (3, dict(N=12), "rollout", 2),
])
def test_c_port_matches_numpy_oracle(wl, shape, start, iters):
    insts = [syn.make_instance(wl, i, **shape) for i in range(3)]
    starts = [syn.initial_iterate(d, start) for d in insts]
    out = cport.solve_batch(insts, starts, iters, threads=2)
    for i in range(3):
        ref = orc.solve(insts[i], *starts[i], iters)
        for k, key in enumerate(("X", "U", "V", "Yeq", "Ybox")):
            b = ref[-1][key]
            if b.size:
                assert np.max(np.abs(out[k][i] - b)) / max(1.0, np.max(np.abs(b))) < 1e-11, key
        assert np.allclose(out[5][i], [r["alpha"] for r in ref], rtol=0, atol=1e-14)


def test_fp32_build_is_fp32_arithmetic():
    """The FP32 build of the restatement (the error model of the device's FP32
    path) differs from FP64 at FP32 rounding level, neither 0 nor large."""
    insts = [syn.make_instance(3, i, n=10, m=4, N=12, p=3) for i in range(4)]
    starts = [syn.initial_iterate(d, "rollout") for d in insts]
    a = cport.solve_batch(insts, starts, 2)
    b = cport.solve_batch(insts, starts, 2, f32=True)
    for x, y in zip(a[:3], b[:3]):
        err = np.max(np.abs(x - y)) / max(1.0, np.max(np.abs(x)))
        assert 1e-9 < err < 1e-3, err
