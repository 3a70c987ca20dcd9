"""The ``optimizers`` mirror: the reference's call signature
(mpx/utils/mpc_wrapper.py:150-158) maps onto the instance layout the oracle
reads (CPU), and one ``mpc`` call equals one oracle SQP step (GPU, FP64)."""
import numpy as np
import pytest

torch = pytest.importorskip("torch")

from mpx_amd import synthetic as syn  # noqa: E402
from mpx_amd.models import Constraints, LTVDynamics, TanhDynamics, TrackingCost  # noqa: E402
from mpx_amd.optimizers import _instances  # noqa: E402
from oracle import pdilqr as orc  # noqa: E402


def _parts(d):
    if d["model"] == "tanh":
        dyn = TanhDynamics(A=d["A"], B=d["B"], W=d["W"])
    else:
        dyn = LTVDynamics(A=d["A"], B=d["B"])
    cost = TrackingCost(Q=d["Q"][0] if d["Q"].shape[0] == 1 else d["Q"],
                        R=d["R"][0] if d["R"].shape[0] == 1 else d["R"], QN=d["QN"])
    cons = None
    if d["p"] > 0 or d.get("ulo") is not None:
        cons = Constraints(D=d["D"] if d["p"] > 0 else None, E=d["E"] if d["p"] > 0 else None,
                           e=d["e"] if d["p"] > 0 else None, u_lo=d.get("ulo"), u_hi=d.get("uhi"))
    return dyn, cost, cons


def _same(a, b):
    keys = ("A", "B", "Q", "R", "QN", "xref", "uref", "D", "E", "e", "x0")
    assert (a["n"], a["m"], a["N"], a["p"], a["model"]) == (b["n"], b["m"], b["N"], b["p"], b["model"])
    for k in keys + (("W",) if a["model"] == "tanh" else ()):
        assert np.array_equal(np.asarray(a[k]), np.asarray(b[k])), k
    for k in ("ulo", "uhi"):
        assert (a.get(k) is None) == (b.get(k) is None)
        if a.get(k) is not None:
            assert np.array_equal(a[k], b[k])


@pytest.mark.parametrize("wl,shape", [(1, dict(n=5, m=2, N=6)), (2, dict(n=6, m=3, N=7)),
                                      (3, dict(n=6, m=3, N=5, p=2))])
def test_instances_unbatched(wl, shape):
    d = syn.make_instance(wl, 0, **shape)
    dyn, cost, cons = _parts(d)
    insts, batched = _instances(cost, dyn, None, (d["xref"], d["uref"]), d["x0"], cons)
    assert not batched and len(insts) == 1
    _same(insts[0], d)


def test_instances_weight_matrix_and_batch():
    """cost=None takes Q, R from W's diagonal blocks; a leading batch axis on
    x0 and on the model arrays gives one instance per row."""
    ds = [syn.make_instance(2, i, n=5, m=2, N=4) for i in range(3)]
    for d in ds:
        d["QN"] = d["Q"][0].copy()         # from_weights: terminal = stage Q
    n, m = 5, 2
    Wt = np.zeros((n + m, n + m))
    Wt[:n, :n], Wt[n:, n:] = ds[0]["Q"][0], ds[0]["R"][0]
    for d in ds[1:]:
        d["Q"], d["R"], d["QN"] = ds[0]["Q"], ds[0]["R"], ds[0]["QN"]
    dyn = TanhDynamics(A=np.stack([d["A"] for d in ds]), B=np.stack([d["B"] for d in ds]),
                       W=np.stack([d["W"] for d in ds]))
    box = Constraints(u_lo=ds[0]["ulo"], u_hi=ds[0]["uhi"])
    insts, batched = _instances(None, dyn, Wt, np.stack([d["xref"] for d in ds]),
                                np.stack([d["x0"] for d in ds]), box)
    assert batched and len(insts) == 3
    for a, b in zip(insts, ds):
        _same(a, b)


@pytest.mark.gpu
@pytest.mark.parametrize("wl,shape", [(2, dict(n=8, m=4, N=12)), (3, dict(n=8, m=4, N=10, p=2)),
                                      (4, dict(n=8, m=6, N=8, p=2))])
def test_mpc_matches_oracle_step(wl, shape):
    from mpx_amd import optimizers
    d = syn.make_instance(wl, 1, **shape)
    X, U, V, Ye, Yb = syn.initial_iterate(d, syn.WORKLOADS[wl]["start"])
    dyn, cost, cons = _parts(d)
    out = optimizers.mpc(cost, dyn, None, False, (d["xref"], d["uref"]), None, None, d["x0"],
                         X, U, V, constraints=cons, V_eq=Ye if cons else None,
                         V_ineq=Yb if cons else None, dtype=torch.float64)
    ref = orc.sqp_step(d, X, U, V, Ye, Yb)
    names = ("X", "U", "V", "Yeq", "Ybox")[:len(out)]
    for name, got in zip(names, out):
        want = ref[name]
        assert got.shape == want.shape, name
        if want.size:
            assert np.max(np.abs(got - want)) <= 1e-9 * max(1.0, np.max(np.abs(want))), name


@pytest.mark.gpu
def test_mpc_batched_equals_unbatched():
    from mpx_amd import optimizers
    ds = [syn.make_instance(2, i, n=6, m=3, N=9, box=False) for i in range(3)]
    starts = [syn.initial_iterate(d, "rollout") for d in ds]
    dyn = TanhDynamics(A=np.stack([d["A"] for d in ds]), B=np.stack([d["B"] for d in ds]),
                       W=np.stack([d["W"] for d in ds]))
    cost = TrackingCost(Q=ds[0]["Q"][0], R=ds[0]["R"][0], QN=ds[0]["QN"])
    for d in ds[1:]:
        d["Q"], d["R"], d["QN"] = ds[0]["Q"], ds[0]["R"], ds[0]["QN"]
    Xb, Ub, Vb = optimizers.mpc(cost, dyn, None, False, np.stack([d["xref"] for d in ds]), None, None,
                                np.stack([d["x0"] for d in ds]), np.stack([s[0] for s in starts]),
                                np.stack([s[1] for s in starts]), np.stack([s[2] for s in starts]),
                                dtype=torch.float64)
    for i, (d, s) in enumerate(zip(ds, starts)):
        ref = orc.sqp_step(d, *s)
        assert np.max(np.abs(Xb[i] - ref["X"])) <= 1e-9 * max(1.0, np.max(np.abs(ref["X"])))
        assert np.max(np.abs(Ub[i] - ref["U"])) <= 1e-9 * max(1.0, np.max(np.abs(ref["U"])))


@pytest.mark.gpu
def test_model_evaluator_helper():
    from mpx_amd import optimizers
    d = syn.make_instance(3, 2, n=7, m=3, N=6, p=0)
    X, U, _, _, _ = syn.initial_iterate(d, "zero")
    X = X + 0.1 * np.random.default_rng(3).standard_normal(X.shape)
    dyn, cost, _ = _parts(d)
    g, c = optimizers.model_evaluator_helper(cost, dyn, d["x0"], X, U, reference=d["xref"])
    g_ref, c_ref = orc.model_evaluator_helper(d, X, U)
    assert np.allclose(g, g_ref, rtol=1e-12, atol=1e-12)
    assert np.allclose(c, c_ref, rtol=1e-12, atol=1e-12)


def test_mpc_refuses_dropped_hessian_and_parameter():
    """The device model families take no caller Hessian and no stage
    parameters: passing them is an error, not a silently different problem."""
    from mpx_amd import optimizers
    d = syn.make_instance(2, 0, n=5, m=2, N=4)
    X, U, V, _, _ = syn.initial_iterate(d, "rollout")
    dyn, cost, _ = _parts(d)
    with pytest.raises(ValueError, match="Hessian"):
        optimizers.mpc(cost, dyn, lambda *a: None, False, (d["xref"], d["uref"]), None, None, d["x0"], X, U, V)
    with pytest.raises(ValueError, match="parameter"):
        optimizers.mpc(cost, dyn, None, False, (d["xref"], d["uref"]), np.zeros((5, 3)), None, d["x0"], X, U, V)


def _reference_style(inst):
    """The caller problem as the reference's callables: dynamics(x, u, t,
    parameter), cost(W, reference, x, u, t), hessian_approx(W, reference, x,
    u, t) (mpx/utils/models.py:6, mpx/utils/objectives.py:15, :75)."""
    dyn = lambda x, u, t, parameter: inst["f"](x, u, t)
    cost = lambda W, reference, x, u, t: inst["cost"](x, u, t)
    hess = lambda W, reference, x, u, t: inst["hess"](x, u, t)
    jac = lambda x, u, t, parameter: inst["fjac"](x, u, t)
    grad = lambda W, reference, x, u, t: inst["grad"](x, u, t)
    return dyn, cost, hess, jac, grad


@pytest.mark.gpu
@pytest.mark.parametrize("exact", [True, False])
def test_mpc_callable_dynamics_matches_oracle(exact):
    """The reference's own callables (nonlinear dynamics, non-quadratic cost,
    cross term M) through ``mpc``: equal to the oracle's step on the same
    functions, to 1e-9 with exact derivatives and to 1e-5 with the central
    differences used when none are given."""
    from caller_problem import make_caller_instance, start
    from mpx_amd import optimizers
    from mpx_amd.models import Constraints
    inst = make_caller_instance(1, n=7, m=3, N=8, p=2, box=True)
    X, U, V, Ye, Yb = start(inst)
    dyn, cost, hess, jac, grad = _reference_style(inst)
    cons = Constraints(D=inst["D"], E=inst["E"], e=inst["e"], u_lo=inst["ulo"], u_hi=inst["uhi"])
    out = optimizers.mpc(cost, dyn, hess, False, None, None, None, inst["x0"], X, U, V, constraints=cons,
                         V_eq=Ye, V_ineq=Yb, dtype=torch.float64,
                         dynamics_jac=jac if exact else None, cost_grad=grad if exact else None)
    ref = orc.sqp_step(inst, X, U, V, Ye, Yb)
    tol = 1e-9 if exact else 1e-5
    for name, got in zip(("X", "U", "V", "Yeq", "Ybox"), out):
        want = ref[name]
        assert np.max(np.abs(got - want)) <= tol * max(1.0, np.max(np.abs(want))), name
