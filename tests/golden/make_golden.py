"""Writes the FP64 oracle fixtures in this folder (run from the repo root:
``python tests/golden/make_golden.py``).

Each ``sqp_w<k>.npz`` holds one small seeded instance of workload k (its
arrays, prefixed ``in_``), a starting iterate, and the oracle's iterates after
each of ``iters`` SQP steps with their line-search diagnostics (prefixed
``out<j>_``).  ``lqr_w1.npz`` holds one LQR subproblem (the linearisation of
workload 1 at its start) with its dense-KKT solution.  The fixtures freeze the
oracle's outputs so that a later change of the oracle or of the device path
shows up as a mismatch against committed data.
"""
import os
import sys

import numpy as np

ROOT = os.path.abspath(os.path.join(os.path.dirname(__file__), "..", ".."))
sys.path.insert(0, ROOT)

from mpx_amd import synthetic as syn  # noqa: E402
from oracle import pdilqr as orc  # noqa: E402

CASES = {
    1: (dict(n=6, m=3, N=9), 1),
    2: (dict(n=8, m=4, N=12), 4),
    3: (dict(n=8, m=4, N=10, p=2), 3),
    4: (dict(n=8, m=6, N=8, p=2), 6),
}
INST_KEYS = ("A", "B", "W", "Q", "R", "QN", "xref", "uref", "D", "E", "e", "ulo", "uhi", "x0")
ITER_KEYS = ("X", "U", "V", "Yeq", "Ybox")
DIAG_KEYS = ("alpha", "rho", "phi0", "dphi", "trials", "status")


def case_arrays(wl):
    shape, iters = CASES[wl]
    inst = syn.make_instance(wl, 7, seed=11, **shape)
    start = syn.initial_iterate(inst, syn.WORKLOADS[wl]["start"])
    out = {"meta": np.array([wl, inst["n"], inst["m"], inst["N"], inst["p"], iters,
                             1 if inst["model"] == "tanh" else 0], dtype=np.int64),
           "h_kappa": np.array([inst["h"], inst["kappa"]])}
    for k in INST_KEYS:
        if inst.get(k) is not None:
            out["in_" + k] = np.asarray(inst[k], dtype=np.float64)
    for k, v in zip(ITER_KEYS, start):
        out["start_" + k] = v
    for j, st in enumerate(orc.solve(inst, *start, iters)):
        for k in ITER_KEYS:
            out[f"out{j}_{k}"] = st[k]
        for k in DIAG_KEYS:
            out[f"out{j}_{k}"] = np.asarray(st[k], dtype=np.float64)
    return out


def lqr_arrays():
    shape, _ = CASES[1]
    inst = syn.make_instance(1, 7, seed=11, **shape)
    X, U, V, Ye, Yb = syn.initial_iterate(inst, "tile")
    L = orc.linearize(inst, X, U, Ye, Yb, orc.DEFAULT_PARAMS["rho_c"])
    sol = orc.lqr_dense(L)
    out = {"L_" + k: np.asarray(v, dtype=np.float64) for k, v in L.items()
           if isinstance(v, np.ndarray) or np.isscalar(v)}
    out.update({"sol_" + k: v for k, v in sol.items()})
    return out


def load_instance(z):
    """Instance dict (the layout of mpx_amd.synthetic) from a fixture."""
    wl, n, m, N, p, _, tanh = (int(v) for v in z["meta"])
    inst = dict(n=n, m=m, N=N, p=p, model="tanh" if tanh else "ltv",
                h=float(z["h_kappa"][0]), kappa=float(z["h_kappa"][1]))
    for k in INST_KEYS:
        inst[k] = z["in_" + k] if ("in_" + k) in z else None
    return inst


if __name__ == "__main__":
    here = os.path.dirname(os.path.abspath(__file__))
    for wl in CASES:
        np.savez_compressed(os.path.join(here, f"sqp_w{wl}.npz"), **case_arrays(wl))
    np.savez_compressed(os.path.join(here, "lqr_w1.npz"), **lqr_arrays())
    print("wrote", sorted(f for f in os.listdir(here) if f.endswith(".npz")))
