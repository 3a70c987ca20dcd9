"""ctypes driver of the FP64 C restatement (oracle/pdilqr_c.c).

TEST INFRASTRUCTURE and CPU baseline only: used by tests/ (checked against
the numpy oracle) and by bench.py's cpu_baseline leg.
"""
from __future__ import annotations

import ctypes as C
import hashlib
import os
import shutil
import subprocess
import tempfile
import time

import numpy as np

from . import pdilqr as orc

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(_HERE, "_build", "libpdilqr_c.so")
LIB32 = os.path.join(_HERE, "_build", "libpdilqr_c_f32.so")
SRC = os.path.join(_HERE, "pdilqr_c.c")
_lib = None
_lib_path = None
_lib32 = None


def available():
    return os.path.exists(LIB) or shutil.which("gcc") is not None


def native_build(f32=False):
    """Compile the restatement with -march=native for the host this runs on
    (into a temp folder keyed by the source hash); returns the path or None.
    f32: the FP32-arithmetic build."""
    gcc = shutil.which("gcc")
    if gcc is None or not os.path.exists(SRC):
        return None
    with open(SRC, "rb") as f:
        tag = hashlib.sha1(f.read()).hexdigest()[:12]
    out_dir = os.path.join(tempfile.gettempdir(), f"mpx_oracle_{os.getuid()}_{tag}")
    out = os.path.join(out_dir, "libpdilqr_c_native_f32.so" if f32 else "libpdilqr_c_native.so")
    if not os.path.exists(out):
        os.makedirs(out_dir, exist_ok=True)
        tmp = out + f".{os.getpid()}.tmp"
        r = subprocess.run([gcc, "-O3", "-march=native", "-fopenmp", "-fPIC", "-std=c11", "-shared",
                            *(["-DPDILQR_F32"] if f32 else []), "-o", tmp, SRC, "-lm"], capture_output=True)
        if r.returncode != 0:
            return None
        os.replace(tmp, out)
    return out


def _load():
    global _lib, _lib_path
    if _lib is None:
        path = native_build() or LIB
        _lib = C.CDLL(path)
        _lib.pdilqr_solve_batch.restype = C.c_int
        _lib_path = path
    return _lib


def _load32():
    global _lib32
    if _lib32 is None:
        path = native_build(f32=True) or LIB32
        _lib32 = C.CDLL(path)
        _lib32.pdilqr_solve_batch.restype = C.c_int
    return _lib32


def build_flags():
    _load()
    return "-O3 -march=native" if _lib_path != LIB else "-O3 -march=x86-64-v3"


def _arr(a, dt=np.float64):
    return np.ascontiguousarray(a, dtype=dt)


def _ptr(a):
    return a.ctypes.data_as(C.c_void_p) if a is not None else C.c_void_p(0)


def solve_batch(insts, starts, iters, threads=0, params=None, f32=False):
    """iters SQP iterations for same-shaped instances; returns (X, U, V, Yeq, Ybox, alphas)
    (float64 arrays).  f32: the same restatement in FP32 arithmetic, inputs and
    iterates rounded to FP32."""
    if f32:
        return _solve_batch(insts, starts, iters, threads, params, np.float32, _load32())
    return _solve_batch(insts, starts, iters, threads, params, np.float64, _load())


def _solve_batch(insts, starts, iters, threads, params, dt, lib):
    _arr_d = lambda a: _arr(a, dt)
    prm = dict(orc.DEFAULT_PARAMS)
    if params:
        prm.update(params)
    i0 = insts[0]
    n, m, N, p = i0["n"], i0["m"], i0["N"], i0["p"]
    B = len(insts)
    st = lambda k: _arr_d(np.stack([d[k] for d in insts]))
    A, Bm = st("A"), st("B")
    W = st("W") if i0["model"] == "tanh" else None
    Q, R = st("Q"), st("R")
    q_tv, r_tv = int(Q.shape[1] > 1), int(R.shape[1] > 1)
    D = st("D") if p else None
    E = st("E") if p else None
    e = st("e") if p else None
    box = i0.get("ulo") is not None
    ulo = st("ulo") if box else None
    uhi = st("uhi") if box else None
    X = _arr_d(np.stack([s[0] for s in starts]))
    U = _arr_d(np.stack([s[1] for s in starts]))
    V = _arr_d(np.stack([s[2] for s in starts]))
    Yeq = _arr_d(np.stack([s[3] for s in starts]))
    Ybox = _arr_d(np.stack([s[4] for s in starts]))
    pv = _arr_d([prm["rho0"], prm["rho_max"], prm["rho_c"], prm["mu_b"], prm["tau"], prm["beta"],
               prm["c1"], prm["armijo_slack"], prm["max_trials"], prm["hess_floor"]])
    alphas = np.zeros((B, iters), dtype=dt)
    threads = threads or os.cpu_count() or 1
    real = C.c_float if dt == np.float32 else C.c_double
    lib.pdilqr_solve_batch(
        B, n, m, N, p, 0 if i0["model"] == "ltv" else 1, int(box), q_tv, r_tv, real(i0["h"]),
        real(i0["kappa"]), _ptr(A), _ptr(Bm), _ptr(W), _ptr(Q), _ptr(R), _ptr(st("QN")),
        _ptr(st("xref")), _ptr(st("uref")), _ptr(D), _ptr(E), _ptr(e), _ptr(ulo), _ptr(uhi), _ptr(st("x0")),
        _ptr(X), _ptr(U), _ptr(V), _ptr(Yeq), _ptr(Ybox), _ptr(pv), iters, _ptr(alphas), threads)
    f64 = lambda a: np.asarray(a, dtype=np.float64)
    return f64(X), f64(U), f64(V), f64(Yeq), f64(Ybox), f64(alphas)


def bench(insts, starts, iters=3, budget_s=15.0, threads=0):
    """Instance-iterations per second of the C restatement on all host cores,
    on a sample of the instances grown until it takes about budget_s."""
    threads = threads or os.cpu_count() or 1
    k = max(threads, 8)
    while True:
        sample = min(k, len(insts))
        t0 = time.perf_counter()
        solve_batch(insts[:sample], starts[:sample], iters, threads)
        dt = time.perf_counter() - t0
        if dt > budget_s / 4 or sample == len(insts):
            break
        k *= 4
    return dict(inst_iters_per_s=sample * iters / dt, cores=threads, kind="port",
                sample=f"{sample} instances x {iters} SQP iterations of the FP64 C restatement "
                       f"(oracle/pdilqr_c.c, {build_flags()}, OpenMP), {dt:.2f} s")
