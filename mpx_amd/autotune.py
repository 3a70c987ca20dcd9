"""Launch-shape autotuner for the batched SQP iteration.

For one (n, m, N, p, box, model, dtype, batch) it times every candidate of
  segments      time-parallel scan segments per instance (1 = sequential Riccati;
                more segments cut the latency of a small batch, cost more flops),
  block_threads threads per block of the per-instance kernels,
  trial_group   line-search trials evaluated together (1 = stop at the first
                accepted trial; all = one pass over every trial)
on clones of a real iterate and keeps the fastest.  Results are memoised per
process and can be saved to / loaded from a JSON table.
"""
from __future__ import annotations

import itertools
import sys
import json
import os

import torch

from . import _lib as L
from .problem import SQPSolver

_CACHE = {}


def key_of(batch):
    bucket = 1
    while bucket < batch.B:
        bucket *= 2
    return f"n{batch.n}_m{batch.m}_N{batch.N}_p{batch.p}_box{int(batch.box)}_{batch.model}_" \
           f"{'f64' if batch.dtype == torch.float64 else 'f32'}_B{bucket}"


def candidates(batch):
    segs = [s for s in (1, 2, 4, 8, 16) if s <= max(1, (batch.N + 1) // 2)]
    if batch.B >= 1024:
        segs = [s for s in segs if s <= 2]      # large batches fill the chip without time parallelism
    threads = (64, 128, 256, 512)
    groups = (1, 0)
    return [dict(segments=s, block_threads=t, trial_group=g)
            for s, t, g in itertools.product(segs, threads, groups)]


def _time(solver, it0, iters):
    it = {k: v.clone() for k, v in it0.items()}
    solver.iterate(it)                      # warm-up (also first-touch of the workspace)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(iters):
        solver.iterate(it)
    e1.record()
    torch.cuda.synchronize()
    return e0.elapsed_time(e1) / iters


def tune(batch, it0, iters=3, cands=None, verbose=False):
    """Best tuning dict for ``batch`` (timed on clones of ``it0``)."""
    k = key_of(batch)
    if k in _CACHE:
        return _CACHE[k]
    best, best_ms, table = None, float("inf"), []
    for c in cands or candidates(batch):
        try:
            solver = SQPSolver(batch, **c)
            ms = _time(solver, it0, iters)
        except L.KKTError:
            continue                         # e.g. LDS does not fit this shape
        table.append((ms, c))
        if verbose:
            print(f"  tune {k} {c}: {ms:.3f} ms", file=sys.stderr, flush=True)
        if ms < best_ms:
            best, best_ms = c, ms
        del solver
    if best is None:
        raise L.KKTError(f"no launch configuration runs for {k}")
    _CACHE[k] = dict(best, ms=best_ms)
    return _CACHE[k]


def save(path):
    with open(path, "w") as f:
        json.dump(_CACHE, f, indent=1, sort_keys=True)


def load(path):
    if os.path.exists(path):
        with open(path) as f:
            _CACHE.update(json.load(f))
