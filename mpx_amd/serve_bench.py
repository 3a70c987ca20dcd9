"""Serving-stream workload: Poisson arrivals of MPC ticks over a mix of shapes.

Each arrival draws its shape from (n, m, N) = (12, 12, 25), (36, 12, 50),
(76, 32, 100) with probabilities 0.5, 0.35, 0.15 and is one tick of an idle
controller of that shape (a new controller opens when none is idle and the
HBM budget allows; a controller closes after ``life`` ticks); every tick needs
1-8 SQP iterations.  An arrival is turned away when the server's queue of
unlaunched ticks is full (``max_pending``) or when it would need a new
controller the budget has no room for; turned-away arrivals count as misses of
infinite latency in ``p99_ms_all_arrivals``.  Latency is measured on the host from the scheduled arrival time
to the moment the tick's first control is back on the host; the loop admits
arrivals, hands finished ticks to their plants and launches pending ticks
without waiting for the device.  The plant of
every controller is its own model, advanced on the host with the returned
control.
"""
from __future__ import annotations

import time

import numpy as np
import torch

from . import synthetic as syn
from .serving import MPCServer


def _plant_step(inst, x, u):
    return x + inst["h"] * (inst["A"] @ x + inst["B"] @ u + inst["kappa"] * np.tanh(inst["W"] @ x))


def run(rate, duration, seed=0, life=200, budget_gb=64.0, max_group=256, n_streams=1, warm=True, log=None,
        max_wall=None, tune="auto", pool_size=32, max_pending=2048):
    """Serve ``rate`` ticks/s for ``duration`` s; returns a summary dict.  With
    ``max_wall`` the run stops after that many seconds even if arrivals are
    still queued; they are reported as ``backlog`` (an overloaded server).
    ``tune``: launch parameters of every bucket ("auto": the autotuner's, chosen
    when the bucket opens, before the clock starts)."""
    rng = np.random.default_rng(seed)
    server = MPCServer(budget_gb * 2**30, max_group=max_group, n_streams=n_streams, tune=tune,
                       max_pending=max_pending)
    shapes, probs = syn.SERVING_SHAPES, syn.SERVING_PROBS
    protos = [syn.make_instance(5, 10**6 + si, seed, n=n, m=m, N=N) for si, (n, m, N) in enumerate(shapes)]
    # controllers draw their model from a pool generated before the clock
    # starts (a fresh initial state each): generating a model in the loop is
    # the plant's business, not the server's
    pool = [[syn.make_instance(5, 1000 * si + j, seed, n=n, m=m, N=N) for j in range(pool_size)]
            for si, (n, m, N) in enumerate(shapes)]
    if warm:
        # build every bucket (and its solvers) before the clock starts
        for inst in protos:
            cid = server.open(inst)
            server.request(cid, inst["x0"], 1)
            server.run_round()
            server.close(cid)
    n_arr = rng.poisson(rate * duration)
    t_rel = np.sort(rng.uniform(0.0, duration, size=n_arr))
    idle, state = {k: [] for k in range(len(shapes))}, {}
    latencies, by_shape, waits, turned_by_shape = [], {}, 0, {}
    t0 = time.perf_counter()
    i = 0
    serial = 0
    while i < n_arr or server.busy:
        now = time.perf_counter() - t0
        if max_wall is not None and now > max_wall:
            break
        while i < n_arr and t_rel[i] <= now:
            k = int(rng.choice(len(shapes), p=probs))
            iters = int(rng.integers(1, 9))
            if not server.accepting or (not idle[k] and not server.room_for(protos[k])):
                # turned away before any set-up work: the queue is full, or no
                # idle controller and no room to open one
                waits += 1
                turned_by_shape[shapes[k]] = turned_by_shape.get(shapes[k], 0) + 1
                i += 1
                continue
            if idle[k]:
                cid = idle[k].pop()
            else:
                n, m, N = shapes[k]
                serial += 1
                inst = dict(pool[k][serial % pool_size])
                inst["x0"] = rng.standard_normal(n)
                cid = server.open(inst)
                state[cid] = dict(inst=inst, x=inst["x0"].copy(), ticks=0, shape=(n, m, N), k=k)
            server.request(cid, state[cid]["x"], iters, t_arrive=t0 + t_rel[i])
            i += 1
        done = server.step()
        if not done:
            if not server.in_flight and i < n_arr:
                time.sleep(max(0.0, min(1e-4, t_rel[i] - (time.perf_counter() - t0))))
            continue
        for r in done:
            lat = r.t_done - r.t_arrive
            latencies.append(lat)
            s = state[r.cid]
            by_shape.setdefault(s["shape"], []).append(lat)
            s["x"] = _plant_step(s["inst"], s["x"], np.clip(r.u0, -1.0, 1.0))
            s["ticks"] += 1
            if s["ticks"] >= life:
                server.close(r.cid)
                del state[r.cid]
            else:
                idle[s["k"]].append(r.cid)
    wall = time.perf_counter() - t0
    torch.cuda.synchronize()     # launches left in flight at the wall-clock cap finish before the buffers go
    lat = np.array(latencies) * 1e3
    backlog = (n_arr - i) + len(server.pending) + sum(len(c[1]) for ln in server.in_flight for c in ln.cohorts)
    # p99 over every arrival: an arrival turned away (or left in the backlog)
    # counts as a miss of infinite latency
    every = np.concatenate([lat, np.full(waits + backlog, np.inf)])
    p99_all = float(np.percentile(every, 99)) if len(every) else None
    out = dict(rate=rate, duration=duration, arrivals=int(n_arr), served=len(lat),
               solves_per_s=len(lat) / wall, backlog=backlog,
               p50_ms=float(np.percentile(lat, 50)) if len(lat) else None,
               p99_ms=float(np.percentile(lat, 99)) if len(lat) else None,
               p99_ms_all_arrivals=p99_all if p99_all is None or np.isfinite(p99_all) else "inf",
               turned_away=waits, turned_away_budget=server.turned_away["budget"],
               turned_away_slots=server.turned_away["slots"], max_pending=max_pending,
               controllers_opened=serial,
               device_bytes_held=server.used,
               by_shape={f"{k[0]}x{k[1]}x{k[2]}": dict(n=len(v), p50_ms=float(np.percentile(v, 50) * 1e3),
                                                      p99_ms=float(np.percentile(v, 99) * 1e3),
                                                      turned_away=turned_by_shape.get(k, 0))
                         for k, v in by_shape.items()})
    if log:
        log(out)
    return out
