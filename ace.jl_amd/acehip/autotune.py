"""Launch-shape autotuner of the energy (+ forces) path.

Free shape parameters, all timed on the caller's batch:

* pooling (``k_pool``): neighbours staged per chunk (8, 16 or 32) and, for
  bases of at most 32 pooled values, one or two environments per wave --
  timed through the library's HIP events around the pooling launch;
* product DAG: waves per workgroup (every wave interprets its share of the
  run programs for the same tile of 64 or 32 environments; the pack carries a
  task assignment for each supported value) and the number of workgroups a
  tile's tasks are split over (1 to 4; record kernel), which shortens the
  last, partly empty round of workgroups.  Parts 0 is the library's per-call
  choice, which also gives the tiles of the last round one more part than the
  rest.

Each shape is run once untimed, then ``reps`` times; the fastest is kept.
Results are cached per (basis size, batch size, mode).
"""
from __future__ import annotations

import torch

_CACHE: dict = {}


def _time_call(fn, reps):
    fn()
    torch.cuda.synchronize()
    best = None
    for _ in range(reps):
        s, e = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        s.record()
        fn()
        e.record()
        e.synchronize()
        t = s.elapsed_time(e)
        best = t if best is None else min(best, t)
    return best


def tune_pool(ev, batch, forces: bool = True, reps: int = 3):
    """Pooling chunk width and environments per wave: the fastest pooling
    launch (HIP events around it).  Returns ((chunk, epw), {shape: ms})."""
    fn = (lambda: ev.energy_forces(batch)) if forces else (lambda: ev.energy(batch))
    shapes = [(c, 0) for c in (8, 16, 32)]
    if ev.npool <= 32:
        shapes += [(c, 1) for c in (8, 16, 32)]
    res = {}
    for ch, epw in shapes:
        ev.set_pool_shape(ch, epw)
        fn()
        torch.cuda.synchronize()
        best = None
        for _ in range(reps):
            fn()
            t = ev.last_timing()["pool_ms"]
            best = t if best is None else min(best, t)
        res[(ch, epw)] = best
    ch, epw = min(res, key=res.get)
    ev.set_pool_shape(ch, epw)
    return (ch, epw), res


def tune(ev, batch, forces: bool = True, reps: int = 3, candidates=None, parts=None):
    key = (ev.nAA, ev.nA, batch.nenv, forces)
    if key in _CACHE:
        W, P = _CACHE[key][0]
        ch, epw = _CACHE[key][2][0]
        ev.set_pool_shape(ch, epw)
        ev.set_tile(W)
        ev.set_parts(P)
        return _CACHE[key]
    pool = tune_pool(ev, batch, forces, reps)
    cands = list(candidates) if candidates else [4, 8, 16]
    fn = (lambda: ev.energy_forces(batch)) if forces else (lambda: ev.energy(batch))
    fn()
    torch.cuda.synchronize()
    records = ev.last_dag_shape()[2] in (0, 1)
    pcands = list(parts) if parts else ([0, 1, 2, 3, 4] if records else [1])
    results = {}
    for W in cands:
        for P in pcands:
            ev.set_tile(W)
            ev.set_parts(P)
            results[(W, P)] = _time_call(fn, reps)
    Wbest, Pbest = min(results, key=results.get)
    ev.set_tile(Wbest)
    ev.set_parts(Pbest)
    _CACHE[key] = ((Wbest, Pbest), results, pool)
    return _CACHE[key]
