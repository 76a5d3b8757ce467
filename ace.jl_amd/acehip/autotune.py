"""Tile-shape autotuner for the product-DAG kernel.

Two shape parameters are free: the number of waves per workgroup (every wave
interprets its share of the run programs for the same 64-environment tile;
the pack carries a task assignment for each supported value) and the number
of workgroups a tile's tasks are split over (1 to 4; Invariant bases),
which shortens the last, partly empty round of workgroups.  Parts 0 is the
library's per-call choice, which also gives the tiles of the last round one
more part than the rest.  The tuner times
every (waves, parts) pair on the caller's batch with events on the launch
stream and keeps the fastest; results are cached per (basis size, batch size,
mode).
"""
from __future__ import annotations

import torch

_CACHE: dict = {}


def tune(ev, batch, forces: bool = True, reps: int = 3, candidates=None, parts=None):
    key = (ev.nAA, ev.nA, batch.nenv, forces)
    if key in _CACHE:
        W, P = _CACHE[key][0]
        ev.set_tile(W)
        ev.set_parts(P)
        return _CACHE[key]
    cands = list(candidates) if candidates else [4, 8, 16]
    pcands = list(parts) if parts else ([0, 1, 2, 3, 4] if ev.prop == 0 else [1])
    results = {}
    for W in cands:
        for P in pcands:
            ev.set_tile(W)
            ev.set_parts(P)
            fn = (lambda: ev.energy_forces(batch)) if forces else (lambda: ev.energy(batch))
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
            results[(W, P)] = best
    Wbest, Pbest = min(results, key=results.get)
    ev.set_tile(Wbest)
    ev.set_parts(Pbest)
    _CACHE[key] = ((Wbest, Pbest), results)
    return _CACHE[key]
