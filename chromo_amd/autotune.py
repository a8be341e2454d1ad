"""Launch-shape autotuner for the persistent Monte Carlo kernel.

The persistent launch (``chromo_mc_run_tuned``) runs each replica's chain with
W wavefronts: W = 1 evaluates one proposal at a time; W = 2 or 4 evaluate that
many proposals of a controller block at once and resolve them in order.  All
shapes give the same chains, so the choice is purely a matter of speed, and the
best one depends on the replica mix (chain lengths, grids, move amplitudes) and
on how many replicas share a compute unit.  ``autotune`` times each candidate on
the live replicas and stores the fastest in ``DeviceReplicas.waves``.

The timed sweeps advance the replicas like any other sweeps (they are ordinary
Monte Carlo steps of the same chains).
"""
from __future__ import annotations

from typing import Dict, Optional, Sequence

import torch

from .engine import DeviceReplicas


def autotune(dr: DeviceReplicas, *, sweeps: int = 2, candidates: Optional[Sequence[int]] = None,
             repeats: int = 1) -> Dict[int, float]:
    """Times ``sweeps`` sweeps per candidate W (seconds per sweep, best of
    ``repeats``), sets ``dr.waves`` to the fastest and returns the timings."""
    cands = list(candidates or dr.WAVE_CHOICES)
    times: Dict[int, float] = {}
    for w in cands:
        dr.run(1, waves=w)                       # first launch of this shape: not timed
        best = float("inf")
        for _ in range(repeats):
            start = torch.cuda.Event(enable_timing=True)
            stop = torch.cuda.Event(enable_timing=True)
            start.record()
            dr.run(sweeps, sync=False, waves=w)
            stop.record()
            torch.cuda.synchronize(dr.device)
            best = min(best, start.elapsed_time(stop) / 1000.0 / sweeps)
        dr.raise_on_error()
        times[w] = best
    dr.waves = min(times, key=times.get)
    return times
