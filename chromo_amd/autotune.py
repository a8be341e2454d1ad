"""Launch-shape autotuner for the persistent Monte Carlo kernel.

Two knobs set how the replicas' chains are laid on the GPU; every setting gives
the same chains bit for bit, so the choice is purely one of speed:

* ``waves`` -- wavefronts per replica (``chromo_mc_run_tuned``): 1 evaluates one
  proposal at a time; 2 evaluates two proposals of a controller block at once
  and resolves them in order.  A replica is one workgroup, so with 512 replicas
  on 256 compute units this also sets the wavefronts resident per SIMD.
* ``sweeps_per_launch`` -- sweeps of every replica in one kernel launch (one
  replay of a captured HIP graph).  One sweep per launch ends every launch on
  the slowest replica's sweep; several let each replica run ahead through
  its own sweeps, so the launch ends on the slowest replica's sum instead.

The best point depends on the replica mix (chain lengths, grids, move
amplitudes).  ``autotune`` times every pair on the live replicas and stores the
fastest in ``DeviceReplicas.waves`` / ``DeviceReplicas.sweeps_per_launch``.
The timed sweeps are ordinary Monte Carlo steps of the same chains.
"""
from __future__ import annotations

from typing import Dict, Optional, Sequence, Tuple

import torch

from .engine import DeviceReplicas


def autotune(dr: DeviceReplicas, *, sweeps: int = 20, waves: Optional[Sequence[int]] = None,
             sweeps_per_launch: Sequence[int] = (1,), repeats: int = 1) -> Dict[Tuple[int, int], float]:
    """Times about ``sweeps`` sweeps per (waves, sweeps_per_launch) pair, each as
    replays of a captured graph of that many sweeps; returns seconds per sweep
    (best of ``repeats``) and keeps the fastest pair on ``dr``."""
    cands_w = list(waves or dr.WAVE_CHOICES)
    times: Dict[Tuple[int, int], float] = {}
    for w in cands_w:
        for s in sweeps_per_launch:
            s = int(s)
            g = dr.capture(s, waves=w)
            g.replay()                                  # first replay of this graph: not timed
            n = max(1, sweeps // s)
            best = float("inf")
            for _ in range(repeats):
                start = torch.cuda.Event(enable_timing=True)
                stop = torch.cuda.Event(enable_timing=True)
                start.record()
                for _ in range(n):
                    g.replay()
                stop.record()
                torch.cuda.synchronize(dr.device)
                best = min(best, start.elapsed_time(stop) / 1000.0 / (n * s))
            dr.raise_on_error()
            times[(w, s)] = best
            del g
    w, s = min(times, key=times.get)
    dr.waves, dr.sweeps_per_launch = w, s
    return times
