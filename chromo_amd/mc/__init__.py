"""Monte Carlo driver (chromo/mc/__init__.py:36-332) on the GPU inner loop.

``polymer_in_field`` keeps the reference's seeding (a master numpy stream seeded
with ``random_seed``, one ``inner_seed = randint(0, 1e9)`` per save point) and
its CSV snapshots; each save point is one ``mc_sim`` call.
"""
from __future__ import annotations

from pathlib import Path
from time import process_time
from typing import Callable, Dict, List, Optional, Tuple

import numpy as np

from .mc_controller import (Controller, NoControl, SimpleControl, all_moves,
                            all_moves_except_binding_state, specific_move)
from .mc_sim import mc_sim, mc_step, seed_libc
from .moves import Bounds, MCAdapter


def get_amplitude_bounds(polymers: List) -> Tuple[Bounds, Bounds]:
    """Bead and move amplitude bounds from chain length and bead spacing
    (chromo/mc/__init__.py:295-332)."""
    poly_len = np.min([polymer.r.shape[0] for polymer in polymers])
    min_spacing = np.min([np.min(polymer.bead_length) for polymer in polymers])
    bead_amp_bounds = Bounds("bead_amp_bounds", {
        "crank_shaft": (min(30, poly_len), min(150, poly_len)),
        "slide": (min(10, poly_len), min(150, poly_len)),
        "end_pivot": (min(50, poly_len / 4), min(150, int(poly_len / 2))),
        "tangent_rotation": (1, poly_len),
        "change_binding_state": (1, 1),
    })
    move_amp_bounds = Bounds("move_amp_bounds", {
        "crank_shaft": (0.1 * np.pi, 0.25 * np.pi),
        "slide": (0.2 * min_spacing, 0.3 * min_spacing),
        "end_pivot": (0.2 * np.pi, 0.25 * np.pi),
        "tangent_rotation": (0.05 * np.pi, 0.2 * np.pi),
        "change_binding_state": (0, 0),
    })
    return bead_amp_bounds, move_amp_bounds


def polymer_in_field(polymers: List, binders, field, num_save_mc: int, num_saves: int,
                     bead_amp_bounds: Bounds, move_amp_bounds: Bounds,
                     mc_move_controllers: Optional[List[Controller]] = None,
                     random_seed: Optional[int] = 0, mu_schedule=None,
                     output_dir: Optional[str] = '.', launch: str = "persistent",
                     write_snapshots: bool = True, **kwargs):
    """Save-point loop of chromo (mc/__init__.py:36-161): one mc_sim per save point."""
    np.random.seed(random_seed)
    if mc_move_controllers is None:
        mc_move_controllers = all_moves(
            log_dir=output_dir, bead_amp_bounds=bead_amp_bounds.bounds,
            move_amp_bounds=move_amp_bounds.bounds, controller=SimpleControl)
    t0 = process_time()
    for mc_count in range(num_saves):
        if mu_schedule is not None:
            mu_adjust_factor = mu_schedule.function(mc_count, num_saves)
        else:
            mu_adjust_factor = 1
        inner_seed = np.random.randint(0, 1E9)
        mc_sim(polymers, binders, num_save_mc, mc_move_controllers, field, mu_adjust_factor,
               inner_seed, launch=launch)
        if write_snapshots:
            for poly in polymers:
                poly.to_csv(str(Path(output_dir) / Path(f"{poly.name}-{mc_count}.csv")))
    polymers[0].runtime_s = process_time() - t0
    return polymers


__all__ = ["mc_sim", "mc_step", "seed_libc", "polymer_in_field", "get_amplitude_bounds", "Bounds",
           "MCAdapter", "Controller", "NoControl", "SimpleControl", "all_moves",
           "all_moves_except_binding_state", "specific_move"]
