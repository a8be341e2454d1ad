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
import pandas as pd

from .mc_controller import (Controller, NoControl, SimpleControl, all_moves,
                            all_moves_except_binding_state, specific_move)
from .mc_sim import mc_sim, mc_step, seed_libc
from .moves import Bounds, MCAdapter
from ..util.poly_stat import (find_polymers_in_output_dir, get_latest_configuration,
                              get_latest_simulation)


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


def _unique_subfolder(root: Path) -> Path:
    """``<root><i>`` and its ``acceptance_trackers`` folder for the first free i
    (mkdir is atomic, so concurrent runs get different folders;
    chromo/util/reproducibility.py:325-359)."""
    i = 1
    while True:
        folder = Path(f"{root}{i}")
        try:
            folder.mkdir(parents=True)
            (folder / "acceptance_trackers").mkdir()
            return folder
        except FileExistsError:
            i += 1


def polymer_in_field(polymers: List, binders, field, num_save_mc: int, num_saves: int,
                     bead_amp_bounds: Bounds, move_amp_bounds: Bounds,
                     mc_move_controllers: Optional[List[Controller]] = None,
                     random_seed: Optional[int] = 0, mu_schedule=None,
                     output_dir: Optional[str] = '.', launch: str = "persistent",
                     write_snapshots: bool = True, **kwargs):
    """Save-point loop of chromo (mc/__init__.py:36-161): one mc_sim per save point.

    As the reference's ``make_reproducible`` wrapper, each call writes into a new
    ``sim_<i>`` folder of ``output_dir``: the field description, one CSV snapshot
    per polymer and save point, and per controller and save point the
    acceptance-tracker file (``acceptance_trackers/<move>_snap_<k>.csv``; the
    reference's mc_sim logs no individual moves into it, so it holds the header).
    Returns the polymers; the run's folder is ``polymers[0].output_dir``.
    """
    sim_dir = _unique_subfolder(Path(output_dir) / "sim_")
    np.random.seed(random_seed)
    if mc_move_controllers is None:
        mc_move_controllers = all_moves(
            log_dir=str(sim_dir), bead_amp_bounds=bead_amp_bounds.bounds,
            move_amp_bounds=move_amp_bounds.bounds, controller=SimpleControl)
    if hasattr(field, "to_file"):
        field.to_file(str(sim_dir / "field"))
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
                poly.to_csv(str(sim_dir / Path(f"{poly.name}-{mc_count}.csv")))
        for controller in mc_move_controllers:
            tracker = controller.move.acceptance_tracker
            if Path(tracker.log_dir).is_dir():
                tracker.create_log_file(mc_count)
                tracker.save_move_log(snapshot=mc_count)
    for poly in polymers:
        poly.output_dir = str(sim_dir)
    polymers[0].runtime_s = process_time() - t0
    with open(sim_dir / "__completion_time__", "w") as f:
        f.write(str(pd.Timestamp.now()))
    return polymers


def continue_polymer_in_field_simulation(polymer_class, binders, field, output_dir: str,
                                         num_save_mc: int, num_saves: int,
                                         mc_move_controllers: Optional[List[Controller]] = None,
                                         random_seed: Optional[int] = 0, **kwargs):
    """Resume from the latest snapshots of the latest run in ``output_dir``
    (chromo/mc/__init__.py:167-227): the polymers are read back from their CSV
    snapshots, the field is rebuilt around them (its densities recomputed), and
    a new run folder continues with fresh controllers and streams seeded by
    ``random_seed``, as the reference does.  (``DeviceReplicas.checkpoint`` /
    ``restore`` continue the streams and controllers exactly instead.)"""
    latest = get_latest_simulation(output_dir)
    run_dir = f"{output_dir}/{latest}"
    names = find_polymers_in_output_dir(run_dir)
    paths = [get_latest_configuration(polymer_prefix=n, directory=run_dir) for n in names]
    polymers = [polymer_class.from_file(p, n) for p, n in zip(paths, names)]
    field.polymers = polymers
    field.n_polymers = len(polymers)
    if hasattr(field, "update_all_densities_for_all_polymers"):
        field.update_all_densities_for_all_polymers()
    bead_amp_bounds, move_amp_bounds = get_amplitude_bounds(polymers)
    return polymer_in_field(polymers, binders, field, num_save_mc, num_saves, bead_amp_bounds,
                            move_amp_bounds, mc_move_controllers=mc_move_controllers,
                            random_seed=random_seed, output_dir=output_dir, **kwargs)


__all__ = ["mc_sim", "mc_step", "seed_libc", "polymer_in_field", "continue_polymer_in_field_simulation",
           "get_amplitude_bounds", "Bounds",
           "MCAdapter", "Controller", "NoControl", "SimpleControl", "all_moves",
           "all_moves_except_binding_state", "specific_move"]
