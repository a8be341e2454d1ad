"""Adaptable moves (chromo/mc/moves.pyx:58-369) and the five move kinds.

The move functions are named markers: the proposals themselves
(chromo/mc/move_funcs.pyx) run inside the HIP kernels, selected by name.
"""
from __future__ import annotations

import numpy as np
import pandas as pd

from ..util import mc_stat


class _MoveKind:
    """Identifies one of chromo's moves; ``__name__`` matches the reference function."""

    def __init__(self, name: str, index: int):
        self.__name__ = name
        self.index = index

    def __call__(self, *args, **kwargs):
        raise TypeError(f"{self.__name__} proposals run on the GPU inside mc_sim / mc_step")

    def __repr__(self):
        return f"<move {self.__name__}>"


crank_shaft = _MoveKind("crank_shaft", 0)
end_pivot = _MoveKind("end_pivot", 1)
slide = _MoveKind("slide", 2)
tangent_rotation = _MoveKind("tangent_rotation", 3)
change_binding_state = _MoveKind("change_binding_state", 4)
move_list = [crank_shaft, end_pivot, slide, tangent_rotation, change_binding_state]


class MCAdapter:
    """Amplitudes, counters and acceptance tracking of one move (moves.pyx:58-130)."""

    def __init__(self, log_dir: str, log_file_prefix: str, move_func, moves_in_average: float,
                 init_amp_bead: int, init_amp_move: float):
        self.name = move_func.__name__
        self.move_func = move_func
        self.amp_move = float(init_amp_move)
        self.num_per_cycle = 1
        self.amp_bead = int(init_amp_bead)
        self.num_attempt = 0
        self.num_success = 0
        self.move_on = 1
        self.last_amp_move = 0
        self.last_amp_bead = 0
        # `moves_in_average` is a C float in the reference's signature
        self.acceptance_tracker = mc_stat.AcceptanceTracker(
            log_dir, log_file_prefix, float(np.float32(moves_in_average)))

    def __str__(self):
        return f"MCAdapter<{self.name}>"

    def to_file(self, path):
        pass


class Bounds:
    """Named amplitude bounds per move (moves.pyx:302-342)."""

    def __init__(self, name: str, bounds: dict):
        self.name = name
        self.bounds = bounds

    def to_dataframe(self):
        names = self.bounds.keys()
        arr = np.atleast_2d(np.array(list(self.bounds.values())).flatten())
        cols = pd.MultiIndex.from_product([names, ('lower_bound', 'upper_bound')])
        return pd.DataFrame(arr, columns=cols)

    def to_csv(self, path):
        return self.to_dataframe().to_csv(path)

    def to_file(self, path):
        return self.to_csv(path)


def all_moves(log_dir: str):
    """MCAdapters for all five moves (moves.pyx:350-369)."""
    return [MCAdapter(log_dir + '/' + m.__name__, "", m, 20, 1, 0.05) for m in move_list]
