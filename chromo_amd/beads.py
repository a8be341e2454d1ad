"""Beads of a polymer (chromo/beads.py:19-446).

Only the two attributes the hot path reads are modelled: ``vol``
(4/3 pi rad^3, the bead volume in the field's volume fraction) and ``binders``
(the registered ReaderProtein objects for the polymer's binder names).
"""
from __future__ import annotations

from typing import Optional, Sequence

import numpy as np

from .binders import get_by_name


class Bead:
    """Bead with position, orientation and binder states (chromo/beads.py:19-80)."""

    def __init__(self, id_: int, r: np.ndarray, t3: Optional[np.ndarray] = None,
                 t2: Optional[np.ndarray] = None, states: Optional[np.ndarray] = None,
                 binder_names: Optional[Sequence[str]] = None):
        self.id = id_
        self.r = r
        self.t3 = t3
        self.t2 = t2
        self.states = states
        self.binder_names = binder_names
        if binder_names is not None:
            self.binders = [get_by_name(name) for name in binder_names]
        else:
            self.binders = None


class GhostBead(Bead):
    """Non-interacting spherical bead (chromo/beads.py:102-170)."""

    def __init__(self, id_: int, r: np.ndarray, *, t3=None, t2=None, states=None,
                 binder_names=None, rad: Optional[float] = 5, **kwargs):
        super().__init__(id_, r, t3, t2, states, binder_names)
        self.rad = rad
        self.vol = (4 / 3) * np.pi * rad ** 3
        self.kwargs = kwargs


class Nucleosome(Bead):
    """Nucleosome bead with a spherical excluded volume (chromo/beads.py:369-446)."""

    def __init__(self, id_: int, r: np.ndarray, *, t3: np.ndarray, t2: np.ndarray,
                 rad: Optional[float] = 5, states=None, binder_names=None):
        super().__init__(id_, r, t3, t2, states, binder_names)
        self.rad = rad
        self.vol = (4 / 3) * np.pi * rad ** 3
