"""Reader proteins (chromo/binders.pyx:19-238).

``ReaderProtein`` carries the parameters that enter the hot path:
sites_per_bead, bind_energy_mod / bind_energy_no_mod (single-site binding free
energy), interaction_energy and interaction_radius (field self-interaction),
chemical_potential and the cross-talk dictionary.  As in the reference, the
polymer's binding dE reads these attributes from the registered objects (looked
up by name), while the field reads the table built by
``make_binder_collection``.
"""
from __future__ import annotations

from typing import Dict, List

import numpy as np
import pandas as pd

_REGISTRY: Dict[str, "Binder"] = {}


class Binder:
    """A component binding the polymer (chromo/binders.pyx:19-48)."""

    def __init__(self, name: str, sites_per_bead: int) -> None:
        self.name = name
        self.sites_per_bead = int(sites_per_bead)
        self.binding_seq = np.zeros((self.sites_per_bead,), dtype=int)


class ReaderProtein(Binder):
    """Chemical properties of a reader protein (chromo/binders.pyx:51-130).

    A protein is registered under its name when first constructed, so that
    ``get_by_name`` (used by beads) finds it; the reference finds only objects
    defined in its own module, which made new binder types impossible without
    editing it.
    """

    def __init__(self, name: str, sites_per_bead: int, bind_energy_mod: float,
                 bind_energy_no_mod: float, interaction_energy: float,
                 chemical_potential: float, interaction_radius: float,
                 cross_talk_interaction_energy: dict = None) -> None:
        super().__init__(name, sites_per_bead)
        self.bind_energy_mod = float(bind_energy_mod)
        self.bind_energy_no_mod = float(bind_energy_no_mod)
        self.interaction_energy = float(interaction_energy)
        self.chemical_potential = float(chemical_potential)
        self.interaction_radius = float(interaction_radius)
        self.interaction_volume = (4.0 / 3.0) * np.pi * self.interaction_radius ** 3
        self.field_energy_prefactor = 0.0
        self.interaction_energy_intranucleosome = 0.0
        self.cross_talk_interaction_energy = (
            {} if cross_talk_interaction_energy is None else cross_talk_interaction_energy)
        self.cross_talk_field_energy_prefactor = {}
        if name not in _REGISTRY:
            _REGISTRY[name] = self

    def dict(self) -> dict:
        return {
            "name": self.name,
            "sites_per_bead": self.sites_per_bead,
            "bind_energy_mod": self.bind_energy_mod,
            "bind_energy_no_mod": self.bind_energy_no_mod,
            "interaction_energy": self.interaction_energy,
            "chemical_potential": self.chemical_potential,
            "interaction_radius": self.interaction_radius,
            "interaction_volume": self.interaction_volume,
            "field_energy_prefactor": self.field_energy_prefactor,
            "interaction_energy_intranucleosome": self.interaction_energy_intranucleosome,
            "cross_talk_interaction_energy": self.cross_talk_interaction_energy,
            "cross_talk_field_energy_prefactor": self.cross_talk_field_energy_prefactor,
        }


null_reader = ReaderProtein(
    'null_reader', sites_per_bead=0, bind_energy_mod=0, bind_energy_no_mod=0,
    interaction_energy=0, chemical_potential=0, interaction_radius=0)

hp1 = ReaderProtein(
    'HP1', sites_per_bead=2, bind_energy_mod=-0.01, bind_energy_no_mod=1.52,
    interaction_energy=-4, chemical_potential=-1, interaction_radius=3,
    cross_talk_interaction_energy={'PRC1': 0})

prc1 = ReaderProtein(
    'PRC1', sites_per_bead=2, bind_energy_mod=-0.01, bind_energy_no_mod=1.52,
    interaction_energy=-4, chemical_potential=-1, interaction_radius=3,
    cross_talk_interaction_energy={'HP1': 0})


def register(binder: Binder) -> Binder:
    """Make ``binder`` the object ``get_by_name(binder.name)`` returns."""
    _REGISTRY[binder.name] = binder
    return binder


def get_by_name(name: str) -> Binder:
    """Look up a registered binder (chromo/binders.pyx:185-207)."""
    if name not in _REGISTRY:
        raise ValueError(f"No binders found in {__name__} with name: {name}")
    return _REGISTRY[name]


def make_binder_collection(binders) -> pd.DataFrame:
    """Summary table of binders (chromo/binders.pyx:210-238), one row per binder."""
    if binders is None:
        return None
    if isinstance(binders, (str, Binder)):
        binders = [binders]
    rows: List[dict] = []
    for b in binders:
        if isinstance(b, str):
            b = get_by_name(b)
        rows.append(b.dict())
    base = ['name', 'sites_per_bead', 'bind_energy', 'interaction_energy', 'chemical_potential']
    cols = list(base)
    for r in rows:
        for k in r:
            if k not in cols:
                cols.append(k)
    df = pd.DataFrame(rows, columns=cols)
    return df


def binding_free_energy_table(binder: ReaderProtein, n_mark_levels: int) -> np.ndarray:
    """-log single-site partition function F[marks, state] (polymers.pyx:1485-1517).

    F[m, s] = -log(sum_{i=0..s} C(m, i) C(Nn - m, s - i) exp(-(i e_mod + (s - i) e_nomod)))
    with Nn = sites_per_bead, evaluated with the same numpy / scipy calls as the
    reference so that every table entry is the value the reference computes.
    """
    from scipy.special import comb
    Nn = binder.sites_per_bead
    out = np.zeros((n_mark_levels, Nn + 1))
    with np.errstate(divide='ignore'):
        for Nm in range(n_mark_levels):
            for s in range(Nn + 1):
                i = np.arange(s + 1)
                out[Nm, s] = -np.log(np.sum(
                    comb(Nm, i) * comb(Nn - Nm, s - i) *
                    (np.exp(-(i * binder.bind_energy_mod + (s - i) * binder.bind_energy_no_mod)))))
    return out
