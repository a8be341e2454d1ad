"""Polymers (chromo/polymers.pyx:43-1886): PolymerBase, SSWLC, Chromatin.

Same constructors, attributes and class methods as the reference.  Coordinates
and binder states live on the GPU while a simulation runs; ``r``, ``t3``, ``t2``
and ``states`` are properties that copy the device state back on access (and
upload on assignment), so scripts that read or edit them keep working.
"""
from __future__ import annotations

from typing import Callable, Optional

import numpy as np
import pandas as pd

from . import beads as _beads
from .binders import binding_free_energy_table
from .engine import ChainSpec
from .util import dss_params
from .util import poly_paths as paths

E_HUGE = 1E25


def sswlc_bond_params(bead_length, lp: float):
    """dssWLC parameters of each bond (polymers.pyx:1578-1601).

    Returns (delta, eps_bend, gamma, eps_par, eps_perp, eta), each with one value
    per bond, with the reference's arithmetic: delta = l / lp, eps_bend =
    interp(col 1) / delta, gamma = interp(col 2) * delta * lp, eps_par =
    interp(col 3) / (delta * lp**2), eps_perp = interp(col 4) / (delta * lp**2),
    eta = interp(col 5) / lp.
    """
    bead_length = np.asarray(bead_length, dtype=np.double).reshape(-1)
    n = len(bead_length)
    out = [np.zeros(n) for _ in range(6)]
    uniq, inv = np.unique(bead_length, return_inverse=True)
    for k, bl in enumerate(uniq):
        d = bl / lp
        sel = inv == k
        out[0][sel] = d
        out[1][sel] = np.interp(d, dss_params[:, 0], dss_params[:, 1]) / d
        out[2][sel] = np.interp(d, dss_params[:, 0], dss_params[:, 2]) * d * lp
        out[3][sel] = np.interp(d, dss_params[:, 0], dss_params[:, 3]) / (d * lp ** 2)
        out[4][sel] = np.interp(d, dss_params[:, 0], dss_params[:, 4]) / (d * lp ** 2)
        out[5][sel] = np.interp(d, dss_params[:, 0], dss_params[:, 5]) / lp
    return tuple(out)


def binding_tables(binders, marks) -> np.ndarray:
    """F[b, marks, state] tables for a list of ReaderProteins (padded to a common shape)."""
    B = len(binders)
    sites = [b.sites_per_bead for b in binders]
    marks = np.asarray(marks)
    levels = int(max(marks.max(initial=0) + 1, max(sites) + 1 if sites else 1))
    smax = max(sites) if sites else 0
    F = np.zeros((B, levels, smax + 1))
    for b, binder in enumerate(binders):
        t = binding_free_energy_table(binder, levels)
        F[b, :, :t.shape[1]] = t
    return F


class TransformedObject:
    """Holds a 4x4 homogeneous transformation matrix (polymers.pyx:43-62)."""

    def __init__(self):
        self.transformation_mat = np.identity(4, dtype=np.double)


class _LazyBeads(dict):
    """``polymer.beads`` as a mapping id -> bead, built on first access per bead.

    The reference builds one Python bead object per bead up front; only
    ``beads[i].vol`` and ``beads[i].binders`` enter the hot path, so beads are
    created lazily here (a 10^6-bead chain would otherwise need 10^6 objects).
    """

    def __init__(self, poly, factory):
        super().__init__()
        self._poly = poly
        self._factory = factory

    def __missing__(self, i):
        if not 0 <= i < self._poly.num_beads:
            raise KeyError(i)
        b = self._factory(i)
        self[i] = b
        return b

    def __len__(self):
        return self._poly.num_beads

    def __iter__(self):
        return iter(range(self._poly.num_beads))


class PolymerBase(TransformedObject):
    """Base polymer (polymers.pyx:65-831)."""

    def __init__(self, name: str, r=None, *, log_path: str = "", t3=None, t2=None,
                 bead_length=None, lp: float = 0, states=None, binder_names=None,
                 chemical_mods=None, chemical_mod_names=None, max_binders: int = -1):
        super().__init__()
        self._device = None          # (DeviceReplicas, index) while a simulation owns the state
        self._host_stale = False
        self.name = name
        r = np.empty((0, 0)) if r is None else np.asarray(r, dtype=np.double)
        self._r = np.ascontiguousarray(r)
        self._t3 = np.ascontiguousarray(np.empty((0, 0)) if t3 is None else np.asarray(t3, dtype=np.double))
        self._t2 = np.ascontiguousarray(np.empty((0, 0)) if t2 is None else np.asarray(t2, dtype=np.double))
        self.bead_length = np.empty((0,)) if bead_length is None else np.asarray(bead_length, dtype=np.double)
        self._states = (np.empty((0, 0), dtype=int) if states is None
                        else np.ascontiguousarray(np.asarray(states, dtype=np.int64)))
        self.max_binders = max_binders
        self.binder_names = np.empty((0,)) if binder_names is None else np.asarray(binder_names)
        self.chemical_mods = (np.empty((0, 0), dtype=int) if chemical_mods is None
                              else np.ascontiguousarray(np.asarray(chemical_mods, dtype=np.int64)))
        self.chemical_mod_names = (np.empty((0,), dtype=str) if chemical_mod_names is None
                                   else np.asarray(chemical_mod_names))
        self.num_beads = self.get_num_beads()
        self.fill_missing_arguments()
        self.all_inds = np.arange(0, self.num_beads, 1)
        self.num_binders = self.get_num_binders()
        self.n_binders_p1 = self.num_binders + 1
        self.log_path = log_path
        self.check_binders(self._states, self.binder_names)
        self.construct_beads()
        self.lp = lp
        self.required_attrs = np.array(["name", "r", "t3", "t2", "states", "binder_names",
                                        "num_binders", "beads", "num_beads", "lp", "bead_length"])
        self._arrays = np.array(['r', 't3', 't2', 'states', 'chemical_mods', 'bead_length'])
        self._3d_arrays = np.array(['r', 't3', 't2'])
        self._single_values = np.array(["name", "lp", "lt", "bp_wrap"])
        self.last_amp_bead = 0
        self.last_amp_move = 0
        self.direction = np.zeros((3,), dtype='d')
        self.point = np.zeros((3,), dtype='d')
        self.mu_adjust_factor = 1

    # -- device synchronisation ---------------------------------------------------
    def _sync_from_device(self):
        if self._device is not None and self._host_stale:
            dr, i = self._device
            self._r, self._t3, self._t2 = dr.positions(i)
            self._states = dr.binder_states(i)
            self._host_stale = False

    def _push_to_device(self):
        if self._device is not None:
            dr, i = self._device
            dr.set_positions(i, self._r, self._t3, self._t2)
            dr.set_binder_states(i, self._states)

    def _attach(self, dr, i):
        self._device = (dr, i)
        self._host_stale = False

    def _mark_device_newer(self):
        self._host_stale = True

    def _prop(name):
        def get(self):
            self._sync_from_device()
            return getattr(self, "_" + name)

        def set_(self, value):
            self._sync_from_device()
            arr = np.ascontiguousarray(np.asarray(value, dtype=np.int64 if name == "states" else np.double))
            setattr(self, "_" + name, arr)
            if self._device is not None:
                self._push_to_device()
                dr, _ = self._device
                dr.density_stale = True
        return property(get, set_)

    r = _prop("r")
    t3 = _prop("t3")
    t2 = _prop("t2")
    states = _prop("states")
    del _prop

    # -- reference methods ------------------------------------------------------
    def fill_missing_arguments(self):
        """polymers.pyx:302-323."""
        if np.size(self._t3) == 0:
            self._t3 = np.zeros((self.num_beads, 3), dtype='d')
        if np.size(self._t2) == 0:
            self._t2 = np.zeros((self.num_beads, 3), dtype='d')
        if np.size(self._states) == 0:
            self._states = np.zeros((self.num_beads, 1), dtype=np.int64)
        if np.size(self.binder_names) == 0:
            self.binder_names = np.array(["null_reader"])
        if np.size(self.chemical_mods) == 0:
            self.chemical_mods = np.zeros((self.num_beads, 1), dtype=np.int64)
        if np.size(self.chemical_mod_names) == 0:
            self.chemical_mod_names = np.array(["null_mod"])

    def check_binders(self, states, binder_names):
        """polymers.pyx:402-420."""
        if states.shape[1] != 0:
            if states.shape[1] != len(binder_names):
                raise ValueError("Each chemical state must be given a name.")
            if states.shape[0] != len(self._r):
                raise ValueError("Initial epigenetic state of wrong length.")

    def check_attrs(self):
        for attr in self.required_attrs:
            if not hasattr(self, str(attr)):
                raise NotImplementedError("Polymer subclass missing required attribute: " + attr)
        if self.lp == 0:
            raise ValueError("Specify the persistence length in the subclass of Polymer")

    def construct_beads(self):
        self.beads = {}

    def get_num_binders(self) -> int:
        return self._states.shape[1]

    def get_num_beads(self) -> int:
        return self._r.shape[0]

    def is_field_active(self) -> bool:
        return True

    def __str__(self):
        return f"Polymer<{self.name}, nbeads={self.num_beads}, nbinder={self.num_binders}>"

    # -- snapshot format (polymers.pyx:575-792) ---------------------------------
    def to_dataframe(self) -> pd.DataFrame:
        self._sync_from_device()
        vec = {k: getattr(self, "_" + k) for k in ("r", "t3", "t2")}
        vector_arr = np.concatenate(list(vec.values()), axis=1)
        vector_index = pd.MultiIndex.from_product([vec.keys(), ('x', 'y', 'z')])
        vector_df = pd.DataFrame(vector_arr, columns=vector_index)
        if len(self.chemical_mod_names) > 0:
            states_index = pd.MultiIndex.from_tuples([('states', n) for n in self.binder_names])
            chem_index = pd.MultiIndex.from_tuples(
                [('chemical_mods', n) for n in self.chemical_mod_names])
            states_df = pd.DataFrame(np.asarray(self._states), columns=states_index)
            chem_df = pd.DataFrame(np.asarray(self.chemical_mods), columns=chem_index, dtype=int)
            df = pd.concat([vector_df, states_df, chem_df], axis=1)
        else:
            df = vector_df
        df["bead_length"] = np.append(self.bead_length, 0)[:len(df.index)]
        if "max_binders" in self._arrays:
            df["max_binders"] = np.full(len(df.index), self.max_binders)
        for name in self._single_values:
            if hasattr(self, str(name)):
                col = np.full(len(df.index), "", dtype=object)
                col[0] = str(getattr(self, str(name)))
                df[name] = col
        return df

    def to_csv(self, path: str):
        return self.to_dataframe().to_csv(path)

    def to_file(self, path: str):
        return self.to_csv(path)

    @classmethod
    def from_dataframe(cls, df, name=None, **kwargs):
        kwnames = np.unique(df.columns.get_level_values(0))
        kw = {n: np.ascontiguousarray(df[n].to_numpy()) for n in kwnames}
        if 'states' in df:
            kw['binder_names'] = df['states'].columns.to_numpy()
        if 'chemical_mods' in df:
            kw['chemical_mod_names'] = df['chemical_mods'].columns.to_numpy()
        if 'max_binders' in df:
            kw['max_binders'] = int(np.asarray(kw['max_binders']).reshape(-1)[0])
        if "name" in df and name is None:
            kw['name'] = np.asarray(kw['name']).reshape(-1)[0]
        elif name is None:
            kw['name'] = "unnamed"
        else:
            kw['name'] = name
        for key in ("lp", "lt", "bp_wrap"):
            if key in df:
                kw[key] = float(np.asarray(kw[key]).reshape(-1)[0])
        if "bead_length" in df:
            kw['bead_length'] = np.asarray(kw['bead_length'][:-1]).flatten().astype(float)
        for key in ("r", "t3", "t2"):
            if key in kw:
                kw[key] = np.ascontiguousarray(kw[key], dtype=np.double)
        for key in ("states", "chemical_mods"):
            if key in kw:
                kw[key] = np.ascontiguousarray(kw[key], dtype=np.int64)
        kw.update(kwargs)
        return cls(**kw)

    @classmethod
    def from_csv(cls, csv_file: str):
        return cls.from_dataframe(pd.read_csv(csv_file))

    @classmethod
    def from_file(cls, path, name=None, **kwargs):
        if name is None:
            name = path.split("/")[-1].split(".")[0]
        # round_trip parsing: snapshots restore coordinates bit for bit
        df = pd.read_csv(path, header=[0, 1], index_col=0, float_precision="round_trip")
        return cls.from_dataframe(df, name, **kwargs)


class SSWLC(PolymerBase):
    """Stretchable, shearable wormlike chain (polymers.pyx:912-1819)."""

    def __init__(self, name: str, r, *, bead_length, lp: float, bead_rad: float = 5, t3=None,
                 t2=None, states=None, binder_names=None, chemical_mods=None,
                 chemical_mod_names=None, log_path: str = "", max_binders: int = -1, **kwargs):
        self.bead_rad = bead_rad
        super().__init__(name, r, t3=t3, t2=t2, states=states, binder_names=binder_names,
                         bead_length=bead_length, lp=lp, log_path=log_path,
                         chemical_mods=chemical_mods, chemical_mod_names=chemical_mod_names,
                         max_binders=max_binders)
        self.bead_rad = bead_rad
        self.construct_beads()
        bl = np.asarray(self.bead_length, dtype=np.double)
        if bl.ndim == 0 or bl.size == 1 and self.num_beads > 2:
            bl = np.full(self.num_beads - 1, float(bl.reshape(-1)[0]))
            self.bead_length = bl
        self._find_parameters(self.bead_length)
        self.required_attrs = np.array(["name", "r", "t3", "t2", "states", "binder_names",
                                        "num_binders", "beads", "num_beads", "lp", "bead_rad"])
        self._arrays = np.array(['r', 't3', 't2', 'states', 'bead_length', 'chemical_mods',
                                 'max_binders'])
        self.check_attrs()
        self.mu_adjust_factor = 1

    def construct_beads(self):
        names = self.binder_names

        def make(i):
            return _beads.GhostBead(id_=i, r=self._r[i], t3=self._t3[i], t2=self._t2[i],
                                    states=self._states[i], binder_names=names,
                                    rad=getattr(self, "bead_rad", 5))
        self.beads = _LazyBeads(self, make)

    def _find_parameters(self, bead_length):
        """Interpolate dssWLC parameters per bond (polymers.pyx:1545-1601)."""
        p = sswlc_bond_params(bead_length, self.lp)
        self.delta, self.eps_bend, self.gamma, self.eps_par, self.eps_perp, self.eta = p

    def bond_params(self) -> np.ndarray:
        return np.stack([self.eps_bend, self.eps_par, self.eps_perp, self.gamma, self.eta])

    def chain_spec(self) -> ChainSpec:
        """Host description of the chain for the device engine."""
        self._sync_from_device()
        B = self.num_binders
        binders = [self.beads[0].binders[b] for b in range(B)]
        sites = [b.sites_per_bead for b in binders]
        mu = [b.chemical_potential for b in binders]
        mods = np.asarray(self.chemical_mods, dtype=np.int64)
        if mods.shape != (self.num_beads, B):
            mods = np.zeros((self.num_beads, B), dtype=np.int64)
        F = binding_tables(binders, mods)
        return ChainSpec(r=self._r, t3=self._t3, t2=self._t2, states=self._states, marks=mods,
                         bonds=self.bond_params(), sites=sites, mu=mu, F=F,
                         max_binders=int(self.max_binders), mu_adjust=float(self.mu_adjust_factor))

    def compute_E(self) -> float:
        """SSWLC.compute_E (polymers.pyx:1348-1381), evaluated on the GPU."""
        from .energy import elastic_energy
        return elastic_energy(self)

    @classmethod
    def straight_line_in_x(cls, name: str, step_sizes: np.ndarray, **kwargs):
        csum = np.cumsum(step_sizes)
        num_beads = len(step_sizes) + 1
        r = np.zeros((num_beads, 3))
        r[1:, 0] = csum.copy()
        t3 = np.zeros((num_beads, 3))
        t3[:, 0] = 1
        t2 = np.zeros((num_beads, 3))
        t2[:, 1] = 1
        return cls(name, r, t3=t3, t2=t2, bead_length=step_sizes, **kwargs)

    @classmethod
    def gaussian_walk_polymer(cls, name: str, num_beads: int, step_lengths: np.ndarray, **kwargs):
        r = paths.gaussian_walk(num_beads - 1, step_lengths)
        t3, t2 = paths.estimate_tangents_from_coordinates(r)
        return cls(name, r, t3=t3, t2=t2, bead_length=step_lengths, **kwargs)


# polymers.pyx:38-40: base pair length (nm) and the bare natural twist per base pair
LENGTH_BP = 0.332
NATURAL_TWIST_BARE = 2 * np.pi / 10.5


class SSTWLC(SSWLC):
    """SSWLC with twist (polymers.pyx:1889-2320).

    Each bond adds 0.5 eps_twist (omega - omega_0)^2, omega the twist angle
    between the two beads' (t2, t3) frames (compute_twist_angle_omega,
    polymers.pyx:3427-3461), omega_0 = bead_length 2 pi / (10.5 * 0.332 nm) the
    natural twist of DNA, the difference wrapped into [-pi, pi);
    eps_twist = lt / (delta lp).  Moves run with one wavefront per replica.
    """

    def __init__(self, name: str, r, *, bead_length, lp: float, lt: float, bead_rad: float = 5,
                 t3=None, t2=None, states=None, binder_names=None, chemical_mods=None,
                 chemical_mod_names=None, log_path: str = "", max_binders: int = -1, **kwargs):
        self.lt = float(lt)
        super().__init__(name, r, bead_length=bead_length, lp=lp, bead_rad=bead_rad, t3=t3, t2=t2,
                         states=states, binder_names=binder_names, chemical_mods=chemical_mods,
                         chemical_mod_names=chemical_mod_names, log_path=log_path,
                         max_binders=max_binders)
        self.required_attrs = np.array(["name", "r", "t3", "t2", "states", "binder_names",
                                        "num_binders", "beads", "num_beads", "lp", "lt", "bead_rad"])

    def _find_parameters(self, bead_length):
        """SSTWLC._find_parameters (polymers.pyx:1956-2001): the dssWLC parameters
        and eps_twist = lt / (delta lp) per bond."""
        super()._find_parameters(bead_length)
        self.eps_twist = self.lt / (np.asarray(self.delta) * self.lp)

    def natural_twist(self) -> np.ndarray:
        """bead_length * NATURAL_TWIST_BARE / LENGTH_BP per bond (polymers.pyx:2088-2090)."""
        bl = np.asarray(self.bead_length, dtype=np.double).reshape(-1)
        return bl * NATURAL_TWIST_BARE / LENGTH_BP

    def bond_params(self) -> np.ndarray:
        return np.vstack([super().bond_params(), self.eps_twist, self.natural_twist()])

    def chain_spec(self) -> ChainSpec:
        cs = super().chain_spec()
        cs.twist = True
        cs.bead_length = np.asarray(self.bead_length, dtype=np.double).reshape(-1).copy()
        return cs

    def compute_E(self) -> float:
        """SSTWLC.compute_E (polymers.pyx:2250-2286), evaluated on the GPU."""
        from .energy import elastic_energy
        return elastic_energy(self)

    def compute_E_no_twist(self) -> float:
        """SSTWLC.compute_E_no_twist (polymers.pyx:2288-2320): the SSWLC energy of
        the same chain, evaluated on the GPU."""
        from .energy import elastic_energy
        return elastic_energy(self, twist=False)


class Chromatin(SSWLC):
    """SSWLC with the persistence length of DNA, lp = 53 nm (polymers.pyx:1821-1886)."""

    def __init__(self, name: str, r, *, bead_length, bead_rad: float = 5, t3=None, t2=None,
                 states=None, binder_names=None, chemical_mods=None, chemical_mod_names=None,
                 log_path: str = "", max_binders: int = -1, **kwargs):
        kwargs.pop("lp", None)
        super().__init__(name, r, bead_length=bead_length, bead_rad=bead_rad, lp=53.0, t3=t3,
                         t2=t2, states=states, binder_names=binder_names, log_path=log_path,
                         chemical_mods=chemical_mods, chemical_mod_names=chemical_mod_names,
                         max_binders=max_binders)

    def construct_beads(self):
        names = self.binder_names

        def make(i):
            return _beads.Nucleosome(id_=i, r=self._r[i], t3=self._t3[i], t2=self._t2[i],
                                     rad=getattr(self, "bead_rad", 5), states=self._states[i],
                                     binder_names=names)
        self.beads = _LazyBeads(self, make)
