"""Fields (chromo/fields.pyx:46-2376): FieldBase, NullField, UniformDensityField.

``UniformDensityField`` keeps the reference's constructor, grid attributes and
binder table (with the derived 'field_energy_prefactor',
'interaction_energy_intranucleosome' and 'cross_talk_field_energy_prefactor'
columns).  Its voxel densities live on the GPU (``density`` copies them back);
the initial densities come from the deterministic HIP scatter, bit-identical to
the reference's bead-order loop (fields.pyx:2041-2106).
"""
from __future__ import annotations

from typing import List

import numpy as np
import pandas as pd

from .engine import FieldSpec

_CONFINE = {"": 0, "Spherical": 1, "Cubical": 2}

_field_descriptors = ['x_width', 'nx', 'y_width', 'ny', 'z_width', 'nz', 'confine_type',
                      'confine_length', 'chi', 'assume_fully_accessible', 'vf_limit', 'fast_field']


def field_prefactors(binders, vol_bin: float):
    """(pref, intra, cross) of init_field_energy_prefactors (fields.pyx:687-712).

    pref_i = 0.5 eps_i v_i vol_bin, intra_i = eps_i (1 - v_i / vol_bin),
    cross_ij = eps_ij v_i vol_bin when binder j is in i's cross-talk dictionary,
    else 0 (v_i = 4/3 pi r_int^3).
    """
    B = len(binders)
    names = [b.name for b in binders]
    pref = [0.5 * b.interaction_energy * b.interaction_volume * vol_bin for b in binders]
    intra = [b.interaction_energy * (1 - b.interaction_volume / vol_bin) for b in binders]
    cross = np.zeros((B, B))
    for i, b in enumerate(binders):
        for j, nb in enumerate(names):
            if nb in b.cross_talk_interaction_energy.keys():
                cross[i, j] = b.cross_talk_interaction_energy[nb] * b.interaction_volume * vol_bin
    return pref, intra, cross


class FieldBase:
    """A discretisation of space (fields.pyx:46-202)."""

    def __init__(self, polymers: List = None, binders=None):
        self.polymers = [] if polymers is None else list(polymers)
        self.n_polymers = len(self.polymers)
        self.binders = binders
        self.confine_type = ""
        self.confine_length = 0.0

    def __contains__(self, poly):
        return poly in self.polymers

    def __str__(self):
        return "FieldBase"


class NullField(FieldBase):
    """A field that contributes nothing: only the elastic energy acts."""

    def __str__(self):
        return "NullField"

    def field_spec(self) -> FieldSpec:
        return FieldSpec(field_on=False)


class UniformDensityField(FieldBase):
    """Rectilinear voxel density field (fields.pyx:324-2376)."""

    def __init__(self, polymers, binders, x_width, nx, y_width, ny, z_width, nz,
                 confine_type="", confine_length=0.0, chi=1.0, assume_fully_accessible=1,
                 vf_limit=0.5, fast_field=0, n_points=1000):
        super().__init__(polymers=polymers, binders=binders)
        self._field_descriptors = _field_descriptors
        for poly in self.polymers:
            if poly.num_binders != len(binders):
                raise NotImplementedError("For now, all polymers must use all of the same binders.")
        if len(self.polymers) > 1:
            raise NotImplementedError(
                "UniformDensityField supports one polymer (as does the reference's field)")
        if fast_field:
            raise NotImplementedError("fast_field sub-bin tables are not implemented")
        if not assume_fully_accessible:
            raise NotImplementedError("accessible-volume estimation is not implemented")
        if confine_type not in _CONFINE:
            raise ValueError("Confinement type " + str(confine_type) + " not found.")
        self.x_width = float(x_width)
        self.y_width = float(y_width)
        self.z_width = float(z_width)
        self.nx, self.ny, self.nz = int(nx), int(ny), int(nz)
        self.init_grid()
        self.num_binders = len(binders)
        self.init_field_energy_prefactors()
        self.confine_type = confine_type
        self.confine_length = float(confine_length)
        self.assume_fully_accessible = assume_fully_accessible
        self.chi = float(chi)
        self.vf_limit = float(np.float32(vf_limit))
        self.binder_dict = self.binders.to_dict(orient='records')
        self.fast_field = fast_field
        self.n_points = n_points
        self._engine = None
        self.update_all_densities_for_all_polymers()
        self.dict_ = self.get_dict()

    def init_grid(self):
        """fields.pyx:536-575."""
        self.dx = self.x_width / self.nx
        self.dy = self.y_width / self.ny
        self.dz = self.z_width / self.nz
        self.dxyz = np.array([self.dx, self.dy, self.dz])
        self.n_bins = self.nx * self.ny * self.nz
        self.vol_bin = self.x_width * self.y_width * self.z_width / self.n_bins
        self.width_xyz = np.array([self.x_width, self.y_width, self.z_width])
        self.half_width_xyz = np.array([0.5 * self.x_width, 0.5 * self.y_width, 0.5 * self.z_width])
        self.half_step_xyz = np.array([0.5 * self.dx, 0.5 * self.dy, 0.5 * self.dz])
        self.n_xyz_m1 = np.array([self.nx - 1, self.ny - 1, self.nz - 1])

    def init_field_energy_prefactors(self):
        """fields.pyx:687-712, on the binder table (pandas, as in the reference)."""
        names = [self.binders.loc[i, "name"] for i in range(self.num_binders)]
        for i in range(self.num_binders):
            row = self.binders.iloc[i]
            self.binders.at[i, 'field_energy_prefactor'] = (
                0.5 * row.interaction_energy * row.interaction_volume * self.vol_bin)
            self.binders.at[i, 'interaction_energy_intranucleosome'] = (
                row.interaction_energy * (1 - row.interaction_volume / self.vol_bin))
            cross = self.binders.iloc[i].cross_talk_field_energy_prefactor
            talk = self.binders.iloc[i].cross_talk_interaction_energy
            for nb in names:
                if nb in talk.keys():
                    cross[nb] = talk[nb] * self.binders.iloc[i].interaction_volume * self.vol_bin
                else:
                    cross[nb] = 0

    def get_dict(self):
        return {k: getattr(self, k) for k in _field_descriptors if hasattr(self, k)}

    def __str__(self):
        return "UniformDensityField"

    # -- device-facing description ---------------------------------------------
    def field_spec(self) -> FieldSpec:
        B = self.num_binders
        pref = [float(self.binder_dict[i]['field_energy_prefactor']) for i in range(B)]
        intra = [float(self.binder_dict[i]['interaction_energy_intranucleosome']) for i in range(B)]
        names = [self.binder_dict[i]['name'] for i in range(B)]
        cross = np.zeros((B, B))
        for i in range(B):
            for j in range(B):
                cross[i, j] = float(self.binder_dict[i]["cross_talk_field_energy_prefactor"][names[j]])
        bead_vol = self.polymers[0].beads[0].vol if self.polymers else 1.0
        return FieldSpec(field_on=True, nx=self.nx, ny=self.ny, nz=self.nz, x_width=self.x_width,
                         y_width=self.y_width, z_width=self.z_width, vf_limit=self.vf_limit,
                         chi=self.chi, bead_vol=float(bead_vol), pref=pref, intra=intra,
                         cross=cross, confine_type=_CONFINE[self.confine_type],
                         confine_length=self.confine_length)

    def update_all_densities_for_all_polymers(self):
        """Initial voxel densities (fields.pyx:2041-2106), computed on the GPU."""
        if not self.polymers:
            return
        from .mc.mc_sim import engine_for
        eng = engine_for(self, rebuild=True)
        eng.dr.density_init()

    @property
    def density(self) -> np.ndarray:
        from .mc.mc_sim import engine_for
        eng = engine_for(self)
        return eng.dr.density_of(0)

    def compute_E(self, poly) -> float:
        """UniformDensityField.compute_E (fields.pyx:1939-1966), evaluated on the GPU."""
        from .energy import field_energy
        return field_energy(self, poly)

    # -- descriptor file (fields.pyx:953-1036) -----------------------------------
    def to_file(self, path):
        s = pd.Series({k: getattr(self, k) for k in _field_descriptors if hasattr(self, k)})
        s.to_csv(path, header=False)
        return s
