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
_int_field_descriptors = ['nx', 'ny', 'nz', 'n_points']
_str_field_descriptors = ['confine_type']
_float_field_descriptors = ['x_width', 'y_width', 'z_width', 'confine_length', 'vf_limit', 'chi']
_bool_field_descriptors = ['assume_fully_accessible', 'fast_field']


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


class _VolumeMap(dict):
    """``access_vols`` of a large grid: the reference's bin -> volume mapping,
    read from an array instead of holding one Python float per voxel."""

    def __init__(self, vols):
        super().__init__()
        self._v = vols

    def __getitem__(self, k):
        return float(self._v[k])

    def __len__(self):
        return len(self._v)

    def __iter__(self):
        return iter(range(len(self._v)))

    def items(self):
        return ((i, float(x)) for i, x in enumerate(self._v))


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
                "this build runs one polymer per field (the reference's field sums the "
                "densities of all its polymers, fields.pyx:2063-2106)")
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
        self.access_vols = self.get_accessible_volumes(
            n_side=20, assume_fully_accessible=assume_fully_accessible)
        self.chi = float(chi)
        self.vf_limit = float(np.float32(vf_limit))
        self.binder_dict = self.binders.to_dict(orient='records')
        self.fast_field = fast_field
        # init_fast_field (fields.pyx:577-671) makes n_points even; the sub-bin
        # weights are evaluated in the kernels from the same expressions
        self.n_points = n_points + (n_points % 2) if fast_field else n_points
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

    def get_accessible_volumes(self, n_side: int, assume_fully_accessible) -> dict:
        """Accessible volume of every voxel (fields.pyx:714-951): vol_bin, except, in
        a spherical confinement, voxels whose centre lies within sqrt(2)/4 of the
        largest voxel side of the sphere's surface; those get vol_bin times the
        fraction of an n_side^3 grid of points (from the voxel's lower corner, steps
        of side / n_side) inside the sphere, with the reference's numpy arithmetic."""
        vols = np.full(self.n_bins, self.vol_bin)
        if not assume_fully_accessible and self.confine_type == "Spherical":
            nx, ny = self.nx, self.ny
            i = np.arange(self.n_bins)
            xyz = np.stack([i % nx, (i // nx) % ny, i // (nx * ny)], axis=1)
            nxyz = np.array([self.nx, self.ny, self.nz])
            coords = (xyz - (nxyz - 1) / 2) * np.array([self.dx, self.dy, self.dz])
            buffer_dist = np.sqrt(2) / 4 * max(self.dx, self.dy, self.dz)
            dist = np.sqrt(coords[:, 0] ** 2 + coords[:, 1] ** 2 + coords[:, 2] ** 2)
            split = ~((dist < self.confine_length - buffer_dist) |
                      (dist > self.confine_length + buffer_dist))
            total = n_side ** 3
            sub = np.reshape(np.mgrid[0:n_side, 0:n_side, 0:n_side].T, (total, 3)).astype(float)
            sub[:, 0] *= self.dx / n_side
            sub[:, 1] *= self.dy / n_side
            sub[:, 2] *= self.dz / n_side
            half = np.array([self.dx / 2, self.dy / 2, self.dz / 2])
            for v in np.nonzero(split)[0]:
                pts = sub + (coords[v] - half)
                inside = np.count_nonzero(np.linalg.norm(pts, axis=1) < self.confine_length)
                vols[v] = self.vol_bin * (float(inside) / float(total))
        self._access_vol_array = vols
        return {int(k): float(v) for k, v in enumerate(vols)} if self.n_bins <= 1 << 18 else \
            _VolumeMap(vols)

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
                         access_vol=self.access_vol_array(),
                         fast_field=bool(self.fast_field), n_points=int(self.n_points),
                         confine_length=self.confine_length)

    def access_vol_array(self):
        """Per-voxel accessible volumes for the kernels, or None when every voxel
        has the full volume vol_bin."""
        a = getattr(self, "_access_vol_array", None)
        if a is None or np.all(a == self.vol_bin):
            return None
        return a

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
        """The field's descriptors, then one row per polymer ('polymer') and per
        binder ('binder'), as key,value lines (fields.pyx:953-974)."""
        rows = {k: self.dict_[k] for k in _field_descriptors}
        for poly in self.polymers:
            rows[poly.name] = 'polymer'
        for _, binder in self.binders.iterrows():
            rows[binder['name']] = 'binder'
        return pd.Series(rows).to_csv(path, header=False)

    @classmethod
    def from_file(cls, path, polymers, binders):
        """A field saved with ``to_file`` around ``polymers`` and ``binders``, which
        must be the ones the file lists (fields.pyx:976-1036).  As in the
        reference, the two flags come back as 0: the file stores 'True'/'False'
        (or 1/0) and the reference's parser compares the key, not the value, with
        'True'."""
        series = pd.read_csv(path, header=None, index_col=0)[1]
        kwargs = series[_field_descriptors].to_dict()
        for key in kwargs:
            if key in _int_field_descriptors:
                kwargs[key] = int(kwargs[key])
            elif key in _float_field_descriptors:
                kwargs[key] = float(kwargs[key])
            elif key in _str_field_descriptors and pd.isna(kwargs[key]):
                kwargs[key] = ""
            elif key in _bool_field_descriptors:
                kwargs[key] = (key == 'True')
        polymer_names = series[series == 'polymer'].index.values
        binder_names = series[series == 'binder'].index.values
        prefix = f"Tried to instantiate class:{cls} from file:{path} with "
        if len(polymers) != len(polymer_names):
            raise ValueError(prefix + f"{len(polymers)} polymers, but  there are "
                             f"{len(polymer_names)} listed.")
        for poly in polymers:
            if poly.name not in polymer_names:
                raise ValueError(prefix + f"polymer:{poly.name}, but  this polymer was not present in file.")
        if len(binders) != len(binder_names):
            raise ValueError(prefix + f"{len(binders)} binders, but  there are {len(binder_names)} listed.")
        for _, binder in binders.iterrows():
            if binder['name'] not in binder_names:
                raise ValueError(prefix + f"binder:{binder}, but  this binder was not present in file.")
        return cls(polymers=polymers, binders=binders, **kwargs)
