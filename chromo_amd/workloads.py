"""Synthetic systems of the benchmark configurations (seeded, reproducible).

Every system is a chromatin chain (bead spacing 25 nm, lp = 53 nm, bead radius
5 nm) started from a Gaussian random walk (chromo/util/poly_paths.py:269-295)
with tangents from chromo/util/poly_paths.py:506-545, in a periodic cubic
UniformDensityField of 30 nm voxels, driven by ``all_moves`` with SimpleControl
and the amplitude bounds of ``get_amplitude_bounds``.  Binders are HP1-like
(2 sites, e_mod = -0.01, e_nomod = 1.52, r_int = 3 nm); their self interaction
eps, chemical potential mu and the chi parameter are set per system.

Configurations (sizes as in the benchmark description):
    config_plain_cpu   1,000 beads, 30^3 voxels of 30 nm, B = 2, Bernoulli(0.3) marks
    config_plain_gpu   100,000 beads, 64^3 voxels, B = 2, two-state Markov marks (domain 20)
    config_batched     512 replicas, N ~ U{2000..50000}, 32^3..64^3 voxels, Philox streams
    config_adjoint     20,000 beads, 48^3 voxels, B = 3
    config_large       16 replicas of 1,000,000 beads, 128^3 voxels, B = 4
"""
from __future__ import annotations

from dataclasses import dataclass
from typing import List, Optional, Sequence

import numpy as np

from . import _lib as L
from .binders import ReaderProtein
from .engine import ChainSpec, CtrlSpec, FieldSpec, field_grid
from .fields import field_prefactors
from .polymers import LENGTH_BP, NATURAL_TWIST_BARE, binding_tables, sswlc_bond_params

BEAD_SPACING = 25.0
LP_CHROMATIN = 53.0
BEAD_RAD = 5.0
VOXEL = 30.0
BEAD_VOL = (4 / 3) * np.pi * BEAD_RAD ** 3


@dataclass
class System:
    chain: ChainSpec
    field: FieldSpec
    ctrls: List[CtrlSpec]
    binders: List[ReaderProtein]


def make_binders(B: int, eps, mu, cross_eps: float = 0.0) -> List[ReaderProtein]:
    """B HP1-like reader proteins with self interaction eps[b] and potential mu[b]."""
    eps = np.broadcast_to(np.asarray(eps, dtype=float), (B,))
    mu = np.broadcast_to(np.asarray(mu, dtype=float), (B,))
    names = [f"R{b}" for b in range(B)]
    out = []
    for b in range(B):
        talk = {names[j]: cross_eps for j in range(B) if j != b}
        out.append(ReaderProtein(names[b], 2, -0.01, 1.52, float(eps[b]), float(mu[b]), 3, talk))
    return out


def gaussian_walk_chain(rs: np.random.RandomState, N: int, spacing: float = BEAD_SPACING):
    """Random walk + tangents with the reference's arithmetic on a private stream."""
    steps = rs.standard_normal((N - 1, 3))
    mag = np.linalg.norm(steps, axis=1)
    pos = np.cumsum(np.divide(steps, mag[:, None]) * np.full(N - 1, spacing)[:, None], axis=0)
    r = np.vstack((np.zeros(3), pos))
    t3 = np.zeros((N, 3))
    d = np.empty((N, 3))
    d[0] = r[1] - r[0]
    d[-1] = r[-1] - r[-2]
    d[1:-1] = r[2:] - r[:-2]
    t3 = d / np.linalg.norm(d, axis=1)[:, None]
    t2 = np.cross(t3, np.array([1, 0, 0]))
    bad = np.all(t2 == 0, axis=1)
    if bad.any():
        t2[bad] = np.cross(t3[bad], np.array([0, 1, 0]))
    t2 = t2 / np.linalg.norm(t2, axis=1)[:, None]
    return r, t3, t2


def bernoulli_marks(rs, N, B, p=0.3):
    return (rs.random_sample((N, B)) < p).astype(np.int64)


def markov_marks(rs, N, B, domain=20.0):
    """Per binder a two-state Markov chain with switch probability 1 / domain."""
    marks = np.zeros((N, B), dtype=np.int64)
    for b in range(B):
        flips = rs.random_sample(N) < 1.0 / domain
        flips[0] = rs.random_sample() < 0.5
        marks[:, b] = np.cumsum(flips) % 2
    return marks


def markov_one_hot_marks(rs, N, B, domain=50.0):
    """A B-state Markov chain (mean domain length ``domain``); state k marks binder k."""
    state = np.empty(N, dtype=np.int64)
    s = rs.randint(B)
    switch = rs.random_sample(N) < 1.0 / domain
    new = rs.randint(B, size=N)
    for i in range(N):
        if switch[i]:
            s = new[i]
        state[i] = s
    marks = np.zeros((N, B), dtype=np.int64)
    marks[np.arange(N), state] = 1
    return marks


def default_controllers(N: int, spacing: float = BEAD_SPACING,
                        per_cycle=(30, 1, 60, 60, 10)) -> List[CtrlSpec]:
    """all_moves + SimpleControl with get_amplitude_bounds (chromo/mc/__init__.py:295-332)."""
    bead = {
        L.CRANK_SHAFT: (min(30, N), min(150, N)),
        L.SLIDE: (min(10, N), min(150, N)),
        L.END_PIVOT: (min(50, N / 4), min(150, int(N / 2))),
        L.TANGENT_ROTATION: (1, N),
        L.CHANGE_BINDING_STATE: (1, 1),
    }
    move = {
        L.CRANK_SHAFT: (0.1 * np.pi, 0.25 * np.pi),
        L.SLIDE: (0.2 * spacing, 0.3 * spacing),
        L.END_PIVOT: (0.2 * np.pi, 0.25 * np.pi),
        L.TANGENT_ROTATION: (0.05 * np.pi, 0.2 * np.pi),
        L.CHANGE_BINDING_STATE: (0, 0),
    }
    order = [L.CRANK_SHAFT, L.END_PIVOT, L.SLIDE, L.TANGENT_ROTATION, L.CHANGE_BINDING_STATE]
    out = []
    for m, npc in zip(order, per_cycle):
        out.append(CtrlSpec(move=m, num_per_cycle=int(npc), amp_move=float(move[m][0]),
                            amp_bead=int(bead[m][0]), bead_lo=float(bead[m][0]),
                            bead_hi=float(bead[m][1]), move_lo=float(move[m][0]),
                            move_hi=float(move[m][1]),
                            alpha=2 / (float(np.float32(20)) + 1)))
    return out


def make_system(N: int, B: int, n_vox: int, *, seed: int, chi: float = 1.0, eps=-4.0, mu=-0.4,
                marks: str = "bernoulli", voxel: float = VOXEL, cross_eps: float = 0.0,
                mu_adjust: float = 1.0, per_cycle=(30, 1, 60, 60, 10)) -> System:
    rs = np.random.RandomState(seed)
    r, t3, t2 = gaussian_walk_chain(rs, N)
    if marks == "bernoulli":
        mk = bernoulli_marks(rs, N, B)
    elif marks == "markov":
        mk = markov_marks(rs, N, B)
    elif marks == "onehot":
        mk = markov_one_hot_marks(rs, N, B)
    else:
        raise ValueError(marks)
    binders = make_binders(B, eps, mu, cross_eps)
    bl = np.full(N - 1, BEAD_SPACING)
    p = sswlc_bond_params(bl, LP_CHROMATIN)
    bonds = np.stack([p[1], p[3], p[4], p[2], p[5]])
    chain = ChainSpec(r=r, t3=t3, t2=t2, states=np.zeros((N, B), dtype=np.int64), marks=mk,
                      bonds=bonds, sites=[b.sites_per_bead for b in binders],
                      mu=[b.chemical_potential for b in binders], F=binding_tables(binders, mk),
                      max_binders=-1, mu_adjust=mu_adjust)
    W = n_vox * voxel
    fs = FieldSpec(field_on=True, nx=n_vox, ny=n_vox, nz=n_vox, x_width=W, y_width=W, z_width=W,
                   vf_limit=0.5, chi=chi, bead_vol=BEAD_VOL)
    _, _, vol_bin, _, _, _ = field_grid(fs)
    fs.pref, fs.intra, fs.cross = field_prefactors(binders, vol_bin)
    return System(chain, fs, default_controllers(N, per_cycle=per_cycle), binders)


def make_twist_system(N: int, B: int, n_vox: int, *, seed: int, lt: float = 100.0, **kw) -> System:
    """make_system with an SSTWLC chain (polymers.pyx:1889-2320): twist persistence
    length lt (nm), eps_twist = lt / (delta lp) and the natural twist
    bead_length 2 pi / (10.5 * 0.332) on every bond."""
    s = make_system(N, B, n_vox, seed=seed, **kw)
    bl = np.full(N - 1, BEAD_SPACING)
    delta = sswlc_bond_params(bl, LP_CHROMATIN)[0]
    et = lt / (delta * LP_CHROMATIN)
    tw0 = bl * NATURAL_TWIST_BARE / LENGTH_BP
    s.chain.bonds = np.vstack([s.chain.bonds, et, tw0])
    s.chain.twist = True
    s.chain.bead_length = bl
    return s


def config_plain_cpu(seed: int = 1) -> System:
    return make_system(1000, 2, 30, seed=seed, chi=1.0, eps=-4.0, mu=-0.4, marks="bernoulli")


def config_plain_gpu(seed: int = 2) -> System:
    return make_system(100_000, 2, 64, seed=seed, chi=1.0, eps=-4.0, mu=-0.4, marks="markov")


def batched_grid(N: int) -> int:
    """Voxels per side of a replica: 32 + 32 * ((N - 2000) / 48000)^(1/3), clamped to [32, 64]."""
    return int(np.clip(np.round(32 + 32 * ((N - 2000) / 48000) ** (1 / 3)), 32, 64))


def config_batched(n_replicas: int = 512, seed: int = 3, n_min: int = 2000,
                   n_max: int = 50000) -> List[System]:
    rs = np.random.RandomState(seed)
    Ns = rs.randint(n_min, n_max + 1, size=n_replicas)
    chis = rs.uniform(0, 2, size=n_replicas)
    epss = rs.uniform(-5, -1, size=(n_replicas, 2))
    mus = rs.uniform(-2, 1, size=(n_replicas, 2))
    out = []
    for i in range(n_replicas):
        out.append(make_system(int(Ns[i]), 2, batched_grid(int(Ns[i])), seed=seed * 100003 + i,
                               chi=float(chis[i]), eps=epss[i], mu=mus[i], marks="bernoulli"))
    return out


def config_adjoint(seed: int = 4) -> System:
    return make_system(20_000, 3, 48, seed=seed, chi=1.0, eps=-4.0, mu=-0.4, marks="bernoulli")


def config_large(n_replicas: int = 16, seed: int = 5, N: int = 1_000_000) -> List[System]:
    rs = np.random.RandomState(seed)
    out = []
    for i in range(n_replicas):
        out.append(make_system(N, 4, 128, seed=seed * 100003 + i, chi=float(rs.uniform(0, 2)),
                               eps=rs.uniform(-5, -1, size=4), mu=rs.uniform(-2, 1, size=4),
                               marks="onehot"))
    return out
