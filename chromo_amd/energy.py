"""Energies, their adjoint and parameter gradients by covariance reweighting.

* ``EnergyModel`` describes one chain in one field for the energy kernels.
* ``energy(model, r, theta, grad=...)`` evaluates the dE-consistent energy
  (the energy whose differences are the reference's move dE) and, through the
  hand-written adjoint kernel, dE/dr and dE/dtheta, with
  theta = (chi, eps_ij for i <= j, mu_j).
* ``ChromoEnergy`` wraps it as a ``torch.autograd.Function``: forward returns E,
  backward recomputes the voxel densities and returns dE/dr, dE/dtheta.
* ``elastic_energy`` / ``field_energy`` are the reference's own
  ``SSWLC.compute_E`` and ``UniformDensityField.compute_E`` on the GPU.
* ``ReweightingAccumulator`` turns sampled (O, dE/dtheta) pairs into
  d<O>/dtheta = <dO/dtheta> - (<O dE/dtheta> - <O><dE/dtheta>)  (kT = 1).
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass
from typing import List, Optional, Sequence

import numpy as np
import torch

from . import _lib as L
from .engine import ChainSpec, FieldSpec, _current_stream, field_grid


def n_theta(B: int) -> int:
    return 1 + B * (B + 1) // 2 + B


def eps_index(B: int, i: int, j: int) -> int:
    i, j = min(i, j), max(i, j)
    return 1 + i * B - (i * (i - 1)) // 2 + (j - i)


def theta_of(binders, chi: float) -> np.ndarray:
    """(chi, eps_ij, mu_j) of a list of ReaderProteins: eps_ii = interaction_energy,
    eps_ij = the cross-talk energy i lists for j (or j for i)."""
    B = len(binders)
    th = np.zeros(n_theta(B))
    th[0] = chi
    names = [b.name for b in binders]
    for i, b in enumerate(binders):
        th[eps_index(B, i, i)] = b.interaction_energy
        for j in range(i + 1, B):
            e = b.cross_talk_interaction_energy.get(names[j])
            if e is None:
                e = binders[j].cross_talk_interaction_energy.get(names[i], 0.0)
            th[eps_index(B, i, j)] = e
        th[1 + B * (B + 1) // 2 + i] = b.chemical_potential
    return th


class EnergyModel:
    """Device-resident description of one chain in one field."""

    def __init__(self, chain: ChainSpec, field: FieldSpec, binders, device=None):
        if not field.field_on:
            raise ValueError("the energy model needs a UniformDensityField")
        if getattr(field, "access_vol", None) is not None:
            raise NotImplementedError("the energy adjoint takes every voxel at its full volume "
                                      "(no per-voxel accessible volumes)")
        self.lib = L.lib()
        dev = device or torch.device("cuda", torch.cuda.current_device())
        self.device = dev
        N, B = chain.N, chain.B
        if not 1 <= B <= L.MAX_BINDERS:
            raise NotImplementedError(f"binder count {B} outside 1..{L.MAX_BINDERS}")
        self.N, self.B = N, B
        if N >= 2 ** 31 - 1:
            raise NotImplementedError("chains of 2^31 - 1 beads or more")
        st, mk = np.asarray(chain.states), np.asarray(chain.marks)
        if st.size and (st.min() < 0 or st.max() > L.MAX_SITES):
            raise ValueError(f"binding states outside 0..{L.MAX_SITES}")
        if mk.size and (mk.min() < 0 or mk.max() >= L.MAX_MARK):
            raise ValueError(f"chemical_mods outside 0..{L.MAX_MARK - 1}")
        self.t3 = torch.as_tensor(np.ascontiguousarray(chain.t3, dtype=np.float64), device=dev)
        self.states = torch.as_tensor(np.ascontiguousarray(np.asarray(chain.states).T, dtype=np.int8), device=dev)
        self.marks = torch.as_tensor(np.ascontiguousarray(np.asarray(chain.marks).T, dtype=np.int8), device=dev)
        self.bonds = torch.as_tensor(np.ascontiguousarray(chain.bonds, dtype=np.float64), device=dev)
        m = L.ChromoEnergyModel()
        m.N, m.B = N, B
        m.twist = 1 if getattr(chain, "twist", False) else 0
        m.t3, m.states, m.marks = self.t3.data_ptr(), self.states.data_ptr(), self.marks.data_ptr()
        m.bonds = self.bonds.data_ptr()
        m.nx, m.ny, m.nz = int(field.nx), int(field.ny), int(field.nz)
        m.x_width, m.y_width, m.z_width = field.x_width, field.y_width, field.z_width
        m.bead_vol = float(field.bead_vol)
        m.vf_limit = float(np.float32(field.vf_limit))
        m.mu_adjust = float(chain.mu_adjust)
        names = [b.name for b in binders]
        for i, b in enumerate(binders):
            m.v_int[i] = float(b.interaction_volume)
            for j, nj in enumerate(names):
                m.cross_mask[i * L.MAX_BINDERS + j] = 1 if nj in b.cross_talk_interaction_energy else 0
        F = np.asarray(chain.F, dtype=np.float64)
        for a in range(B):
            for lv in range(min(F.shape[1], L.MAX_MARK)):
                for s in range(min(F.shape[2], L.MAX_SITES + 1)):
                    m.F[(a * L.MAX_MARK + lv) * (L.MAX_SITES + 1) + s] = float(F[a, lv, s])
        self.m = m
        self.field = field
        self.binders = list(binders)
        self.theta0 = theta_of(binders, field.chi)


def energy(model: EnergyModel, r: torch.Tensor, theta: torch.Tensor, grad: bool = True):
    """(E[4] = total, elastic, field, binding; dE/dr or None; dE/dtheta or None)."""
    r = r.detach().to(model.device, torch.float64).contiguous()
    theta = theta.detach().to(model.device, torch.float64).contiguous()
    if r.shape != (model.N, 3) or theta.numel() != n_theta(model.B):
        raise ValueError("r must be (N, 3) and theta (1 + B(B+1)/2 + B,)")
    E = torch.zeros(4, dtype=torch.float64, device=model.device)
    gr = torch.empty_like(r) if grad else None
    gt = torch.empty_like(theta) if grad else None
    L.check(model.lib.chromo_energy_adjoint(ctypes.byref(model.m), r.data_ptr(), theta.data_ptr(),
                                            E.data_ptr(), gr.data_ptr() if grad else None,
                                            gt.data_ptr() if grad else None, _current_stream()),
            "chromo_energy_adjoint")
    return E, gr, gt


class ChromoEnergy(torch.autograd.Function):
    """E(r, theta) with the adjoint kernel as backward (densities recomputed)."""

    @staticmethod
    def forward(ctx, r, theta, model):
        E, _, _ = energy(model, r, theta, grad=False)
        ctx.model = model
        ctx.save_for_backward(r, theta)
        return E[0]

    @staticmethod
    def backward(ctx, gout):
        r, theta = ctx.saved_tensors
        _, gr, gt = energy(ctx.model, r, theta, grad=True)
        return gout * gr.to(r.dtype), gout * gt.to(theta.dtype), None


def chromo_energy(r: torch.Tensor, theta: torch.Tensor, model: EnergyModel) -> torch.Tensor:
    return ChromoEnergy.apply(r, theta, model)


# ---------------------------------------------------------------------------
# The reference's absolute energies on the GPU
# ---------------------------------------------------------------------------

def _reference_energies(poly, field, twist: bool = True):
    from .fields import UniformDensityField
    from .mc.mc_sim import engine_for
    chain = poly.chain_spec()
    if chain.twist and not twist:
        chain.twist = False
        chain.bonds = np.ascontiguousarray(np.asarray(chain.bonds)[:5])
    if isinstance(field, UniformDensityField):
        fs = field.field_spec()
        binders = [poly.beads[0].binders[b] for b in range(poly.num_binders)]
    else:
        fs = FieldSpec(field_on=True, nx=1, ny=1, nz=1, x_width=1.0, y_width=1.0, z_width=1.0,
                       bead_vol=1.0, pref=[0.0] * chain.B, intra=[0.0] * chain.B)
        binders = [poly.beads[0].binders[b] for b in range(poly.num_binders)]
    model = EnergyModel(chain, fs, binders)
    dev = model.device
    pos = np.concatenate([chain.r, chain.t3, chain.t2], axis=1)     # N x 9
    pos_t = torch.as_tensor(np.ascontiguousarray(pos), device=dev)
    pref = torch.as_tensor(np.asarray(fs.pref, dtype=np.float64), device=dev)
    intra = torch.as_tensor(np.asarray(fs.intra, dtype=np.float64), device=dev)
    out = torch.zeros(2, dtype=torch.float64, device=dev)
    dens = torch.zeros(fs.n_bins * (chain.B + 1), dtype=torch.float64, device=dev)
    L.check(model.lib.chromo_energy_reference(ctypes.byref(model.m), pos_t.data_ptr(), 0,
                                              pref.data_ptr(), intra.data_ptr(), float(fs.chi),
                                              dens.data_ptr(), out.data_ptr(), _current_stream()),
            "chromo_energy_reference")
    torch.cuda.synchronize(dev)
    return out.cpu().numpy(), dens


def elastic_energy(poly, twist: bool = True) -> float:
    """SSWLC.compute_E (polymers.pyx:1348-1381); for a twisting chain
    SSTWLC.compute_E (polymers.pyx:2250-2286), or with twist=False
    SSTWLC.compute_E_no_twist (polymers.pyx:2288-2320)."""
    from .fields import NullField
    out, _ = _reference_energies(poly, NullField(), twist)
    return float(out[0])


def field_energy(field, poly) -> float:
    """UniformDensityField.compute_E (fields.pyx:1939-1966).  Like the reference,
    this replaces the field's densities by a fresh scatter without the 1e-18
    clean-up of the initial densities."""
    out, dens = _reference_energies(poly, field)
    from .mc.mc_sim import engine_for
    eng = engine_for(field)
    dr = eng.dr
    dr.density[int(dr.o_den[0]):int(dr.o_den[1])].copy_(dens)
    return float(out[1])


# ---------------------------------------------------------------------------
# Covariance reweighting
# ---------------------------------------------------------------------------

class ReweightingAccumulator:
    """d<O>/dtheta from samples of (O, dO/dtheta, dE/dtheta) drawn at theta (kT = 1):

        d<O>/dtheta = <dO/dtheta> - ( <O dE/dtheta> - <O> <dE/dtheta> ).
    """

    def __init__(self, n_theta: int):
        self.n = 0
        self.sO = 0.0
        self.sdO = np.zeros(n_theta)
        self.sdE = np.zeros(n_theta)
        self.sOdE = np.zeros(n_theta)

    def add(self, O: float, dO: np.ndarray, dE: np.ndarray) -> None:
        self.n += 1
        self.sO += float(O)
        self.sdO += dO
        self.sdE += dE
        self.sOdE += float(O) * dE

    def gradient(self) -> np.ndarray:
        n = self.n
        mO, mdO, mdE, mOdE = self.sO / n, self.sdO / n, self.sdE / n, self.sOdE / n
        return mdO - (mOdE - mO * mdE)

    def mean(self) -> float:
        return self.sO / self.n


def field_energy_dtheta(model: EnergyModel, r: torch.Tensor, theta: torch.Tensor):
    """(E_field, dE_field/dtheta, dE/dtheta) at one configuration.  E_field depends
    on theta through chi and eps only; the binding energy carries the mu terms."""
    E, _, gt = energy(model, r, theta, grad=True)
    g = gt.cpu().numpy()
    B = model.B
    dO = g.copy()
    dO[1 + B * (B + 1) // 2:] = 0.0      # d E_field / d mu = 0
    return float(E[2].item()), dO, g
