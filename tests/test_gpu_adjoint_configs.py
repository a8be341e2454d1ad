"""The energy adjoint at the benchmark sizes and against the reference's own
formulas.

* One 1,000,000-bead replica of the large workload (B = 4, 128^3 voxels, whose
  density comes from the block-ranked gather): energy, dE/dr and dE/dtheta
  against the FP64 PyTorch transcription.
* The elastic part of dE/dr against the force on each bead written in the
  reference's documentation (docs/source/poly_models.rst:290-317), with every
  field interaction off.
* Python's round(phi, 2) > vf_limit inside the reference compute_E kernel,
  on volume fractions one ulp either side of the two-decimal halfway points.
* Covariance reweighting at the 20,000-bead / B = 3 / 48^3 size on 200
  configurations sampled every 20 sweeps.
"""
import copy
import ctypes
import dataclasses
import math
import sys

import numpy as np
import pytest
import torch

from chromo_amd import _lib as L
from chromo_amd import workloads as W

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def test_large_replica_energy_and_gradients_match_torch():
    from test_gpu_adjoint import torch_energy
    from chromo_amd.energy import EnergyModel, energy
    say("adjoint vs torch, 1e6 beads, B=4, 128^3")
    s = W.config_large(1)[0]
    s.chain.states = np.random.RandomState(5).randint(0, 3, size=(s.chain.N, s.chain.B))
    model = EnergyModel(s.chain, s.field, s.binders)
    r = torch.tensor(s.chain.r, dtype=torch.float64)
    th = torch.tensor(model.theta0, dtype=torch.float64)
    E, gr, gt = energy(model, r.cuda(), th.cuda())
    say("  device done; torch transcription")
    rt = r.clone().requires_grad_(True)
    tht = th.clone().requires_grad_(True)
    Et, el, fi, bi = torch_energy(s, rt, tht)
    Et.backward()
    E = E.cpu().numpy()
    assert np.allclose(E, [Et.item(), el.item(), fi.item(), bi.item()], rtol=1e-11, atol=1e-7)
    assert np.allclose(gr.cpu().numpy(), rt.grad.numpy(), rtol=1e-9, atol=1e-9)
    assert np.allclose(gt.cpu().numpy(), tht.grad.numpy(), rtol=1e-10, atol=1e-6)


def _doc_forces(r, t3, bonds):
    """f^(n) of docs/source/poly_models.rst:290-317 (tangents held fixed), with
    eps/Delta there = the per-bond eps here and tau_b = eps_b (t3' - t3 - eta dr_perp)."""
    eb, epar, eperp, gam, eta = bonds
    dr = r[1:] - r[:-1]
    par = np.sum(dr * t3[:-1], axis=1)
    perp = dr - par[:, None] * t3[:-1]
    tau = eb[:, None] * (t3[1:] - t3[:-1] - eta[:, None] * perp)
    tdot = np.sum(tau * t3[:-1], axis=1)
    # terms of bond n acting on bead n + 1 (the "(n)" terms of the formula for bead n+1
    # appear with the (n-1) index there) and on bead n
    on_next = (-eta[:, None] * tau + (eta * tdot)[:, None] * t3[:-1] +
               (epar * (par - gam))[:, None] * t3[:-1] + eperp[:, None] * perp)
    f = np.zeros_like(r)
    f[:-1] += on_next          # bond n pulls bead n along +(...)
    f[1:] -= on_next           # and bead n + 1 along -(...)
    return f


def test_elastic_gradient_equals_documented_force():
    from chromo_amd.energy import EnergyModel, energy
    say("elastic gradient vs documented force")
    s = W.make_system(3000, 2, 12, seed=71, chi=0.0, eps=0.0, mu=0.0)
    s.field.pref = [0.0, 0.0]
    s.field.intra = [0.0, 0.0]
    model = EnergyModel(s.chain, s.field, s.binders)
    th = torch.tensor(model.theta0, dtype=torch.float64, device="cuda")
    E, gr, _ = energy(model, torch.tensor(s.chain.r, device="cuda"), th)
    assert abs(E[2].item()) < 1e-12                      # no field energy
    f = _doc_forces(s.chain.r, s.chain.t3, s.chain.bonds)
    g = gr.cpu().numpy()
    assert np.allclose(g, -f, rtol=1e-10, atol=1e-9 * np.abs(f).max())


def _ref_field_E(model, r9):
    from chromo_amd.engine import _current_stream
    dev = model.device
    pos = torch.tensor(r9, device=dev)
    z = torch.zeros(model.B, dtype=torch.float64, device=dev)
    out = torch.zeros(2, dtype=torch.float64, device=dev)
    L.check(model.lib.chromo_energy_reference(ctypes.byref(model.m), pos.data_ptr(), 0, z.data_ptr(),
                                              z.data_ptr(), float(model.field.chi), None, out.data_ptr(),
                                              _current_stream()), "chromo_energy_reference")
    return float(out[1].item())


def test_reference_compute_E_rounds_volume_fractions_like_python(oracle):
    """Two beads at the centre of a one-voxel grid: vf = (2 / vol_bin) * V exactly,
    with V stepped one ulp at a time across vf = k/100 and (k + 1/2)/100 near
    vf_limit = 0.5.  The E_HUGE branch is taken iff Python's round(vf, 2) > 0.5."""
    from chromo_amd.energy import EnergyModel
    say("device round(vf, 2)")
    s = W.make_system(2, 1, 1, seed=72, chi=1.0, eps=0.0, mu=0.0)
    s.chain.r = np.zeros((2, 3))
    vb = s.field.x_width ** 3
    r9 = np.concatenate([s.chain.r, s.chain.t3, s.chain.t2], 1)
    checked = flips = 0
    for target in (0.495, 0.5, 0.505, 0.5049999999999999, 0.515, 0.51):
        V0 = target * vb / 2
        for k in range(-6, 7):
            V = V0
            for _ in range(abs(k)):
                V = np.nextafter(V, np.inf if k > 0 else -np.inf)
            f = dataclasses.replace(s.field, bead_vol=float(V))
            vf = (2.0 / vb) * float(V)           # the density of two unit weights, times V
            model = EnergyModel(s.chain, f, s.binders)
            Eg = _ref_field_E(model, r9)
            huge = round(vf, 2) > 0.5
            assert (Eg > 1e90) == huge, (target, k, vf, round(vf, 2), Eg)
            Eo = oracle.OracleSim(s.chain, f).field_E()
            assert Eg == Eo or abs(Eg - Eo) < 1e-12 * abs(Eo), (vf, Eg, Eo)
            checked += 1
            flips += huge
    assert 0 < flips < checked


def test_covariance_reweighting_full_size():
    """20,000 beads, B = 3, 48^3: 200 configurations every 20 sweeps.  The
    covariance formula equals the derivative of the importance-reweighted
    average of E_field (central differences in theta) on the same samples."""
    from chromo_amd.energy import EnergyModel, ReweightingAccumulator, energy
    from chromo_amd.engine import DeviceReplicas
    say("reweighting, 20k beads")
    s = W.config_adjoint()
    dr = DeviceReplicas([s.chain], [s.field], [copy.deepcopy(s.ctrls)], rng_mode=L.RNG_PHILOX,
                        seeds=[31], streams=[0])
    dr.run(100, waves=2)
    model0 = EnergyModel(s.chain, s.field, s.binders)
    th = torch.tensor(model0.theta0, dtype=torch.float64, device="cuda")
    B = s.chain.B
    acc = ReweightingAccumulator(th.numel())
    samples = []
    for k in range(200):
        dr.run(20, waves=2)
        r, t3, _ = dr.positions(0)
        c = dataclasses.replace(s.chain, r=r, t3=t3, states=dr.binder_states(0))
        m = EnergyModel(c, s.field, s.binders)
        rr = torch.tensor(r, device="cuda")
        E, _, g = energy(m, rr, th)
        g = g.cpu().numpy()
        dO = g.copy()
        dO[1 + B * (B + 1) // 2:] = 0
        acc.add(E[2].item(), dO, g)
        samples.append((m, rr, E[0].item()))
    say("  sampled")
    grad = acc.gradient()
    h = 1e-6
    E0 = np.array([e for _, _, e in samples])
    for k in range(th.numel()):
        vals = []
        for sgn in (1, -1):
            tk = th.clone()
            tk[k] += sgn * h
            Es1 = np.array([energy(m, rr, tk, grad=False)[0].cpu().numpy() for m, rr, _ in samples])
            a = -(Es1[:, 0] - E0)
            w = np.exp(a - a.max())
            vals.append(np.sum(w * Es1[:, 2]) / np.sum(w))
        fd = (vals[0] - vals[1]) / (2 * h)
        assert abs(fd - grad[k]) < 1e-5 + 1e-5 * abs(grad[k]), (k, fd, grad[k])


def test_adjoint_repeats_bit_for_bit():
    """Dense buckets (20,000 beads on an 8^3 grid, ~40 beads per cell): the atomic
    placement leaves a bucket's beads in a different order from call to call, and
    the exact fixed-point density sums give the same energy and gradients bit for
    bit, and agree with the torch transcription."""
    from test_gpu_adjoint import torch_energy
    from chromo_amd.energy import EnergyModel, energy
    say("adjoint repeatability, dense buckets")
    s = W.make_system(20000, 3, 8, seed=11, eps=[-4.0, -2.5, -3.0], mu=-0.5)
    # ~40 beads per voxel exceed vf_limit = 0.5, whose 1e99 branch the
    # transcription leaves out
    s.field.vf_limit = 1e9
    s.chain.states = np.random.RandomState(6).randint(0, 3, size=(s.chain.N, s.chain.B))
    model = EnergyModel(s.chain, s.field, s.binders)
    r = torch.tensor(s.chain.r, dtype=torch.float64, device="cuda")
    th = torch.tensor(model.theta0, dtype=torch.float64, device="cuda")
    outs = [energy(model, r, th) for _ in range(3)]
    for E, gr, gt in outs[1:]:
        assert torch.equal(E, outs[0][0]) and torch.equal(gr, outs[0][1]) and torch.equal(gt, outs[0][2])
    rt = r.cpu().clone().requires_grad_(True)
    tht = th.cpu().clone().requires_grad_(True)
    Et, el, fi, bi = torch_energy(s, rt, tht)
    Et.backward()
    E = outs[0][0].cpu().numpy()
    assert np.allclose(E, [Et.item(), el.item(), fi.item(), bi.item()], rtol=1e-11, atol=1e-8)
    assert np.allclose(outs[0][1].cpu().numpy(), rt.grad.numpy(), rtol=1e-9, atol=1e-9)
    assert np.allclose(outs[0][2].cpu().numpy(), tht.grad.numpy(), rtol=1e-10, atol=1e-8)
