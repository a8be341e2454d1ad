"""Adjoint of field + elastic + binding energy (HIP) against an FP64 PyTorch
transcription (autograd), central finite differences, the oracle's move dE and
the reference's compute_E; covariance reweighting against importance-reweighted
finite differences on the same samples."""
import copy

import numpy as np
import pytest
import torch

from chromo_amd import workloads as W

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


def system(N=400, B=3, n_vox=8, seed=0, cross=0.6):
    s = W.make_system(N, B, n_vox, seed=seed, eps=[-4.0, -2.5, -3.0, -1.5][:B],
                      mu=[-0.4, 0.3, -1.0, 0.2][:B], cross_eps=cross, chi=1.3)
    s.chain.states = np.random.RandomState(seed).randint(0, 3, size=(N, B))
    return s


def torch_energy(s, r, theta):
    """Plain float64 PyTorch transcription of the dE-consistent energy."""
    from chromo_amd.energy import eps_index
    c, f = s.chain, s.field
    B = c.B
    n = torch.tensor([f.nx, f.ny, f.nz], dtype=torch.float64)
    Wd = torch.tensor([f.x_width, f.y_width, f.z_width], dtype=torch.float64)
    dx = Wd / n
    vb = float(f.x_width * f.y_width * f.z_width / (f.nx * f.ny * f.nz))
    x = torch.remainder(r + Wd / 2, Wd) - dx / 2
    q = x / dx
    lo = torch.floor(q)
    w = 1 - (q - lo.detach())
    lo = lo.long()
    st = torch.as_tensor(np.asarray(c.states), dtype=torch.float64)
    rho = torch.zeros(f.nx * f.ny * f.nz, B + 1, dtype=torch.float64)
    ni = n.long()
    for l in range(8):
        o = torch.tensor([l & 1, (l >> 1) & 1, (l >> 2) & 1])
        idx = torch.remainder(lo + o, ni)
        wt = torch.prod(torch.where(o == 0, w, 1 - w), dim=1) / vb
        v = idx[:, 0] + idx[:, 1] * ni[0] + idx[:, 2] * ni[0] * ni[1]
        contrib = torch.cat([wt[:, None], wt[:, None] * st], dim=1)
        rho = rho.index_add(0, v, contrib)
    chi = theta[0]
    vint = [b.interaction_volume for b in s.binders]
    names = [b.name for b in s.binders]
    E = torch.zeros((), dtype=torch.float64)
    for j in range(B):
        e = theta[eps_index(B, j, j)]
        E = E + 0.5 * e * vint[j] * vb * torch.sum(rho[:, j + 1] ** 2)
        E = E + e * (1 - vint[j] / vb) * float(np.sum(np.asarray(c.states)[:, j] == 2))
        for k in range(B):
            if names[k] in s.binders[j].cross_talk_interaction_energy:
                E = E + theta[eps_index(B, j, k)] * vint[j] * vb * torch.sum(rho[:, j + 1] * rho[:, k + 1])
    phi = rho[:, 0] * f.bead_vol
    E = E + torch.sum(chi * (vb / f.bead_vol) * phi ** 2)
    Efield = E
    # elastic
    t3 = torch.as_tensor(c.t3)
    bd = torch.as_tensor(c.bonds)
    dr = r[1:] - r[:-1]
    par = torch.sum(t3[:-1] * dr, dim=1)
    perp = dr - t3[:-1] * par[:, None]
    bend = t3[1:] - t3[:-1] - perp * bd[4][:, None]
    Eel = torch.sum(0.5 * bd[0] * torch.sum(bend * bend, 1) + 0.5 * bd[1] * (par - bd[3]) ** 2 +
                    0.5 * bd[2] * torch.sum(perp * perp, 1))
    # binding
    Eb = torch.zeros((), dtype=torch.float64)
    F = np.asarray(c.F)
    stn = np.asarray(c.states)
    mk = np.asarray(c.marks)
    for j in range(B):
        mu = theta[1 + B * (B + 1) // 2 + j]
        Eb = Eb + float(np.sum(F[j, mk[:, j], stn[:, j]])) - mu * float(stn[:, j].sum())
    return Eel + Efield + Eb, Eel, Efield, Eb


def test_energy_and_gradients_match_torch():
    say("adjoint vs torch")
    from chromo_amd.energy import EnergyModel, energy
    s = system()
    model = EnergyModel(s.chain, s.field, s.binders)
    r = torch.tensor(s.chain.r, dtype=torch.float64)
    th = torch.tensor(model.theta0, dtype=torch.float64)
    E, gr, gt = energy(model, r.cuda(), th.cuda())
    rt = r.clone().requires_grad_(True)
    tht = th.clone().requires_grad_(True)
    Et, el, fi, bi = torch_energy(s, rt, tht)
    Et.backward()
    E = E.cpu().numpy()
    assert np.allclose(E, [Et.item(), el.item(), fi.item(), bi.item()], rtol=1e-11, atol=1e-9)
    assert np.allclose(gr.cpu().numpy(), rt.grad.numpy(), rtol=1e-9, atol=1e-9)
    assert np.allclose(gt.cpu().numpy(), tht.grad.numpy(), rtol=1e-10, atol=1e-9)


def test_energy_and_gradients_match_torch_full_size():
    """The adjoint configuration at full size: 20,000 beads, B = 3, 48^3 voxels."""
    say("adjoint vs torch, 20k beads")
    from chromo_amd.energy import EnergyModel, energy
    s = W.config_adjoint()
    s.chain.states = np.random.RandomState(4).randint(0, 3, size=(s.chain.N, s.chain.B))
    model = EnergyModel(s.chain, s.field, s.binders)
    r = torch.tensor(s.chain.r, dtype=torch.float64)
    th = torch.tensor(model.theta0, dtype=torch.float64)
    E, gr, gt = energy(model, r.cuda(), th.cuda())
    rt = r.clone().requires_grad_(True)
    tht = th.clone().requires_grad_(True)
    Et, el, fi, bi = torch_energy(s, rt, tht)
    Et.backward()
    E = E.cpu().numpy()
    assert np.allclose(E, [Et.item(), el.item(), fi.item(), bi.item()], rtol=1e-11, atol=1e-8)
    assert np.allclose(gr.cpu().numpy(), rt.grad.numpy(), rtol=1e-9, atol=1e-9)
    assert np.allclose(gt.cpu().numpy(), tht.grad.numpy(), rtol=1e-10, atol=1e-8)


def test_gradients_central_differences():
    say("adjoint vs central FD")
    from chromo_amd.energy import EnergyModel, energy
    s = system(N=300, B=3, n_vox=7, seed=2)
    model = EnergyModel(s.chain, s.field, s.binders)
    r = torch.tensor(s.chain.r, dtype=torch.float64, device="cuda")
    th = torch.tensor(model.theta0, dtype=torch.float64, device="cuda")
    _, gr, gt = energy(model, r, th)
    gr, gt = gr.cpu().numpy(), gt.cpu().numpy()
    h = 1e-6
    dx = s.field.x_width / s.field.nx
    q = (np.mod(s.chain.r + s.field.x_width / 2, s.field.x_width) - dx / 2) / dx
    away = np.all(np.abs(q - np.round(q)) > 1e-3, axis=1)
    rs = np.random.RandomState(0)
    picks = rs.choice(np.nonzero(away)[0], size=25, replace=False)
    for b in picks:
        for c in range(3):
            rp, rm = r.clone(), r.clone()
            rp[b, c] += h
            rm[b, c] -= h
            Ep = energy(model, rp, th, grad=False)[0][0].item()
            Em = energy(model, rm, th, grad=False)[0][0].item()
            fd = (Ep - Em) / (2 * h)
            assert abs(fd - gr[b, c]) < 1e-5 + 1e-6 * abs(gr[b, c]), (b, c, fd, gr[b, c])
    for k in range(th.numel()):
        tp, tm = th.clone(), th.clone()
        tp[k] += h
        tm[k] -= h
        fd = (energy(model, r, tp, grad=False)[0][0].item() -
              energy(model, r, tm, grad=False)[0][0].item()) / (2 * h)
        assert abs(fd - gt[k]) < 1e-5 + 1e-6 * abs(gt[k]), (k, fd, gt[k])


def test_autograd_function():
    say("autograd function")
    from chromo_amd.energy import EnergyModel, chromo_energy
    s = system(N=60, B=2, n_vox=5, seed=3)
    model = EnergyModel(s.chain, s.field, s.binders)
    r = torch.tensor(s.chain.r, dtype=torch.float64, device="cuda", requires_grad=True)
    th = torch.tensor(model.theta0, dtype=torch.float64, device="cuda", requires_grad=True)
    E = chromo_energy(r, th, model)
    (2.5 * E).backward()
    rt = torch.tensor(s.chain.r, dtype=torch.float64, requires_grad=True)
    tht = torch.tensor(model.theta0, dtype=torch.float64, requires_grad=True)
    (2.5 * torch_energy(s, rt, tht)[0]).backward()
    assert np.allclose(r.grad.cpu().numpy(), rt.grad.numpy(), rtol=1e-9, atol=1e-9)
    assert np.allclose(th.grad.cpu().numpy(), tht.grad.numpy(), rtol=1e-10, atol=1e-9)


def test_energy_differences_are_move_dE(oracle):
    """E_field(after) - E_field(before) equals the reference field dE of the move."""
    say("adjoint vs oracle dE")
    from chromo_amd.energy import EnergyModel, energy
    s = system(N=300, B=2, n_vox=8, seed=4)
    model = EnergyModel(s.chain, s.field, s.binders)
    o = oracle.OracleSim(s.chain, s.field)
    rs = np.random.RandomState(1)
    th = torch.tensor(model.theta0, dtype=torch.float64, device="cuda")
    r0 = torch.tensor(s.chain.r, dtype=torch.float64, device="cuda")
    E0 = energy(model, r0, th, grad=False)[0][2].item()
    for n in (1, 7, 40):
        ind0 = rs.randint(0, 300 - n)
        inds = np.arange(ind0, ind0 + n)
        a = rs.normal(size=3)
        T = oracle.axis_rotation(a / np.linalg.norm(a), s.chain.r[ind0], 0.9)
        rt, _, _ = oracle.transform(T, s.chain.r[inds], s.chain.t3[inds], s.chain.t2[inds])
        dE, _, _ = o.field_dE_tape(inds, rt, s.chain.states[inds], 0)
        r1 = r0.clone()
        r1[inds] = torch.tensor(rt, device="cuda")
        E1 = energy(model, r1, th, grad=False)[0][2].item()
        assert abs((E1 - E0) - dE) < 1e-8 * max(1.0, abs(E0))


def test_reference_compute_E(oracle):
    say("reference compute_E")
    import ctypes

    from chromo_amd import _lib as L
    from chromo_amd.energy import EnergyModel
    from chromo_amd.engine import _current_stream
    s = system(N=500, B=2, n_vox=9, seed=5)
    model = EnergyModel(s.chain, s.field, s.binders)
    pos = torch.tensor(np.concatenate([s.chain.r, s.chain.t3, s.chain.t2], 1), device="cuda")
    pref = torch.tensor(np.asarray(s.field.pref), device="cuda")
    intra = torch.tensor(np.asarray(s.field.intra), device="cuda")
    out = torch.zeros(2, dtype=torch.float64, device="cuda")
    L.check(model.lib.chromo_energy_reference(ctypes.byref(model.m), pos.data_ptr(), 0,
                                              pref.data_ptr(), intra.data_ptr(), s.field.chi, None,
                                              out.data_ptr(), _current_stream()))
    o = oracle.OracleSim(s.chain, s.field)
    el, fi = out.cpu().numpy()
    assert abs(el - o.elastic_E()) < 1e-11 * abs(o.elastic_E())
    assert abs(fi - o.field_E()) < 1e-11 * max(1.0, abs(o.field_E()))


def test_covariance_reweighting_identity():
    """On one set of samples, the covariance formula equals the derivative of the
    importance-reweighted average <O>_theta+h (central difference)."""
    say("covariance reweighting")
    from chromo_amd import _lib as L
    from chromo_amd.energy import EnergyModel, ReweightingAccumulator, energy
    from chromo_amd.engine import DeviceReplicas
    s = system(N=400, B=2, n_vox=8, seed=6)
    dr = DeviceReplicas([s.chain], [s.field], [copy.deepcopy(s.ctrls)], rng_mode=L.RNG_PHILOX,
                        seeds=[3], streams=[0])
    model = EnergyModel(s.chain, s.field, s.binders)
    th = torch.tensor(model.theta0, dtype=torch.float64, device="cuda")
    confs = []
    acc = ReweightingAccumulator(th.numel())
    for k in range(12):
        dr.run(5)
        r, t3, _ = dr.positions(0)
        st = dr.binder_states(0)
        c = copy.copy(s.chain)
        c.r, c.t3, c.states = r, t3, st
        m = EnergyModel(c, s.field, s.binders)
        rr = torch.tensor(r, device="cuda")
        E, _, g = energy(m, rr, th)
        dO = g.cpu().numpy().copy()
        dO[1 + 2 * 3 // 2:] = 0
        acc.add(E[2].item(), dO, g.cpu().numpy())
        confs.append((m, rr))
    grad = acc.gradient()
    h = 1e-6
    for k in range(th.numel()):
        vals = []
        for sgn in (1, -1):
            tk = th.clone()
            tk[k] += sgn * h
            Es0 = np.array([energy(m, rr, th, grad=False)[0][0].item() for m, rr in confs])
            Es1 = np.array([energy(m, rr, tk, grad=False)[0].cpu().numpy() for m, rr in confs])
            w = np.exp(-(Es1[:, 0] - Es0) - np.max(-(Es1[:, 0] - Es0)))
            vals.append(np.sum(w * Es1[:, 2]) / np.sum(w))
        fd = (vals[0] - vals[1]) / (2 * h)
        assert abs(fd - grad[k]) < 1e-6 + 1e-5 * abs(grad[k]), (k, fd, grad[k])
