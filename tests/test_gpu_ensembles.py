"""Ensemble checks of the HIP Monte Carlo against closed forms.

The draw order of each move is the oracle's restatement of the reference
(no reference trajectory exists to pin it), so these checks pin what any correct
order must give: the equilibrium the chain samples.

* Binding with every interaction off (chi = 0, eps = 0): each bead's binder count
  s follows P(s) ~ Z_m(s) exp(mu_eff s), with Z_m(s) = sum_i C(m,i) C(n-m,s-i)
  exp(-(i e_mod + (s-i) e_nomod)) over its m marked of n sites; for fully marked
  beads this is the Langmuir isotherm of the reference's validation notebook
  (simulations/validations/validate_isotherms.ipynb, cell 14):
  theta = exp(mu - e) / (1 + exp(mu - e)).  mu_eff = mu f for mu <= 0 and
  mu (2 - f) for mu > 0 (chromo/polymers.pyx:1522-1537).
* The SSWLC chain with the field off: the mean squared distance of beads n apart
  follows the wormlike chain, <R^2> / (2 lp)^2 = x - 1/2 + exp(-2x) / 2 with
  x = n l / (2 lp) (simulations/validations/validate_end_to_end_distances.py:133).
"""
import copy
import math
import sys

import numpy as np
import pytest

from chromo_amd import _lib as L
from chromo_amd import workloads as W
from chromo_amd.engine import ChainSpec, CtrlSpec, FieldSpec

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _exact_mean_bound(m, n, e_mod, e_nomod, mu_eff):
    w = []
    for s in range(n + 1):
        z = sum(math.comb(m, i) * math.comb(n - m, s - i) * math.exp(-(i * e_mod + (s - i) * e_nomod))
                for i in range(0, min(m, s) + 1) if s - i <= n - m)
        w.append(z * math.exp(mu_eff * s))
    w = np.array(w) / sum(w)
    return float(np.dot(np.arange(n + 1), w))


def test_binding_isotherms():
    """32 replicas scanning mu in [-3, 2] (four with mu_adjust_factor 0.4 or 1.6),
    200 beads each, marks 0/1/2 of two sites, binding moves only."""
    from chromo_amd.engine import DeviceReplicas
    n_rep, N = 32, 200
    mus = np.linspace(-3.0, 2.0, n_rep)
    systems = []
    for k in range(n_rep):
        s = W.make_system(N, 1, 6, seed=300 + k, chi=0.0, eps=0.0, mu=float(mus[k]),
                          per_cycle=(0, 0, 0, 0, 200))
        s.chain.marks = np.random.RandomState(k).randint(0, 3, size=(N, 1))
        from chromo_amd.polymers import binding_tables
        s.chain.F = binding_tables(s.binders, s.chain.marks)
        systems.append(s)
    f_adj = np.ones(n_rep)
    f_adj[[3, 10]] = 0.4
    f_adj[[25, 30]] = 1.6
    dr = DeviceReplicas([s.chain for s in systems], [s.field for s in systems],
                        [copy.deepcopy(s.ctrls) for s in systems], rng_mode=L.RNG_PHILOX,
                        seeds=[17] * n_rep, streams=list(range(n_rep)))
    dr.set_mu_adjust(f_adj)
    say("isotherms: equilibrate")
    dr.run(40, waves=2)
    n_snap = 100
    sums = np.zeros((n_snap, n_rep, 3))
    counts = np.zeros((n_rep, 3))
    for k in range(n_rep):
        for m in range(3):
            counts[k, m] = np.sum(systems[k].chain.marks[:, 0] == m)
    for t in range(n_snap):
        dr.run(8, waves=2)
        for k in range(n_rep):
            st = dr.binder_states(k)[:, 0]
            mk = systems[k].chain.marks[:, 0]
            for m in range(3):
                sums[t, k, m] = st[mk == m].mean()
    say("isotherms: sampled")
    b = systems[0].binders[0]
    worst = 0.0
    for k in range(n_rep):
        mu = mus[k]
        mu_eff = mu * (2 - f_adj[k]) if mu > 0 else mu * f_adj[k]
        for m in range(3):
            exact = _exact_mean_bound(m, 2, b.bind_energy_mod, b.bind_energy_no_mod, mu_eff)
            est = sums[:, k, m].mean()
            se = sums[:, k, m].std(ddof=1) / math.sqrt(n_snap)
            worst = max(worst, abs(est - exact) / (se + 1e-4))
            assert abs(est - exact) < 5 * se + 2e-3, (k, m, mu, f_adj[k], est, exact, se)
        # fully marked beads: the notebook's Langmuir isotherm per site
        theta = math.exp(mu_eff - b.bind_energy_mod) / (1 + math.exp(mu_eff - b.bind_energy_mod))
        se2 = sums[:, k, 2].std(ddof=1) / 2 / math.sqrt(n_snap)
        assert abs(sums[:, k, 2].mean() / 2 - theta) < 5 * se2 + 2e-3
    say(f"isotherms: worst deviation {worst:.2f} standard errors")


def _persistent_walk(rs, N, lp, b):
    """A discrete wormlike chain: bond i along t_i with |b|, consecutive tangents at
    angle theta with p(cos theta) ~ exp(kappa cos theta), <cos theta> = exp(-b / lp).
    The MC run then starts near equilibrium at every scale it measures."""
    target = math.exp(-b / lp)
    lo, hi = 1e-6, 1e3
    for _ in range(200):                       # coth(k) - 1/k = target
        k = 0.5 * (lo + hi)
        if 1 / math.tanh(k) - 1 / k < target:
            lo = k
        else:
            hi = k
    t = np.zeros((N, 3))
    t[0] = [0.0, 0.0, 1.0]
    for i in range(1, N):
        c = 1 + math.log(1 - rs.uniform() * (1 - math.exp(-2 * k))) / k
        ph = rs.uniform(0, 2 * math.pi)
        a = t[i - 1]
        e1 = np.cross(a, [1.0, 0, 0] if abs(a[0]) < 0.9 else [0, 1.0, 0])
        e1 /= np.linalg.norm(e1)
        e2 = np.cross(a, e1)
        sn = math.sqrt(max(0.0, 1 - c * c))
        t[i] = c * a + sn * (math.cos(ph) * e1 + math.sin(ph) * e2)
        t[i] /= np.linalg.norm(t[i])
    r = np.vstack([np.zeros(3), np.cumsum(b * t[:-1], axis=0)])
    t2 = np.cross(t, [1.0, 0.0, 0.0])
    bad = np.linalg.norm(t2, axis=1) < 1e-6
    t2[bad] = np.cross(t[bad], [0.0, 1.0, 0.0])
    t2 /= np.linalg.norm(t2, axis=1)[:, None]
    return r, t, t2


def _wlc_system(N, lp, seed):
    from chromo_amd.polymers import sswlc_bond_params
    rs = np.random.RandomState(seed)
    r, t3, t2 = _persistent_walk(rs, N, lp, W.BEAD_SPACING)
    p = sswlc_bond_params(np.full(N - 1, W.BEAD_SPACING), lp)
    bonds = np.stack([p[1], p[3], p[4], p[2], p[5]])
    chain = ChainSpec(r=r, t3=t3, t2=t2, states=np.zeros((N, 1), dtype=np.int64),
                      marks=np.zeros((N, 1), dtype=np.int64), bonds=bonds, sites=[0], mu=[0.0],
                      F=np.zeros((1, 1, 1)), max_binders=-1)
    ctrls = W.default_controllers(N, per_cycle=(30, 1, 60, 60, 0))
    return chain, FieldSpec(field_on=False), ctrls


def test_wlc_end_to_end_distances():
    """64 chains of 500 beads (bead spacing 25 nm, lp = 100 nm, as the reference's
    validation), field off, crank-shaft / end-pivot / slide / tangent moves,
    started from a discrete wormlike chain and run 1,500 sweeps before sampling.
    Bound: 6 % (the discretised SSWLC differs from the continuous chain by a few
    per cent at the bond scale) plus five standard errors."""
    from chromo_amd.engine import DeviceReplicas
    n_rep, N, lp = 64, 500, 100.0
    parts = [_wlc_system(N, lp, 500 + k) for k in range(n_rep)]
    dr = DeviceReplicas([p[0] for p in parts], [p[1] for p in parts], [p[2] for p in parts],
                        rng_mode=L.RNG_PHILOX, seeds=[23] * n_rep, streams=list(range(n_rep)))
    say("wlc: equilibrate")
    dr.run(1500, waves=2)
    seps = np.array([1, 2, 4, 8, 16, 32, 64])
    n_snap = 40
    acc = np.zeros((n_snap, len(seps)))
    for t in range(n_snap):
        dr.run(25, waves=2)
        vals = np.zeros(len(seps))
        for k in range(n_rep):
            r = dr.positions(k)[0]
            for j, n in enumerate(seps):
                vals[j] += np.mean(np.sum((r[n:] - r[:-n]) ** 2, axis=1))
        acc[t] = vals / n_rep / (2 * lp) ** 2
    say("wlc: sampled")
    x = seps * W.BEAD_SPACING / (2 * lp)
    theory = x - 0.5 + np.exp(-2 * x) / 2
    est = acc.mean(axis=0)
    se = acc.std(axis=0, ddof=1) / math.sqrt(n_snap)
    say("  x", x.tolist())
    say("  <R^2>/(2lp)^2 sim", est.tolist())
    say("  theory", theory.tolist())
    rel = np.abs(est - theory) / theory
    assert np.all(rel < 0.06 + 5 * se / theory), (rel, se / theory)
