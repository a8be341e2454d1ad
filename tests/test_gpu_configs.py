"""The HIP path against the oracle on every benchmark configuration and on the
kernel branches the other parity tests do not reach.

* the 100,000-bead / 64^3 / Markov-mark chain on the plain path (one launch per
  proposal) for more than 10^4 proposals;
* four full-size replicas of the 512-replica workload, tangent rotations on,
  with the two-wavefront launch the benchmark uses;
* a 1,000-bead / 30^3 chain for 500 sweeps (80,500 proposals) in the glibc
  stream: every decision, window and voxel count equal over a long horizon;
* voxels over vf_limit (the 1e99 volume-fraction penalty), spherical and
  cubical confinement, max_binders (the 1e25 binder penalty) and
  mu_adjust_factor != 1 with chemical potentials of both signs;
* a twisting chain (SSTWLC) on every evaluator (one-wavefront segments,
  64-bead segments, chunked large moves, the plain path) and its absolute
  energy.

Bar as in test_gpu_parity.py: windows, voxel counts, decisions, Metropolis
draws and binder states equal; dE within 1e-9 kT absolute or 1e-12 relative
(1e-9 relative where a 1e25 or 1e99 penalty enters, whose rounding residue
depends on the summation order); positions within 1e-9 nm.
"""
import copy
import sys

import numpy as np
import pytest

from chromo_amd import _lib as L
from chromo_amd import workloads as W

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _dev(systems, **kw):
    from chromo_amd.engine import DeviceReplicas
    return DeviceReplicas([s.chain for s in systems], [s.field for s in systems],
                          [copy.deepcopy(s.ctrls) for s in systems], **kw)


def _lower_voxels(field, r):
    n = np.array([field.nx, field.ny, field.nz])
    Wd = np.array([field.x_width, field.y_width, field.z_width])
    dx = Wd / n
    return np.floor((np.mod(r + Wd / 2, Wd) - dx / 2) / dx).astype(np.int64)


def _dE_close(a, b, rel=1e-12):
    """1e-9 kT, or rel of the larger value; a dE above 1e50 kT carries 1e99 volume
    fraction or confinement terms whose cancellation residue (up to ~1e84 kT at
    volume fractions near 1) depends on the summation order."""
    big = np.maximum(np.abs(a), np.abs(b))
    return np.abs(a - b) <= 1e-9 + rel * big + np.where(big > 1e50, 1e85, 0.0)


def _compare(lg, lo, rel=1e-12):
    assert len(lg) == len(lo), (len(lg), len(lo))
    for f in ("move", "n", "ind0", "indf", "n_bins", "accepted"):
        bad = np.nonzero(lg[f] != lo[f])[0]
        assert bad.size == 0, (f, bad[:3], lg[bad[:1]], lo[bad[:1]])
    assert np.array_equal(lg["u"], lo["u"])
    for f in ("dE_poly", "dE_field"):
        bad = np.nonzero(~_dE_close(lg[f], lo[f], rel))[0]
        assert bad.size == 0, (f, bad[:3], lg[f][bad[:3]], lo[f][bad[:3]])


def _compare_state(dr, i, o, s):
    r, t3, t2 = dr.positions(i)
    ro, t3o, t2o, sto = o.state()
    assert np.array_equal(dr.binder_states(i), sto)
    assert np.max(np.abs(r - ro)) < 1e-9 and np.max(np.abs(t3 - t3o)) < 1e-9
    assert np.max(np.abs(t2 - t2o)) < 1e-9
    if s.field.field_on:
        assert np.array_equal(_lower_voxels(s.field, r), _lower_voxels(s.field, ro))
        one_bead = s.field.n_bins / (s.field.x_width * s.field.y_width * s.field.z_width)
        assert np.max(np.abs(dr.density_of(i) - o.density())) < 1e-9 * one_bead


def _run_pair(oracle, s, steps, *, rng_mode=L.RNG_GLIBC, seed=1, stream=0, inner_seed=123,
              waves=1, launch="run", mu_adjust=1.0, rel=1e-12):
    ctrls_o = copy.deepcopy(s.ctrls)
    o = oracle.OracleSim(s.chain, s.field, rng_mode=rng_mode, seed=seed, stream=stream)
    o.np_seed(inner_seed)
    lo = o.mc_sim(steps, ctrls_o, mu_adjust_factor=mu_adjust, max_logs=2 * 10 ** 6)
    dr = _dev([s], rng_mode=rng_mode, seeds=[seed], streams=[stream], log_cap=len(lo) + 100)
    dr.np_seed([inner_seed])
    if mu_adjust != 1.0:
        dr.set_mu_adjust([mu_adjust])
    say(f"  {launch} waves={waves} proposals={len(lo)}")
    if launch == "run":
        dr.run(steps, waves=waves)
    else:
        dr.run_plain(steps)
    _compare(dr.logs(0), lo, rel)
    _compare_state(dr, 0, o, s)
    for a, b in zip(dr.controllers(0), ctrls_o):
        assert a.amp_bead == b.amp_bead and a.num_success == b.num_success and a.rate == b.rate
    cnt = dr.counter_array()[0]
    assert cnt[L.CNT_ERROR] == 0 and cnt[L.CNT_PROPOSALS] == len(lo)
    return lo, dr


# ------------------------------------------------------ benchmark configs ---

def test_plain_path_full_size_markov_chain(oracle):
    """100,000 beads, 64^3 voxels of 30 nm, B = 2, two-state Markov marks (mean
    domain 20): 63 sweeps of the reference schedule, one launch per proposal."""
    say("plain path, 100k beads, 64^3")
    s = W.config_plain_gpu()
    lo, _ = _run_pair(oracle, s, 63, launch="plain")
    assert len(lo) >= 10_000
    assert lo["accepted"].sum() > 1000


@pytest.mark.parametrize("rep", [0, 1, 2, 3])
def test_batched_workload_replicas_full_size(oracle, batched_systems, rep):
    """Replicas of the 512-replica workload at full size (Philox streams, tangent
    rotations on), advanced with two wavefronts each as the benchmark does."""
    s = batched_systems[rep]
    say(f"batched replica {rep}: N={s.chain.N}, grid {s.field.nx}^3")
    lo, _ = _run_pair(oracle, s, 30, rng_mode=L.RNG_PHILOX, seed=20240601, stream=rep, waves=2)
    assert np.any(lo["move"] == L.TANGENT_ROTATION)


@pytest.fixture(scope="module")
def batched_systems():
    systems = W.config_batched(512, seed=3)
    # the four replicas with the shortest and the longest chains and two between
    order = np.argsort([s.chain.N for s in systems])
    return [systems[i] for i in (order[0], order[-1], order[170], order[340])]


def test_long_horizon_config1(oracle):
    """1,000 beads on 30^3 voxels (B = 2, Bernoulli(0.3) marks) for 500 sweeps
    in the reference's glibc stream: no decision, window or voxel count drifts
    over 80,500 proposals (device sin/cos/acos/log10 may differ from glibc by an
    ulp; a flip of a floor or a truncation would show here)."""
    say("config 1, 500 sweeps")
    s = W.config_plain_cpu()
    lo, dr = _run_pair(oracle, s, 500, waves=2)
    assert len(lo) > 70_000
    c = dr.counter_array()[0]
    say(f"  near ties: Metropolis {c[L.CNT_NEAR_TIES]}, voxel floor {c[L.CNT_TIE_VOXEL]}")


# --------------------------------------------------------- kernel branches ---

def test_volume_fraction_limit_branch(oracle):
    """Ten groups of 35 beads within 1.5 nm of voxel centres put ten isolated
    voxels above vf_limit = 0.5 (35 x ~0.85 x 523.6 nm^3 / 27,000 nm^3 ~ 0.58):
    moves into them pay 1e99 per volume fraction and moves out gain it.

    The groups are not exactly coincident: a bead that a rotation leaves exactly
    where it was (coincident with the fulcrum) gives a voxel over the limit with
    no density change, where the reference's running sum +1e99 vf - 1e99 vf drops
    the terms that CPython's set order put before it; the device counts such
    voxels (CHROMO_CNT_TIE_HUGE) and sums in its own order there."""
    say("vf_limit branch")
    s = W.make_system(1500, 2, 16, seed=61)
    f = s.field
    dx = f.x_width / f.nx
    rs = np.random.RandomState(7)
    cells = []
    while len(cells) < 10:                         # groups at least 3 voxels apart (periodic)
        c = rs.randint(0, f.nx, size=3)
        if all(np.max(np.minimum(np.abs(c - d), f.nx - np.abs(c - d))) >= 3 for d in cells):
            cells.append(c)
    for g in range(10):
        # voxel centres sit at -W/2 + (i + 1/2) dx
        centre = -f.x_width / 2 + (cells[g] + 0.5) * dx
        s.chain.r[100 + 140 * g: 135 + 140 * g] = centre + rs.uniform(-1.5, 1.5, size=(35, 3))
    o = oracle.OracleSim(s.chain, s.field)
    phi = o.density()[:, 0] * f.bead_vol
    assert np.sum(phi > 0.5) >= 10
    lo, dr = _run_pair(oracle, s, 8, waves=2, rel=1e-9)
    assert np.any(np.abs(lo["dE_field"]) > 1e90)
    say(f"  voxels over the limit before and after a move: {dr.counter_array()[0][L.CNT_TIE_HUGE]}")


@pytest.mark.parametrize("kind,length", [(1, 300.0), (2, 500.0)], ids=["spherical", "cubical"])
@pytest.mark.parametrize("waves", [1, 2])
def test_confinement_branch(oracle, kind, length, waves):
    """A chain that starts partly outside a sphere (cube) that holds most of it:
    moves out pay 1e99 per bead (per coordinate), moves back in gain it
    (spherical) or not (cubical: the current configuration never counts)."""
    say("confinement", kind, waves)
    s = W.make_system(800, 2, 24, seed=62)
    s.chain.r -= s.chain.r.mean(axis=0)
    s.field.confine_type = kind
    s.field.confine_length = length
    out = (np.linalg.norm(s.chain.r, axis=1) > length if kind == 1
           else np.any(np.abs(s.chain.r) > length / 2, axis=1))
    assert 0 < out.sum() < 800
    lo, _ = _run_pair(oracle, s, 8, waves=waves, rel=1e-9)
    assert np.any(lo["dE_field"] > 1e90) and not np.all(lo["accepted"] == 0)


@pytest.mark.parametrize("waves", [1, 2])
def test_max_binders_branch(oracle, waves):
    """max_binders = 2 with two 2-site binders and favourable potentials: states
    with three or four bound binders pay 1e25 per binder over the limit."""
    say("max_binders", waves)
    s = W.make_system(600, 2, 10, seed=63, mu=[1.5, 0.8], per_cycle=(30, 1, 60, 60, 80))
    s.chain.max_binders = 2
    lo, dr = _run_pair(oracle, s, 10, waves=waves, rel=1e-9)
    binding = lo[lo["move"] == L.CHANGE_BINDING_STATE]
    assert np.any(binding["dE_poly"] > 1e24)
    assert dr.binder_states(0).sum(axis=1).max() <= 2


@pytest.mark.parametrize("f", [0.35, 1.6])
@pytest.mark.parametrize("rng_mode", [L.RNG_GLIBC, L.RNG_PHILOX], ids=["glibc", "philox"])
def test_mu_adjust_factor_branch(oracle, f, rng_mode):
    """mu_adjust_factor != 1 scales a negative mu by f and a positive one by 2 - f
    (chromo/polymers.pyx:1522-1537)."""
    say("mu_adjust", f, rng_mode)
    s = W.make_system(700, 3, 10, seed=64, mu=[-0.7, 0.9, 0.2], eps=[-3.0, -2.0, -4.0],
                      per_cycle=(30, 1, 60, 60, 60))
    lo, _ = _run_pair(oracle, s, 8, waves=2, rng_mode=rng_mode, seed=5, mu_adjust=f)
    assert np.sum(lo["move"] == L.CHANGE_BINDING_STATE) > 100


def test_capped_exponential_truncation_matches_glibc(oracle):
    """Every bead window of the default amplitude bounds (windows 1..150) draws
    through int(-log10(x / RAND_MAX + 1e-5) * w * 0.45 + 1.0001).  On the draws
    within 40 of every integer crossing -- the only draws where an ulp of log10
    could change the truncation -- the device value equals glibc's, so the
    device's bead windows equal the reference's for all 2^31 draws."""
    import torch
    from cexp_crossings import crossing_draws
    say("capped exponential vs glibc, windows 1..150")
    d, w = crossing_draws(range(1, 151), radius=40)
    host = oracle.cexp_trunc(d, w)
    dd = torch.tensor(d, device="cuda")
    dw = torch.tensor(w, device="cuda")
    out = torch.zeros(len(d), dtype=torch.int64, device="cuda")
    lib = L.lib()
    L.check(lib.chromo_capped_exponential_tape(dd.data_ptr(), dw.data_ptr(), len(d), out.data_ptr(),
                                               torch.cuda.current_stream().cuda_stream),
            "chromo_capped_exponential_tape")
    dev = out.cpu().numpy()
    bad = np.nonzero(dev != host)[0]
    say(f"  {len(d)} draws, {bad.size} differ")
    assert bad.size == 0, (d[bad[:5]], w[bad[:5]], dev[bad[:5]], host[bad[:5]])


def test_accessible_volumes_spherical_confinement(oracle):
    """Spherical confinement with assume_fully_accessible = 0: the voxels the
    sphere's surface splits get the fraction of their 20^3 subgrid inside
    (chromo/fields.pyx:714-951), which divides every bead weight in them and sets
    their chi volume.  The initial densities bit for bit, the tape (delta
    densities bit for bit) and a run of every move type against the oracle."""
    from chromo_amd.binders import hp1, make_binder_collection
    from chromo_amd.fields import UniformDensityField
    say("accessible volumes")
    s = W.make_system(1200, 2, 24, seed=65)
    s.chain.r -= s.chain.r.mean(axis=0)
    R = 300.0
    f = s.field
    udf = UniformDensityField([], make_binder_collection([hp1]), f.x_width, f.nx, f.y_width, f.ny,
                              f.z_width, f.nz, confine_type="Spherical", confine_length=R,
                              assume_fully_accessible=0)
    f.access_vol = udf.access_vol_array()
    f.confine_type, f.confine_length = 1, R
    assert f.access_vol is not None and np.sum(f.access_vol < udf.vol_bin) > 100
    o = oracle.OracleSim(s.chain, s.field)
    dr = _dev([s])
    assert np.array_equal(dr.density_of(0), o.density())
    rs = np.random.RandomState(3)
    for n in (1, 9, 40):
        ind0 = rs.randint(0, 1200 - n)
        inds = np.arange(ind0, ind0 + n)
        rt = s.chain.r[inds] + rs.normal(scale=15.0, size=(n, 3))
        st = s.chain.states[inds]
        dEo, bins_o, delta_o = o.field_dE_tape(inds, rt, st, 0)
        dEg, bins_g, delta_g = dr.field_dE_tape(0, inds, rt, st.astype(np.int8), False)
        assert sorted(bins_g) == sorted(bins_o)
        assert np.array_equal(delta_g, delta_o[bins_g])
        assert _dE_close(np.array([dEg]), np.array([dEo]))[0]
    _run_pair(oracle, s, 8, waves=2, rel=1e-9)


@pytest.mark.parametrize("npts,waves", [(1000, 2), (6, 1), (6, 2)])
def test_fast_field(oracle, npts, waves):
    """fast_field = 1: the move dE from sub-bin weights (chromo/fields.pyx:577-671,
    1235-1368) with no 1e-18 rules, the stored densities from the exact scatter.
    Tape delta densities bit for bit, then a run of every move type."""
    say("fast field", npts, waves)
    s = W.make_system(1500, 3, 16, seed=66, eps=[-4, -3, -2], cross_eps=0.5, mu=[-0.4, 0.2, -1])
    s.field.fast_field, s.field.n_points = True, npts
    o = oracle.OracleSim(s.chain, s.field)
    dr = _dev([s])
    assert np.array_equal(dr.density_of(0), o.density())
    rs = np.random.RandomState(4)
    for n in (1, 12, 45):
        ind0 = rs.randint(0, 1500 - n)
        inds = np.arange(ind0, ind0 + n)
        rt = s.chain.r[inds] + rs.normal(scale=8.0, size=(n, 3))
        st = s.chain.states[inds]
        dEo, bins_o, delta_o = o.field_dE_tape(inds, rt, st, 0)
        dEg, bins_g, delta_g = dr.field_dE_tape(0, inds, rt, st.astype(np.int8), False)
        assert sorted(bins_g) == sorted(bins_o)
        assert np.array_equal(delta_g, delta_o[bins_g])
        assert _dE_close(np.array([dEg]), np.array([dEo]))[0]
    _run_pair(oracle, s, 8, waves=waves)


# ------------------------------------------------------------------ twist ---

@pytest.mark.parametrize("N,launch", [(1500, "run"), (300, "run"), (1500, "plain")])
def test_twist_chain(oracle, N, launch):
    """SSTWLC: E_pair_with_twist in the forward and reverse pairs of every move
    (chromo/polymers.pyx:2003-2248).  The speculative evaluator has no twist
    terms, so run(waves=2) runs one wavefront and gives the same chain."""
    say("twist", N, launch)
    s = W.make_twist_system(N, 2, 12, seed=71 + N, lt=100.0, eps=[-4, -2], mu=[-0.4, 0.3])
    _run_pair(oracle, s, 10, waves=2, launch=launch)


def test_twist_absolute_energy(oracle):
    """SSTWLC.compute_E and compute_E_no_twist on the GPU against the oracle; the
    energy adjoint refuses a twisting chain."""
    import torch
    from chromo_amd.energy import EnergyModel, energy
    from chromo_amd.engine import FieldSpec
    from chromo_amd.polymers import SSTWLC
    say("twist energy")
    s = W.make_twist_system(2000, 2, 12, seed=5, lt=100.0)
    c = s.chain
    p = SSTWLC("tw", c.r, t3=c.t3, t2=c.t2, bead_length=c.bead_length, lp=W.LP_CHROMATIN, lt=100.0)
    assert np.array_equal(p.bond_params(), c.bonds)
    null = FieldSpec(field_on=False)
    Eo = oracle.OracleSim(c, null).elastic_E()
    c0 = copy.deepcopy(c)
    c0.twist, c0.bonds = False, c0.bonds[:5]
    Eo0 = oracle.OracleSim(c0, null).elastic_E()
    assert abs(p.compute_E() - Eo) <= 1e-12 * abs(Eo)
    assert abs(p.compute_E_no_twist() - Eo0) <= 1e-12 * abs(Eo0)
    assert Eo > Eo0
    m = EnergyModel(c, s.field, s.binders)
    with pytest.raises(L.ChromoError, match="twist"):
        energy(m, torch.as_tensor(c.r), torch.as_tensor(m.theta0))


# ------------------------------------------------ long windows, two wavefronts ---

@pytest.mark.parametrize("rng_mode", [L.RNG_PHILOX, L.RNG_GLIBC], ids=["philox", "glibc"])
def test_long_window_blocks_two_wavefronts(oracle, rng_mode):
    """Steady-state amplitudes (crank-shaft, slide and end-pivot windows of 100-150
    beads) with two wavefronts: about a third of those proposals exceed the
    speculative evaluator and run as sequential slots between speculative rounds,
    evaluated by both wavefronts (the 64-bead and chunked evaluators); the chain
    is the oracle's."""
    say("long windows, W=2", rng_mode)
    s = W.make_system(6000, 2, 24, seed=97, eps=[-4, -2], mu=[-0.4, 0.3])
    for c in s.ctrls:
        if c.move in (L.CRANK_SHAFT, L.SLIDE, L.END_PIVOT):
            c.amp_bead, c.bead_lo, c.bead_hi = 130, 100.0, 150.0
    lo, _ = _run_pair(oracle, s, 6, rng_mode=rng_mode, waves=2, rel=1e-9)
    assert np.any(lo["n"] > 64)


@pytest.mark.parametrize("B", [1, 3, 4])
def test_long_window_blocks_two_wavefronts_binder_counts(oracle, B):
    """The two-wavefront sequential evaluation (33-64-bead and chunked windows)
    in the launches of the other binder counts, against the oracle."""
    say("long windows, W=2, B =", B)
    s = W.make_system(2500, B, 16, seed=60 + B, eps=-3.0, cross_eps=0.3 if B > 1 else 0.0, mu=-0.5)
    for c in s.ctrls:
        if c.move in (L.CRANK_SHAFT, L.SLIDE, L.END_PIVOT):
            c.amp_bead, c.bead_lo, c.bead_hi = 120, 40.0, 150.0
    lo, _ = _run_pair(oracle, s, 5, rng_mode=L.RNG_PHILOX, waves=2, rel=1e-9)
    n = lo["n"]
    assert np.any((n > 32) & (n <= 64)) and np.any(n > 64)
