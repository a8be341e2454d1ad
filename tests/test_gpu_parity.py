"""Parity of the HIP path (through the C ABI) with the CPU oracle.

Bar: voxel indices, unique-voxel sets, binder states, proposal bead ranges and
accept/reject decisions identical; per-voxel delta densities and the initial
density grid bit-identical; dE within 1e-9 kT (absolute, and relative for
E_HUGE penalties); positions within 1e-9 nm (device sin/cos/acos/log10 may
differ from glibc's by an ulp).
"""
import copy
import sys

import numpy as np
import pytest

from chromo_amd import _lib as L
from chromo_amd import workloads as W

pytestmark = pytest.mark.gpu

# Eight and four wavefronts per replica: an earlier source of those shapes faulted the
# GPU (8) and hung and gave wrong voxel counts (4), with no cause found.  The current
# source cannot build them: the speculative round's flags have one wavefront on each
# side, and the per-wavefront LDS blocks are sized for two.  The entry point rejects
# both; these cases stay as skips until a design for more wavefronts exists.
W8_SKIP = "waves_per_replica = 8 is not built (two-wavefront round protocol); an earlier build faulted"
W4_SKIP = "waves_per_replica = 4 is not built (two-wavefront round protocol); an earlier build hung"
P4 = pytest.param(4, marks=pytest.mark.skip(reason=W4_SKIP))


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _dev(systems, **kw):
    from chromo_amd.engine import DeviceReplicas
    return DeviceReplicas([s.chain for s in systems], [s.field for s in systems],
                          [copy.deepcopy(s.ctrls) for s in systems], **kw)


def _lower_voxels(field, r):
    """Lower-corner voxel index of every bead (chromo/fields.pyx:1434-1457)."""
    n = np.array([field.nx, field.ny, field.nz])
    Wd = np.array([field.x_width, field.y_width, field.z_width])
    dx = Wd / n
    q = (np.mod(r + Wd / 2, Wd) - dx / 2) / dx
    return np.floor(q).astype(np.int64)


def _dE_close(a, b):
    tol = 1e-9 + 1e-12 * np.maximum(np.abs(a), np.abs(b))
    return np.abs(a - b) <= tol


# ---------------------------------------------------------------- density ---

# 96^3 voxels reach the gather that ranks voxels inside 1024-voxel blocks (grids of
# at least 512 such blocks); the others use 256-voxel blocks
@pytest.mark.parametrize("N,nvox,seed", [(300, 7, 0), (2000, 16, 1), (5000, 12, 2), (64, 1, 3),
                                         (64, 2, 4), (1000, 5, 5), (200000, 96, 6)])
def test_density_init_bitwise(oracle, N, nvox, seed):
    say("density_init", N, nvox, seed)
    s = W.make_system(N, 2, nvox, seed=seed)
    s.chain.states = np.random.RandomState(seed).randint(0, 3, size=(N, 2))
    dr = _dev([s])
    g = dr.density_of(0)
    o = oracle.OracleSim(s.chain, s.field).density()
    assert np.array_equal(g, o)


def test_density_init_edge_positions(oracle):
    say("density_init edges")
    s = W.make_system(8, 2, 4, seed=0)
    Wd = s.field.x_width
    dx = Wd / 4
    s.chain.r = np.array([[-Wd / 2, 0, 0], [Wd / 2, Wd / 2, Wd / 2], [-Wd / 2 + 0.5 * dx] * 3,
                          [0, 0, 0], [Wd / 2 - 1e-12, -Wd / 2 + 1e-12, 3 * Wd], [dx, -dx, 2 * dx],
                          [-7 * Wd, 5.5 * Wd, 0.25], [1e-300, -1e-300, 0.0]])
    s.chain.states = np.array([[1, 2]] * 8)
    dr = _dev([s])
    assert np.array_equal(dr.density_of(0), oracle.OracleSim(s.chain, s.field).density())


# ------------------------------------------------------------------- tape ---

@pytest.mark.parametrize("n_move", [1, 3, 17, 32, 33, 80, 150])
@pytest.mark.parametrize("seed", [0, 1])
def test_field_tape(oracle, n_move, seed):
    say("field_tape", n_move, seed)
    N = 400
    s = W.make_system(N, 2, 9, seed=seed, eps=[-3.0, -2.0], cross_eps=0.7, chi=1.4)
    rs = np.random.RandomState(seed)
    s.chain.states = rs.randint(0, 3, size=(N, 2))
    o = oracle.OracleSim(s.chain, s.field)
    dr = _dev([s])
    for trial in range(4):
        ind0 = rs.randint(0, N - n_move)
        inds = np.arange(ind0, ind0 + n_move)
        a = rs.normal(size=3)
        T = oracle.axis_rotation(a / np.linalg.norm(a), s.chain.r[ind0], rs.uniform(-1.5, 1.5))
        rt, _, _ = oracle.transform(T, s.chain.r[inds], s.chain.t3[inds], s.chain.t2[inds])
        st = s.chain.states[inds]
        dEo, bins_o, delta_o = o.field_dE_tape(inds, rt, st, 0)
        dEg, bins_g, delta_g = dr.field_dE_tape(0, inds, rt, st.astype(np.int8), False)
        assert sorted(bins_g) == sorted(bins_o)
        assert np.array_equal(delta_g, delta_o[bins_g])           # per-voxel fold, bit for bit
        assert _dE_close(dEg, dEo)
        # binding (state change) on one bead
        k = rs.randint(0, N)
        st2 = s.chain.states[[k]].copy()
        st2[0, trial % 2] = (st2[0, trial % 2] + 1) % 3
        dEo, bins_o, delta_o = o.field_dE_tape([k], s.chain.r[[k]], st2, 1)
        dEg, bins_g, delta_g = dr.field_dE_tape(0, [k], s.chain.r[[k]], st2.astype(np.int8), True)
        assert sorted(bins_g) == sorted(bins_o)
        assert np.array_equal(delta_g, delta_o[bins_g])
        assert _dE_close(dEg, dEo)


# ---------------------------------------------------------------- mc_sim ---

def _compare_run(oracle, s, steps, inner_seed=123, rng_mode=0, seed=1, stream=0, launch="run",
                 waves=1):
    ctrls_o = copy.deepcopy(s.ctrls)
    o = oracle.OracleSim(s.chain, s.field, rng_mode=rng_mode, seed=seed, stream=stream)
    o.np_seed(inner_seed)
    lo = o.mc_sim(steps, ctrls_o, max_logs=10 ** 6)
    dr = _dev([s], rng_mode=rng_mode, seeds=[seed], streams=[stream], log_cap=len(lo) + 100)
    dr.np_seed([inner_seed])
    say("  waves", waves, "launch", launch, "proposals", len(lo))
    if launch == "run":
        dr.run(steps, waves=waves)
    else:
        dr.run_plain(steps)
    lg = dr.logs(0)
    assert len(lg) == len(lo)
    for f in ("move", "n", "ind0", "indf", "n_bins"):
        assert np.array_equal(lg[f], lo[f]), f
    assert np.array_equal(lg["accepted"], lo["accepted"])
    assert np.all(_dE_close(lg["dE_poly"], lo["dE_poly"]))
    assert np.all(_dE_close(lg["dE_field"], lo["dE_field"]))
    assert np.array_equal(lg["u"], lo["u"])
    r, t3, t2 = dr.positions(0)
    ro, t3o, t2o, sto = o.state()
    assert np.array_equal(dr.binder_states(0), sto)
    # unit tangents: the same 1e-9 bound (ulp differences of the device's
    # sin/cos/acos accumulate over repeated rotations of one bead)
    assert np.max(np.abs(r - ro)) < 1e-9 and np.max(np.abs(t3 - t3o)) < 1e-9
    assert np.max(np.abs(t2 - t2o)) < 1e-9
    assert np.array_equal(_lower_voxels(s.field, r), _lower_voxels(s.field, ro))
    # densities: per-voxel folds of weights whose positions may differ by ulps;
    # tolerance 1e-9 of one bead's density per voxel (1 / vol_bin)
    one_bead = 1.0 / (s.field.x_width * s.field.y_width * s.field.z_width / s.field.n_bins)
    assert np.max(np.abs(dr.density_of(0) - o.density())) < 1e-9 * one_bead
    cg = dr.controllers(0)
    for a, b in zip(cg, ctrls_o):
        assert a.amp_bead == b.amp_bead and a.num_attempt == b.num_attempt
        assert a.num_success == b.num_success and a.rate == b.rate
        assert abs(a.amp_move - b.amp_move) <= 1e-15 * abs(b.amp_move)
    cnt = dr.counter_array()[0]
    assert cnt[L.CNT_PROPOSALS] == len(lo) and cnt[L.CNT_ERROR] == 0
    return lo


@pytest.mark.parametrize("waves", [1, 2, P4])
@pytest.mark.parametrize("N,nvox,seed", [(300, 8, 11), (1000, 30, 12), (5000, 20, 13)])
def test_mc_sim_glibc_parity(oracle, N, nvox, seed, waves):
    say("mc_sim glibc", N, nvox, seed)
    s = W.make_system(N, 2, nvox, seed=seed)
    lo = _compare_run(oracle, s, steps=15, waves=waves)
    assert lo["accepted"].sum() > 100


@pytest.mark.parametrize("waves", [1, 2])
@pytest.mark.parametrize("rng_mode", [0, 1], ids=["glibc", "philox"])
@pytest.mark.parametrize("per_cycle", [(120, 0, 0, 0, 0), (0, 0, 120, 0, 0), (0, 40, 0, 0, 0)],
                         ids=["crank_shaft", "slide", "end_pivot"])
def test_single_move_chains(oracle, per_cycle, rng_mode, waves):
    """One move type per chain.  Every proposal reads the window's two neighbour
    beads and bonds (the crank-shaft axis, the end-pivot fulcrum, the polymer dE
    at the window's ends), so a neighbour read that misses its load shows in the
    first few proposals."""
    say("single move", per_cycle, rng_mode, waves)
    s = W.make_system(4000, 2, 20, seed=17, per_cycle=per_cycle)
    lo = _compare_run(oracle, s, steps=12, rng_mode=rng_mode, seed=99, stream=3, waves=waves)
    assert lo["accepted"].sum() > 10


@pytest.mark.parametrize("waves", [1, 2, pytest.param(8, marks=pytest.mark.skip(reason=W8_SKIP))])
def test_mc_sim_philox_parity(oracle, waves):
    say("mc_sim philox")
    s = W.make_system(800, 2, 12, seed=21, chi=0.7, eps=[-4.5, -1.5], mu=[-1.0, 0.5])
    _compare_run(oracle, s, steps=12, rng_mode=1, seed=777, stream=5, waves=waves)


@pytest.mark.parametrize("waves", [1, 2, P4])
def test_mc_sim_three_binders_and_cross_talk(oracle, waves):
    say("mc_sim B=3")
    s = W.make_system(600, 3, 10, seed=22, eps=[-4, -3, -2], cross_eps=0.5, mu=[-0.4, 0.2, -1])
    _compare_run(oracle, s, steps=10, waves=waves)


@pytest.mark.parametrize("B,waves", [(1, 2), (4, 2), pytest.param(1, 4, marks=pytest.mark.skip(reason=W4_SKIP)),
                                     pytest.param(4, 8, marks=pytest.mark.skip(reason=W8_SKIP))])
def test_mc_sim_one_and_four_binders(oracle, B, waves):
    say("mc_sim B", B)
    s = W.make_system(700, B, 10, seed=40 + B, eps=-3.0, cross_eps=0.3 if B > 1 else 0.0,
                      mu=-0.5)
    _compare_run(oracle, s, steps=8, waves=waves)


@pytest.mark.parametrize("waves", [1, 2, P4])
def test_mc_sim_large_moves(oracle, waves):
    """Crank, slide and end-pivot windows up to 150 beads: moves above 32 beads run
    one at a time, up to 64 beads in the wide evaluator block and longer ones in
    the chunked LDS evaluator."""
    say("mc_sim large moves")
    s = W.make_system(1500, 2, 16, seed=23)
    for c in s.ctrls:
        if c.move in (L.CRANK_SHAFT, L.SLIDE, L.END_PIVOT):
            c.amp_bead = int(c.bead_hi)
            c.bead_lo = c.bead_hi
    lo = _compare_run(oracle, s, steps=6, waves=waves)
    assert lo["n"].max() > 32


@pytest.mark.parametrize("nvox,voxel,waves", [(1, W.VOXEL, 1), (2, W.VOXEL, 2), (40, 10.0, 1), (40, 10.0, 2),
                                              (64, 2.0, 2)])
def test_mc_sim_long_moves_tiny_grids_and_overflow(oracle, nvox, voxel, waves):
    """Windows up to 150 beads on grids of one and two voxels per side (corners
    coincide or wrap), and on 10 nm and 2 nm voxels, where long moves touch
    hundreds to thousands of voxels: those beyond what the LDS evaluator holds
    go through the global arena."""
    say("mc_sim long moves", nvox, voxel)
    s = W.make_system(1500, 2, nvox, seed=28, chi=0.1, voxel=voxel)
    s.field.vf_limit = 1e9
    for c in s.ctrls:
        if c.move in (L.CRANK_SHAFT, L.SLIDE, L.END_PIVOT):
            c.amp_bead = int(c.bead_hi)
            c.bead_lo = c.bead_hi
    lo = _compare_run(oracle, s, steps=4, waves=waves)
    assert lo["n"].max() > 64
    if voxel < W.VOXEL:
        assert lo["n_bins"].max() > 1000


@pytest.mark.parametrize("waves", [2, P4, pytest.param(8, marks=pytest.mark.skip(reason=W8_SKIP))])
def test_mc_sim_short_chain_conflicts(oracle, waves):
    """A 40-bead chain on a 3^3 grid: proposals of one round overlap in beads and
    voxels all the time, so most rounds re-evaluate."""
    say("mc_sim short chain")
    s = W.make_system(40, 2, 3, seed=25)
    lo = _compare_run(oracle, s, steps=25, waves=waves)
    assert lo["accepted"].sum() > 50


@pytest.mark.parametrize("nvox,waves", [(1, 2), (2, 2), (1, 1),
                                        pytest.param(1, 4, marks=pytest.mark.skip(reason=W4_SKIP))])
def test_mc_sim_tiny_grids(oracle, nvox, waves):
    """Grids of one or two voxels per side: corners wrap onto the same voxel."""
    say("mc_sim tiny grid", nvox)
    s = W.make_system(120, 2, nvox, seed=26, chi=0.1)
    s.field.vf_limit = 1e9
    _compare_run(oracle, s, steps=6, waves=waves)


@pytest.mark.parametrize("waves,rng_mode", [(1, 0), (2, 0), (2, 1),
                                            pytest.param(4, 0, marks=pytest.mark.skip(reason=W4_SKIP)),
                                            pytest.param(8, 1, marks=pytest.mark.skip(reason=W8_SKIP))])
def test_mc_sim_coincident_beads_redraw_crank_axis(oracle, waves, rng_mode):
    """Beads 200..260 sit on one point, so crank-shaft axes there are shorter
    than 1e-5 and the reference draws a random axis (two extra draws); with
    several wavefronts such a proposal is redone one at a time from its RNG state."""
    say("mc_sim coincident beads")
    s = W.make_system(600, 2, 12, seed=27)
    s.chain.r[200:261] = s.chain.r[200]
    # 61 beads in one voxel exceed vf_limit; keep the 1e99 penalty (whose
    # cancellation residue depends on summation order) out of this test
    s.field.vf_limit = 1e9
    lo = _compare_run(oracle, s, steps=10, waves=waves, rng_mode=rng_mode, seed=3)
    crank = lo[lo["move"] == L.CRANK_SHAFT]
    assert np.any((crank["ind0"] >= 200) & (crank["indf"] <= 261))


def test_all_wave_counts_identical():
    """Every enabled launch shape gives the same chains bit for bit: logs (dE included),
    positions, states, densities, RNG streams and controllers."""
    say("waves 1/2/4/8 identical")
    systems = [W.make_system(n, 2, g, seed=50 + i, chi=0.4 + 0.3 * i)
               for i, (n, g) in enumerate([(3000, 16), (800, 10), (12000, 24)])]
    runs = []
    for w in (1, 2):   # 4 and 8: disabled, see W4_SKIP / W8_SKIP
        dr = _dev(systems, rng_mode=1, seeds=[9] * 3, streams=[0, 1, 2], log_cap=20000)
        dr.run(12, waves=w)
        runs.append(dr)
    base = runs[0]
    for dr in runs[1:]:
        for i in range(3):
            la, lb = base.logs(i), dr.logs(i)
            assert la.tobytes() == lb.tobytes()
            for a, b in zip(base.positions(i), dr.positions(i)):
                assert np.array_equal(a, b)
            assert np.array_equal(base.binder_states(i), dr.binder_states(i))
            assert np.array_equal(base.density_of(i), dr.density_of(i))
            assert np.array_equal(base.rng_words(i), dr.rng_words(i))
            for ca, cb in zip(base.controllers(i), dr.controllers(i)):
                assert ca == cb
        assert np.array_equal(base.counter_array(), dr.counter_array())


def test_plain_equals_persistent(oracle):
    say("plain vs persistent")
    s = W.make_system(500, 2, 10, seed=24)
    _compare_run(oracle, s, steps=4, launch="plain")


def test_batched_equals_single(oracle):
    """Replicas advanced together in one launch give the same chains as alone."""
    say("batched vs single")
    systems = [W.make_system(n, 2, g, seed=30 + i, chi=0.5 + 0.2 * i)
               for i, (n, g) in enumerate([(300, 8), (900, 12), (2000, 16), (450, 9)])]
    together = _dev(systems, rng_mode=1, seeds=[5] * 4, streams=[0, 1, 2, 3])
    together.run(8)
    for i, s in enumerate(systems):
        alone = _dev([s], rng_mode=1, seeds=[5], streams=[i])
        alone.run(8)
        for a, b in zip(together.positions(i), alone.positions(0)):
            assert np.array_equal(a, b)
        assert np.array_equal(together.binder_states(i), alone.binder_states(0))
        assert np.array_equal(together.density_of(i), alone.density_of(0))


def test_density_init_many_beads_per_voxel(oracle):
    """10^5 beads on a 4^3 grid (about 1,600 beads in every bucket): the buckets
    are put in bead order by heapsort, and the scatter stays bitwise the
    reference's bead-order loop."""
    say("density_init 1e5 beads on 4^3")
    s = W.make_system(100000, 2, 4, seed=7)
    s.chain.states = np.random.RandomState(7).randint(0, 3, size=(100000, 2))
    dr = _dev([s])
    assert np.array_equal(dr.density_of(0), oracle.OracleSim(s.chain, s.field).density())
