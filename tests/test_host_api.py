"""Host-side API (no GPU): reference class/function names, snapshot format,
field prefactors, amplitude bounds, controllers."""
import copy
import ctypes

import numpy as np
import pytest


def _chromatin(N=50, seed=0):
    from chromo_amd.polymers import Chromatin
    np.random.seed(seed)
    rs = np.random.RandomState(seed)
    return Chromatin.gaussian_walk_polymer(
        "chr", N, np.full(N - 1, 25.0), states=rs.randint(0, 3, size=(N, 2)),
        binder_names=np.array(["HP1", "PRC1"]),
        chemical_mods=(rs.random_sample((N, 2)) < 0.3).astype(int),
        chemical_mod_names=np.array(["H3K9me3", "H3K27me3"]))


def test_chromatin_attributes():
    p = _chromatin()
    assert p.lp == 53.0 and p.num_beads == 50 and p.num_binders == 2 and p.n_binders_p1 == 3
    assert p.bead_rad == 5
    assert abs(p.beads[0].vol - (4 / 3) * np.pi * 125) < 1e-12
    assert [b.name for b in p.beads[3].binders] == ["HP1", "PRC1"]
    assert p.eps_bend.shape == (49,) and np.all(p.eps_bend > 0)
    assert len(p.beads) == 50
    with pytest.raises(ValueError):
        from chromo_amd.polymers import Chromatin
        Chromatin("x", np.zeros((5, 3)), bead_length=np.ones(4), states=np.zeros((4, 2), int),
                  binder_names=np.array(["HP1", "PRC1"]))


def test_snapshot_roundtrip(tmp_path):
    from chromo_amd.polymers import Chromatin
    p = _chromatin(40, 1)
    path = str(tmp_path / "chr-0.csv")
    p.to_csv(path)
    q = Chromatin.from_file(path)
    assert q.name == "chr-0"
    assert np.array_equal(q.r, p.r) and np.array_equal(q.t3, p.t3) and np.array_equal(q.t2, p.t2)
    assert np.array_equal(q.states, p.states) and np.array_equal(q.chemical_mods, p.chemical_mods)
    assert list(q.binder_names) == ["HP1", "PRC1"]
    assert np.allclose(q.bead_length, p.bead_length)


def test_sswlc_parameters_interpolate_table():
    from chromo_amd.polymers import sswlc_bond_params
    from chromo_amd.util import dss_params
    d, eb, g, ep, eperp, eta = sswlc_bond_params(np.array([25.0, 25.0, 10.0]), 53.0)
    D = 25.0 / 53.0
    assert d[0] == D
    assert eb[0] == np.interp(D, dss_params[:, 0], dss_params[:, 1]) / D
    assert eta[2] == np.interp(10 / 53, dss_params[:, 0], dss_params[:, 5]) / 53.0
    assert eb[0] == eb[1] and eb[0] != eb[2]


def test_field_prefactors_formula():
    from chromo_amd.binders import hp1, prc1
    from chromo_amd.fields import field_prefactors
    vb = 27000.0
    pref, intra, cross = field_prefactors([hp1, prc1], vb)
    v = (4.0 / 3.0) * np.pi * 3 ** 3
    assert pref[0] == 0.5 * -4 * v * vb
    assert intra[1] == -4 * (1 - v / vb)
    assert np.all(cross == 0)        # HP1 <-> PRC1 cross-talk defaults to 0


def test_binder_collection_and_lookup():
    from chromo_amd.binders import get_by_name, hp1, make_binder_collection
    df = make_binder_collection(["HP1", "PRC1"])
    assert list(df["name"]) == ["HP1", "PRC1"] and df["sites_per_bead"].tolist() == [2, 2]
    assert get_by_name("HP1") is hp1
    with pytest.raises(ValueError):
        get_by_name("no such binder")


def test_amplitude_bounds():
    from chromo_amd import mc
    p = _chromatin(1000, 2)
    b, m = mc.get_amplitude_bounds([p])
    assert b.bounds["crank_shaft"] == (30, 150) and b.bounds["end_pivot"] == (50, 150)
    assert b.bounds["tangent_rotation"] == (1, 1000)
    assert m.bounds["slide"] == (0.2 * 25.0, 0.3 * 25.0)
    small = _chromatin(40, 3)
    b2, _ = mc.get_amplitude_bounds([small])
    assert b2.bounds["end_pivot"] == (10.0, 20) and b2.bounds["crank_shaft"] == (30, 40)


def test_all_moves_schedule():
    from chromo_amd.mc import mc_controller as ctl
    from chromo_amd import mc
    p = _chromatin(200, 4)
    b, m = mc.get_amplitude_bounds([p])
    cs = ctl.all_moves(".", b.bounds, m.bounds, controller=ctl.SimpleControl)
    assert [c.move.name for c in cs] == ["crank_shaft", "end_pivot", "slide", "tangent_rotation",
                                         "change_binding_state"]
    assert [c.move.num_per_cycle for c in cs] == [30, 1, 60, 60, 10]
    assert cs[0].move.amp_bead == 30 and cs[0].move.amp_move == 0.1 * np.pi
    assert cs[0].move.acceptance_tracker.alpha == 2 / 21
    cs2 = ctl.all_moves_except_binding_state(".", b.bounds, m.bounds)
    assert [c.move.num_per_cycle for c in cs2] == [5, 5, 5, 5]


@pytest.mark.parametrize("rate,amp,ab", [(0.3, 1.0, 10), (0.3, 0.52, 10), (0.7, 1.95, 10),
                                         (0.7, 1.0, 19), (0.5, 1.0, 10), (0.3, 0.52, 5)])
def test_host_controller_matches_oracle(oracle, rate, amp, ab):
    """SimpleControl.update_amplitudes on the host equals the device/oracle recurrence."""
    from chromo_amd.mc import mc_controller as ctl
    from chromo_amd.mc.moves import MCAdapter, crank_shaft
    ad = MCAdapter(".", "", crank_shaft, 20, ab, amp)
    ad.acceptance_tracker.acceptance_rate = rate
    c = ctl.SimpleControl(ad, (5, 20), (0.5, 2.0))
    c.update_amplitudes()
    o = oracle.OrCtrl(0, 1, 1, 1, amp, ab, rate, 2 / 21, 5, 20, 0.5, 2.0, 0, 0)
    oracle.lib().or_update_amplitudes(ctypes.byref(o))
    assert (ad.amp_move, ad.amp_bead) == (o.amp_move, o.amp_bead)


def test_moves_are_device_markers():
    from chromo_amd.mc import moves
    assert moves.crank_shaft.__name__ == "crank_shaft"
    with pytest.raises(TypeError):
        moves.slide(None, 1.0, 1)


def test_mu_schedules():
    """The annealing factors at the start, middle and end of a run, against the
    closed forms of chromo/util/mu_schedules.py."""
    import math
    from chromo_amd.util import mu_schedules as M
    N = 120
    assert M.linear_1(0, N) == -1 and M.linear_1(N, N) == 1
    assert M.linear_2(30, N) == -1 and M.linear_2(31, N) == 2 * (1 / 90) - 1 and M.linear_2(N, N) == 1
    assert M.linear_3(20, N) == -1 and M.linear_4(15, N) == -1 and M.linear_4(N, N) == 1
    assert math.isclose(M.tanh_1(0, N), math.tanh(-3)) and math.isclose(M.tanh_2(N, N), math.tanh(2))
    assert M.tanh_3(20, N) == math.tanh(-1) and math.isclose(M.tanh_4(N, N), math.tanh(2))
    assert math.isclose(M.log_1(0, N), -1) and math.isclose(M.log_2(N, N), math.log(4.4) - 0.5)
    assert math.isclose(M.log_3(0, N), math.log(0.5) - 0.098)
    assert M.sawtooth_1(N - 1, N) == 1 and math.isclose(M.sawtooth_1(40, N), math.log(40 / 12 + 0.5) - 0.34)
    assert M.linear_2_for_negative_cp(0, N) == 4.0 and M.linear_2_for_negative_cp(N, N) == 1.0
    assert M.linear_step_for_negative_cp(24, N) == 4.0 and M.linear_step_for_negative_cp(96, N) == 1.0
    assert math.isclose(M.linear_step_for_negative_cp(60, N), 4 - 3 * 36 / 72)
    assert M.linear_step_for_negative_cp_mild(0, N) == 2.0 and M.linear_step_for_negative_cp_mild(N, N) == 1.0
    s = M.Schedule(M.tanh_3)
    assert s.name == "tanh_3" and s.function(0, N) == math.tanh(-1)
