"""mc_sim of the oracle: schedule, controllers, faithful vs lean build, invariants."""
import copy
import ctypes
import math

import numpy as np
import pytest

from chromo_amd import workloads as W


def test_faithful_equals_lean(oracle):
    """Dropping the O(n_bins) sweeps (fields.pyx:1223-1224, 1971-1975) changes nothing."""
    s = W.make_system(150, 2, 8, seed=4)
    a = oracle.OracleSim(s.chain, s.field, faithful=True)
    b = oracle.OracleSim(s.chain, s.field, faithful=False)
    ca, cb = copy.deepcopy(s.ctrls), copy.deepcopy(s.ctrls)
    a.np_seed(123)
    b.np_seed(123)
    la = a.mc_sim(20, ca, max_logs=10000)
    lb = b.mc_sim(20, cb, max_logs=10000)
    assert len(la) == len(lb) > 1000
    assert np.array_equal(la, lb)
    for x, y in zip(a.state(), b.state()):
        assert np.array_equal(x, y)
    assert np.array_equal(a.density(), b.density())


def test_density_stays_consistent_with_positions(oracle):
    """After many accepted moves the incrementally updated density equals a
    fresh scatter of the final positions (to rounding)."""
    s = W.make_system(200, 2, 8, seed=5)
    sim = oracle.OracleSim(s.chain, s.field)
    sim.np_seed(9)
    logs = sim.mc_sim(30, copy.deepcopy(s.ctrls), max_logs=100000)
    assert logs["accepted"].sum() > 100
    r, t3, t2, st = sim.state()
    c = copy.copy(s.chain)
    c.r, c.t3, c.t2, c.states = r, t3, t2, st
    fresh = oracle.OracleSim(c, s.field).density()
    assert np.allclose(sim.density(), fresh, rtol=1e-9, atol=1e-15)
    # tangents stay unit vectors, bond lengths stay physical
    assert np.allclose(np.linalg.norm(t3, axis=1), 1, atol=1e-9)


def test_schedule_counts(oracle):
    s = W.make_system(120, 2, 6, seed=6)
    sim = oracle.OracleSim(s.chain, s.field)
    ctrls = copy.deepcopy(s.ctrls)
    logs = sim.mc_sim(7, ctrls, max_logs=100000)
    # every proposal increments num_attempt; evaluated proposals appear in the log
    assert [c.num_attempt for c in ctrls] == [7 * c.num_per_cycle for c in s.ctrls]
    for k, c in enumerate(ctrls):
        sel = logs["move"] == c.move
        assert c.num_success == int(logs["accepted"][sel].sum())
    assert set(np.unique(logs["move"])) <= {0, 1, 2, 3, 4}
    # tangent rotations never touch the field; binding moves touch 8 voxels per bead
    assert np.all(logs["n_bins"][logs["move"] == 3] == 0)
    assert np.all(logs["n_bins"][logs["move"] == 4] == 8 * logs["n"][logs["move"] == 4])


def test_metropolis_uses_u_below_exp(oracle):
    s = W.make_system(100, 2, 6, seed=7)
    sim = oracle.OracleSim(s.chain, s.field)
    logs = sim.mc_sim(10, copy.deepcopy(s.ctrls), max_logs=100000)
    dE = logs["dE_poly"] + logs["dE_field"]
    assert np.array_equal(logs["accepted"] == 1, logs["u"] < np.exp(-dE))
    assert np.all((logs["u"] >= 0) & (logs["u"] <= 1))


class _C:
    pass


def _ctrl(oracle, rate, amp, ab, blo, bhi, mlo, mhi):
    c = oracle.OrCtrl(0, 1, 1, 1, amp, ab, rate, 2 / 21, blo, bhi, mlo, mhi, 0, 0)
    oracle.lib().or_update_amplitudes(ctypes.byref(c))
    return c.amp_move, c.amp_bead


@pytest.mark.parametrize("case", [
    # (rate, amp, amp_bead, bead bounds, move bounds) -> (amp, amp_bead); mc_controller.py:148-213
    ((0.3, 1.0, 10, 5, 20, 0.5, 2.0), (0.95, 10)),
    ((0.3, 0.52, 10, 5, 20, 0.5, 2.0), (2.0, 9)),          # amp floor hit: amp reset high, bead - 1
    ((0.3, 0.52, 5, 5, 20, 0.5, 2.0), (2.0, 5)),           # bead floor
    ((0.7, 1.0, 10, 5, 20, 0.5, 2.0), (1.0 / 0.95, 10)),
    ((0.7, 1.95, 10, 5, 20, 0.5, 2.0), (0.5, 11)),         # amp ceiling: amp reset low, bead + 1
    ((0.7, 1.95, 20, 5, 20, 0.5, 2.0), (0.5, 20)),         # bead ceiling
    ((0.5, 1.0, 10, 5, 20, 0.5, 2.0), (1.0, 10)),          # at setpoint: unchanged
    ((0.2, 0.0, 1, 1, 1, 0.0, 0.0), (0.0, 1)),             # binding move bounds (0, 0)
])
def test_simple_control(oracle, case):
    (rate, amp, ab, blo, bhi, mlo, mhi), (eamp, eab) = case
    got = _ctrl(oracle, rate, amp, ab, blo, bhi, mlo, mhi)
    assert got[1] == eab and math.isclose(got[0], eamp, rel_tol=0, abs_tol=0)


def test_ewma_recurrence(oracle):
    s = W.make_system(80, 2, 6, seed=8)
    sim = oracle.OracleSim(s.chain, s.field)
    ctrls = copy.deepcopy(s.ctrls)
    for c in ctrls:
        c.controlled = 0
    logs = sim.mc_sim(3, ctrls, max_logs=100000)
    for c in ctrls:
        rate = 0.0
        for acc in logs["accepted"][logs["move"] == c.move]:
            rate = (c.alpha * float(acc)) + (1 - c.alpha) * rate
        assert rate == c.rate


def test_crank_draw_order(oracle):
    """First crank proposal from srand(1): the proposal's bead range follows from
    the draws ang, b0, side, capped exponential (move_funcs.pyx:80-84)."""
    libc = ctypes.CDLL("libc.so.6")
    libc.srand(1)
    s = W.make_system(500, 2, 6, seed=9, per_cycle=(1, 0, 0, 0, 0))
    ctrls = copy.deepcopy(s.ctrls)
    sim = oracle.OracleSim(s.chain, s.field)
    logs = sim.mc_sim(1, ctrls, max_logs=10)
    d = [libc.rand() for _ in range(4)]
    N = 500
    w = ctrls[0].amp_bead
    b0 = int((d[1] / 2147483647) * N)
    side = d[2] % 2
    r = int(-math.log10(d[3] / 2147483647 + 1e-5) * (min(w, b0) if side == 0 else min(w, N - b0))
            * 0.45 + 1.0001)
    b1 = (b0 - r) if side == 0 else (r + b0)
    lo, hi = min(b0, max(b1, 1)), max(b0, max(b1, 1))
    if lo == hi:
        hi = lo + 1
    assert (logs["ind0"][0], logs["indf"][0]) == (lo, hi)
