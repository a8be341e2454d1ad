"""The large configurations at full size: 16 replicas of 10^6 beads, B = 4,
128^3 voxels, and the 512-replica batched workload (N ~ U{2000..50000}, 32^3-64^3
voxels).  The oracle cannot run these in seconds, so the checks are properties
that hold at any size:

* the logged dE of the accepted moves adds up to the change of the
  dE-consistent total energy (elastic + field + binding), for every move type
  but tangent rotation (the reference evaluates two adjacent rotated beads with
  mixed trial and current tangents, so its dE is not an energy difference);
* the density grid kept up to date move by move equals a fresh deterministic
  density initialisation from the final chain;
* tangents stay unit vectors, states stay in range, no replica reports an error.
"""
import copy
import dataclasses

import numpy as np
import pytest
import torch

from chromo_amd import _lib as L
from chromo_amd import workloads as W

pytestmark = pytest.mark.gpu

N_REP = 16
CHECKED = (0, 7, 15)


def say(*a):
    print(*a, flush=True)


@pytest.fixture(scope="module")
def large():
    say("building 16 x 1e6-bead systems")
    return W.config_large(N_REP)


def _final_chain(dr, i, chain):
    r, t3, t2 = dr.positions(i)
    states = dr.binder_states(i).astype(np.int64)
    return dataclasses.replace(chain, r=np.asarray(r), t3=np.asarray(t3), t2=np.asarray(t2), states=states)


def _total_energy(chain, s):
    from chromo_amd.energy import EnergyModel, energy
    model = EnergyModel(chain, s.field, s.binders)
    th = torch.tensor(model.theta0, dtype=torch.float64, device="cuda")
    r = torch.tensor(chain.r, dtype=torch.float64, device="cuda")
    return float(energy(model, r, th, grad=False)[0][0].item())


def _check_state(dr, i, s):
    r, t3, t2 = dr.positions(i)
    n3 = np.linalg.norm(np.asarray(t3), axis=1)
    n2 = np.linalg.norm(np.asarray(t2), axis=1)
    assert np.all(np.abs(n3 - 1) < 1e-9) and np.all(np.abs(n2 - 1) < 1e-9)
    st = dr.binder_states(i)
    assert st.min() >= 0 and st.max() <= max(s.chain.sites)
    # incremental densities against a fresh scatter of the final chain; the two
    # sum the same contributions in different orders
    inc = dr.density_of(i)
    return inc


def test_large_config_energy_and_density(large):
    from chromo_amd.engine import DeviceReplicas
    sweeps = 3
    ctrls = []
    for s in large:
        cs = copy.deepcopy(s.ctrls)
        for c in cs:
            if c.move == L.TANGENT_ROTATION:
                c.num_per_cycle = 0
        ctrls.append(cs)
    say("device replicas")
    dr = DeviceReplicas([s.chain for s in large], [s.field for s in large], ctrls,
                        rng_mode=L.RNG_PHILOX, seeds=[11] * N_REP, streams=list(range(N_REP)),
                        log_cap=161 * sweeps + 16)
    E0 = {i: _total_energy(large[i].chain, large[i]) for i in CHECKED}
    say("run", sweeps, "sweeps, 2 wavefronts")
    dr.run(sweeps, waves=2)
    dr.raise_on_error()
    cnt = dr.counter_array()
    assert np.all(cnt[:, L.CNT_PROPOSALS] > 0)
    for i in CHECKED:
        lg = dr.logs(i)
        acc = lg["accepted"] != 0
        assert acc.sum() > 10
        dE = float(np.sum(lg["dE_poly"][acc]) + np.sum(lg["dE_field"][acc]))
        E1 = _total_energy(_final_chain(dr, i, large[i].chain), large[i])
        assert abs(E0[i]) < 1e90 and abs(E1) < 1e90     # no volume-fraction penalty in play
        say(f"  replica {i}: E0 {E0[i]:.6f} E1 {E1:.6f} E1 - E0 - sum dE = {(E1 - E0[i]) - dE:.3e}")
        assert abs((E1 - E0[i]) - dE) < 1e-12 * (abs(E0[i]) + abs(E1)) + 1e-7
    inc = {i: _check_state(dr, i, large[i]) for i in CHECKED}
    dr.density_init()
    for i in CHECKED:
        f = large[i].field
        one_bead = 1.0 / (f.x_width * f.y_width * f.z_width / (f.nx * f.ny * f.nz))
        assert np.max(np.abs(inc[i] - dr.density_of(i))) < 1e-9 * one_bead


def test_large_config_all_moves(large):
    """Every move type, with tangent rotations, at both launch shapes."""
    from chromo_amd.engine import DeviceReplicas
    for waves in (2, 1):
        say("all moves, waves", waves)
        dr = DeviceReplicas([s.chain for s in large], [s.field for s in large],
                            [copy.deepcopy(s.ctrls) for s in large], rng_mode=L.RNG_PHILOX,
                            seeds=[12] * N_REP, streams=list(range(N_REP)))
        dr.run(2, waves=waves)
        dr.raise_on_error()
        cnt = dr.counter_array()
        # proposals with an empty bead window stop before dE and are not counted
        assert np.all((cnt[:, L.CNT_PROPOSALS] > 2 * 100) & (cnt[:, L.CNT_PROPOSALS] <= 2 * 161))
        inc = {i: _check_state(dr, i, large[i]) for i in CHECKED}
        dr.density_init()
        for i in CHECKED:
            f = large[i].field
            one_bead = 1.0 / (f.x_width * f.y_width * f.z_width / (f.nx * f.ny * f.nz))
            assert np.max(np.abs(inc[i] - dr.density_of(i))) < 1e-9 * one_bead
        del dr
        torch.cuda.empty_cache()


def test_batched_512_energy_and_density():
    """The 512-replica workload of the benchmark, 5 sweeps at two wavefronts."""
    from chromo_amd.engine import DeviceReplicas
    say("building the 512-replica workload")
    systems = W.config_batched(512, seed=3)
    sweeps = 5
    ctrls = []
    for s in systems:
        cs = copy.deepcopy(s.ctrls)
        for c in cs:
            if c.move == L.TANGENT_ROTATION:
                c.num_per_cycle = 0
        ctrls.append(cs)
    dr = DeviceReplicas([s.chain for s in systems], [s.field for s in systems], ctrls,
                        rng_mode=L.RNG_PHILOX, seeds=[20240601] * 512, streams=list(range(512)),
                        log_cap=161 * sweeps + 16)
    picks = (0, 101, 255, 388, 511)
    E0 = {i: _total_energy(systems[i].chain, systems[i]) for i in picks}
    say("run", sweeps, "sweeps")
    dr.run(sweeps, waves=2)
    dr.raise_on_error()
    inc = {}
    for i in picks:
        lg = dr.logs(i)
        acc = lg["accepted"] != 0
        dE = float(np.sum(lg["dE_poly"][acc]) + np.sum(lg["dE_field"][acc]))
        E1 = _total_energy(_final_chain(dr, i, systems[i].chain), systems[i])
        assert abs(E0[i]) < 1e90 and abs(E1) < 1e90
        say(f"  replica {i}: N {systems[i].chain.N} accepted {int(acc.sum())} "
            f"E1 - E0 - sum dE = {(E1 - E0[i]) - dE:.3e}")
        assert abs((E1 - E0[i]) - dE) < 1e-12 * (abs(E0[i]) + abs(E1)) + 1e-7
        inc[i] = _check_state(dr, i, systems[i])
    dr.density_init()
    for i in picks:
        f = systems[i].field
        one_bead = 1.0 / (f.x_width * f.y_width * f.z_width / (f.nx * f.ny * f.nz))
        assert np.max(np.abs(inc[i] - dr.density_of(i))) < 1e-9 * one_bead
