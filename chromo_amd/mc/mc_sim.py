"""mc_sim / mc_step (chromo/mc/mc_sim.pyx:26-182) on the GPU.

``mc_sim`` hands the whole schedule (steps x controllers x num_per_cycle
proposals, then the amplitude updates) to one persistent kernel launch, or --
with ``launch="plain"`` -- to one kernel launch per proposal.  Around the launch
it moves three pieces of process state the reference keeps implicitly:

* the libc ``rand()`` stream, which the reference never seeds (== srand(1)) and
  which persists across calls in the process: kept here as ``_LIBC`` words and
  loaded into / read back from the replica's RNG block;
* numpy's global MT19937 stream, reseeded by ``np.random.seed(random_seed)``
  (mc_sim.pyx:81) and consumed by the binding-state draws: copied to the device
  before the run and back into numpy after it, so later host draws (e.g.
  ``polymer_in_field``'s next inner seed) see the same stream as with the
  reference;
* the controllers' amplitudes, counters and acceptance rates.
"""
from __future__ import annotations

from typing import List, Optional

import numpy as np

from .. import _lib as L
from ..engine import CtrlSpec, DeviceReplicas, FieldSpec
from ..fields import FieldBase, UniformDensityField
from . import mc_controller as ctl

_LIBC: Optional[np.ndarray] = None     # 33 words: glibc state (31), front, rear


def _srand_words(seed: int) -> np.ndarray:
    """glibc srandom_r (TYPE_3) state after seeding and the 310 discarded outputs."""
    seed = int(seed) & 0xffffffff
    if seed == 0:
        seed = 1
    word = seed - (1 << 32) if seed >= (1 << 31) else seed
    st = [word]
    for _ in range(1, 31):
        hi = int(word / 127773)          # C truncating division
        lo = word - hi * 127773
        word = 16807 * lo - 2836 * hi
        if word < 0:
            word += 2147483647
        st.append(word)
    st = [x & 0xffffffff for x in st]
    f, r = 3, 0
    for _ in range(310):
        st[f] = (st[f] + st[r]) & 0xffffffff
        f += 1
        if f >= 31:
            f = 0
            r += 1
        else:
            r += 1
            if r >= 31:
                r = 0
    return np.array(st + [f, r], dtype=np.uint32)


def seed_libc(seed: int) -> None:
    """srand(seed) for the process-wide rand() stream used by the moves."""
    global _LIBC
    _LIBC = _srand_words(seed)


def libc_state() -> np.ndarray:
    global _LIBC
    if _LIBC is None:
        _LIBC = _srand_words(1)
    return _LIBC.copy()


class _Engine:
    """The device replica behind one (polymer, field) pair."""

    def __init__(self, poly, field):
        self.poly = poly
        self.field = field
        fs = field.field_spec() if hasattr(field, "field_spec") else FieldSpec(field_on=False)
        if isinstance(field, UniformDensityField) and not (poly in field and poly.is_field_active()):
            fs = FieldSpec(field_on=False)
        self.field_on = fs.field_on
        self.dr = DeviceReplicas([poly.chain_spec()], [fs], [[]], rng_mode=L.RNG_GLIBC,
                                 seeds=[1], caps=[(min(poly.num_beads, 151), min(poly.num_beads, 151))])
        self.dr.density_stale = False
        poly._attach(self.dr, 0)


def engine_for(obj, rebuild: bool = False) -> _Engine:
    """The cached engine of a field (or of a polymer run without a field)."""
    field = obj
    if isinstance(obj, FieldBase):
        polys = obj.polymers
        if len(polys) != 1:
            raise NotImplementedError("exactly one polymer per field")
        poly = polys[0]
    else:
        poly, field = obj, None
    holder = field if field is not None else poly
    eng = getattr(holder, "_engine", None)
    if eng is None or rebuild or eng.poly is not poly:
        if eng is not None:
            poly._sync_from_device()
        eng = _Engine(poly, field)
        holder._engine = eng
    return eng


def _ctrl_spec(c) -> CtrlSpec:
    m = c.move
    return CtrlSpec(move=L.MOVE_INDEX[m.name], num_per_cycle=int(m.num_per_cycle),
                    amp_move=float(m.amp_move), amp_bead=int(m.amp_bead),
                    bead_lo=float(c.bead_amp_bounds[0]), bead_hi=float(c.bead_amp_bounds[1]),
                    move_lo=float(c.move_amp_bounds[0]), move_hi=float(c.move_amp_bounds[1]),
                    on=int(m.move_on), controlled=1 if isinstance(c, ctl.SimpleControl) else 0,
                    rate=float(m.acceptance_tracker.acceptance_rate),
                    alpha=float(m.acceptance_tracker.alpha), num_attempt=int(m.num_attempt),
                    num_success=int(m.num_success))


def _apply_ctrl(c, s: CtrlSpec) -> None:
    m = c.move
    m.amp_move = s.amp_move
    m.amp_bead = int(s.amp_bead)
    m.num_attempt = int(s.num_attempt)
    m.num_success = int(s.num_success)
    m.acceptance_tracker.acceptance_rate = s.rate


def _load_streams(dr: DeviceReplicas) -> None:
    words = dr.rng_words(0)
    lc = libc_state()
    words[L.RNG_GLIBC_STATE:L.RNG_GLIBC_STATE + 31] = lc[:31]
    words[L.RNG_GLIBC_F] = lc[31]
    words[L.RNG_GLIBC_R] = lc[32]
    st = np.random.get_state()
    words[L.RNG_MT_STATE:L.RNG_MT_STATE + 624] = np.asarray(st[1], dtype=np.uint32)
    words[L.RNG_MT_INDEX] = int(st[2])
    dr.set_rng_words(0, words)


def _store_streams(dr: DeviceReplicas) -> None:
    global _LIBC
    words = dr.rng_words(0)
    _LIBC = np.concatenate([words[L.RNG_GLIBC_STATE:L.RNG_GLIBC_STATE + 31],
                            words[[L.RNG_GLIBC_F, L.RNG_GLIBC_R]]]).astype(np.uint32)
    key = words[L.RNG_MT_STATE:L.RNG_MT_STATE + 624].astype(np.uint32)
    st = np.random.get_state()
    np.random.set_state((st[0], key, int(words[L.RNG_MT_INDEX]), st[3], st[4]))


def mc_sim(polymers: List, readerproteins, num_mc_steps: int, mc_move_controllers: List,
           field, mu_adjust_factor: float, random_seed: int, *, launch: str = "persistent"):
    """Run ``num_mc_steps`` MC steps of every active controller (mc_sim.pyx:26-103).

    ``launch="persistent"`` runs the whole call in one kernel launch;
    ``launch="plain"`` issues one launch per proposal.
    """
    np.random.seed(random_seed)
    if len(polymers) != 1:
        raise NotImplementedError("mc_sim on the GPU takes one polymer per field")
    poly = polymers[0]
    poly.mu_adjust_factor = mu_adjust_factor
    eng = engine_for(field if field is not None and poly in getattr(field, "polymers", []) else poly)
    dr = eng.dr
    dr.set_controllers(0, [_ctrl_spec(c) for c in mc_move_controllers])
    dr.set_mu_adjust([mu_adjust_factor])
    _load_streams(dr)
    if launch == "persistent":
        dr.run(int(num_mc_steps))
    elif launch == "plain":
        dr.run_plain(int(num_mc_steps))
    else:
        raise ValueError("launch must be 'persistent' or 'plain'")
    for c, s in zip(mc_move_controllers, dr.controllers(0)):
        _apply_ctrl(c, s)
    _store_streams(dr)
    poly._mark_device_newer()


def mc_step(adaptible_move, poly, readerproteins, field, active_field: bool = True,
            update_distances: bool = False):
    """One proposal of ``adaptible_move`` (an MCAdapter) on ``poly`` (mc_sim.pyx:106-182)."""
    eng = engine_for(field if field is not None and poly in getattr(field, "polymers", []) else poly)
    dr = eng.dr
    c = ctl.NoControl(adaptible_move, (1, max(1, poly.num_beads)), (0.0, 0.0))
    spec = _ctrl_spec(c)
    dr.set_controllers(0, [spec])
    _load_streams(dr)
    dr.step(0)
    dr.raise_on_error()
    s = dr.controllers(0)[0]
    adaptible_move.num_attempt = int(s.num_attempt)
    adaptible_move.num_success = int(s.num_success)
    adaptible_move.acceptance_tracker.acceptance_rate = s.rate
    _store_streams(dr)
    poly._mark_device_newer()
