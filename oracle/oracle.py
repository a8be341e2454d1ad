"""ctypes wrapper of the CPU oracle (oracle/liboracle.so) -- TEST INFRASTRUCTURE ONLY.

Used by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg as the
checker / the timed CPU baseline.  The product (chromo_amd) never imports it.

Inputs are duck-typed: a chain object with r, t3, t2 (N x 3), states, marks
(N x B), bonds (5 x N-1), sites, mu, F (B x levels x sites+1), max_binders; a
field object with the UniformDensityField numbers (field_on, nx.., x_width..,
vf_limit, chi, bead_vol, pref, intra, cross, confine_type, confine_length); and
controller objects with the chromo_ctrl fields.
"""
from __future__ import annotations

import ctypes
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB = HERE / "liboracle.so"

OR_MAXB, OR_MAXLEV, OR_MAXS = 8, 8, 8
RNG_GLIBC, RNG_PHILOX = 0, 1

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int64)


class OrSpec(ctypes.Structure):
    _fields_ = [
        ("N", ctypes.c_int64), ("B", ctypes.c_int64),
        ("r", _dp), ("t3", _dp), ("t2", _dp), ("states", _ip), ("marks", _ip),
        ("eps_bend", _dp), ("eps_par", _dp), ("eps_perp", _dp), ("gamma", _dp), ("eta", _dp),
        ("max_binders", ctypes.c_int64), ("sites", ctypes.c_int64 * OR_MAXB),
        ("mu", ctypes.c_double * OR_MAXB), ("n_mark_levels", ctypes.c_int64),
        ("F", ctypes.c_double * (OR_MAXB * OR_MAXLEV * (OR_MAXS + 1))),
        ("field_on", ctypes.c_int32), ("nx", ctypes.c_int64), ("ny", ctypes.c_int64),
        ("nz", ctypes.c_int64), ("x_width", ctypes.c_double), ("y_width", ctypes.c_double),
        ("z_width", ctypes.c_double), ("vf_limit", ctypes.c_float), ("chi", ctypes.c_double),
        ("bead_vol", ctypes.c_double), ("pref", ctypes.c_double * OR_MAXB),
        ("intra", ctypes.c_double * OR_MAXB), ("cross", ctypes.c_double * (OR_MAXB * OR_MAXB)),
        ("confine_type", ctypes.c_int32), ("confine_length", ctypes.c_double),
        ("access_vols", _dp),
        ("fast_field", ctypes.c_int32), ("n_points", ctypes.c_int64),
        ("twist", ctypes.c_int32), ("eps_twist", _dp), ("bead_length", _dp),
    ]


class OrCtrl(ctypes.Structure):
    _fields_ = [
        ("move", ctypes.c_int32), ("on", ctypes.c_int32), ("controlled", ctypes.c_int32),
        ("num_per_cycle", ctypes.c_int32), ("amp_move", ctypes.c_double),
        ("amp_bead", ctypes.c_int64), ("rate", ctypes.c_double), ("alpha", ctypes.c_double),
        ("bead_lo", ctypes.c_double), ("bead_hi", ctypes.c_double), ("move_lo", ctypes.c_double),
        ("move_hi", ctypes.c_double), ("num_attempt", ctypes.c_int64),
        ("num_success", ctypes.c_int64),
    ]


class OrLog(ctypes.Structure):
    _fields_ = [
        ("move", ctypes.c_int32), ("accepted", ctypes.c_int32), ("n", ctypes.c_int64),
        ("ind0", ctypes.c_int64), ("indf", ctypes.c_int64), ("dE_poly", ctypes.c_double),
        ("dE_field", ctypes.c_double), ("u", ctypes.c_double), ("exp_dE", ctypes.c_double),
        ("n_bins", ctypes.c_int64),
    ]


LOG_DTYPE = np.dtype([("move", "<i4"), ("accepted", "<i4"), ("n", "<i8"), ("ind0", "<i8"),
                      ("indf", "<i8"), ("dE_poly", "<f8"), ("dE_field", "<f8"), ("u", "<f8"),
                      ("exp_dE", "<f8"), ("n_bins", "<i8")])

_lib = None


def lib():
    global _lib
    if _lib is None:
        if not LIB.exists():
            from .build import build as _build
            _build()
        h = ctypes.CDLL(str(LIB))
        V = ctypes.c_void_p
        h.or_new.restype = V
        h.or_new.argtypes = [ctypes.POINTER(OrSpec), ctypes.c_int32, ctypes.c_uint32,
                             ctypes.c_uint32, ctypes.c_int32]
        h.or_free.argtypes = [V]
        h.or_np_seed.argtypes = [V, ctypes.c_uint32]
        h.or_mc_sim.restype = ctypes.c_int64
        h.or_mc_sim.argtypes = [V, ctypes.c_int64, ctypes.POINTER(OrCtrl), ctypes.c_int32,
                                ctypes.c_double, V, ctypes.c_int64, _ip]
        h.or_mc_step.restype = ctypes.c_int
        h.or_mc_step.argtypes = [V, ctypes.POINTER(OrCtrl), V]
        h.or_update_amplitudes.argtypes = [ctypes.POINTER(OrCtrl)]
        h.or_get_state.argtypes = [V, _dp, _dp, _dp, _ip]
        h.or_set_positions.argtypes = [V, _dp, _dp, _dp]
        h.or_get_density.argtypes = [V, _dp]
        h.or_n_bins.restype = ctypes.c_int64
        h.or_n_bins.argtypes = [V]
        h.or_density_init.argtypes = [V]
        h.or_field_dE_tape.restype = ctypes.c_double
        h.or_field_dE_tape.argtypes = [V, _ip, ctypes.c_int64, _dp, _ip, ctypes.c_int32, _dp, _ip, _ip]
        h.or_elastic_dE_tape.restype = ctypes.c_double
        h.or_elastic_dE_tape.argtypes = [V, ctypes.c_int32, _ip, ctypes.c_int64, _dp, _dp, _dp]
        h.or_twist_omega.restype = ctypes.c_double
        h.or_twist_omega.argtypes = [_dp, _dp, _dp, _dp]
        h.or_binding_dE_tape.restype = ctypes.c_double
        h.or_binding_dE_tape.argtypes = [V, ctypes.c_int64, ctypes.c_int64, _ip]
        h.or_elastic_E.restype = ctypes.c_double
        h.or_elastic_E.argtypes = [V]
        h.or_field_E.restype = ctypes.c_double
        h.or_field_E.argtypes = [V]
        h.or_cexp_trunc.argtypes = [ctypes.POINTER(ctypes.c_int32), _ip, ctypes.c_int64, _ip]
        h.or_py_round2.restype = ctypes.c_double
        h.or_py_round2.argtypes = [ctypes.c_double]
        h.or_set_mu_adjust.argtypes = [V, ctypes.c_double]
        h.or_glibc_stream.argtypes = [ctypes.c_uint32, ctypes.POINTER(ctypes.c_int32), ctypes.c_int64]
        h.or_mt_randint_stream.argtypes = [ctypes.c_uint32, ctypes.c_int64, _ip, ctypes.c_int64]
        h.or_philox4x32_10.argtypes = [ctypes.POINTER(ctypes.c_uint32)] * 3
        h.or_axis_rotation.argtypes = [_dp, _dp, ctypes.c_double, _dp]
        h.or_translation.argtypes = [_dp, _dp]
        h.or_transform.argtypes = [_dp, _dp, _dp, _dp, ctypes.c_int64, _dp, _dp, _dp]
        h.or_select_stream.restype = ctypes.c_int64
        h.or_select_stream.argtypes = [ctypes.c_uint32, ctypes.c_int32, ctypes.c_int64,
                                       ctypes.c_int64, ctypes.c_int64, _ip, ctypes.c_int64]
        h.or_check_bead_bounds.argtypes = [ctypes.c_int64] * 3 + [_ip, _ip]
        h.or_pyset_order.restype = ctypes.c_int64
        h.or_pyset_order.argtypes = [_ip, ctypes.c_int64, _ip]
        h.or_run_batch.restype = ctypes.c_int64
        h.or_run_batch.argtypes = [ctypes.POINTER(V), ctypes.c_int32, ctypes.POINTER(OrCtrl),
                                   ctypes.c_int32, ctypes.c_int64, ctypes.c_int32, _dp]
        _lib = h
    return _lib


def _d(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _i(a):
    return np.ascontiguousarray(a, dtype=np.int64)


def _dptr(a):
    return a.ctypes.data_as(_dp)


def _iptr(a):
    return a.ctypes.data_as(_ip)


def ctrl_struct(c) -> OrCtrl:
    return OrCtrl(int(c.move), int(c.on), int(c.controlled), int(c.num_per_cycle),
                  float(c.amp_move), int(c.amp_bead), float(c.rate), float(c.alpha),
                  float(c.bead_lo), float(c.bead_hi), float(c.move_lo), float(c.move_hi),
                  int(c.num_attempt), int(c.num_success))


class OracleSim:
    """One replica on the CPU oracle."""

    def __init__(self, chain, field, rng_mode=RNG_GLIBC, seed=1, stream=0, faithful=False):
        h = lib()
        N, B = chain.r.shape[0], chain.states.shape[1]
        self._keep = []
        keep = self._keep.append
        arrs = {k: _d(getattr(chain, k)) for k in ("r", "t3", "t2")}
        for a in arrs.values():
            keep(a)
        st, mk = _i(chain.states), _i(chain.marks)
        keep(st)
        keep(mk)
        bonds = _d(chain.bonds)
        keep(bonds)
        sp = OrSpec()
        sp.N, sp.B = N, B
        sp.r, sp.t3, sp.t2 = _dptr(arrs["r"]), _dptr(arrs["t3"]), _dptr(arrs["t2"])
        sp.states, sp.marks = _iptr(st), _iptr(mk)
        rows = [np.ascontiguousarray(bonds[k]) for k in range(5)]
        for a in rows:
            keep(a)
        sp.eps_bend, sp.eps_par, sp.eps_perp, sp.gamma, sp.eta = [_dptr(a) for a in rows]
        if getattr(chain, "twist", False):
            # eps_twist from the chain's bond table, the natural twist from the bead
            # lengths (the oracle recomputes it as E_pair_with_twist does)
            et = np.ascontiguousarray(bonds[5])
            bl = _d(chain.bead_length)
            keep(et)
            keep(bl)
            sp.twist = 1
            sp.eps_twist, sp.bead_length = _dptr(et), _dptr(bl)
        sp.max_binders = int(chain.max_binders)
        F = np.asarray(chain.F, dtype=np.float64)
        sp.n_mark_levels = F.shape[1]
        for b in range(B):
            sp.sites[b] = int(chain.sites[b])
            sp.mu[b] = float(chain.mu[b])
            for m in range(F.shape[1]):
                for s in range(F.shape[2]):
                    sp.F[(b * OR_MAXLEV + m) * (OR_MAXS + 1) + s] = F[b, m, s]
        sp.field_on = 1 if field.field_on else 0
        sp.nx, sp.ny, sp.nz = int(field.nx), int(field.ny), int(field.nz)
        sp.x_width, sp.y_width, sp.z_width = field.x_width, field.y_width, field.z_width
        sp.vf_limit = float(field.vf_limit)
        sp.chi, sp.bead_vol = float(field.chi), float(field.bead_vol)
        cross = np.zeros((B, B)) if field.cross is None else np.asarray(field.cross, dtype=float)
        for a in range(B):
            sp.pref[a] = float(field.pref[a]) if len(field.pref) else 0.0
            sp.intra[a] = float(field.intra[a]) if len(field.intra) else 0.0
            for b in range(B):
                sp.cross[a * B + b] = float(cross[a, b])
        sp.confine_type = int(field.confine_type)
        sp.confine_length = float(field.confine_length)
        sp.fast_field = 1 if getattr(field, "fast_field", False) else 0
        sp.n_points = int(getattr(field, "n_points", 1000))
        av = getattr(field, "access_vol", None)
        if av is not None and field.field_on:
            av = _d(av).reshape(-1)
            keep(av)
            sp.access_vols = _dptr(av)
        self.N, self.B = N, B
        self.h = h
        self.p = h.or_new(ctypes.byref(sp), int(rng_mode), int(seed) & 0xffffffff,
                          int(stream) & 0xffffffff, 1 if faithful else 0)

    def __del__(self):
        if getattr(self, "p", None):
            self.h.or_free(self.p)
            self.p = None

    def np_seed(self, seed):
        self.h.or_np_seed(self.p, int(seed) & 0xffffffff)

    def mc_sim(self, steps, ctrls, mu_adjust_factor=1.0, max_logs=0):
        arr = (OrCtrl * len(ctrls))(*[ctrl_struct(c) for c in ctrls])
        logs = np.zeros(max_logs, dtype=LOG_DTYPE)
        nl = ctypes.c_int64(0)
        rc = self.h.or_mc_sim(self.p, int(steps), arr, len(ctrls), float(mu_adjust_factor),
                              logs.ctypes.data if max_logs else None, int(max_logs),
                              ctypes.byref(nl))
        if rc < 0:
            raise RuntimeError(f"oracle mc_sim error {rc}")
        for c, a in zip(ctrls, arr):
            for f, _ in OrCtrl._fields_:
                setattr(c, f, getattr(a, f))
        return logs[:nl.value]

    def state(self):
        N, B = self.N, self.B
        r, t3, t2 = np.zeros((N, 3)), np.zeros((N, 3)), np.zeros((N, 3))
        st = np.zeros((N, max(B, 1)), dtype=np.int64)
        self.h.or_get_state(self.p, _dptr(r), _dptr(t3), _dptr(t2), _iptr(st))
        return r, t3, t2, st

    def density(self):
        nb = self.h.or_n_bins(self.p)
        d = np.zeros((nb, self.B + 1))
        self.h.or_get_density(self.p, _dptr(d))
        return d

    def field_dE_tape(self, inds, r_trial, states_trial, state_change):
        inds = _i(inds)
        rt = _d(r_trial)
        st = _i(states_trial)
        nb = self.h.or_n_bins(self.p)
        delta = np.zeros((nb, self.B + 1))
        bins = np.zeros(16 * len(inds), dtype=np.int64)
        k = ctypes.c_int64(0)
        dE = self.h.or_field_dE_tape(self.p, _iptr(inds), len(inds), _dptr(rt), _iptr(st),
                                     int(state_change), _dptr(delta), _iptr(bins), ctypes.byref(k))
        return dE, bins[:k.value].copy(), delta

    def elastic_dE_tape(self, move, inds, r_trial, t3_trial, t2_trial=None):
        inds = _i(inds)
        t2 = None if t2_trial is None else _dptr(_d(t2_trial))
        return self.h.or_elastic_dE_tape(self.p, int(move), _iptr(inds), len(inds), _dptr(_d(r_trial)),
                                         _dptr(_d(t3_trial)), t2)

    def binding_dE_tape(self, ind0, indf, states_trial):
        return self.h.or_binding_dE_tape(self.p, int(ind0), int(indf), _iptr(_i(states_trial)))

    def set_mu_adjust(self, f):
        self.h.or_set_mu_adjust(self.p, float(f))

    def elastic_E(self):
        return self.h.or_elastic_E(self.p)

    def field_E(self):
        return self.h.or_field_E(self.p)


def cexp_trunc(draws, windows):
    d = np.ascontiguousarray(draws, dtype=np.int32)
    w = _i(windows)
    out = np.zeros(len(d), dtype=np.int64)
    lib().or_cexp_trunc(d.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)), _iptr(w), len(d), _iptr(out))
    return out


def twist_omega(t2_0, t3_0, t2_1, t3_1):
    """compute_twist_angle_omega (chromo/polymers.pyx:3427-3461)."""
    a = [_d(v) for v in (t2_0, t3_0, t2_1, t3_1)]
    return lib().or_twist_omega(*[_dptr(v) for v in a])


def py_round2(x):
    return lib().or_py_round2(float(x))


def glibc_stream(seed, n):
    out = np.zeros(n, dtype=np.int32)
    lib().or_glibc_stream(seed, out.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)), n)
    return out


def mt_randint_stream(seed, high, n):
    out = np.zeros(n, dtype=np.int64)
    lib().or_mt_randint_stream(seed, high, _iptr(out), n)
    return out


def philox4x32_10(ctr, key):
    c = np.ascontiguousarray(ctr, dtype=np.uint32)
    k = np.ascontiguousarray(key, dtype=np.uint32)
    o = np.zeros(4, dtype=np.uint32)
    P = ctypes.POINTER(ctypes.c_uint32)
    lib().or_philox4x32_10(c.ctypes.data_as(P), k.ctypes.data_as(P), o.ctypes.data_as(P))
    return o


def axis_rotation(axis, point, ang):
    T = np.zeros(16)
    lib().or_axis_rotation(_dptr(_d(axis)), _dptr(_d(point)), float(ang), _dptr(T))
    return T.reshape(4, 4)


def translation(delta):
    T = np.zeros(16)
    lib().or_translation(_dptr(_d(delta)), _dptr(T))
    return T.reshape(4, 4)


def transform(T, r, t3, t2):
    r, t3, t2 = _d(r), _d(t3), _d(t2)
    n = r.shape[0]
    ro, t3o, t2o = np.zeros((n, 3)), np.zeros((n, 3)), np.zeros((n, 3))
    lib().or_transform(_dptr(_d(np.asarray(T).reshape(-1))), _dptr(r), _dptr(t3), _dptr(t2), n,
                       _dptr(ro), _dptr(t3o), _dptr(t2o))
    return ro, t3o, t2o


def select_stream(seed, which, window, nbeads, ind0, n):
    """which: 0 from_left, 1 from_right, 2 from_point, 3 capped_exponential(window, nbeads)."""
    out = np.zeros(n, dtype=np.int64)
    rc = lib().or_select_stream(seed, which, window, nbeads, ind0, _iptr(out), n)
    if rc < 0:
        raise ValueError("Bead selection window size must be less than polymer length")
    return out


def check_bead_bounds(b0, b1, nb):
    a, b = ctypes.c_int64(), ctypes.c_int64()
    lib().or_check_bead_bounds(b0, b1, nb, ctypes.byref(a), ctypes.byref(b))
    return a.value, b.value


def pyset_order(keys):
    k = _i(keys)
    out = np.zeros(len(k), dtype=np.int64)
    m = lib().or_pyset_order(_iptr(k), len(k), _iptr(out))
    return out[:m]


def run_batch(sims, ctrl_lists, steps, threads):
    n = len(sims)
    nc = len(ctrl_lists[0])
    ptrs = (ctypes.c_void_p * n)(*[s.p for s in sims])
    flat = (OrCtrl * (n * nc))(*[ctrl_struct(c) for cl in ctrl_lists for c in cl])
    sec = ctypes.c_double(0)
    props = lib().or_run_batch(ptrs, n, flat, nc, int(steps), int(threads), ctypes.byref(sec))
    return props, sec.value
