"""ctypes binding of libchromo_hip.so (the C ABI declared in include/chromo_hip.h).

The library is built in-tree (``chromo_amd/libchromo_hip.so``) by
``chromo_amd.build.build()``.  Loading fails loudly when it is missing: there is
no CPU fallback for any operation routed through this module.
"""
from __future__ import annotations

import ctypes
import os
from pathlib import Path

_HERE = Path(__file__).resolve().parent
LIB_PATH = _HERE / "libchromo_hip.so"

MAX_BINDERS = 4
MAX_MARK = 4
MAX_SITES = 3
RNG_WORDS = 704

CRANK_SHAFT, END_PIVOT, SLIDE, TANGENT_ROTATION, CHANGE_BINDING_STATE = range(5)
MOVE_NAMES = ("crank_shaft", "end_pivot", "slide", "tangent_rotation", "change_binding_state")
MOVE_INDEX = {n: i for i, n in enumerate(MOVE_NAMES)}
RNG_GLIBC, RNG_PHILOX = 0, 1

RNG_GLIBC_STATE, RNG_GLIBC_F, RNG_GLIBC_R = 0, 31, 32
RNG_MT_STATE, RNG_MT_INDEX = 64, 688
RNG_PHILOX_KEY, RNG_PHILOX_CTR, RNG_PHILOX_BUF, RNG_PHILOX_IDX = 690, 692, 694, 698

(CNT_PROPOSALS, CNT_ACCEPTED, CNT_SUM_N, CNT_SUM_BINS, CNT_ERROR, CNT_NEAR_TIES, CNT_BYTES,
 CNT_LOGGED, CNT_TIE_VOXEL, CNT_TIE_HUGE) = range(10)
CNT_WORDS = 10

ERROR_TEXT = {
    1: "bead selection window larger than the polymer (ValueError in bead_selection)",
    2: "move touches more beads than the replica's scratch arena holds (seg_cap)",
    3: "tangent rotation asked for more distinct beads than the polymer has",
    4: "binder count outside 1..4, or not in the launch's binder_mask",
    5: "a rejection-sampling or probe loop exceeded its iteration bound (amplitude below 1?)",
    6: "internal move-arena index check failed",
    7: "a twisting chain (SSTWLC) runs with one wavefront per replica (waves=1)",
}


class ChromoCtrl(ctypes.Structure):
    _fields_ = [
        ("move", ctypes.c_int32), ("on", ctypes.c_int32), ("controlled", ctypes.c_int32),
        ("num_per_cycle", ctypes.c_int32), ("amp_move", ctypes.c_double),
        ("amp_bead", ctypes.c_int64), ("rate", ctypes.c_double), ("alpha", ctypes.c_double),
        ("bead_lo", ctypes.c_double), ("bead_hi", ctypes.c_double),
        ("move_lo", ctypes.c_double), ("move_hi", ctypes.c_double),
        ("num_attempt", ctypes.c_int64), ("num_success", ctypes.c_int64),
    ]


class ChromoMoveLog(ctypes.Structure):
    _fields_ = [
        ("move", ctypes.c_int32), ("accepted", ctypes.c_int32), ("n", ctypes.c_int64),
        ("ind0", ctypes.c_int64), ("indf", ctypes.c_int64), ("dE_poly", ctypes.c_double),
        ("dE_field", ctypes.c_double), ("u", ctypes.c_double), ("exp_dE", ctypes.c_double),
        ("n_bins", ctypes.c_int64),
    ]


class ChromoReplica(ctypes.Structure):
    _fields_ = [
        ("pos", ctypes.c_void_p), ("states", ctypes.c_void_p), ("marks", ctypes.c_void_p),
        ("bonds", ctypes.c_void_p), ("density", ctypes.c_void_p), ("rng", ctypes.c_void_p),
        ("ctrl", ctypes.c_void_p), ("log", ctypes.c_void_p), ("counters", ctypes.c_void_p),
        ("scratch", ctypes.c_void_p),
        ("N", ctypes.c_int64), ("log_cap", ctypes.c_int64), ("seg_cap", ctypes.c_int64),
        ("bead_cap", ctypes.c_int64),
        ("B", ctypes.c_int32), ("n_ctrl", ctypes.c_int32), ("rng_mode", ctypes.c_int32),
        ("field_on", ctypes.c_int32),
        ("nx", ctypes.c_int64), ("ny", ctypes.c_int64), ("nz", ctypes.c_int64),
        ("width", ctypes.c_double * 3), ("dxyz", ctypes.c_double * 3),
        ("half_width", ctypes.c_double * 3), ("half_step", ctypes.c_double * 3),
        ("vol_bin", ctypes.c_double), ("chi", ctypes.c_double), ("bead_vol", ctypes.c_double),
        ("vf_limit", ctypes.c_double),
        ("pref", ctypes.c_double * MAX_BINDERS), ("intra", ctypes.c_double * MAX_BINDERS),
        ("cross", ctypes.c_double * (MAX_BINDERS * MAX_BINDERS)),
        ("mu", ctypes.c_double * MAX_BINDERS), ("mu_adjust", ctypes.c_double),
        ("max_binders", ctypes.c_int64), ("sites", ctypes.c_int32 * MAX_BINDERS),
        ("n_mark_levels", ctypes.c_int32), ("confine_type", ctypes.c_int32),
        ("confine_length", ctypes.c_double),
        ("F", ctypes.c_double * (MAX_BINDERS * MAX_MARK * (MAX_SITES + 1))),
        ("access_vol", ctypes.c_void_p),
        ("fast_field", ctypes.c_int32), ("pad_ff", ctypes.c_int32),
        ("n_sub", ctypes.c_int64 * 3), ("sub_width", ctypes.c_double * 3),
        ("half_points", ctypes.c_double),
        ("twist", ctypes.c_int32), ("pad_tw", ctypes.c_int32),
    ]


class ChromoEnergyModel(ctypes.Structure):
    _fields_ = [
        ("N", ctypes.c_int64), ("B", ctypes.c_int32), ("twist", ctypes.c_int32),
        ("t3", ctypes.c_void_p), ("states", ctypes.c_void_p), ("marks", ctypes.c_void_p),
        ("bonds", ctypes.c_void_p),
        ("nx", ctypes.c_int64), ("ny", ctypes.c_int64), ("nz", ctypes.c_int64),
        ("x_width", ctypes.c_double), ("y_width", ctypes.c_double), ("z_width", ctypes.c_double),
        ("bead_vol", ctypes.c_double), ("vf_limit", ctypes.c_double), ("mu_adjust", ctypes.c_double),
        ("v_int", ctypes.c_double * MAX_BINDERS),
        ("cross_mask", ctypes.c_int32 * (MAX_BINDERS * MAX_BINDERS)),
        ("F", ctypes.c_double * (MAX_BINDERS * MAX_MARK * (MAX_SITES + 1))),
    ]


_c_void_p = ctypes.c_void_p
_c_i32 = ctypes.c_int32
_c_i64 = ctypes.c_int64

# name -> (restype, argtypes)
_SIGNATURES = {
    "chromo_last_error": (ctypes.c_int, [ctypes.c_char_p, ctypes.c_size_t]),
    "chromo_abi_info": (ctypes.c_int, [ctypes.POINTER(_c_i64)] * 4),
    "chromo_scratch_bytes": (_c_i64, [_c_i64, _c_i64, _c_i32]),
    "chromo_scratch_init": (ctypes.c_int, [_c_void_p, _c_i32, _c_void_p]),
    "chromo_rng_seed": (ctypes.c_int, [_c_void_p, _c_i32, _c_void_p, _c_void_p, _c_i32,
                                       _c_void_p]),
    "chromo_np_seed": (ctypes.c_int, [_c_void_p, _c_i32, _c_void_p, _c_void_p]),
    "chromo_density_init": (ctypes.c_int, [_c_void_p, _c_i32, _c_void_p]),
    "chromo_mc_run": (ctypes.c_int, [_c_void_p, _c_i32, _c_i64, _c_void_p]),
    "chromo_mc_run_tuned": (ctypes.c_int, [_c_void_p, _c_i32, _c_i64, _c_i32, _c_i32, _c_void_p]),
    "chromo_mc_run_plain": (ctypes.c_int, [_c_void_p, _c_i32, _c_i64, _c_void_p]),
    "chromo_mc_step": (ctypes.c_int, [_c_void_p, _c_i32, _c_i32, _c_void_p]),
    "chromo_update_amplitudes": (ctypes.c_int, [_c_void_p, _c_i32, _c_i32, _c_void_p]),
    "chromo_field_dE_tape": (ctypes.c_int, [_c_void_p, _c_i32, _c_void_p, _c_i64, _c_void_p,
                                            _c_void_p, _c_i32, _c_void_p, _c_void_p, _c_void_p,
                                            _c_void_p, _c_void_p]),
    "chromo_checkpoint": (ctypes.c_int, [_c_void_p, _c_i32, ctypes.c_char_p, _c_void_p]),
    "chromo_restore": (ctypes.c_int, [_c_void_p, _c_i32, ctypes.c_char_p, _c_void_p]),
    "chromo_capped_exponential_tape": (ctypes.c_int, [_c_void_p, _c_void_p, _c_i64, _c_void_p,
                                                      _c_void_p]),
    "chromo_energy_adjoint": (ctypes.c_int, [ctypes.POINTER(ChromoEnergyModel), _c_void_p,
                                             _c_void_p, _c_void_p, _c_void_p, _c_void_p,
                                             _c_void_p]),
    "chromo_energy_adjoint_noalloc": (ctypes.c_int, [ctypes.POINTER(ChromoEnergyModel), _c_void_p,
                                                     _c_void_p, _c_void_p, _c_void_p, _c_void_p,
                                                     _c_void_p]),
    "chromo_energy_reference": (ctypes.c_int, [ctypes.POINTER(ChromoEnergyModel), _c_void_p,
                                               _c_i32, _c_void_p, _c_void_p, ctypes.c_double,
                                               _c_void_p, _c_void_p, _c_void_p]),
}

_lib = None


class ChromoError(RuntimeError):
    """Raised when a libchromo_hip.so entry point reports a failure."""


def lib() -> ctypes.CDLL:
    """Load libchromo_hip.so (once).  Raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        raise ImportError(
            f"{LIB_PATH} is missing: build it with `python -c 'import __graft_entry__ as g; "
            "g.build()'` (hipcc --offload-arch=gfx950).  chromo_amd has no CPU fallback.")
    h = ctypes.CDLL(str(LIB_PATH), mode=os.RTLD_LOCAL | getattr(os, "RTLD_NOW", 2))
    for name, (res, args) in _SIGNATURES.items():
        f = getattr(h, name)
        f.restype = res
        f.argtypes = args
    v, rb, cb, lb = _c_i64(), _c_i64(), _c_i64(), _c_i64()
    h.chromo_abi_info(ctypes.byref(v), ctypes.byref(rb), ctypes.byref(cb), ctypes.byref(lb))
    sizes = (rb.value, cb.value, lb.value)
    expect = (ctypes.sizeof(ChromoReplica), ctypes.sizeof(ChromoCtrl), ctypes.sizeof(ChromoMoveLog))
    if sizes != expect:
        raise ImportError(f"ABI layout mismatch: library {sizes} vs python {expect}")
    _lib = h
    return h


def exported_symbols():
    return list(_SIGNATURES)


def check(rc: int, what: str = "") -> None:
    """Raise ChromoError with the library's message when rc != 0."""
    if rc == 0:
        return
    buf = ctypes.create_string_buffer(1024)
    lib().chromo_last_error(buf, 1024)
    raise ChromoError(f"{what or 'chromo'} failed ({rc}): {buf.value.decode(errors='replace')}")
