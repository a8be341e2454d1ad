"""The C ABI library loads without a GPU, exports every entry point that
include/chromo_hip.h declares, and its struct layouts match the Python side."""
import ctypes
import re
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]


def _declared():
    text = (ROOT / "include" / "chromo_hip.h").read_text()
    return sorted(set(re.findall(r"^\s*(?:int|int64_t)\s+(chromo_\w+)\s*\(", text, re.M)))


def test_header_declares_entry_points():
    names = _declared()
    assert "chromo_mc_run" in names and "chromo_density_init" in names and len(names) >= 12


def test_library_exports_every_declared_symbol():
    from chromo_amd import _lib
    h = _lib.lib()
    for name in _declared():
        assert hasattr(h, name), name
    assert set(_declared()) <= set(_lib.exported_symbols())


def test_struct_layouts():
    from chromo_amd import _lib
    h = _lib.lib()
    v, rb, cb, lb = (ctypes.c_int64() for _ in range(4))
    h.chromo_abi_info(ctypes.byref(v), ctypes.byref(rb), ctypes.byref(cb), ctypes.byref(lb))
    assert rb.value == ctypes.sizeof(_lib.ChromoReplica)
    assert cb.value == ctypes.sizeof(_lib.ChromoCtrl) == 96
    assert lb.value == ctypes.sizeof(_lib.ChromoMoveLog) == 72


def test_scratch_sizes_and_errors():
    from chromo_amd import _lib
    h = _lib.lib()
    assert h.chromo_scratch_bytes(0, 0, 2) == 0
    a = h.chromo_scratch_bytes(151, 151, 2)
    b = h.chromo_scratch_bytes(1000, 151, 2)
    assert 0 < a < b and a % 256 == 0
    assert h.chromo_scratch_bytes(10, 20, 2) == -1       # field cap above bead cap
    assert h.chromo_scratch_bytes(10, 10, 9) == -1       # too many binders
    # error reporting without a device
    assert h.chromo_mc_run(None, 1, 1, None) != 0
    buf = ctypes.create_string_buffer(256)
    h.chromo_last_error(buf, 256)
    assert b"NULL" in buf.value


def test_replica_error_codes_agree():
    """The header's replica error codes, the device constants and the host's
    messages name the same codes."""
    from chromo_amd import _lib
    hdr = (ROOT / "include" / "chromo_hip.h").read_text()
    block = hdr[hdr.index("enum chromo_replica_error"):]
    block = block[:block.index("};")]
    header = {int(v): n for n, v in re.findall(r"CHROMO_ERR_(\w+)\s*=\s*(\d+)", block)}
    dev = (ROOT / "chromo_amd" / "csrc" / "chromo_device.hpp").read_text()
    device = {int(v): n for n, v in re.findall(r"\bERR_(\w+)\s*=\s*(\d+)", dev)}
    assert header == device
    assert set(header) == set(_lib.ERROR_TEXT) == set(range(1, 8))
