"""CPU-side checks of the C ABI: the libraries load and export every entry
point the headers declare (no device calls)."""
import ctypes
import os
import re

import pytest

from afquantumsim_amd import hip

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared(header):
    text = open(os.path.join(ROOT, "include", header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(aqs_[a-z0-9_]+)\s*\(", text)))


def test_hip_library_exports_every_declared_symbol():
    names = declared("aqs_hip.h")
    assert len(names) > 30
    lib = ctypes.CDLL(hip.LIB_PATH)
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    # the Python binding covers the same surface
    assert set(hip.EXPORTS) <= set(names)


def test_op_struct_layout():
    # int32 kind, ntargets, targets[5], ncontrols, controls[3], flags | double m[8] | pointer | uint64 mask
    assert ctypes.sizeof(hip.Op) == 12 * 4 + 8 * 8 + 8 + 8
    assert hip.Op.m.offset == 48 and hip.Op.u.offset == 112 and hip.Op.more_controls.offset == 120


def test_more_than_three_controls_pack_into_the_mask():
    arr, keep = hip.pack_ops([{"kind": hip.OP_X, "targets": (0,), "controls": (1, 2, 3, 5, 9)}])
    assert arr[0].ncontrols == 3 and list(arr[0].controls) == [1, 2, 3]
    assert arr[0].more_controls == (1 << 5) | (1 << 9)


def test_missing_library_fails_loudly(tmp_path):
    with pytest.raises(RuntimeError):
        hip._lib = None
        try:
            hip.load(str(tmp_path / "nope.so"))
        finally:
            hip._lib = None
