"""The C ABI library loads and exports every entry point include/*.h declares.

CPU-only: loads liblunarscout_hip.so and resolves symbols; no compute call.
"""

from __future__ import annotations

import ctypes
import re
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]


def declared_entry_points() -> list[str]:
    names = []
    for header in sorted((ROOT / "include").glob("*.h")):
        text = header.read_text()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        names += re.findall(r"^\s*(?:const\s+)?\w+\**\s+\**(lsh_\w+)\s*\(", text, flags=re.M)
    return sorted(set(names))


def test_header_declares_entry_points():
    names = declared_entry_points()
    assert "lsh_horizon_tile" in names and "lsh_session_create" in names
    assert len(names) >= 15


def test_library_exports_every_declared_entry_point():
    from lunarscout_amd import _lib

    lib = _lib.load_library()
    missing = [n for n in declared_entry_points() if not hasattr(lib, n)]
    assert not missing, missing
    # every declared entry point has a ctypes signature in the binding
    unbound = [n for n in declared_entry_points() if n not in _lib.SIGNATURES]
    assert not unbound, unbound


def test_version_and_device_count_without_gpu():
    from lunarscout_amd import _lib

    lib = _lib.load_library()
    assert lib.lsh_version().decode().count(".") == 2
    n = ctypes.c_int32(-1)
    assert lib.lsh_device_count(ctypes.byref(n)) == 0
    assert n.value >= 0


def test_invalid_arguments_are_rejected_before_any_device_work():
    from lunarscout_amd import _lib

    lib = _lib.load_library()
    count = ctypes.c_int32()
    mips = ctypes.c_int64()
    assert lib.lsh_pyramid_layout(0, 5, None, 0, ctypes.byref(count), ctypes.byref(mips)) == _lib.LSH_EINVAL
    assert "positive" in _lib.last_error()
    assert lib.lsh_pyramid_layout(1024, 1024, None, 0, ctypes.byref(count), ctypes.byref(mips)) == 0
    assert count.value == 6 and mips.value == 69905


def test_session_without_device_reports_unavailable(monkeypatch):
    from lunarscout_amd import _lib
    from lunarscout_amd.errors import HipBackendError

    if _lib.device_count() > 0:
        pytest.skip("a HIP device is visible")
    with pytest.raises(HipBackendError):
        _lib.Session(0)
