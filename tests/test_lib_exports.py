"""libkrt.so loads and exports every entry point include/krt.h declares (no GPU calls)."""
import ctypes
import re
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]


def declared():
    text = (ROOT / "include" / "krt.h").read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(krt_[a-z_0-9]+)\s*\(", text)))


def test_header_declares_entry_points():
    names = declared()
    assert "krt_armada_render" in names and "krt_accel_build" in names and "krt_trace_closest" in names


def test_library_exports_all_declared():
    from artemis_amd import _lib

    if not _lib.LIB_PATH.exists():
        pytest.fail("libkrt.so is not built (run __graft_entry__.build())")
    lib = ctypes.CDLL(str(_lib.LIB_PATH))
    missing = [n for n in declared() if not hasattr(lib, n)]
    assert not missing, missing
    assert set(declared()) == set(_lib.EXPORTED)


def test_last_error_without_device():
    from artemis_amd import _lib

    lib = _lib.lib()
    assert isinstance(lib.krt_last_error(), bytes)
    # null arguments are rejected before any HIP call
    assert lib.krt_scene_upload(None, None, None) == _lib.KRT_ERR_INVALID
    assert b"null" in lib.krt_last_error()
    assert lib.krt_armada_render(None, 0, None, 1) == _lib.KRT_ERR_INVALID
