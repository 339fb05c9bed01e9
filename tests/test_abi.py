"""The C-ABI library loads, exports every entry point that include/sqc_hip.h
declares, and agrees with the binding on the config layout.  No compute call
is made (this runs without a GPU)."""

import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_functions():
    text = open(os.path.join(ROOT, "include", "sqc_hip.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(sqc_(?:hip|dump)_[a-z_]+)\s*\(", text)))


def test_library_exports_header():
    from sqcalc_amd import _lib
    lib = _lib.load()
    names = header_functions()
    assert len(names) >= 23
    assert "sqc_dump_read_frames" in names
    for n in names:
        assert hasattr(lib, n), n
        assert n in _lib.SIGNATURES, n


def test_config_layout():
    from sqcalc_amd import _lib
    lib = _lib.load()
    assert lib.sqc_hip_config_size() == ctypes.sizeof(_lib.Config)
    assert _lib.Config.cell.offset % 8 == 0


def test_error_strings_and_bad_arguments():
    from sqcalc_amd import _lib
    lib = _lib.load()
    assert lib.sqc_hip_error_string(0) == b"success"
    assert b"ABI" in lib.sqc_hip_error_string(30)
    plan = ctypes.c_void_p()
    cfg = _lib.Config()
    cfg.abi_version = 99
    assert lib.sqc_hip_plan_create(ctypes.byref(plan), ctypes.byref(cfg)) == 30
    assert b"ABI" in lib.sqc_hip_last_error()
    cfg.abi_version = _lib.ABI_VERSION
    cfg.precision = 7
    assert lib.sqc_hip_plan_create(ctypes.byref(plan), ctypes.byref(cfg)) == 20
    assert not plan.value
    # calls on no plan are refused before anything touches a device
    assert lib.sqc_hip_set_tiles(None, 8, 2, 8, 2) == 20
    assert b"plan" in lib.sqc_hip_last_error()


def test_product_has_no_cpu_fallback(monkeypatch):
    from sqcalc_amd import _lib
    monkeypatch.setattr(_lib, "_LIB", None)
    monkeypatch.setattr(_lib, "LIB_PATH", "/nonexistent/libsqc_hip.so")
    with pytest.raises(RuntimeError, match="not built"):
        _lib.load()


def test_product_does_not_import_oracle():
    for dirpath, _, files in os.walk(os.path.join(ROOT, "sqcalc_amd")):
        for f in files:
            if f.endswith((".py", ".cpp", ".hip", ".hpp", ".f90")):
                text = open(os.path.join(dirpath, f)).read()
                assert "oracle" not in text.replace("oracle/", ""), f
