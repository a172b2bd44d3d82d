"""The C-ABI library loads without a GPU, exports every symbol include/mhd_amd.h declares,
and reports errors through return codes (no compute calls here)."""
import ctypes as C
import os
import re

import pytest

from mhdturbulence_amd import capi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    text = open(os.path.join(ROOT, "include", "mhd_amd.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return re.findall(r"\b(mhd_[a-z_0-9]+)\s*\(", text)


def test_header_symbols_exported():
    L = capi.load()
    names = _declared()
    assert len(names) >= 25
    for n in names:
        assert hasattr(L, n), n
    assert set(names) == set(capi.EXPORTS)


def test_version_and_error_strings():
    L = capi.load()
    assert L.mhd_version() == 1
    for code, text in ((0, b"ok"), (-1, b"invalid argument"), (-2, b"HIP runtime error"), (-3, b"device allocation failed")):
        assert L.mhd_error_string(code) == text


def test_bad_config_rejected_before_device_use():
    L = capi.load()
    cfg = capi.MhdConfig()
    cfg.nx = cfg.ny = cfg.nz = 2  # too small
    h = C.c_void_p()
    rc = L.mhd_create(C.byref(cfg), C.byref(h))
    assert rc == -1 and not h.value
    assert b"at least 4" in L.mhd_last_error()
    cfg.nx = cfg.ny = cfg.nz = 8
    for q in range(6):
        cfg.bc[q] = 1
    cfg.bc[2] = 7
    assert L.mhd_create(C.byref(cfg), C.byref(h)) == -1
    assert L.mhd_create(None, C.byref(h)) == -1
    assert L.mhd_create(C.byref(cfg), None) == -1
    # stages: 0 or 1 (forward Euler) or 2 (two-stage Runge-Kutta)
    cfg.bc[2] = 2
    cfg.gam, cfg.cfl = 1.4, 0.05
    for q in range(3):
        cfg.xmin[q], cfg.xmax[q] = 0.0, 1.0
    for bad in (3, -1):
        cfg.stages = bad
        assert L.mhd_create(C.byref(cfg), C.byref(h)) == -1 and not h.value
        assert b"stages" in L.mhd_last_error()
    assert L.mhd_batch_create(C.byref(cfg), 2, 1 << 20, C.byref(h)) == -1
    assert b"stages" in L.mhd_last_error()


def test_null_context_calls_fail_cleanly():
    L = capi.load()
    d = C.c_double()
    assert L.mhd_control_timestep(None, C.byref(d)) == -1
    assert L.mhd_set_boundary(None) == -1
    assert L.mhd_numerical_flux(None, 1, 1.0) == -1
    assert L.mhd_destroy(None) == 0


def test_missing_library_fails_loudly(tmp_path, monkeypatch):
    monkeypatch.setattr(capi, "_lib", None)
    with pytest.raises(capi.MhdError):
        capi.load(str(tmp_path / "nope.so"))
