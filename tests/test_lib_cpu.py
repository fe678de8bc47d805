"""CPU-side checks of the C ABI library: it loads, exports every declared symbol,
and fails cleanly (no crash, negative status, message) without a GPU."""
import ctypes
import os
import re

import pytest

import gsk

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    hdr = open(os.path.join(ROOT, "include", "gsk.h")).read()
    return sorted(set(re.findall(r"^\s*(?:const\s+)?\w+\*?\s+\*?(gsk_\w+)\s*\(", hdr, re.M)))


def test_header_declares_core_entry_points():
    syms = declared_symbols()
    for s in ("gsk_create", "gsk_destroy", "gsk_build_btensor", "gsk_set_field", "gsk_get_field", "gsk_step",
              "gsk_reduce", "gsk_last_error", "gsk_noise_field", "gsk_fft_r2c", "gsk_fft_c2r"):
        assert s in syms


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(gsk.LIB_PATH)
    missing = [s for s in declared_symbols() if not hasattr(lib, s)]
    assert not missing, missing


def test_config_defaults_follow_reference():
    c = gsk.default_config(gsk.PHYS_PREC)
    # src/dynamics_precipitate.cu:21-29
    assert list(c.a)[1:] == [54.0, -17.0, 7.0, 2.5, 0.2, 0.2, 0.2]
    assert (c.kappa, c.kappa_c, c.M, c.L, c.xi, c.c1, c.c2, c.dt) == (300.0, 300.0, 0.4, 0.4, 400.0, 0.1, 0.44, 0.1)
    assert c.S == 0.0
    m = gsk.default_config(gsk.PHYS_MART)
    assert (m.A, m.B, m.C, m.kappa, m.L, m.dt, m.nvar) == (0.2, 3.0, 3.0, 1.0, 1.0, 0.01, 3)


def test_invalid_arguments_return_errors():
    lib = gsk.lib()
    with pytest.raises(gsk.GskError):
        gsk.default_config(7)
    c = gsk.default_config(gsk.PHYS_MART, n=100)
    h = ctypes.c_void_p()
    rc = lib.gsk_create(ctypes.byref(c), ctypes.byref(h))
    assert rc == -5 and b"unsupported grid size" in lib.gsk_last_error()
    c = gsk.reference_config(gsk.PHYS_PREC, path=gsk.PATH_PLAIN64)
    assert lib.gsk_create(ctypes.byref(c), ctypes.byref(h)) == -5 and b"FP32 fused path" in lib.gsk_last_error()
    c = gsk.default_config(gsk.PHYS_MART, nvar=0)
    assert lib.gsk_create(ctypes.byref(c), ctypes.byref(h)) == -2


def test_create_without_gpu_fails_cleanly():
    from conftest import gpu_available
    if gpu_available():
        pytest.skip("a GPU is present")
    c = gsk.default_config(gsk.PHYS_MART, n=16)
    h = ctypes.c_void_p()
    rc = gsk.lib().gsk_create(ctypes.byref(c), ctypes.byref(h))
    assert rc < 0 and not h
    assert gsk.lib().gsk_last_error()


def test_host_philox_known_answers():
    # Random123 known-answer vectors for Philox4x32-10 (Salmon et al., SC'11)
    assert [hex(x) for x in gsk.philox4x32([0, 0, 0, 0], 0)] == ["0x6627e8d5", "0xe169c58d", "0xbc57ac4c", "0x9b00dbd8"]
    ones = gsk.philox4x32([0xFFFFFFFF] * 4, 0xFFFFFFFFFFFFFFFF)
    assert [hex(x) for x in ones] == ["0x408f276d", "0x41c83b0e", "0xa20bc7c6", "0x6d5451fd"]
    pi = gsk.philox4x32([0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344], 0x299F31D0A4093822)
    assert [hex(x) for x in pi] == ["0xd16cfe09", "0x94fdcceb", "0x5001e420", "0x24126ea1"]
    u = gsk.philox_uniform(3, 0, 0)
    assert 0.0 <= u < 1.0
