"""CPU-side checks of the C ABI: the HIP library loads, exports every
function include/brkga_hip.h declares, and reports the no-device case
instead of falling back to a CPU path."""
import ctypes
import os
import re

import pytest

import brkga_amd

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    names = set()
    inc = os.path.join(ROOT, "include")
    for fn in os.listdir(inc):
        if fn.endswith(".h"):
            with open(os.path.join(inc, fn)) as f:
                text = f.read()
            names.update(re.findall(r"^\s*(?:brkga_status|const char\*)\s+(brkga_\w+)\s*\(", text, re.M))
    return names


def test_header_declares_entry_points():
    names = declared_symbols()
    assert {"brkga_evaluate", "brkga_population_create", "brkga_evolve"} <= names


def test_library_exports_every_declared_symbol():
    if not os.path.exists(brkga_amd.LIB_PATH):
        pytest.fail("engine library not built; run __graft_entry__.build()")
    lib = ctypes.CDLL(brkga_amd.LIB_PATH)
    missing = [n for n in sorted(declared_symbols()) if not hasattr(lib, n)]
    assert not missing, missing


def test_binding_covers_header():
    assert declared_symbols() <= set(brkga_amd.exported_symbols())


def test_version_and_last_error():
    assert b"gfx950" in brkga_amd.lib().brkga_version()
    assert isinstance(brkga_amd.lib().brkga_last_error(), bytes)


def test_no_device_is_reported_not_faked():
    if brkga_amd.device_count() > 0:
        pytest.skip("a GPU is visible")
    with pytest.raises(brkga_amd.BrkgaError) as e:
        brkga_amd.Engine(0)
    assert e.value.status == brkga_amd.ERR_NO_DEVICE


def test_null_arguments_are_rejected():
    L = brkga_amd.lib()
    assert L.brkga_engine_create(0, None) == brkga_amd.ERR_INVALID_ARG
    assert L.brkga_population_create(None, None, None) == brkga_amd.ERR_INVALID_ARG
    assert L.brkga_evolve(None, 1, None) == brkga_amd.ERR_INVALID_ARG
