"""CPU-side checks of the C ABI library: it loads and exports every symbol of include/cppe_hip.h."""
import ctypes
import os
import re

import pytest

from cppe_amd import capi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_symbols():
    txt = open(os.path.join(ROOT, "include", "cppe_hip.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(cppe_hip_[a-z0-9_]+)\s*\(", txt)))


def test_header_matches_binding_list():
    assert header_symbols() == sorted(capi.EXPORTS)


@pytest.mark.skipif(not os.path.exists(capi.LIB_PATH), reason="libcppe_hip.so not built")
def test_library_exports_every_symbol():
    lib = ctypes.CDLL(capi.LIB_PATH)
    for name in header_symbols():
        assert hasattr(lib, name), name


@pytest.mark.skipif(not os.path.exists(capi.LIB_PATH), reason="libcppe_hip.so not built")
def test_no_device_errors_cleanly():
    # On a host without a GPU the library still loads and reports zero devices.
    n = capi.device_count()
    assert n >= 0
