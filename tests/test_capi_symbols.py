"""CPU checks of the C ABI library: it loads without a GPU, exports every
function declared in include/honei_hip.h, and the Python binding declares
exactly those."""
import os
import re
import subprocess

import pytest

import honei_amd as H

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "honei_hip.h")


def declared():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(honei_hip_\w+)\s*\(", text)))


def test_library_loads_without_gpu():
    L = H.lib()
    assert L.honei_hip_version().startswith(b"honei-hip")


def test_all_declared_symbols_exported():
    out = subprocess.run(["nm", "-D", "--defined-only", H.LIB_PATH], capture_output=True, text=True,
                         check=True).stdout
    exported = set(re.findall(r"\b[TW] (honei_hip_\w+)", out))
    missing = [s for s in declared() if s not in exported]
    assert not missing, missing


def test_binding_covers_header():
    assert sorted(H.declared_symbols()) == declared()


def test_no_device_reports_cleanly():
    if H.device_count() > 0:
        pytest.skip("GPU present")
    import ctypes as C
    n = C.c_int(-1)
    rc = H.lib().honei_hip_device_count(C.byref(n))
    assert rc == 7 and n.value == 0
    assert H.lib().honei_hip_init(0) == 7
