"""The C++ operator API (honei_amd/honei/) through tests/cpp/api_test.cpp.

CPU: the test program compiles and links against the backend library.
GPU: every C++ test passes on the device (HONEI-style checks including the
operator-list W-cycle multigrid against FEAT's poisson/sol_6)."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CPP = os.path.join(ROOT, "tests", "cpp")
BIN = os.path.join(CPP, "_build", "api_test")
GOLDEN = os.path.join(ROOT, "tests", "golden", "honei_testdata")


def test_cpp_api_builds():
    r = subprocess.run(["make", "-s", "-C", CPP], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-4000:]
    assert os.path.exists(BIN)


@pytest.mark.gpu
def test_cpp_api_on_gpu():
    if not os.path.exists(BIN):
        pytest.skip("C++ API test not built (run __graft_entry__.build())")
    r = subprocess.run([BIN, GOLDEN], capture_output=True, text=True, timeout=240)
    print(r.stdout)
    assert r.returncode == 0, r.stdout[-4000:] + r.stderr[-2000:]
    assert "FAIL" not in r.stdout
