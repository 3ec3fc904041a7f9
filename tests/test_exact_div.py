"""div3_rn (honei_amd/csrc/exact_div.hh), the correctly rounded x / 3 of the
matrix-free Q1 weights, against the division x / 3.0 on the host: random
normal doubles over the whole exponent range, runs of consecutive
significands (binade edges, quotients next to powers of two) and the
weights' own operands (lognormal coefficients and their sums), bit for bit.
The check runs tests/cpp/div3_check, built from the same header the device
code includes."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
TOOL = os.path.join(ROOT, "tests", "cpp", "_build", "div3_check")


@pytest.fixture(scope="module", autouse=True)
def _built():
    r = subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "tests", "cpp")], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-4000:]


def test_div3_rn_matches_division():
    r = subprocess.run([TOOL, "20000000"], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    assert "agrees" in r.stdout
