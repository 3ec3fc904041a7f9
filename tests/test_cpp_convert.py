"""The C++ mirror's own SparseMatrix -> SparseMatrixELL conversion
(honei_amd/honei/la/sparse_matrix_ell.hh, following la/sparse_matrix_ell-impl.hh:
75-122) pinned bit for bit -- index arrays, values, stride, width and row
lengths -- against the reference's own data file (5pt_10x10.ell, written by
clients/io_tools/mtxtoell.cc through the same conversion), against the
reference's code built from its sources (oracle/_ref, skipped where absent)
and against the oracle's restatement, for threads-per-row 1..16.

The conversion runs in tests/cpp/convert_tool (host code only)."""
import os
import subprocess

import numpy as np
import pytest

import oracle
from oracle import ref
from tests.patterns import honei_pattern_csr, random_band_csr

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
TOOL = os.path.join(ROOT, "tests", "cpp", "_build", "convert_tool")


@pytest.fixture(scope="module", autouse=True)
def _built():
    r = subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "tests", "cpp")], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-4000:]


def mirror_convert(rows, cols, rp, col, val, threads):
    """Run the mirror's conversion on the entries of CSR arrays (assigned in
    row order, columns ascending)."""
    rp = np.asarray(rp, np.int64)
    n = int(rp[-1])
    r = np.repeat(np.arange(rows, dtype=np.uint64), np.diff(rp))
    rec = np.zeros(n, dtype=[("r", "<u8"), ("c", "<u8"), ("v", "<f8")])
    rec["r"], rec["c"], rec["v"] = r, np.asarray(col, np.uint64), np.asarray(val, np.float64)
    inp = np.array([rows, cols, threads, n], np.uint64).tobytes() + rec.tobytes()
    out = subprocess.run([TOOL], input=inp, capture_output=True, check=True).stdout
    stride, ncpr = np.frombuffer(out[:16], np.uint64)
    size = int(stride * ncpr)
    o = 16
    Aj = np.frombuffer(out[o:o + 8 * size], np.uint64)
    o += 8 * size
    Ax = np.frombuffer(out[o:o + 8 * size], np.float64)
    o += 8 * size
    Arl = np.frombuffer(out[o:o + 8 * rows], np.uint64)
    assert o + 8 * rows == len(out)
    return oracle.Ell(rows, cols, int(stride), int(ncpr), threads, Aj.copy(), Ax.copy(), Arl.copy())


def _same(A, B):
    assert (A.rows, A.cols, A.stride, A.ncpr, A.threads) == (B.rows, B.cols, B.stride, B.ncpr, B.threads)
    assert np.array_equal(A.Aj, B.Aj)
    assert np.array_equal(A.Ax.view(np.uint64), B.Ax.astype(np.float64).view(np.uint64))
    assert np.array_equal(A.Arl, B.Arl)


def test_mirror_conversion_reproduces_reference_file(golden):
    """5pt_10x10.mtx -> SparseMatrix -> SparseMatrixELL(T = 1) gives the
    arrays of the reference's 5pt_10x10.ell."""
    rows, cols, rp, col, val = oracle.read_mtx(os.path.join(golden, "5pt_10x10.mtx"))
    M = mirror_convert(rows, cols, rp, col, val, 1)
    F = oracle.read_ell(os.path.join(golden, "5pt_10x10.ell"))
    _same(M, F)
    assert (M.stride, M.ncpr) == (128, 5)


@pytest.mark.parametrize("threads", [1, 2, 4, 8, 16])
@pytest.mark.parametrize("case", ["pattern", "band", "5pt"])
def test_mirror_conversion_bitexact(golden, case, threads):
    """Threads-per-row layouts: the mirror's arrays equal the oracle's and
    (where built) those of the reference's own SparseMatrixELL."""
    if case == "pattern":
        rows = cols = 500
        rp, col, val = honei_pattern_csr(rows, cols, plus_one=True)
    elif case == "band":
        rows = cols = 2000
        rp, col, val = random_band_csr(rows, seed=threads, max_len=27, band=60)
    else:
        rows, cols, rp, col, val = oracle.read_mtx(os.path.join(golden, "5pt_10x10.mtx"))
    M = mirror_convert(rows, cols, rp, col, val, threads)
    _same(M, oracle.ell_from_sm(rows, cols, rp, col, val, threads))
    if ref.available():
        _same(M, ref.ell_from_sm(rows, cols, np.asarray(rp, np.uint64), np.asarray(col, np.uint64),
                                 np.asarray(val, np.float64), threads).to_ell())


def test_mirror_conversion_drops_stored_zeros():
    """An entry assigned the value 0 counts toward the row length but is not
    stored (la/sparse_matrix_ell-impl.hh:88-118), as the reference does."""
    rp = np.array([0, 3, 4], np.uint64)
    col = np.array([0, 2, 5, 1], np.uint64)
    val = np.array([1.5, 0.0, -2.0, 3.0])
    for T in (1, 2):
        M = mirror_convert(2, 6, rp, col, val, T)
        _same(M, oracle.ell_from_sm(2, 6, rp, col, val, T))
        if ref.available():
            _same(M, ref.ell_from_sm(2, 6, rp, col, val, T).to_ell())
