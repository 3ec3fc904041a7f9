"""Symmetric positive definite test systems for the Jacobi-preconditioned CG
(test infrastructure: built with the oracle's assembly)."""
import numpy as np

import oracle
from tests.patterns import lognormal_cells


def diag_inverse(A):
    """dd_inverted[i] = 1 / A(i, i) in A's value type, as the reference's
    PCG-Jacobi tests build it (math/dune_regression_TEST.cc:78-82)."""
    d = np.zeros(A.rows, A.dtype)
    rows = np.arange(A.rows, dtype=np.uint64)
    for k in range(A.Ax.size // A.stride):
        for t in range(A.threads):
            s = rows * np.uint64(A.threads) + np.uint64(t) + np.uint64(k * A.stride)
            hit = (A.Aj[s] == rows) & (A.Ax[s] != 0)
            d[hit] = A.Ax[s][hit]
    return (A.dtype.type(1) / d).astype(A.dtype)


def lognormal_q1_spd(side, seed, dtype=np.float64, threads=1, sigma=1.0):
    """Variable-coefficient Q1 operator (lognormal cell coefficients) made
    symmetric positive definite, the setting CG assumes: the couplings of
    interior rows to Dirichlet columns are dropped, so Dirichlet rows and
    columns are both identity."""
    kappa = lognormal_cells(seed, (side - 1) ** 2, sigma)
    rp, col, val = oracle.banded_to_csr(oracle.q1_coeff_bands(side, kappa, dtype))
    keep0 = val != 0
    r0 = np.repeat(np.arange(side * side), np.diff(rp).astype(np.int64))
    col, val = col[keep0], val[keep0]
    rp = np.concatenate([[0], np.cumsum(np.bincount(r0[keep0], minlength=side * side))]).astype(np.uint64)
    g = np.arange(side * side)
    bd = (g < side) | (g >= side * (side - 1)) | (g % side == 0) | (g % side == side - 1)
    c = col.astype(np.int64)
    rows = np.repeat(g, np.diff(rp).astype(np.int64))
    keep = ~bd[c] | (c == rows)
    rp2 = np.concatenate([[0], np.cumsum(np.bincount(rows[keep], minlength=g.size))]).astype(np.uint64)
    return oracle.ell_from_sm(g.size, g.size, rp2, col[keep], val[keep], threads)
