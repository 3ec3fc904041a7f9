"""CPU model of the batched row-wise transfers (k_b_restrict_rows,
k_b_prolong_rows in honei_amd/csrc/batch_kernels.hip) against the oracle's
restatement of the reference's sweeps (math/restriction.hh:40-172,
math/prolongation.hh:41-139, every border Dirichlet), bit for bit.

The batched restriction adds the eight weighted fine terms of an interior
coarse node (r, c) to the centre value in a fixed row-term order with the two
column terms placed by c - r; this restates that placement in numpy (float32
and float64, the same operations one by one) and checks it reproduces the
CPU sweep's accumulation exactly, so a wrong placement (a column term one
slot early or late) fails on the diagonal band where the order changes.
The prolongation's four-case formula is checked the same way."""
import numpy as np
import pytest

import oracle
from tests.patterns import uniform_pm1


def model_restriction(fine, m):
    dt = fine.dtype.type
    nf, nc = 2 * m + 1, m + 1
    F = fine.reshape(nf, nf)
    out = np.zeros((nc, nc), fine.dtype)
    h, q = dt(0.5), dt(0.25)
    for r in range(1, m):
        for c in range(1, m):
            fm, f0r, fp = F[2 * r - 1], F[2 * r], F[2 * r + 1]
            cc = 2 * c
            A, B, C, D = h * fm[cc], h * fp[cc], h * f0r[cc + 1], h * f0r[cc - 1]
            E, Fv, G, Hh = q * fp[cc + 1], q * fp[cc - 1], q * fm[cc + 1], q * fm[cc - 1]
            dd = c - r
            s = f0r[cc]
            if dd <= -1:
                s = dt(s + D)
            if dd <= -2:
                s = dt(s + C)
            s = dt(s + A)
            if dd == -1:
                s = dt(s + C)
            if dd == 0:
                s = dt(s + D)
            s = dt(s + G)
            s = dt(s + Hh)
            s = dt(s + B)
            if dd == 0:
                s = dt(s + C)
            if dd == 1:
                s = dt(s + D)
            s = dt(s + E)
            s = dt(s + Fv)
            if dd >= 2:
                s = dt(s + D)
            if dd >= 1:
                s = dt(s + C)
            out[r, c] = s
    return out.reshape(-1)


def model_prolongation_axpy(x_f, coarse, m):
    dt = coarse.dtype.type
    nf, nc = 2 * m + 1, m + 1
    X = coarse.reshape(nc, nc)
    out = x_f.copy().reshape(nf, nf)
    h, q = dt(0.5), dt(0.25)
    for R in range(nf):
        for Cc in range(nf):
            a, b = X[R >> 1, Cc >> 1], X[R >> 1, (Cc + 1) >> 1]
            c, d = X[(R + 1) >> 1, Cc >> 1], X[(R + 1) >> 1, (Cc + 1) >> 1]
            if R & 1 and Cc & 1:
                f = dt(dt(dt(dt(dt(0) + a * q) + b * q) + c * q) + d * q)
            elif R & 1:
                f = dt(dt(dt(0) + a * h) + c * h)
            elif Cc & 1:
                f = dt(dt(dt(0) + a * h) + b * h)
            else:
                f = a
            if R in (0, nf - 1) or Cc in (0, nf - 1):
                f = dt(0)
            out[R, Cc] = dt(out[R, Cc] + f * dt(1))
    return out.reshape(-1)


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("m", [2, 3, 4, 8, 16])
def test_row_restriction_order_matches_cpu_sweep(dtype, m):
    nf = 2 * m + 1
    fine = uniform_pm1(50 + m, nf * nf).astype(dtype)
    want = oracle.restriction(fine)
    got = model_restriction(fine, m)
    assert np.array_equal(got.view(np.uint8), want.view(np.uint8))


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("m", [1, 2, 4, 8])
def test_row_prolongation_matches_cpu_sweep(dtype, m):
    nf, nc = 2 * m + 1, m + 1
    coarse = uniform_pm1(70 + m, nc * nc).astype(dtype)
    x_f = uniform_pm1(90 + m, nf * nf).astype(dtype)
    p = oracle.prolongation(coarse)
    want = (x_f + p * dtype(1)).astype(dtype)
    got = model_prolongation_axpy(x_f, coarse, m)
    assert np.array_equal(got.view(np.uint8), want.view(np.uint8))
