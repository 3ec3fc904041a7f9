"""Input generators shared by the tests and the bench.

The matrices here restate the patterns HONEI's own tests build
(la/sparse_matrix_ell_TEST.cc:45-55, la/product_TEST.cc:786-803) and the
synthetic workloads of the benchmark configurations (Q1 Laplacian with
x uniform in [-1, 1]; random banded ELL with row lengths uniform in
{1..27}).  Seeds go through SplitMix64 so C++ and Python draw the same
numbers.
"""
from __future__ import annotations

import numpy as np

MASK64 = (1 << 64) - 1


def splitmix64(seed: int, n: int) -> np.ndarray:
    """n outputs of SplitMix64 starting from `seed` (vectorised)."""
    i = np.arange(1, n + 1, dtype=np.uint64)
    with np.errstate(over="ignore"):
        z = np.uint64(seed & MASK64) + i * np.uint64(0x9E3779B97F4A7C15)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    return z


def uniform01(seed: int, n: int) -> np.ndarray:
    """U[0,1) = (u >> 11) * 2^-53."""
    return (splitmix64(seed, n) >> np.uint64(11)).astype(np.float64) * (1.0 / (1 << 53))


def uniform_pm1(seed: int, n: int) -> np.ndarray:
    return 2.0 * uniform01(seed, n) - 1.0


def honei_pattern_csr(rows: int, cols: int, plus_one: bool = False, corner: float | None = None,
                      dtype=np.float64):
    """Element index i = r*cols + c gets i/1.234 (+1) when i % 5 == 0."""
    idx = np.arange(rows * cols, dtype=np.int64)
    sel = idx[idx % 5 == 0]
    val = sel.astype(np.float64) / 1.234 + (1.0 if plus_one else 0.0)
    r = sel // cols
    c = sel % cols
    val = val.astype(dtype)
    if corner is not None:
        if len(sel) and sel[0] == 0:
            val[0] = dtype(corner)
    rp = np.zeros(rows + 1, np.uint64)
    np.add.at(rp, r + 1, 1)
    rp = np.cumsum(rp).astype(np.uint64)
    # value at index 0 is 0/1.234 (+0) = 0 when plus_one is False: SparseMatrix
    # does not store assigned zeros, so drop it unless overridden.
    keep = val != 0
    if not keep.all():
        r, c, val = r[keep], c[keep], val[keep]
        rp = np.zeros(rows + 1, np.uint64)
        np.add.at(rp, r + 1, 1)
        rp = np.cumsum(rp).astype(np.uint64)
    return rp, c.astype(np.uint64), val


def random_band_csr(n: int, seed: int, max_len: int = 27, band: int = 4096, dtype=np.float64):
    """Random ELL workload: row lengths uniform in {1..max_len}, distinct
    columns uniform in [r-band, r+band] clipped to [0, n), values N(0,1)."""
    rng = np.random.default_rng(seed)
    lens = rng.integers(1, max_len + 1, size=n)
    rp = np.zeros(n + 1, np.uint64)
    rp[1:] = np.cumsum(lens)
    nnz = int(rp[-1])
    cols = np.empty(nnz, np.int64)
    rows = np.repeat(np.arange(n, dtype=np.int64), lens)
    # draw offsets with rejection of duplicates per row (vectorised, few passes)
    off = rng.integers(-band, band + 1, size=nnz)
    cols[:] = np.clip(rows + off, 0, n - 1)
    for _ in range(64):
        order = np.lexsort((cols, rows))
        cols_s, rows_s = cols[order], rows[order]
        dup = np.zeros(nnz, bool)
        dup[1:] = (cols_s[1:] == cols_s[:-1]) & (rows_s[1:] == rows_s[:-1])
        if not dup.any():
            break
        d_idx = order[dup]
        cols[d_idx] = np.clip(rows[d_idx] + rng.integers(-band, band + 1, size=d_idx.size), 0, n - 1)
    order = np.lexsort((cols, rows))
    cols = cols[order]
    vals = rng.standard_normal(nnz).astype(dtype)
    vals[vals == 0] = dtype(1.0)
    return rp, cols.astype(np.uint64), vals


def lognormal_cells(seed: int, cells: int, sigma: float = 1.0) -> np.ndarray:
    """exp(sigma * z), z ~ N(0,1) by Box-Muller from draws 2i (u1 = 1 - U)
    and 2i+1 -- the device generator's formula (transcendentals may differ
    from the device's in the last bit)."""
    u = uniform01(seed, 2 * cells)
    z = np.sqrt(-2.0 * np.log(1.0 - u[0::2])) * np.cos(6.283185307179586 * u[1::2])
    return np.exp(sigma * z)


def poisson_system_seeds(base: int, i: int):
    """(coefficient seed, right-hand-side seed) of system i of a workload."""
    return (base << 32) + 2 * i, (base << 32) + 2 * i + 1


def batch_levels(nsys: int, seed: int, kmin: int = 5, kmax: int = 9) -> np.ndarray:
    """Finest level k of each system, uniform over {kmin..kmax} (side 2^k+1,
    log-uniform over the valid geometric-multigrid sides)."""
    u = uniform01(seed, nsys)
    return kmin + np.minimum((u * (kmax - kmin + 1)).astype(np.int64), kmax - kmin)


SERVE_LEVELS = (6, 7, 8, 9, 10)   # sides 65 .. 1025


def serving_requests(seconds: float, rate: float = 2000.0, seed: int = 5, levels=SERVE_LEVELS):
    """Serving stream: Poisson arrivals (exponential gaps -ln(1-U)/rate) over
    `seconds`, side 2^k+1 with k uniform over `levels`; request i carries
    the seeds poisson_system_seeds(seed, i).  Tuples (arrival s, level,
    coefficient seed, right-hand-side seed)."""
    n_max = int(rate * seconds * 1.5) + 100
    gaps = -np.log1p(-uniform01(seed, n_max)) / rate
    t = np.cumsum(gaps)
    n = int(np.searchsorted(t, seconds))
    pick = np.minimum((uniform01((seed << 8) + 1, n) * len(levels)).astype(int), len(levels) - 1)
    reqs = []
    for i in range(n):
        cs, bs = poisson_system_seeds(seed, i)
        reqs.append((float(t[i]), int(levels[pick[i]]), cs, bs))
    return reqs
