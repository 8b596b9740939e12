"""Synthetic workloads of the benchmark and tests (inputs only; no evaluation here).

* pairs_uniform: 4096 (x, y) pairs, |x|, |y| uniform in [0.01, 3.0], isotropic (seed 1).
* pairs_log_uniform: 2^24 (x, y) pairs, |x|, |y| log-uniform in [1e-3, 4], isotropic (seed 2).
* lattice: the sites of an L^4 periodic lattice, n_mu in [-L/2, L/2 - 1], x = a n (a m_mu = 0.0536),
  per-site FP32 tensors T(x)_c = N(0,1) exp(-|x|/2) (seed 3), and 16 fixed y sites (Y_SITES).
* ragged: 256 y points (|y| uniform [0.05, 3]) with x-sets of lognormal size (median 16384,
  sigma 1.0, clipped to [1024, 262144]), |x| log-uniform [1e-3, 4], isotropic (seed 4).
"""
import numpy as np

LATTICE_A = 0.0536

# 16 fixed y sites (lattice vectors n_y; y = LATTICE_A * n_y): on-axis, diagonal and generic
# directions from one to twelve lattice units, so that x = y and |x - y| = a both occur.
Y_SITES = np.array([
    (0, 0, 0, 1), (0, 0, 1, 1), (0, 1, 1, 1), (1, 1, 1, 1),
    (0, 0, 0, 2), (0, 0, 2, 1), (1, 2, 0, 2), (2, 2, 2, 2),
    (0, 0, 0, 4), (0, 3, 0, 3), (1, 2, 3, 4), (4, 4, 0, 0),
    (0, 0, 0, 8), (5, 3, 2, 1), (6, 6, 6, 6), (0, 0, 0, 12),
], dtype=np.float64)


def iso(rng, n):
    v = rng.normal(size=(n, 4))
    return v / np.linalg.norm(v, axis=1, keepdims=True)


def pairs_log_uniform(seed, n, lo=1e-3, hi=4.0):
    rng = np.random.default_rng(seed)
    rx = np.exp(rng.uniform(np.log(lo), np.log(hi), (n, 1)))
    ry = np.exp(rng.uniform(np.log(lo), np.log(hi), (n, 1)))
    return iso(rng, n) * rx, iso(rng, n) * ry


def lattice_sites(L, a=LATTICE_A):
    """x[L^4][4] of the periodic lattice, n_mu in [-L/2, L/2 - 1], lexicographic in (n0, n1, n2, n3)."""
    r = np.arange(-(L // 2), L - L // 2, dtype=np.float64)
    n = np.stack(np.meshgrid(r, r, r, r, indexing="ij"), axis=-1).reshape(-1, 4)
    return a * n


def site_tensors(x, seed=3, chunk=1 << 16):
    """T[nx][384] FP32 = N(0,1) * exp(-|x|/2), generated in chunks of sites (same stream for any chunk)."""
    rng = np.random.default_rng(seed)
    nx = x.shape[0]
    T = np.empty((nx, 384), np.float32)
    for i in range(0, nx, chunk):
        j = min(nx, i + chunk)
        T[i:j] = rng.standard_normal((j - i, 384), dtype=np.float32)
        T[i:j] *= np.exp(-0.5 * np.linalg.norm(x[i:j], axis=1)).astype(np.float32)[:, None]
    return T


def y_sites(a=LATTICE_A):
    return a * Y_SITES


def ragged(seed=4, ny=256, median=16384, sigma=1.0, lo=1024, hi=262144):
    """(y[ny][4], sizes[ny], x[sum sizes][4]): the x-set of y_j is x[off_j : off_j + sizes_j]."""
    rng = np.random.default_rng(seed)
    y = iso(rng, ny) * rng.uniform(0.05, 3.0, (ny, 1))
    sizes = np.clip(np.round(rng.lognormal(np.log(median), sigma, ny)), lo, hi).astype(np.int64)
    n = int(sizes.sum())
    x = iso(rng, n) * np.exp(rng.uniform(np.log(1e-3), np.log(4.0), (n, 1)))
    return y, sizes, x


def pairs_uniform(seed, n, lo=0.01, hi=3.0):
    """(x, y) pairs with |x|, |y| uniform in [lo, hi], isotropic: the CPU L0 benchmark law (4096
    pairs, [0.01, 3.0], seed 1)."""
    rng = np.random.default_rng(seed)
    rx = rng.uniform(lo, hi, (n, 1))
    ry = rng.uniform(lo, hi, (n, 1))
    return iso(rng, n) * rx, iso(rng, n) * ry
