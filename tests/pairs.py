"""Seeded (x, y) pair sets covering every regime of the kernel (test inputs only)."""
import numpy as np


def iso(rng, n):
    v = rng.normal(size=(n, 4))
    return v / np.linalg.norm(v, axis=1, keepdims=True)


def regime_mix(seed, n_each=40):
    """Pairs in each regime of the reference's dispatch (kernels.c:82-115, chnr_d*.c)."""
    rng = np.random.default_rng(seed)
    xs, ys, tags = [], [], []

    def add(x, y, tag):
        xs.append(np.asarray(x, float)); ys.append(np.asarray(y, float)); tags.extend([tag] * len(x))

    # standard: |x|, |y| in [0.05, 3]
    add(iso(rng, n_each) * rng.uniform(0.05, 3.0, (n_each, 1)), iso(rng, n_each) * rng.uniform(0.05, 3.0, (n_each, 1)), "std")
    # Taylor-x: |x| below the first x node (0.0242)
    add(iso(rng, n_each) * rng.uniform(1e-3, 0.024, (n_each, 1)), iso(rng, n_each) * rng.uniform(0.05, 3.0, (n_each, 1)), "taylor")
    # x <-> x-y swap: y close to x
    x = iso(rng, n_each) * rng.uniform(0.1, 3.0, (n_each, 1))
    add(x, x + iso(rng, n_each) * rng.uniform(0.01, 0.3, (n_each, 1)) * np.linalg.norm(x, axis=1, keepdims=True), "swap")
    # small y: below the first y node
    add(iso(rng, n_each) * rng.uniform(0.05, 3.0, (n_each, 1)), iso(rng, n_each) * rng.uniform(1e-3, 8e-3, (n_each, 1)), "smally")
    # both small: Taylor and swap together
    add(iso(rng, n_each) * rng.uniform(1e-3, 0.02, (n_each, 1)), iso(rng, n_each) * rng.uniform(1e-3, 0.02, (n_each, 1)), "small2")
    # special points
    z = np.zeros((4, 4))
    v = iso(rng, 4) * rng.uniform(0.05, 3.0, (4, 1))
    add(z, v, "x0"); add(v, z, "y0"); add(v, v.copy(), "xy"); add(z, z, "00")
    # past the grids: |x| > 6.54 (zero kernel), |y| > 6.54 (top y edge path)
    add(iso(rng, 4) * 7.0, iso(rng, 4) * 1.0, "outx")
    add(iso(rng, 4) * 1.0, iso(rng, 4) * 6.8, "edgey")
    return np.concatenate(xs), np.concatenate(ys), np.array(tags)


def log_uniform(seed, n, lo=1e-3, hi=4.0):
    """Workload of the FP64 batch benchmark: |x|, |y| log-uniform in [lo, hi], isotropic."""
    rng = np.random.default_rng(seed)
    rx = np.exp(rng.uniform(np.log(lo), np.log(hi), (n, 1)))
    ry = np.exp(rng.uniform(np.log(lo), np.log(hi), (n, 1)))
    return iso(rng, n) * rx, iso(rng, n) * ry


def normwise(a, b):
    """Per-pair error |a-b|_max / |b|_max (0 where both are zero)."""
    a = np.asarray(a).reshape(len(a), -1)
    b = np.asarray(b).reshape(len(b), -1)
    den = np.abs(b).max(axis=1)
    num = np.abs(a - b).max(axis=1)
    return np.where(den > 0, num / np.where(den > 0, den, 1), num)
