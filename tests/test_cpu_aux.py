"""CPU checks of the host-side pieces: the surrogate's CPU reference against finite differences,
the synthetic workloads, and the benchmark's command-line contract."""
import os
import subprocess
import sys

import numpy as np

from surrogate_ref import exact, forward_backward

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
from kqed_amd import workloads as W  # noqa: E402

K = (16, 128, 128, 128, 128)
N = (128, 128, 128, 128, 48)


def _layers(rng, scale=0.3):
    out = []
    for k, n in zip(K, N):
        Wt = rng.normal(0, scale / np.sqrt(k), (k, n))
        Wt[3:16, :] = 0.0 if k == 16 else Wt[3:16, :]
        out.append((Wt, rng.normal(0, 0.05, n)))
    return out


def test_surrogate_reference_gradients_match_finite_differences():
    """The backward pass of the CPU reference (exact arithmetic) is the derivative of its loss."""
    rng = np.random.default_rng(1)
    layers = _layers(rng)
    B = 7
    X3 = rng.uniform(-1, 1, (B, 3))
    tgt = rng.normal(0, 0.5, (B, 34))
    wgt = (rng.uniform(size=(B, 34)) < 0.8).astype(float)
    _, grads = forward_backward(layers, X3, tgt, wgt, rnd=exact)
    h = 1e-6
    for l, (i, j) in ((0, (1, 5)), (2, (7, 11)), (4, (3, 20)), (1, (100, 0))):
        Wp = [(w.copy(), b.copy()) for w, b in layers]
        Wm = [(w.copy(), b.copy()) for w, b in layers]
        Wp[l][0][i, j] += h
        Wm[l][0][i, j] -= h
        fd = (forward_backward(Wp, X3, tgt, wgt, rnd=exact)[0] - forward_backward(Wm, X3, tgt, wgt, rnd=exact)[0]) / (2 * h)
        assert abs(fd - grads[l][0][i, j]) <= 1e-6 * max(1.0, abs(fd)), (l, i, j, fd, grads[l][0][i, j])
    for l, j in ((0, 3), (3, 77), (4, 30)):
        bp = [(w.copy(), b.copy()) for w, b in layers]
        bm = [(w.copy(), b.copy()) for w, b in layers]
        bp[l][1][j] += h
        bm[l][1][j] -= h
        fd = (forward_backward(bp, X3, tgt, wgt, rnd=exact)[0] - forward_backward(bm, X3, tgt, wgt, rnd=exact)[0]) / (2 * h)
        assert abs(fd - grads[l][1][j]) <= 1e-6 * max(1.0, abs(fd))


def test_lattice_and_sites():
    x = W.lattice_sites(4)
    assert x.shape == (256, 4)
    n = x / W.LATTICE_A
    assert np.allclose(n, np.round(n)) and n.min() == -2 and n.max() == 1
    assert len(np.unique(n, axis=0)) == 256
    assert (np.abs(x).sum(axis=1) == 0).sum() == 1  # the site x = 0 is on the lattice
    y = W.y_sites()
    assert y.shape == (16, 4) and np.all(np.linalg.norm(y, axis=1) > 0)
    T = W.site_tensors(x, seed=3)
    assert T.dtype == np.float32 and T.shape == (256, 384)
    assert np.array_equal(T, W.site_tensors(x, seed=3))
    # the envelope exp(-|x|/2): the first site's row scaled back is a unit-variance sample
    assert np.array_equal(T[:64], W.site_tensors(x[:64], seed=3)[:64])


def test_ragged_workload_law():
    y, sizes, x = W.ragged(seed=4)
    assert len(y) == 256 and len(sizes) == 256
    assert sizes.min() >= 1024 and sizes.max() <= 262144
    assert x.shape == (sizes.sum(), 4)
    ny = np.linalg.norm(y, axis=1)
    assert ny.min() >= 0.05 and ny.max() <= 3.0
    nx = np.linalg.norm(x, axis=1)
    assert nx.min() >= 1e-3 and nx.max() <= 4.0
    assert 10000 < np.median(sizes) < 27000


def test_bench_refuses_several_gpus():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 2 and "one GPU" in r.stderr
