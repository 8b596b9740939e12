"""Workload configurations, reference identities and concurrency on the GPU path.

* The CPU L0 benchmark law (4096 pairs, |x|, |y| uniform in [0.01, 3.0], seed 1; the reference's
  own timing loop is amu_forlattice.c:1402-1433): every pair against the oracle.
* The ragged batch law (256 y with lognormal x-sets, reduced here to 24 groups): the graph-replayed
  pipeline, FP64 and FP32 outputs, on a uniform sample against the oracle.
* The identities the reference's examples check (amu_forlattice.c examples 3, 9, 11, 15), on the
  GPU's own outputs.
* Thread safety (example 22, amu_forlattice.c:1234-1256): concurrent host callers of the per-pair
  API get bit-identical results; ragged batches on two streams do not share buffers.
"""
import threading

import numpy as np
import pytest

import kqed_amd
from kqed_amd import workloads as W
from pairs import normwise

pytestmark = pytest.mark.gpu


def _dev(a):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a)).to("cuda")


def test_cpu_benchmark_law_all_pairs(gpu_full, oracle_full):
    x, y = W.pairs_uniform(1, 4096)
    got = gpu_full.batch_host(0, x, y)
    ref = oracle_full.batch(0, x, y)
    err = normwise(got, ref)
    assert err.max() <= 1e-12, f"max {err.max():.2e} median {np.median(err):.2e}"


@pytest.mark.parametrize("layout", [kqed_amd.KQED_AOS, kqed_amd.KQED_AOS_F32])
def test_ragged_graph_law_against_oracle(gpu_full, oracle_full, layout):
    import torch
    y, sizes, x = W.ragged(seed=4, ny=24, median=2048, lo=256, hi=8192)
    off = np.concatenate([[0], np.cumsum(sizes)]).astype(np.uintp)
    n = int(off[-1])
    xd, yd = _dev(x), _dev(y)
    f32 = layout == kqed_amd.KQED_AOS_F32
    out = torch.empty((n, 384), dtype=torch.float32 if f32 else torch.float64, device="cuda")
    g = gpu_full.graph_ragged(0, off, yd.data_ptr(), xd.data_ptr(), out.data_ptr(), layout=layout)
    try:
        for _ in range(2):
            out.fill_(float("nan"))
            g.launch(torch.cuda.current_stream().cuda_stream)
            torch.cuda.synchronize()
            got = out.cpu().numpy().astype(np.float64)
            idx = np.sort(np.random.default_rng(6).choice(n, 3000, replace=False))
            ypair = np.repeat(y, sizes, axis=0)
            err = normwise(got[idx], oracle_full.batch(0, x[idx], ypair[idx]))
            assert err.max() <= (1e-5 if f32 else 1e-12), f"max {err.max():.2e} median {np.median(err):.2e}"
    finally:
        g.destroy()


def test_identities_example3(gpu_small):
    """L0(0,0) = L1(x,x) = L2(0,y) = L2(x,0) = L3(x,x) = L3(0,y) = L3(0,x) = L3(y,y) = 0."""
    z = np.zeros(4)
    x = np.array([0.1, 0.2, 0.3, 0.4])
    y = np.array([0.4, 0.3, 0.2, 0.1])
    cases = [(0, z, z), (1, x, x), (2, z, y), (2, x, z), (3, x, x), (3, z, y), (3, z, x), (3, y, y)]
    for which, a, b in cases:
        k = gpu_small.batch_host(which, a[None], b[None])
        assert np.abs(k).max() <= 1e-14, (which, a, b)


def test_identity_example9(gpu_small):
    """L(x,x) = -L_{mu<->lambda}(x,0)."""
    th = np.arange(0.1, 1.0, 0.15)[:, None]
    x = th * np.array([1.0, 2.0, 3.0, 4.0])
    k1 = gpu_small.batch_host(0, x, x).reshape(-1, 6, 4, 4, 4)
    k2 = gpu_small.batch_host(0, x, np.zeros_like(x)).reshape(-1, 6, 4, 4, 4)
    assert np.abs(k1 + k2.transpose(0, 1, 4, 3, 2)).max() <= 1e-14


def test_identity_example11(gpu_small):
    """L_i(x,y) = -L_i(-x,-y) on the reference's 2x2x2x4 toy lattice, i = 0..3."""
    def crd(i):
        c = [i % 2, (i // 2) % 2, (i // 4) % 2, i // 8]
        c[3] = c[3] - 4 if c[3] > 2 else c[3]
        return c
    pts = np.array([crd(i) for i in range(32)], dtype=np.float64)
    X = np.repeat(pts, 32, axis=0)
    Y = np.tile(pts, (32, 1))
    for which in range(4):
        a = gpu_small.batch_host(which, X, Y)
        b = gpu_small.batch_host(which, -X, -Y)
        scale = np.maximum(np.abs(a).max(axis=1, keepdims=True), 1.0)
        assert (np.abs(a + b) / scale).max() <= 1e-14, which


def test_identity_example15(gpu_small):
    """L(x-y,0) = -L(y-x,0) and L(0,x-y) = -L(0,y-x)."""
    rng = np.random.default_rng(15)
    x, y = rng.uniform(-1, 1, (10, 4)), rng.uniform(-1, 1, (10, 4))
    z = np.zeros_like(x)
    assert np.abs(gpu_small.batch_host(0, x - y, z) + gpu_small.batch_host(0, y - x, z)).max() <= 1e-14
    assert np.abs(gpu_small.batch_host(0, z, x - y) + gpu_small.batch_host(0, z, y - x)).max() <= 1e-14


def test_concurrent_per_pair_callers(gpu_full):
    """Eight host threads call QED_kernel_L0 at once, as example 22 does with its OpenMP team
    (x = (1,2,3,4), y = (1,2,2,2), both orders), plus pairs of their own: bit-identical results."""
    base = [(np.array([1.0, 2, 3, 4]), np.array([1.0, 2, 2, 2])), (np.array([1.0, 2, 2, 2]), np.array([1.0, 2, 3, 4]))]
    rng = np.random.default_rng(22)
    own = [[(rng.normal(size=4), rng.normal(size=4)) for _ in range(6)] for _ in range(8)]
    serial = [[gpu_full.QED_kernel_L0(a, b).copy() for a, b in base + own[t]] for t in range(8)]
    results = [None] * 8
    errors = []
    start = threading.Barrier(8)

    def worker(t):
        try:
            start.wait()
            out = []
            for _ in range(3):
                out = [gpu_full.QED_kernel_L0(a, b).copy() for a, b in base + own[t]]
            results[t] = out
        except Exception as e:  # reported below
            errors.append(e)

    th = [threading.Thread(target=worker, args=(t,)) for t in range(8)]
    for h in th:
        h.start()
    for h in th:
        h.join(timeout=120)
    assert not errors, errors
    for t in range(8):
        for k1, k2 in zip(results[t], serial[t]):
            assert np.array_equal(k1, k2)
        assert np.array_equal(results[t][0], results[0][0]) and np.array_equal(results[t][1], results[0][1])


def test_ragged_batches_on_two_streams(gpu_full):
    """Two ragged batches enqueued back to back on two streams from one thread: each equals its
    own plain batch (the offsets and expanded y vectors are per stream)."""
    import torch
    cases = []
    for seed, ny in ((41, 9), (42, 14)):
        y, sizes, x = W.ragged(seed=seed, ny=ny, median=1500, lo=100, hi=4000)
        off = np.concatenate([[0], np.cumsum(sizes)]).astype(np.uintp)
        cases.append((y, off, x))
    streams = [torch.cuda.Stream(), torch.cuda.Stream()]
    outs = []
    keep = []
    for (y, off, x), s in zip(cases, streams):
        xd, yd = _dev(x), _dev(y)
        out = torch.empty((int(off[-1]), 384), dtype=torch.float64, device="cuda")
        keep += [xd, yd]
        gpu_full.batch_ragged(0, off, yd.data_ptr(), xd.data_ptr(), out.data_ptr(), stream=s.cuda_stream)
        outs.append(out)
    torch.cuda.synchronize()
    for (y, off, x), out in zip(cases, outs):
        ypair = np.repeat(y, np.diff(off).astype(np.int64), axis=0)
        assert np.array_equal(out.cpu().numpy(), gpu_full.batch_host(0, x, ypair))
