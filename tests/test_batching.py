"""Ragged groups, graph replay, launch parameters, the autotuner and the FP32 output layouts.

Per-pair results do not depend on the batch around the pair (every pair runs the same code; sorting
only reorders), so ragged, graph and re-tuned runs must equal the plain batch bit for bit.  The FP32
layouts run the assembly maps in FP32: their bar is a normwise relative error of 1e-5 against FP64.
"""
import numpy as np
import pytest

import kqed_amd
from kqed_amd import workloads as W
from pairs import normwise, regime_mix

pytestmark = pytest.mark.gpu


def _dev(a, dtype=None):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a if dtype is None else a.astype(dtype))).to("cuda")


def _ragged_case(seed=4, ny=12):
    y, sizes, x = W.ragged(seed=seed, ny=ny, median=2000, lo=1, hi=6000)
    sizes[3] = 0  # an empty group
    off = np.concatenate([[0], np.cumsum(sizes)]).astype(np.uintp)
    x = x[: off[-1]]
    return y, off, x


def test_ragged_equals_expanded_batch(gpu_full):
    import torch
    y, off, x = _ragged_case()
    n = int(off[-1])
    xd, yd = _dev(x), _dev(y)
    out = torch.empty((n, 384), dtype=torch.float64, device="cuda")
    gpu_full.batch_ragged(0, off, yd.data_ptr(), xd.data_ptr(), out.data_ptr(),
                          stream=torch.cuda.current_stream().cuda_stream)
    torch.cuda.synchronize()
    ypair = np.repeat(y, np.diff(off).astype(np.int64), axis=0)
    ref = gpu_full.batch_host(0, x, ypair)
    assert np.array_equal(out.cpu().numpy(), ref)


def test_graph_replays_match_eager(gpu_full):
    import torch
    y, off, x = _ragged_case(seed=5)
    n = int(off[-1])
    xd, yd = _dev(x), _dev(y)
    out = torch.empty((n, 384), dtype=torch.float64, device="cuda")
    stream = torch.cuda.current_stream().cuda_stream
    g = gpu_full.graph_ragged(2, off, yd.data_ptr(), xd.data_ptr(), out.data_ptr())
    ypair = np.repeat(y, np.diff(off).astype(np.int64), axis=0)
    eager = gpu_full.batch_host(2, x, ypair)
    for rep in range(3):
        out.fill_(np.nan)
        g.launch(stream)
        torch.cuda.synchronize()
        assert np.array_equal(out.cpu().numpy(), eager), f"replay {rep}"
        _ = gpu_full.batch_host(0, x[:100], ypair[:100])  # eager work between replays
    # new contents of the captured input buffer are picked up by the next replay
    x2 = x * 0.9
    xd.copy_(_dev(x2))
    g.launch(stream)
    torch.cuda.synchronize()
    assert np.array_equal(out.cpu().numpy(), gpu_full.batch_host(2, x2, ypair))
    g.destroy()


def test_tune_variants_identical(gpu_full):
    x, y = W.pairs_log_uniform(8, 40000)
    base = gpu_full.batch_host(0, x, y)
    t0 = gpu_full.get_tune()
    try:
        for kw in (dict(chunk=1 << 14), dict(asm_grid=37), dict(sort_mode=kqed_amd.KQED_SORT_XNODE),
                   dict(sort_mode=kqed_amd.KQED_SORT_NONE), dict(chunk=5000, sort_min=100000)):
            gpu_full.set_tune(**kw)
            assert np.array_equal(gpu_full.batch_host(0, x, y), base), kw
            gpu_full.set_tune(t0)
    finally:
        gpu_full.set_tune(t0)


def test_autotune_installs_a_candidate(gpu_full):
    import torch
    x, y = W.pairs_log_uniform(9, 1 << 17)
    xd, yd = _dev(x), _dev(y)
    out = torch.empty((len(x), 384), dtype=torch.float64, device="cuda")
    t0 = gpu_full.get_tune()
    try:
        tu, ms = gpu_full.autotune(0, len(x), xd.data_ptr(), yd.data_ptr(), out.data_ptr(), reps=1,
                                   stream=torch.cuda.current_stream().cuda_stream)
        assert ms > 0 and tu.chunk in (1 << 17, 1 << 18) and tu.sort_mode in (0, 1)
        assert gpu_full.get_tune().as_dict() == tu.as_dict()
    finally:
        gpu_full.set_tune(t0)


@pytest.mark.parametrize("which", [0, 1, 2, 3])
def test_fp32_layout_error(gpu_full, which):
    x, y, _ = regime_mix(40 + which, n_each=24)
    ref = gpu_full.batch_host(which, x, y)
    got = gpu_full.batch_host(which, x, y, layout=kqed_amd.KQED_AOS_F32)
    assert got.dtype == np.float32
    err = normwise(got.astype(np.float64), ref)
    assert err.max() <= 1e-5, f"max {err.max():.2e} median {np.median(err):.2e}"
    soa = gpu_full.batch_host(which, x, y, layout=kqed_amd.KQED_SOA_F32)
    assert np.array_equal(soa.T, got)
