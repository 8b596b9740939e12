"""FP64 HIP path vs the CPU oracle: every regime, every kernel (L0..L3, M-L2).

Bar: per-pair normwise relative error max|gpu - oracle| / max|oracle| <= 1e-12 (FP64 target of the
plain path); pairs the oracle returns as all-zero must be all-zero on the GPU too.
"""
import numpy as np
import pytest

import kqed_amd
from pairs import iso, log_uniform, normwise, regime_mix

TOL = 1e-12

pytestmark = pytest.mark.gpu


def _check(gpu, oracle, which, x, y, M=0.0, tags=None):
    got = gpu.batch_host(which, x, y, M=M)
    ref = oracle.batch(which, x, y, M=M)
    err = normwise(got, ref)
    zero_ref = np.abs(ref).max(axis=1) == 0
    assert np.all(np.abs(got[zero_ref]).max(axis=1, initial=0) == 0), "GPU nonzero where the oracle is zero"
    worst = int(np.argmax(err))
    msg = f"which={which} worst pair {worst} tag={None if tags is None else tags[worst]} err={err[worst]:.3e}"
    assert err.max() <= TOL, msg
    return err


@pytest.mark.parametrize("which", [0, 1, 2, 3])
def test_regimes_small_table(gpu_small, oracle_small, which):
    x, y, tags = regime_mix(11 + which)
    _check(gpu_small, oracle_small, which, x, y, tags=tags)


@pytest.mark.parametrize("M", [0.0, 0.5, 1.2])
def test_mkernel_small_table(gpu_small, oracle_small, M):
    x, y, tags = regime_mix(31)
    _check(gpu_small, oracle_small, kqed_amd.KQED_ML2, x, y, M=M, tags=tags)


@pytest.mark.parametrize("which", [0, 2, 3])
def test_regimes_full_table(gpu_full, oracle_full, which):
    x, y, tags = regime_mix(101 + which, n_each=24)
    _check(gpu_full, oracle_full, which, x, y, tags=tags)


def test_sorted_batch_full_table(gpu_full, oracle_full):
    """A batch above the sort threshold (cell-sorted path) on the benchmark's input law."""
    x, y = log_uniform(7, 6000)
    got = gpu_full.batch_host(0, x, y)
    idx = np.random.default_rng(3).choice(len(x), 600, replace=False)
    ref = oracle_full.batch(0, x[idx], y[idx])
    assert normwise(got[idx], ref).max() <= TOL


def test_per_pair_api_matches_batch(gpu_small):
    x, y, _ = regime_mix(5, n_each=3)
    b = gpu_small.batch_host(0, x, y)
    for i in range(0, len(x), 7):
        k = gpu_small.QED_kernel_L0(x[i], y[i]).reshape(-1)
        assert np.array_equal(k, b[i])
    k2 = gpu_small.QED_kernel_L2(x[0], y[0]).reshape(-1)
    assert np.array_equal(k2, gpu_small.batch_host(2, x[:1], y[:1])[0])
    k3 = gpu_small.QED_Mkernel_L2(0.5, x[0], y[0]).reshape(-1)
    assert np.array_equal(k3, gpu_small.batch_host(4, x[:1], y[:1], M=0.5)[0])


def test_soa_layout(gpu_small):
    x, y, _ = regime_mix(8, n_each=5)
    a = gpu_small.batch_host(1, x, y, layout=kqed_amd.KQED_AOS)
    s = gpu_small.batch_host(1, x, y, layout=kqed_amd.KQED_SOA)
    assert np.array_equal(a, s.T)


def test_soa_layout_sorted(gpu_full):
    """SoA output of a cell-sorted batch: every pair lands in its own column."""
    x, y = log_uniform(12, 5000)
    a = gpu_full.batch_host(0, x, y, layout=kqed_amd.KQED_AOS)
    s = gpu_full.batch_host(0, x, y, layout=kqed_amd.KQED_SOA)
    assert np.array_equal(a, s.T)


def test_multi_pass_batch(gpu_full, oracle_full):
    """A batch longer than one pass (passes of 2^19 pairs here): sorted once, walked in passes;
    results equal the unsorted small-batch path pair by pair and the oracle on a sample."""
    n = (1 << 20) + 3001
    x, y = log_uniform(9, n)
    old = gpu_full.get_tune()
    try:
        gpu_full.set_tune(chunk=1 << 19)
        got = gpu_full.batch_host(0, x, y)
    finally:
        gpu_full.set_tune(old)
    rng = np.random.default_rng(4)
    idx = np.concatenate([rng.choice(n, 1500, replace=False), np.arange((1 << 19) - 8, (1 << 19) + 8),
                          np.arange((1 << 20) - 8, (1 << 20) + 8), [n - 1]])
    small = gpu_full.batch_host(0, x[idx], y[idx])
    assert normwise(got[idx], small).max() <= 1e-15
    ref = oracle_full.batch(0, x[idx[:300]], y[idx[:300]])
    assert normwise(got[idx[:300]], ref).max() <= TOL


def test_two_pair_assembly_bit_identical(gpu_full):
    """Sorted one-slot batches assemble the packed and Taylor classes two pairs per wavefront (the
    special pairs keep the one-pair kernel); every output equals the one-pair kernel's on the
    unsorted batch bit for bit: FP64 and FP32, AoS and SoA, odd sizes, special pairs mixed in, one
    pass and several passes of odd length."""
    x, y = log_uniform(21, 5001)
    x[:3] = 0.0
    y[3:6] = 0.0
    x[6:9] = y[6:9]
    old = gpu_full.get_tune()
    try:
        gpu_full.set_tune(sort_min=1 << 30)  # never sorted: every pair through the one-pair kernel
        ref = gpu_full.batch_host(0, x, y)
        ref32 = gpu_full.batch_host(0, x, y, layout=kqed_amd.KQED_AOS_F32)
        gpu_full.set_tune(old)
        got = gpu_full.batch_host(0, x, y)
        got_s = gpu_full.batch_host(0, x, y, layout=kqed_amd.KQED_SOA)
        got32 = gpu_full.batch_host(0, x, y, layout=kqed_amd.KQED_AOS_F32)
        got32_s = gpu_full.batch_host(0, x, y, layout=kqed_amd.KQED_SOA_F32)
        gpu_full.set_tune(chunk=1001)
        got_p = gpu_full.batch_host(0, x, y)
    finally:
        gpu_full.set_tune(old)
    assert np.array_equal(got, ref)
    assert np.array_equal(got_s, ref.T)
    assert np.array_equal(got_p, ref)
    assert got32.dtype == np.float32 and np.array_equal(got32, ref32)
    assert np.array_equal(got32_s, ref32.T)


def test_two_pair_multislot_assembly(gpu_full, oracle_full):
    """Sorted L1, L2, L3 and M-L2 batches assemble the pairs past the special ones two per wavefront
    with every slot (L0(x,x), L0(0,y), L0(x,0), the M-L2 terms).  M-L2 equals the one-pair kernel's
    output on the unsorted batch bit for bit.  L1, L2 and L3 apply each Tabd map once to the
    combined inputs of its slots, which rounds differently: they are held to the FP64 bar against
    the one-pair kernel and the oracle.  FP64 and FP32, AoS and SoA, an odd number of pairs with
    special pairs mixed in, several passes of odd length, and the FP64 per-pair dot of the lepton
    loop."""
    import torch
    x, y = log_uniform(23, 4001)
    x[:3] = 0.0
    y[3:6] = 0.0
    x[6:9] = y[6:9]
    x[9:12] = -y[9:12]
    dev = torch.device("cuda:0")
    xd, yd = torch.from_numpy(x).to(dev), torch.from_numpy(y).to(dev)

    def loop(which, M):
        out = torch.zeros(len(x), dtype=torch.float64, device=dev)
        gpu_full.lepton_loop_device(which, len(x), xd.data_ptr(), yd.data_ptr(), out.data_ptr(), M=M, Mv=1.5,
                                    stream=torch.cuda.current_stream().cuda_stream)
        torch.cuda.synchronize()
        return out.cpu().numpy()

    old = gpu_full.get_tune()
    cases = ((1, 0.0), (2, 0.0), (3, 0.0), (4, 0.4))
    try:
        gpu_full.set_tune(sort_min=1 << 30)
        ref = {w: gpu_full.batch_host(w, x, y, M=M) for w, M in cases}
        ref32 = {w: gpu_full.batch_host(w, x, y, M=M, layout=kqed_amd.KQED_AOS_F32) for w, M in cases}
        refd = {w: loop(w, M) for w, M in cases}
        gpu_full.set_tune(old)
        gpu_full.set_tune(sort_min=0)
        idx = np.random.default_rng(5).choice(len(x), 400, replace=False)
        for w, M in cases:
            print(f"multi-slot two-pair which={w}", flush=True)
            got = gpu_full.batch_host(w, x, y, M=M)
            got32 = gpu_full.batch_host(w, x, y, M=M, layout=kqed_amd.KQED_AOS_F32)
            assert got32.dtype == np.float32
            if w == 4:
                assert np.array_equal(got, ref[w]), w
                assert np.array_equal(gpu_full.batch_host(w, x, y, M=M, layout=kqed_amd.KQED_SOA), ref[w].T), w
                assert np.array_equal(got32, ref32[w]), w
                assert np.array_equal(gpu_full.batch_host(w, x, y, M=M, layout=kqed_amd.KQED_SOA_F32), ref32[w].T), w
                assert np.array_equal(loop(w, M).view(np.int64), refd[w].view(np.int64)), w  # NaN at x = 0 too
                continue
            assert normwise(got, ref[w]).max() <= TOL, w
            assert normwise(got[idx], oracle_full.batch(w, x[idx], y[idx], M=M)).max() <= TOL, w
            assert np.array_equal(gpu_full.batch_host(w, x, y, M=M, layout=kqed_amd.KQED_SOA), got.T), w
            assert normwise(got32.astype(np.float64), got).max() <= 1e-5, w
            assert np.array_equal(gpu_full.batch_host(w, x, y, M=M, layout=kqed_amd.KQED_SOA_F32), got32.T), w
            d, r = loop(w, M), refd[w]
            fin = np.isfinite(r)
            assert np.array_equal(np.isfinite(d), fin), w
            assert np.all(np.abs(d[fin] - r[fin]) <= 1e-11 * np.maximum(np.abs(r[fin]), np.abs(r[fin]).max() * 1e-6)), w
        gpu_full.set_tune(chunk=999)
        for w, M in cases:
            got = gpu_full.batch_host(w, x, y, M=M)
            if w == 4:
                assert np.array_equal(got, ref[w]), w
            else:
                assert normwise(got, ref[w]).max() <= TOL, w
    finally:
        gpu_full.set_tune(old)


def _on_nodes(rng, nodes, other_lo, other_hi, x_side):
    """Pairs whose |x| (x_side) or |y| is each node, and the doubles on either side of it."""
    r = np.concatenate([nodes, np.nextafter(nodes, 0.0), np.nextafter(nodes, np.inf)])
    a = np.zeros((len(r), 4))
    a[:, 0] = r  # |a| = sqrt(r*r) = r exactly
    b = iso(rng, len(r)) * rng.uniform(other_lo, other_hi, (len(r), 1))
    return (a, b) if x_side else (b, a)


@pytest.mark.parametrize("which,M", [(0, 0.0), (1, 0.0), (2, 0.0), (3, 0.0), (4, 0.7)])
def test_grid_nodes_and_neighbours(gpu_full, oracle_full, which, M):
    """Node searches at their boundaries (reference bisection getff-new.c:199-210): |x| on every x
    node, |y| on every y node and on every node of the Taylor table's y row, and the doubles next to
    each, through the sorted and the unsorted path, for L0..L3 and M-L2 (whose subtraction terms put
    the same nodes into the x = 0, y = 0 and x = y branches)."""
    rng = np.random.default_rng(17)
    XX, YY, TXY = gpu_full.XX, gpu_full.YY, oracle_full.TX[0]
    sets = [_on_nodes(rng, XX, 0.05, 3.0, True),
            _on_nodes(rng, YY, 0.05, 3.0, False),
            _on_nodes(rng, YY, 1e-3, XX[0], False),
            _on_nodes(rng, TXY, 1e-3, XX[0], False)]
    x = np.concatenate([s[0] for s in sets])
    y = np.concatenate([s[1] for s in sets])
    ref = oracle_full.batch(which, x, y, M=M)
    got = gpu_full.batch_host(which, x, y, M=M)
    assert len(x) >= 2048  # above the sort threshold
    assert normwise(got, ref).max() <= TOL
    sub = rng.choice(len(x), 1500, replace=False)
    assert normwise(gpu_full.batch_host(which, x[sub], y[sub], M=M), ref[sub]).max() <= TOL


def test_deterministic_and_order_independent(gpu_full):
    """Each pair's result is independent of the batch around it: the same batch twice, and the batch
    in a shuffled order, give identical per-pair results."""
    x, y = log_uniform(13, 30000)
    a = gpu_full.batch_host(0, x, y)
    b = gpu_full.batch_host(0, x, y)
    assert np.array_equal(a, b)
    perm = np.random.default_rng(5).permutation(len(x))
    c = gpu_full.batch_host(0, x[perm], y[perm])
    assert np.array_equal(c, a[perm])


@pytest.mark.gpu
def test_reg_fixtures_on_the_real_table():
    """The reference's golden outputs (PRECOMP/REGexample*.txt, all columns) through the GPU path,
    when the real form-factor table is supplied (KQED_REAL_TABLE_DIR); validated with
    kqed_tables_info first (shape, monotone grids, finite coefficients)."""
    import json
    import os
    real = os.environ.get("KQED_REAL_TABLE_DIR")
    if not real or not os.path.exists(os.path.join(real, "FFxy_single_cksum.bin")):
        pytest.skip("needs the reference's form-factor table FFxy_single_cksum.bin (set KQED_REAL_TABLE_DIR); "
                    "not available here")
    from test_tables import table_info
    rc, info = table_info(real)
    assert rc == 0 and info.x_increasing and info.y_increasing and info.nonfinite == 0
    cases = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "reg_examples.json")))["cases"]
    t = kqed_amd.QEDKernelTemps.initialise(real)
    try:
        for c in cases:
            got = t.QED_kernel_L0(c["x"], c["y"]).reshape(384)
            exp = np.array(c["expected"])
            bad = np.abs(got - exp) > 5.01e-6 * np.abs(exp) + 1e-300
            assert not bad.any(), (c["example"], c["column"], np.flatnonzero(bad)[:5])
    finally:
        t.free()


@pytest.mark.gpu
@pytest.mark.parametrize("which", [0, 2])
def test_longest_chebyshev_rows(which):
    """A synthetic table with nfx up to 126 (Clenshaw lengths up to 128, the limit of the reference's
    GPU port), as a real table may have: every regime against the oracle."""
    from oracle_ctypes import Oracle
    d = kqed_amd.table_dir("wide")
    o = Oracle(d)
    t = kqed_amd.QEDKernelTemps.initialise(d)
    try:
        x, y, _ = regime_mix(17, n_each=24)
        print(f"wide table which={which} n={len(x)}", flush=True)
        got = t.batch_host(which, x, y)
        ref = o.batch(which, x, y)
        assert normwise(got, ref).max() < 1e-12
        # a sorted batch: cell-uniform wavefronts stage these long node arrays in LDS
        x2, y2 = log_uniform(18, 6000)
        print(f"wide table sorted which={which} n={len(x2)}", flush=True)
        got2 = t.batch_host(which, x2, y2)
        idx = np.random.default_rng(19).choice(len(x2), 300, replace=False)
        ref2 = o.batch(which, x2[idx], y2[idx])
        assert normwise(got2[idx], ref2).max() < 1e-12
    finally:
        t.free()
        o.close()
