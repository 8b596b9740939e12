"""Contraction s(y) = sum_x L(x,y) . T(x) on a small lattice: GPU (FP64 accumulation, linear
shortcut for L0..L3, full per-pair contraction for M-L2) vs the CPU oracle.

Bar: |s_gpu - s_oracle| <= 1e-12 * a, a = sum_x sum_c |L(x,y)_c T(x)_c| (the scale of the rounding
error of any summation order; the shortcut and the oracle add the same terms in different orders).
"""
import numpy as np
import pytest

import kqed_amd
from kqed_amd import workloads as W

pytestmark = pytest.mark.gpu

L_TEST = 6  # 1296 sites: x = 0 and x = y sites included


def _lattice():
    x = W.lattice_sites(L_TEST)
    T = W.site_tensors(x, seed=3)
    y = W.y_sites()[:6]
    return x, T, y


@pytest.mark.parametrize("which", [0, 1, 2, 3])
def test_contract_matches_oracle(gpu_full, oracle_full, which):
    x, T, y = _lattice()
    s = gpu_full.contract(which, x, T, y)
    ref, a = oracle_full.contract(which, x, T, y)
    err = np.abs(s - ref) / a
    assert err.max() <= 1e-12, f"which={which} err={err}"


def test_contract_mkernel(gpu_full, oracle_full):
    x, T, y = _lattice()
    y = y[:3]
    s = gpu_full.contract(kqed_amd.KQED_ML2, x, T, y, M=0.5)
    ref, a = oracle_full.contract(kqed_amd.KQED_ML2, x, T, y, M=0.5)
    assert (np.abs(s - ref) / a).max() <= 1e-12


def test_contract_equals_batch_dot(gpu_full):
    """The shortcut equals the explicit kernel contracted on the host (same pairs, GPU kernels)."""
    x, T, y = _lattice()
    y = y[:2]
    s = gpu_full.contract(2, x, T, y)
    for j in range(len(y)):
        K = gpu_full.batch_host(2, x, np.repeat(y[j:j + 1], len(x), axis=0))
        full = (K * T.astype(np.float64)).sum()
        scale = np.abs(K * T.astype(np.float64)).sum()
        assert abs(s[j] - full) <= 1e-12 * scale


def test_contract_deterministic(gpu_full):
    x, T, y = _lattice()
    a = gpu_full.contract(2, x, T, y[:2])
    b = gpu_full.contract(2, x, T, y[:2])
    assert np.array_equal(a, b)


# ------------------------------------------------------------------ through a kernel combination
M4 = np.array([0.0, 0.4, 0.8, 1.2])


@pytest.mark.parametrize("comb,M", [(kqed_amd.KQED_COMB_ALL, None), (kqed_amd.KQED_COMB_CON_M_L2, M4),
                                    (kqed_amd.KQED_COMB_SYMXY_V2, None), (kqed_amd.KQED_COMB_SYMXY0_V2, None)])
def test_contract_through_combination(gpu_full, oracle_full, comb, M):
    """A lattice driver's per-y contraction with symmetrised / connected kernels: every output kernel
    of the combination against the oracle, 4^4 lattice (x = 0 and x = y sites included), 3 y sites."""
    import torch
    x = W.lattice_sites(4)
    T = W.site_tensors(x, seed=5)
    y = W.y_sites()[:3]
    dev = torch.device("cuda:0")
    xd, Td = torch.from_numpy(x).to(dev), torch.from_numpy(T).to(dev)
    print(f"contract comb={comb} nx={len(x)}", flush=True)
    s = gpu_full.contract_comb_device(comb, len(x), xd.data_ptr(), Td.data_ptr(), y, M=M,
                                      stream=torch.cuda.current_stream().cuda_stream)
    ref, a = oracle_full.contract_comb(comb, x, T, y, M=M)
    err = np.abs(s - ref) / a
    assert err.max() <= 1e-12, err


def test_contract_through_combination_multi_pass(gpu_full, oracle_full):
    """More sites than one combination pass; deterministic across calls."""
    import torch
    x = W.lattice_sites(6)
    T = W.site_tensors(x, seed=6)
    y = W.y_sites()[3:5]
    dev = torch.device("cuda:0")
    xd, Td = torch.from_numpy(x).to(dev), torch.from_numpy(T).to(dev)
    old = gpu_full.get_tune()
    try:
        gpu_full.set_tune(chunk=2048)
        st = torch.cuda.current_stream().cuda_stream
        s1 = gpu_full.contract_comb_device(kqed_amd.KQED_COMB_CON_V2, len(x), xd.data_ptr(), Td.data_ptr(), y, stream=st)
        s2 = gpu_full.contract_comb_device(kqed_amd.KQED_COMB_CON_V2, len(x), xd.data_ptr(), Td.data_ptr(), y, stream=st)
    finally:
        gpu_full.set_tune(old)
    assert np.array_equal(s1, s2)
    ref, a = oracle_full.contract_comb(kqed_amd.KQED_COMB_CON_V2, x, T, y)
    assert (np.abs(s1 - ref) / a).max() <= 1e-12


def test_contract_two_pair_bit_identical(gpu_full):
    """A cell-sorted contraction (the two-pair kernels: the shortcut's dot of the tensor entries with
    the site weights for L0..L3, the full per-pair dot of M-L2) gives the same sums, bit for bit, as
    the unsorted one (one-pair kernel): the per-pair dots are identical and the site sums run in site
    order either way."""
    x, T, y = _lattice()
    y = y[:3]
    old = gpu_full.get_tune()
    cases = ((0, 0.0), (1, 0.0), (2, 0.0), (3, 0.0), (kqed_amd.KQED_ML2, 0.5))
    try:
        gpu_full.set_tune(sort_min=1 << 30)
        ref = [gpu_full.contract(w, x, T, y, M=M) for w, M in cases]
        gpu_full.set_tune(old)
        gpu_full.set_tune(sort_min=0)
        got = [gpu_full.contract(w, x, T, y, M=M) for w, M in cases]
    finally:
        gpu_full.set_tune(old)
    for (w, _), a, b in zip(cases, got, ref):
        assert np.array_equal(a, b), w
