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
