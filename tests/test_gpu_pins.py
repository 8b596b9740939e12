"""The HIP path against mathematics, not only against the oracle (GPU).

* Injected weight functions (kqed_batch_inject): the GPU's chain-rule algebra and assembly, fed the
  analytic stand-ins of tests/autodiff_defs.py, equal the oracle fed the same values (<= 1e-12)
  and the assembly map applied to autodiff of the reference's definitions (chnr_dS.c:47-48,
  chnr_dT.c:182-183, chnr_dV.c:361), standard and swap regimes.  Pairs off the packed path in the
  same batch ignore their rows and equal the oracle's table evaluation.
* The polynomial table (tests/poly_table.py): the GPU's interpolation, algebra and assembly equal
  the closed form end to end.
"""
import numpy as np
import pytest
import torch

import autodiff_defs as ad
import kqed_amd
import poly_table
from oracle_ctypes import assembly_maps
from pairs import normwise, regime_mix

pytestmark = pytest.mark.gpu


def _pairs(o, n, seed, lo, hi):
    rng = np.random.default_rng(seed)
    got = {5: ([], []), 13: ([], [])}
    while min(len(v[0]) for v in got.values()) < n:
        a, b = rng.normal(size=4), rng.normal(size=4)
        x = a / np.linalg.norm(a) * rng.uniform(lo, hi)
        y = b / np.linalg.norm(b) * rng.uniform(lo, hi)
        r = o.regime(x, y)
        if r in got and len(got[r][0]) < n:
            got[r][0].append(x)
            got[r][1].append(y)
    return {r: (np.array(v[0]), np.array(v[1])) for r, v in got.items()}


def _definition_L0(X, Y, swap, W):
    Mstd, _ = assembly_maps()
    return ad.stv_definition_batch(X, Y, swap, W=W) @ Mstd.T


def test_injected_weight_functions(gpu_small, oracle_small):
    P = _pairs(oracle_small, 64, 21, 0.1, 4.0)
    X = np.concatenate([P[5][0], P[13][0]])
    Y = np.concatenate([P[5][1], P[13][1]])
    W = np.concatenate([ad.inject_batch(*P[5], False), ad.inject_batch(*P[13], True)])
    ref_math = np.concatenate([_definition_L0(*P[5], False, ad.WFUN), _definition_L0(*P[13], True, ad.WFUN)])
    # pairs off the packed path ride along with rows of garbage they must ignore
    xo, yo, tags = regime_mix(22, n_each=6)
    keep = np.isin(tags, ["taylor", "small2", "x0", "y0", "xy", "00", "outx", "edgey"])
    xo, yo = xo[keep], yo[keep]
    Xa, Ya = np.concatenate([X, xo]), np.concatenate([Y, yo])
    Wa = np.concatenate([W, np.full((len(xo), 34), 1e30)])
    dev = lambda a: torch.from_numpy(np.ascontiguousarray(a, np.float64)).cuda()
    xd, yd, wd = dev(Xa), dev(Ya), dev(Wa)
    out = torch.empty((len(Xa), 384), dtype=torch.float64, device="cuda")
    gpu_small.batch_inject_device(len(Xa), xd.data_ptr(), yd.data_ptr(), wd.data_ptr(), out.data_ptr(),
                                  stream=torch.cuda.current_stream().cuda_stream)
    got = out.cpu().numpy()
    n = len(X)
    ref_or = np.array([oracle_small.L0_inject(x, y, w) for x, y, w in zip(X, Y, W)])
    assert normwise(got[:n], ref_or).max() <= 1e-12
    assert normwise(got[:n], ref_math).max() <= 2e-12
    assert normwise(got[n:], oracle_small.batch(0, xo, yo)).max() <= 1e-12


@pytest.fixture(scope="module")
def poly_dir(tmp_path_factory):
    return poly_table.table_dir(str(tmp_path_factory.mktemp("tables")))


def test_polynomial_table_end_to_end(poly_dir):
    from oracle_ctypes import Oracle
    o = Oracle(poly_dir)
    t = kqed_amd.QEDKernelTemps.initialise(poly_dir)
    try:
        P = _pairs(o, 96, 23, 0.07, 6.0)
        Wf = poly_table.weight_functions()
        for r, (X, Y) in P.items():
            got = t.batch_host(0, X, Y)
            assert normwise(got, _definition_L0(X, Y, r == 13, Wf)).max() <= 2e-12
            assert normwise(got, o.batch(0, X, Y)).max() <= 1e-12
    finally:
        t.free()
        o.close()
