"""Perturbative lepton loop (the reference's pi_pert.c): iPihat_{rho sigma; mu nu lambda}(x, y).

The oracle's restatement is pinned on the CPU by what the mathematics and the literature fix:
  * the Bessel functions against scipy (the reference's polynomial forms are Abramowitz & Stegun
    9.8.5-9.8.8, good to ~1e-7);
  * every eight-gamma trace of G8 against explicit Euclidean gamma matrices;
  * current conservation: d/dx_mu iPihat_{..mu..}(x, y) = 0 and d/dy_nu iPihat_{..nu..}(x, y) = 0
    away from coincident points (finite differences), which a wrong index or sign in the pihat1 sums
    or in the antisymmetrised combination breaks (a planted transposition is shown to fail);
  * the analytic light-by-light results against the published values of the tau- and electron-loop
    contributions to a_mu (Laporta and Remiddi, PLB 301 (1993) 440; CODATA-era values).
GPU tests compare kqed_pihat / pihat1 / kqed_lepton_loop with the oracle.
"""
import ctypes

import numpy as np
import pytest

ALPHA_PI3 = (1 / 137.035999 / np.pi) ** 3


@pytest.fixture(scope="module")
def olib():
    from oracle_ctypes import _lib
    L = _lib()
    v = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
    L.oq_ipihat.argtypes = [v, v, v]
    L.oq_ipihat_batch.argtypes = [ctypes.c_size_t, v, v, v, ctypes.c_int]
    L.oq_pihat1.argtypes = [v, v, v, v]
    L.oq_bessk.argtypes = [ctypes.c_int, ctypes.c_double, v]
    L.oq_diractrace.argtypes = [ctypes.c_int, np.ctypeslib.ndpointer(np.intc, flags="C_CONTIGUOUS")]
    return L


def ipihat(L, x, y):
    o = np.zeros(384)
    L.oq_ipihat(np.ascontiguousarray(x, float), np.ascontiguousarray(y, float), o)
    return o.reshape(6, 4, 4, 4)


def pihat1(L, x, y):
    v = np.zeros(1024)
    r = np.zeros(256)
    L.oq_pihat1(np.ascontiguousarray(x, float), np.ascontiguousarray(y, float), v, r)
    return v.reshape(4, 4, 4, 4, 4), r.reshape(4, 4, 4, 4)


def antisym_from_pihat1(L, x, y, planted=False):
    """The (rho, sigma) combination of pi_pert.c:196-218 in numpy, from the oracle's pihat1."""
    x = np.asarray(x, float)
    v1, _ = pihat1(L, x, y)
    v2, r2 = pihat1(L, y - x, -x)
    v3, r3 = pihat1(L, -x, y - x)
    # P[rho, mu, nu, la, sig] before antisymmetrisation
    t2 = np.einsum("rnlms->rmnls", v2) if not planted else np.einsum("rlnms->rmnls", v2)
    P = (v1 + t2 + np.einsum("r,nlms->rmnls", x, r2) + np.einsum("rlnms->rmnls", v3)
         + np.einsum("r,lnms->rmnls", x, r3))
    out = np.zeros((6, 4, 4, 4))
    rs = 0
    for rho in range(4):
        for sig in range(rho + 1, 4):
            out[rs] = 0.5 * (P[rho, :, :, :, sig] - P[sig, :, :, :, rho])
            rs += 1
    return out


def divergences(f, x, y, h=1e-5):
    """max |d/dx_mu P_{..mu..}| and max |d/dy_nu P_{..nu..}|, each over the largest derivative."""
    dx = np.zeros((6, 4, 4))
    dy = np.zeros((6, 4, 4))
    sx = sy = 0.0
    for m in range(4):
        e = np.zeros(4)
        e[m] = h
        px = (f(x + e, y) - f(x - e, y)) / (2 * h)
        py = (f(x, y + e) - f(x, y - e)) / (2 * h)
        dx += px[:, m, :, :]
        dy += py[:, :, m, :]
        sx = max(sx, np.abs(px).max())
        sy = max(sy, np.abs(py).max())
    return np.abs(dx).max() / sx, np.abs(dy).max() / sy


def points(seed, n):
    rng = np.random.default_rng(seed)
    return rng.normal(size=(n, 4)) * 0.8, rng.normal(size=(n, 4)) * 0.8


def test_bessk_matches_scipy(olib):
    import kqed_amd
    from scipy.special import kv
    for x in (0.01, 0.3, 1.0, 1.99, 2.0, 2.01, 3.7, 8.0):
        ref = np.array([kv(0, x), kv(1, x), kv(2, x), kv(3, x)])
        b = np.zeros(4)
        olib.oq_bessk(3, x, b)
        assert np.abs(b / ref - 1).max() < 5e-7, (x, b / ref - 1)
        p = np.zeros(4)
        kqed_amd.lib().bessk(3, x, p)  # the library's host function
        assert np.array_equal(p, b)


def test_traces_match_gamma_matrices(olib):
    """G8 = 4 * (1/4) Tr of eight Euclidean gammas ({g_mu, g_nu} = 2 delta), all 65536 index sets."""
    import kqed_amd
    s = [np.array([[0, 1], [1, 0]]), np.array([[0, -1j], [1j, 0]]), np.array([[1, 0], [0, -1]])]
    z = np.zeros((2, 2))
    g = [np.block([[z, -1j * sk], [1j * sk, z]]) for sk in s] + [np.block([[z, np.eye(2)], [np.eye(2), z]])]
    g = np.array(g)
    for a in range(4):
        for b in range(4):
            assert np.allclose(g[a] @ g[b] + g[b] @ g[a], 2 * (a == b) * np.eye(4))
    idx = np.array(np.unravel_index(np.arange(65536), (4,) * 8)).T
    prod = np.broadcast_to(np.eye(4, dtype=complex), (65536, 4, 4))
    for k in range(8):
        prod = prod @ g[idx[:, k]]
    ref = np.trace(prod, axis1=1, axis2=2).real  # = 4 * (1/4) Tr
    got = np.array([4 * olib.oq_diractrace(8, np.ascontiguousarray(i, np.intc)) for i in idx[::97]])
    assert np.array_equal(got, np.rint(ref[::97]).astype(int))
    lib_tr = np.array([4 * kqed_amd.lib().diractrace(8, np.ascontiguousarray(i, np.intc)) for i in idx[::97]])
    assert np.array_equal(lib_tr, got)


def test_pihat_conserves_current(olib):
    """d/dx_mu iPihat_{rs;mu nu la} = 0 and d/dy_nu iPihat_{rs;mu nu la} = 0 (free-lepton loop)."""
    X, Y = points(3, 4)
    for x, y in zip(X, Y):
        dxr, dyr = divergences(lambda a, b: ipihat(olib, a, b), x, y)
        assert dxr < 1e-5 and dyr < 1e-5, (dxr, dyr)


def test_antisym_combination_and_planted_transposition(olib):
    X, Y = points(4, 2)
    for x, y in zip(X, Y):
        assert np.allclose(antisym_from_pihat1(olib, x, y), ipihat(olib, x, y), rtol=1e-13, atol=1e-18)
    x, y = X[0], Y[0]
    dxr, dyr = divergences(lambda a, b: antisym_from_pihat1(olib, a, b, planted=True), x, y)
    assert max(dxr, dyr) > 1e-2, (dxr, dyr)


def test_analytic_lbl_against_published_values():
    import kqed_amd
    L = kqed_amd.lib()
    # electron loop: 20.947 924 89(16) (alpha/pi)^3; tau loop: 0.002 142 83(69) (alpha/pi)^3
    assert abs(L.amu_lightlept(0.51099895 / 105.6583755) / ALPHA_PI3 - 20.94792489) < 2e-7
    assert abs(L.amu_heavlept(1776.86 / 105.6583755) / ALPHA_PI3 - 0.00214283) < 7e-7
    # heavy-mass limit: mh^2 a_mu -> (3/2 zeta3 - 19/16) (alpha/pi)^3
    z3 = 1.2020569031595942854
    assert abs(L.amu_heavlept(1e3) * 1e6 / ALPHA_PI3 - (1.5 * z3 - 19 / 16)) < 1e-4


# ------------------------------------------------------------------ GPU parity


@pytest.mark.gpu
def test_gpu_pihat_matches_oracle(gpu_small, olib):
    X, Y = points(7, 64)
    print("pihat n=64", flush=True)
    got = gpu_small.pihat(X, Y)
    ref = np.zeros((64, 384))
    olib.oq_ipihat_batch(64, np.ascontiguousarray(X), np.ascontiguousarray(Y), ref.reshape(-1), 0)
    err = (np.abs(got - ref).max(axis=1) / np.abs(ref).max(axis=1)).max()
    assert err < 1e-12, err


@pytest.mark.gpu
def test_gpu_pihat1_matches_oracle(gpu_small, olib):
    X, Y = points(8, 3)
    for x, y in zip(X, Y):
        v, r = gpu_small.pihat1(x, y)
        vr, rr = pihat1(olib, x, y)
        assert np.abs(v - vr).max() <= 1e-12 * np.abs(vr).max()
        assert np.abs(r - rr).max() <= 1e-12 * np.abs(rr).max()


@pytest.mark.gpu
def test_gpu_lepton_loop_integrand(gpu_small, oracle_small, olib):
    """sum_c L0(x,y)_c iPihat(Mv x, Mv y)_c on the reference's example1 grid (amu_forlattice.c:82-95)
    plus random pairs, and through L2 / M-L2."""
    import torch
    xs, ys = [], []
    for a in np.arange(0.005, 1, 0.2):
        for b in np.arange(0.005, np.pi, np.pi / 4):
            for c in np.arange(0.005, 1, 0.2):
                xs.append([0, 0, a * np.sin(b), a * np.cos(b)])
                ys.append([0, 0, 0, c])
    X, Y = np.array(xs), np.array(ys)
    Xr, Yr = points(9, 40)
    X, Y = np.concatenate([X, Xr]), np.concatenate([Y, Yr])
    Mv = 2.0
    dev = torch.device("cuda:0")
    xd, yd = torch.from_numpy(X).to(dev), torch.from_numpy(Y).to(dev)
    ref_p = np.zeros((len(X), 384))
    olib.oq_ipihat_batch(len(X), np.ascontiguousarray(Mv * X), np.ascontiguousarray(Mv * Y), ref_p.reshape(-1), 0)
    for which, M in ((0, 0.0), (2, 0.0), (4, 0.5)):
        print(f"lepton loop which={which} n={len(X)}", flush=True)
        out = torch.zeros(len(X), dtype=torch.float64, device=dev)
        gpu_small.lepton_loop_device(which, len(X), xd.data_ptr(), yd.data_ptr(), out.data_ptr(), M=M, Mv=Mv,
                                     stream=torch.cuda.current_stream().cuda_stream)
        torch.cuda.synchronize()
        L = oracle_small.batch(which, X, Y, M=M)
        ref = (L * ref_p).sum(axis=1)
        scale = np.abs(L * ref_p).sum(axis=1) + 1e-300
        err = (np.abs(out.cpu().numpy() - ref) / scale).max()
        assert err < 1e-12, (which, err)


@pytest.mark.gpu
def test_gpu_lepton_loop_passes_and_pieces(gpu_small):
    """A batch longer than one lepton-loop pass (2^18 pairs) gives, pair by pair, what separate calls
    on its first and last 3000 pairs give (each pair's iPihat and its dot with L0 are independent of
    the batch around it; every call is above the sort threshold), and kqed_pihat agrees the same way."""
    import torch
    n = (1 << 18) + 3001
    X, Y = points(21, n)
    dev = torch.device("cuda:0")
    st = torch.cuda.current_stream().cuda_stream

    def loop(sel):
        xd = torch.from_numpy(np.ascontiguousarray(X[sel])).to(dev)
        yd = torch.from_numpy(np.ascontiguousarray(Y[sel])).to(dev)
        out = torch.zeros(len(xd), dtype=torch.float64, device=dev)
        gpu_small.lepton_loop_device(0, len(xd), xd.data_ptr(), yd.data_ptr(), out.data_ptr(), Mv=1.5, stream=st)
        p = torch.zeros((len(xd), 384), dtype=torch.float64, device=dev)
        gpu_small.pihat_device(len(xd), xd.data_ptr(), yd.data_ptr(), p.data_ptr(), stream=st)
        torch.cuda.synchronize()
        return out.cpu().numpy(), p.cpu().numpy()

    print(f"lepton loop passes n={n}", flush=True)
    whole, pw = loop(slice(0, n))
    head, ph = loop(slice(0, 3000))
    tail, pt = loop(slice(n - 3000, n))
    assert np.array_equal(whole[:3000], head) and np.array_equal(whole[-3000:], tail)
    assert np.array_equal(pw[:3000], ph) and np.array_equal(pw[-3000:], pt)
