"""Kernel combinations (the reference's all_kernels.c, con_kernel.c, sub_kernel.c, SYMXY.c, SYMXY0.c).

CPU tests pin the oracle's combination layer against the per-pair kernels L0..L3 and M-L2 (pinned
elsewhere by the reference's golden files and identities): every combination is, by the
reference's own header comments (con_kernel.c:1-10, sub_kernel.c:1-10, SYMXY.c:1-8,
SYMXY0.c:1-15), a signed sum of per-pair kernels at derived arguments with permuted
(mu, nu, lambda).  A wrong index, argument or sign in the restatement fails these.

GPU tests compare the HIP path (kqed_combine) with the oracle on the same pairs.
"""
import numpy as np
import pytest

from pairs import iso

COMB = dict(ALL=0, ALL_M=1, ALL_M_V2=2, CON=3, CON_V2=4, CON_M_L2=5, SUB_M_L2=6, SYMXY=7, SYMXY_V2=8, SYMXY0=9,
            SYMXY0_V2=10)
M4 = np.array([0.0, 0.4, 0.8, 1.2])  # the M set of amu_forlattice.c:1173


def perm(k, order):
    """out[rs, m, n, l] = k[rs, order...]: order names the output index in each slot, e.g. 'nml'."""
    return np.einsum(f"r{order}->rmnl", np.asarray(k).reshape(6, 4, 4, 4)).reshape(384)


def combo_pairs(seed, n=6):
    """Pairs across the regimes the derived arguments reach, plus the special points."""
    rng = np.random.default_rng(seed)
    xs, ys = [], []
    for lo, hi, lo2, hi2 in ((0.1, 2.5, 0.1, 2.5), (1e-3, 0.02, 0.1, 2.0), (0.1, 2.0, 1e-3, 0.02)):
        xs.append(iso(rng, n) * rng.uniform(lo, hi, (n, 1)))
        ys.append(iso(rng, n) * rng.uniform(lo2, hi2, (n, 1)))
    x = iso(rng, n) * rng.uniform(0.2, 2.0, (n, 1))
    xs.append(x)
    ys.append(x + iso(rng, n) * 0.05)                # x - y small: swap regime of (x, x-y)
    v = iso(rng, 2) * 0.7
    z = np.zeros((2, 4))
    xs += [z, v, v, v]
    ys += [v, z, v.copy(), -v]                        # x = 0, y = 0, x = y, x = -y
    return np.concatenate(xs), np.concatenate(ys)


def _check(got, terms, floor, tol=2e-14):
    """|got - sum c*t| within tol of the size of the summands, or of floor (the size of the whole
    combination output) where the summands cancel to exact zero (e.g. L3(x,x) = 0)."""
    ref = sum(c * t for c, t in terms)
    scale = max(sum(abs(c) * np.abs(t) for c, t in terms).max(), floor) + 1e-300
    err = np.abs(got - ref).max() / scale
    assert err < tol, err


def _lk(o, k):
    return [o.L0, o.L1, o.L2, o.L3][k]


@pytest.fixture(scope="module")
def pts():
    return combo_pairs(11)


def test_all_kernels_are_L_of_xy_and_yx(oracle_small, pts):
    o = oracle_small
    for x, y in zip(*pts):
        K = o.combine(COMB["ALL"], x, y)
        F = np.abs(K).max()
        for k in range(4):
            _check(K[2 * k], [(1, _lk(o, k)(x, y))], F)
            _check(K[2 * k + 1], [(1, _lk(o, k)(y, x))], F)


def test_all_mkernels_are_ML2(oracle_small, pts):
    o = oracle_small
    for x, y in zip(*pts):
        K = o.combine(COMB["ALL_M"], x, y, M4)
        K2 = o.combine(COMB["ALL_M_V2"], x, y, M4)
        F = max(np.abs(K).max(), np.abs(K2).max())
        for a in range(4):
            mxy, myx = o.ML2(M4[a], x, y), o.ML2(M4[a], y, x)
            _check(K[2 * a], [(1, mxy)], F)
            _check(K[2 * a + 1], [(1, myx)], F)
            # v2: L(x,y) + L_{nu mu lambda}(y,x) and the (x,-y) kernel (all_kernels.c:286-289)
            _check(K2[2 * a], [(1, mxy), (1, perm(myx, "nml"))], F)
            _check(K2[2 * a + 1], [(1, o.ML2(M4[a], x, -y))], F)


def test_con_kernels_are_A_B_C(oracle_small, pts):
    o = oracle_small
    for x, y in zip(*pts):
        K = o.combine(COMB["CON_V2"], x, y)
        Q = o.combine(COMB["CON"], x, y)
        F = np.abs(K).max()
        for k in range(4):
            f = _lk(o, k)
            A, B, Yx = f(x, y), perm(f(x, x - y), "lnm"), perm(f(y, x), "nml")
            _check(K[k], [(1, A)], F)
            _check(K[4 + k], [(1, B)], F)
            _check(K[8 + k], [(1, A), (1, Yx), (-1, B)], F)
            _check(Q[2 * k], [(1, A), (1, Yx), (-1, B)], F)
            _check(Q[2 * k + 1], [(1, B)], F)


@pytest.mark.parametrize("name,order", [("CON_M_L2", "lnm"), ("SUB_M_L2", "nlm")])
def test_m_l2_combinations(oracle_small, pts, name, order):
    o = oracle_small
    for x, y in zip(*pts):
        K = o.combine(COMB[name], x, y, M4)
        F = np.abs(K).max()
        for a in range(4):
            A = o.ML2(M4[a], x, y)
            B = perm(o.ML2(M4[a], x, x - y), order)
            Yx = perm(o.ML2(M4[a], y, x), "nml")
            _check(K[a], [(1, A)], F)
            _check(K[4 + a], [(1, B)], F)
            _check(K[8 + a], [(1, A), (1, Yx), (-1, B)], F)


def test_symxy(oracle_small, pts):
    o = oracle_small
    for x, y in zip(*pts):
        K = o.combine(COMB["SYMXY_V2"], x, y)
        Q = o.combine(COMB["SYMXY"], x, y)
        F = np.abs(K).max()
        for k in range(4):
            f = _lk(o, k)
            terms = [(0.5, f(x, y)), (0.5, perm(f(y, x), "nml"))]
            _check(K[k], terms, F)
            _check(Q[2 * k], terms, F)
            _check(Q[2 * k + 1], [(0.5, perm(f(x, y), "nml")), (0.5, f(y, x))], F)


def test_symxy0_is_the_six_term_average(oracle_small, pts):
    """SYMXY0.c:1-15: the average over the six argument/index permutations, for every L^k."""
    o = oracle_small
    for x, y in zip(*pts):
        K = o.combine(COMB["SYMXY0_V2"], x, y)
        Q = o.combine(COMB["SYMXY0"], x, y)
        F = np.abs(K).max()
        xmy = x - y
        for k in range(4):
            f = _lk(o, k)
            terms = [(1 / 6, f(x, y)), (1 / 6, perm(f(y, x), "nml")), (-1 / 6, perm(f(xmy, x), "nlm")),
                     (-1 / 6, perm(f(-xmy, y), "mln")), (-1 / 6, perm(f(x, xmy), "lnm")),
                     (-1 / 6, perm(f(y, -xmy), "lmn"))]
            _check(K[k], terms, F)
            _check(Q[2 * k], terms, F)
            _check(Q[2 * k + 1], [(c, perm(t, "nml")) for c, t in terms], F)


def test_oracle_combination_catches_a_planted_transposition(oracle_small, pts):
    """The identity check is sharp: swapping two indices of one term breaks it."""
    o = oracle_small
    x, y = pts[0][0], pts[1][0]
    K = o.combine(COMB["CON_V2"], x, y)
    F = np.abs(K).max()
    A, Yx = o.L0(x, y), perm(o.L0(y, x), "nml")
    wrong_B = perm(o.L0(x, x - y), "nlm")
    with pytest.raises(AssertionError):
        _check(K[8], [(1, A), (1, Yx), (-1, wrong_B)], F)


# ------------------------------------------------------------------ GPU parity


@pytest.mark.gpu
@pytest.mark.parametrize("name", list(COMB))
def test_gpu_combine_matches_oracle(gpu_small, oracle_small, name):
    import kqed_amd
    x, y = combo_pairs(23, n=10)
    M = M4 if name in ("ALL_M", "ALL_M_V2", "CON_M_L2", "SUB_M_L2") else None
    print(f"combine {name} n={len(x)}", flush=True)
    got = gpu_small.combine(getattr(kqed_amd, "KQED_COMB_" + name), x, y, M)
    ref = oracle_small.combine_batch(COMB[name], x, y, M)
    assert got.shape == ref.shape
    scale = np.abs(ref).max(axis=2, keepdims=True) + 1e-300
    err = (np.abs(got - ref) / scale).max()
    assert err < 1e-12, err


@pytest.mark.gpu
def test_gpu_combine_full_table_many_passes(gpu_full, oracle_full):
    """More pairs than one pass of the combination workspace, full-size table, con kernels."""
    import kqed_amd
    from pairs import log_uniform
    t = gpu_full
    old = t.get_tune()
    try:
        t.set_tune(chunk=4096)
        x, y = log_uniform(31, 3000)
        got = t.combine(kqed_amd.KQED_COMB_CON_M_L2, x, y, M4)
    finally:
        t.set_tune(old)
    idx = np.arange(0, 3000, 7)
    ref = oracle_full.combine_batch(COMB["CON_M_L2"], x[idx], y[idx], M4)
    scale = np.abs(ref).max(axis=2, keepdims=True) + 1e-300
    err = (np.abs(got[idx] - ref) / scale).max()
    assert err < 1e-12, err


def test_oracle_contract_comb_is_the_dot_of_the_combination(oracle_small):
    """The oracle's combination contraction equals the dot of its combination outputs with T."""
    rng = np.random.default_rng(3)
    x, _ = combo_pairs(2, n=3)
    T = rng.normal(size=(len(x), 384)).astype(np.float32)
    y = np.array([[0.1, -0.2, 0.3, 0.5], [0.0, 0.0, 0.0, 0.7]])
    s, a = oracle_small.contract_comb(COMB["CON_M_L2"], x, T, y, M=M4)
    for j in range(len(y)):
        K = oracle_small.combine_batch(COMB["CON_M_L2"], x, np.repeat(y[j:j + 1], len(x), axis=0), M4)
        ref = np.einsum("iqc,ic->q", K, T.astype(np.float64))
        assert np.abs(s[j] - ref).max() <= 1e-13 * a[j].max()
