"""Mathematical pins of the oracle's standard, swap and Taylor-x paths (CPU).

The real form-factor table is missing, so no reference output exercises these paths.  They are
pinned instead by what the reference states its code computes:

* chain-rule algebra (chnr_dS.c:50-95, chnr_dT.c:185-293, chnr_dV.c:363-479, swap included): fed
  analytic weight functions, the oracle's S, T and V equal torch autodiff of the definitions at
  chnr_dS.c:47-48, chnr_dT.c:182-183 and chnr_dV.c:361 (tests/autodiff_defs.py);
* interpolation (getff6/getff7, accessv6/7, extractff2, Clenshaw, getff-new.c:14-312,
  simd.c:54-687): on a table whose interpolation is exact (tests/poly_table.py) the interpolated
  scalars equal the closed forms, and S, T, V equal autodiff of the definitions, end to end;
* Taylor-x blend (chnr_dS.c:10-45, chnr_dT.c:16-143, chnr_dV.c:19-174): S, T and the V inputs the
  blend ax = (XX[0] - x)/XX[0] keeps continuous are continuous across x = XX[0].

Each pin is shown to catch a deliberate mistake planted in the oracle (oq_plant).
"""
import numpy as np
import pytest

import autodiff_defs as ad
import poly_table

# normwise per block (S^x, S^y, T^x, T^y, V); measured worst 1.7e-13 (swap: the reference's
# (1 + 1e-14) factor in |x - y|^2, getff-new.c:337, is not in the definitions)
TOL_ALGEBRA = 2e-12
TOL_POLY = 2e-12
TOL_CONT = 1e-5  # jump across XX[0] at x = XX[0](1 -+ 1e-9); measured <= 7e-7


def _pairs(o, n, rng, lo, hi, want):
    """n pairs of each regime in `want` (5 standard, 13 standard with swap), |x|, |y| in [lo, hi]."""
    got = {r: ([], []) for r in want}
    while min(len(v[0]) for v in got.values()) < n:
        a = rng.normal(size=4)
        b = rng.normal(size=4)
        x = a / np.linalg.norm(a) * rng.uniform(lo, hi)
        y = b / np.linalg.norm(b) * rng.uniform(lo, hi)
        r = o.regime(x, y)
        if r in got and len(got[r][0]) < n:
            got[r][0].append(x)
            got[r][1].append(y)
    return {r: (np.array(v[0]), np.array(v[1])) for r, v in got.items()}


@pytest.fixture(scope="module")
def inj_pairs(oracle_small):
    return _pairs(oracle_small, 24, np.random.default_rng(11), 0.1, 4.0, (5, 13))


def _inject_errors(o, pairs):
    worst = 0.0
    for r, (X, Y) in pairs.items():
        sw = r == 13
        W = ad.inject_batch(X, Y, sw)
        R = ad.stv_definition_batch(X, Y, sw)
        for x, y, w, ref in zip(X, Y, W, R):
            rc, got = o.stv_inject(x, y, w)
            assert rc == 0
            worst = max(worst, max(ad.block_errors(got, ref).values()))
    return worst


def test_chain_rule_algebra_matches_autodiff(oracle_small, inj_pairs):
    """S, T, V from injected weight functions = autodiff of the definitions, standard and swap."""
    assert _inject_errors(oracle_small, inj_pairs) <= TOL_ALGEBRA


@pytest.mark.parametrize("plant", [1, 2, 3, 4, 5, 6, 7, 8])
def test_chain_rule_pin_catches_planted_mistakes(oracle_small, inj_pairs, plant):
    oracle_small.plant(plant)
    try:
        err = _inject_errors(oracle_small, inj_pairs)
    finally:
        oracle_small.plant(0)
    assert err > 1e-4, f"planted mistake {plant} not caught ({err:.1e})"


@pytest.fixture(scope="module")
def oracle_poly(tmp_path_factory):
    from oracle_ctypes import Oracle
    o = Oracle(poly_table.table_dir(str(tmp_path_factory.mktemp("tables"))))
    yield o
    o.close()


@pytest.fixture(scope="module")
def poly_pairs(oracle_poly):
    return _pairs(oracle_poly, 24, np.random.default_rng(12), 0.07, 6.0, (5, 13))


def test_polynomial_table_interpolation_is_exact(oracle_poly, poly_pairs):
    """Every interpolated scalar (value and x, y, cb derivatives of the 8 extractff2 calls and the
    d/dy of g3) equals the closed form of the table."""
    for r, (X, Y) in poly_pairs.items():
        E = poly_table.expected_scalars(X, Y, r == 13)
        G = np.array([oracle_poly.scalars(x, y)[1][:37] for x, y in zip(X, Y)])
        scale = np.abs(E).max(axis=0) + 1e-300
        assert (np.abs(G - E) / scale).max() <= TOL_POLY


def _poly_stv_errors(o, pairs):
    W = poly_table.weight_functions()
    worst = 0.0
    for r, (X, Y) in pairs.items():
        R = ad.stv_definition_batch(X, Y, r == 13, W=W)
        for x, y, ref in zip(X, Y, R):
            rc, got = o.stv(x, y)
            assert rc == 0
            worst = max(worst, max(ad.block_errors(got, ref).values()))
    return worst


def test_polynomial_table_end_to_end(oracle_poly, poly_pairs):
    """Table -> interpolation -> chain rule: S, T, V = autodiff of the definitions."""
    assert _poly_stv_errors(oracle_poly, poly_pairs) <= 2e-12


@pytest.mark.parametrize("plant", [11, 12, 13])
def test_interpolation_pin_catches_planted_mistakes(oracle_poly, poly_pairs, plant):
    oracle_poly.plant(plant)
    try:
        err = _poly_stv_errors(oracle_poly, poly_pairs)
    finally:
        oracle_poly.plant(0)
    assert err > 1e-6, f"planted mistake {plant} not caught ({err:.1e})"


# V inputs the Taylor-x blend keeps continuous at x = XX[0] (chnr_dV.c:19-174): every lerp in ax
# whose far end is the standard algebra's value.  dg1dx, dg1dcb, dg2dx, dg2dcb follow the small-x
# form d/dx = cb d2/dxdcb, d/dcb = x d2/dxdcb instead (chnr_dV.c:52-58), which a synthetic table
# does not obey; the g3 terms are not set by the Taylor branch (it hands over g1 directly).
V_CONT = [0, 2, 5, 10, 11, 12, 13, 14, 20, 21, 22, 23, 24]


def _continuity(o, rng, n=60):
    """Worst jump of S, T and the continuous V inputs across x = XX[0] (non-swap for V)."""
    worst = {"S": 0.0, "T": 0.0, "V": 0.0}
    seen = {4: 0, 12: 0}
    tries = 0
    while min(seen.values()) < n and tries < 20000:
        tries += 1
        a = rng.normal(size=4)
        a /= np.linalg.norm(a)
        b = rng.normal(size=4)
        b /= np.linalg.norm(b)
        x0 = o.XX[0]
        yv = b * rng.uniform(0.03, 4.0) if tries % 3 else a * x0 * rng.uniform(0.5, 2.0) + b * x0 * 0.3
        lo, hi = a * x0 * (1 - 1e-9), a * x0 * (1 + 1e-9)
        rl, rh = o.regime(lo, yv), o.regime(hi, yv)
        if rl not in seen or rh != rl + 1 or seen[rl] >= n:
            continue
        (el, sl), (eh, sh) = o.stv(lo, yv), o.stv(hi, yv)
        if el or eh:
            continue
        seen[rl] += 1
        e = ad.block_errors(sl, sh)
        worst["S"] = max(worst["S"], e["Sx"], e["Sy"])
        worst["T"] = max(worst["T"], e["Tx"], e["Ty"])
        if rl == 4:
            vl, vh = o.vterms(lo, yv)[1][V_CONT], o.vterms(hi, yv)[1][V_CONT]
            worst["V"] = max(worst["V"], float(np.max(np.abs(vl - vh) / (np.abs(vh) + 1e-6 * np.abs(vh).max()))))
    assert min(seen.values()) >= n, seen
    return worst


@pytest.mark.parametrize("table", ["small", "poly"])
def test_taylor_x_blend_is_continuous(oracle_small, oracle_poly, table):
    o = oracle_small if table == "small" else oracle_poly
    w = _continuity(o, np.random.default_rng(13))
    assert max(w.values()) <= TOL_CONT, w


@pytest.mark.parametrize("plant,part", [(9, "T"), (10, "S"), (14, "V")])
def test_continuity_pin_catches_planted_mistakes(oracle_poly, plant, part):
    oracle_poly.plant(plant)
    try:
        w = _continuity(oracle_poly, np.random.default_rng(13), n=20)
    finally:
        oracle_poly.plant(0)
    assert w[part] > 1e-3, f"planted mistake {plant} not caught ({w})"
