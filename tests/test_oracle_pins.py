"""Pins of the CPU oracle against what the reference fixes.

- the reference's own golden outputs (PRECOMP/REGexample*.txt, 6 significant digits) for every
  column that depends only on the shipped Taylor tables;
- the shipped tables' checksums;
- closed forms of the Chebyshev-U sums and their derivatives;
- the kernel identities the reference's driver checks (amu_forlattice.c examples 3, 9, 11, 15);
- the reference's hand-unrolled assembly (read as text when the reference is present).
"""
import json
import math
import os

import numpy as np
import pytest

from oracle_ctypes import Oracle, assembly_maps

HERE = os.path.dirname(os.path.abspath(__file__))
REG = json.load(open(os.path.join(HERE, "golden", "reg_examples.json")))["cases"]
REF = "/root/reference/src/LIBFILES"


def _taylor_only(c):
    z = [0.0, 0.0, 0.0, 0.0]
    return c["x"] == z or c["y"] == z


@pytest.mark.parametrize("case", [c for c in REG if _taylor_only(c)],
                         ids=lambda c: f"ex{c['example']}b{c['block']}c{c['column']}")
def test_reg_fixture_taylor_columns(oracle_small, case):
    """L0(0,y) (Tabd_xeq0, TX only) and L0(x,0) (Tabd_yeq0, TY on the reference's x grid)."""
    got = oracle_small.L0(case["x"], case["y"])
    exp = np.array(case["expected"])
    # %lg keeps 6 significant digits: |err| <= 5e-6 |value| (+ printing of tiny values)
    bad = np.abs(got - exp) > 5.01e-6 * np.abs(exp) + 1e-300
    assert not bad.any(), (np.flatnonzero(bad)[:5], got[bad][:5], exp[bad][:5])


REAL = os.environ.get("KQED_REAL_TABLE_DIR")  # directory holding the reference's real three table files
NO_REAL = "needs the reference's form-factor table FFxy_single_cksum.bin (set KQED_REAL_TABLE_DIR); it is " \
          "not available here, so these columns stay parity unpinned and the synthetic table is used instead"


@pytest.fixture(scope="module")
def oracle_real():
    if not REAL or not os.path.exists(os.path.join(REAL, "FFxy_single_cksum.bin")):
        pytest.skip(NO_REAL)
    o = Oracle(REAL)
    yield o
    o.close()


@pytest.mark.parametrize("case", [c for c in REG if not _taylor_only(c)],
                         ids=lambda c: f"ex{c['example']}b{c['block']}c{c['column']}")
def test_reg_fixture_needs_real_table(oracle_real, case):
    """REGexample4/5/6/7 and the standard-regime columns of 2 and 10, on the real table."""
    got = oracle_real.L0(case["x"], case["y"])
    exp = np.array(case["expected"])
    bad = np.abs(got - exp) > 5.01e-6 * np.abs(exp) + 1e-300
    assert not bad.any(), (np.flatnonzero(bad)[:5], got[bad][:5], exp[bad][:5])


def test_shipped_taylor_checksums(small_dir):
    import struct
    for name, want in (("taylorx_cksum.bin", 2860436042), ("taylory_cksum.bin", 4020752886)):
        b = open(os.path.join(small_dir, name), "rb").read()
        a, bb = struct.unpack_from("<II", b, len(b) - 8)
        assert a == bb == want


def _U_poly(k):
    """Chebyshev U_k in the power basis, exact integer coefficients."""
    P = [[1], [0, 2]]
    for n in range(2, k + 1):
        a = [0] + [2 * c for c in P[n - 1]]
        b = P[n - 2] + [0, 0]
        P.append([a[i] - b[i] for i in range(len(a))])
    return P[k]


def _exact(order, k, x):
    """d^order/dx^order U_k(x) at the decimal x, in exact rational arithmetic."""
    from fractions import Fraction
    c = _U_poly(k)
    for _ in range(order):
        c = [i * c[i] for i in range(1, len(c))] or [0]
    X = Fraction(x)
    v = Fraction(0)
    for a in reversed(c):
        v = v * X + a
    return v


@pytest.mark.parametrize("order", [0, 1, 2, 3])
def test_clenshaw_closed_form(oracle_small, order):
    """Both Clenshaw forms (cheby.c and the FMA/ratio-table form of simd.c) against exact sums."""
    from fractions import Fraction
    rng = np.random.default_rng(order)
    for nk in (4, 9, 17, 40):
        co = rng.normal(size=nk)
        for x in (-0.83, -0.2, 0.0, 0.37, 0.91):
            terms = [Fraction(float(co[k])) * _exact(order, k, x) for k in range(nk)]
            want = float(sum(terms))
            scale = float(sum(abs(t) for t in terms)) + 1e-300
            assert abs(oracle_small.chebU(order, co, x) - want) <= 4e-15 * nk * scale
            assert abs(oracle_small.chebU_avx(order, co, x) - want) <= 4e-15 * nk * scale


def test_clenshaw_trig_form(oracle_small):
    """U_k(cos t) = sin((k+1)t)/sin t."""
    t = 0.731
    for k in range(12):
        co = np.zeros(k + 1)
        co[k] = 1.0
        assert abs(oracle_small.chebU(0, co, math.cos(t)) - math.sin((k + 1) * t) / math.sin(t)) < 1e-13


def _nz(k, scale):
    return np.abs(k).max() > 0


def test_identities_example3(oracle_small):
    """L0(0,0) = L1(x,x) = L2(0,y) = L2(x,0) = L3(x,x) = L3(0,y) = L3(0,x) = L3(y,y) = 0."""
    o = oracle_small
    z = [0.0] * 4
    x = [0.1, 0.2, 0.3, 0.4]
    y = [0.4, 0.3, 0.2, 0.1]
    for k in (o.L0(z, z), o.L1(x, x), o.L2(z, y), o.L2(x, z), o.L3(x, x), o.L3(z, y), o.L3(z, x), o.L3(y, y)):
        assert np.abs(k).max() <= 1e-14


def test_identity_example9(oracle_small):
    """L(x,x) = -L_{mu<->lambda}(x,0)."""
    for th in np.arange(0.1, 1.0, 0.15):
        x = [th, 2 * th, 3 * th, 4 * th]
        k1 = oracle_small.L0(x, x).reshape(6, 4, 4, 4)
        k2 = oracle_small.L0(x, [0.0] * 4).reshape(6, 4, 4, 4)
        assert np.abs(k1 + k2.transpose(0, 3, 2, 1)).max() <= 1e-14


def _crds(i):
    c = [i % 2, (i // 2) % 2, (i // 4) % 2, i // 8]
    c[3] = c[3] - 4 if c[3] > 2 else c[3]
    return [float(v) for v in c]


def test_identity_example11(oracle_small):
    """L_i(x,y) = -L_i(-x,-y) on the reference's 2x2x2x4 toy lattice, i = 0..3."""
    o = oracle_small
    pts = [_crds(i) for i in range(32)]
    for yv in pts[::3]:
        for xv in pts[::2]:
            mx, my = [-v for v in xv], [-v for v in yv]
            for f in (o.L0, o.L1, o.L2, o.L3):
                assert np.abs(f(xv, yv) + f(mx, my)).max() <= 1e-14


def test_identity_example15(oracle_small):
    """L(x-y,0) = -L(y-x,0) and L(0,x-y) = -L(0,y-x)."""
    o = oracle_small
    z = [0.0] * 4
    rng = np.random.default_rng(15)
    for _ in range(10):
        x, y = rng.uniform(-1, 1, 4), rng.uniform(-1, 1, 4)
        assert np.abs(o.L0(x - y, z) + o.L0(y - x, z)).max() <= 1e-14
        assert np.abs(o.L0(z, x - y) + o.L0(z, y - x)).max() <= 1e-14


def test_mkernel_M0_is_L2(oracle_small):
    rng = np.random.default_rng(4)
    for _ in range(5):
        x, y = rng.uniform(-1, 1, 4), rng.uniform(-1, 1, 4)
        assert np.array_equal(oracle_small.ML2(0.0, x, y), oracle_small.L2(x, y))


def _sym_projector(pairs):
    """Inputs symmetric under swapping tensor slots: project the map onto that subspace."""
    def proj(M, off, sw):
        out = M.copy()
        for c in range(64):
            idx = [c >> 4, (c >> 2) & 3, c & 3]
            j = idx.copy()
            j[sw[0]], j[sw[1]] = idx[sw[1]], idx[sw[0]]
            cj = 16 * j[0] + 4 * j[1] + j[2]
            out[:, off + c] = 0.5 * (M[:, off + c] + M[:, off + cj])
        return out
    return proj


def _linear_form(expr, env):
    """Value of one right-hand side of the reference's unrolled assembly under `env`, by a strict
    parser of the only grammar those lines use (no eval of the reference's text):
        expr := term (('+' | '-') term)*      term := unary ('*' unary)*
        unary := ('+' | '-') unary | atom     atom := NUMBER | NAME | NAME '[' INT ']' | '(' expr ')'
    NAME must be a key of env; anything else raises ValueError."""
    import re
    toks = re.findall(r"[A-Za-z_]\w*|\d+(?:\.\d*)?|[-+*()\[\]]|\S", expr)
    pos = [0]

    def peek():
        return toks[pos[0]] if pos[0] < len(toks) else None

    def take(want=None):
        t = peek()
        if t is None or (want is not None and t != want):
            raise ValueError(f"unexpected {t!r} in {expr!r}")
        pos[0] += 1
        return t

    def atom():
        t = take()
        if t == "(":
            v = parse_expr()
            take(")")
            return v
        if re.fullmatch(r"\d+(?:\.\d*)?", t):
            return float(t)
        if t not in env:
            raise ValueError(f"unknown name {t!r}")
        if peek() == "[":
            take("[")
            k = take()
            if not k.isdigit():
                raise ValueError(f"bad index {k!r}")
            take("]")
            return env[t][int(k)]
        return env[t]

    def unary():
        if peek() in ("+", "-"):
            sign = -1.0 if take() == "-" else 1.0
            return sign * unary()
        return atom()

    def term():
        v = unary()
        while peek() == "*":
            take("*")
            v = v * unary()
        return v

    def parse_expr():
        v = term()
        while peek() in ("+", "-"):
            v = v + term() if take() == "+" else v - term()
        return v

    v = parse_expr()
    if peek() is not None:
        raise ValueError(f"trailing {peek()!r} in {expr!r}")
    return v


def _parse_unrolled(path, fn):
    import re
    src = open(path).read()
    i = src.index(fn)
    body = src[i:src.index("return", i)]
    N = 200
    e = lambda k: np.eye(N)[k]
    env = {"vv": [e(k) for k in range(64)], "txv": [e(64 + k) for k in range(64)],
           "tyv": [e(128 + k) for k in range(64)], "sxv": [e(192 + k) for k in range(4)],
           "syv": [e(196 + k) for k in range(4)]}
    for k in range(4):
        env[f"sxv{k}"] = env["sxv"][k] / 4.0
        env[f"sxyv{k}"] = (env["sxv"][k] + env["syv"][k]) / 4.0
    M = np.zeros((384, N))
    for idx, expr in re.findall(r"kp\[(\d+)\]\s*=\s*(.*?);", body, re.S):
        M[int(idx)] = _linear_form(expr.replace("\n", " "), env)
    return M


@pytest.mark.skipif(not os.path.isdir(REF), reason="reference sources not present (read as text only)")
def test_assembly_matches_reference_text():
    """The Dirac-trace maps equal the reference's hand-unrolled assemblies (QED_kernel.c:16-423,
    QED_kernel_xy0.c:8-399) on inputs with the tensors' index symmetries."""
    Ms, Me = assembly_maps()
    A = _parse_unrolled(os.path.join(REF, "QED_kernel.c"), "CONSTRUCT_FULL_KERNEL")
    B = _parse_unrolled(os.path.join(REF, "QED_kernel_xy0.c"), "CONSTRUCT_KERNEL_FULL_EQZERO")[:, :192]
    P = _sym_projector(None)
    # standard: Txv, Tyv symmetric in (alpha, delta) = slots (0, 2)
    for off in (64, 128):
        A, Ms = P(A, off, (0, 2)), P(Ms, off, (0, 2))
    assert np.abs(A - Ms).max() == 0.0
    # x or y = 0: T^II symmetric in (beta, delta), T^III in (alpha, delta)
    B, Me = P(B, 64, (1, 2)), P(Me, 64, (1, 2))
    B, Me = P(B, 128, (0, 2)), P(Me, 128, (0, 2))
    assert np.abs(B - Me).max() == 0.0


def _sym_fill(t, off, sw):
    """Give the 64 inputs at off the symmetry under swapping tensor slots sw (copy rep onto twin)."""
    for c in range(64):
        idx = [c >> 4, (c >> 2) & 3, c & 3]
        j = idx.copy()
        j[sw[0]], j[sw[1]] = idx[sw[1]], idx[sw[0]]
        cj = 16 * j[0] + 4 * j[1] + j[2]
        if cj < c:
            t[..., off + c] = t[..., off + cj]
    return t


def test_single_map_factorisations():
    """The identities the L1-L3 assembly kernel runs on (one Tabd map application per pair), checked
    on the oracle's maps, which are built from explicit gamma matrices:
    (i) standard map = Tabd map after folding S/T/V into T^I = Vv, T^II_abd = Txv[b][a][d] +
        1/4 d_bd Sxv[a], T^III_abd = Txv + Tyv + 1/4 d_ad (Sxv + Syv)[b];
    (ii) L(v,v) = -L_{mu<->lambda}(v,0) (reference kernels.c:30-48) = the Tabd map on
        (T^I, T^II, T^III)_abd <- (T^I, T^III, T^II)_bad.
    A dropped 1/4 term, a missed transposition or T^II/T^III left unexchanged fails."""
    Ms, Me = assembly_maps()
    rng = np.random.default_rng(5)
    u = rng.normal(size=(16, 200))
    u = _sym_fill(_sym_fill(u, 64, (0, 2)), 128, (0, 2))
    t = np.zeros((16, 192))
    for c in range(64):
        a, b, d = c >> 4, (c >> 2) & 3, c & 3
        t[:, c] = u[:, c]
        t[:, 64 + c] = u[:, 64 + 16 * b + 4 * a + d] + (0.25 * u[:, 192 + a] if b == d else 0.0)
        t[:, 128 + c] = u[:, 64 + c] + u[:, 128 + c] + (0.25 * (u[:, 192 + b] + u[:, 196 + b]) if a == d else 0.0)
    assert np.abs(u @ Ms.T - t @ Me.T).max() < 1e-12
    tt = rng.normal(size=(16, 192))
    tt = _sym_fill(_sym_fill(tt, 64, (1, 2)), 128, (0, 2))
    s = np.zeros_like(tt)
    for c in range(64):
        a, b, d = c >> 4, (c >> 2) & 3, c & 3
        p = 16 * b + 4 * a + d
        s[:, c], s[:, 64 + c], s[:, 128 + c] = tt[:, p], tt[:, 128 + p], tt[:, 64 + p]
    rows = np.array([64 * (i // 64) + 16 * (i & 3) + 4 * ((i >> 2) & 3) + ((i >> 4) & 3) for i in range(384)])
    mulam = -(tt @ Me.T)[:, rows]
    assert np.abs(mulam - s @ Me.T).max() < 1e-12
    # the unexchanged form is not the same map
    s_bad = s.copy()
    s_bad[:, 64:128], s_bad[:, 128:] = s[:, 128:], s[:, 64:128]
    assert np.abs(mulam - s_bad @ Me.T).max() > 1e-3
