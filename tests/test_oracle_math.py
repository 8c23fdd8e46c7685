"""Pins of the oracle against mathematics (the reference ships no golden vectors).

Each check is chosen so that a plausible mistake -- a dropped term, a sign,
a transposed index, a wrong factorial -- fails it.
"""
import itertools
import math

import numpy as np
import pytest

TRLESS7 = [(a, b, c) for l in range(7) for a in range(l, -1, -1) for b in range(l - a, -1, -1) for c in [l - a - b]
           if c <= 1 or (a, b, c) == (0, 0, 2)]


def test_trless_layout(oc):
    # tensor_trless_sym<T,P>: P^2+1 entries, index (l+m)(l+m+1)/2 + m + P(P+1)/2 [n==1] + P^2 [n==2]
    idx7 = sorted(oc.trless_index(a, b, c, 7) for a, b, c in TRLESS7)
    assert idx7 == list(range(50))
    assert oc.trless_index(0, 0, 0, 7) == 0
    assert oc.trless_index(1, 0, 0, 7) == 1
    assert oc.trless_index(0, 1, 0, 7) == 2
    assert oc.trless_index(0, 0, 1, 7) == 28
    assert oc.trless_index(0, 0, 2, 7) == 49
    t6 = [(a, b, c) for (a, b, c) in TRLESS7 if a + b + c < 6]
    assert sorted(oc.trless_index(a, b, c, 6) for a, b, c in t6) == list(range(37))


@pytest.mark.parametrize("X", [(0.013, -0.021, 0.0071), (-0.2, 0.05, 0.31), (1e-4, 3e-4, -2e-4)])
def test_greens_against_sympy(oc, X):
    sp = pytest.importorskip("sympy")
    x, y, z = sp.symbols("x y z", real=True)
    f = -1 / sp.sqrt(x ** 2 + y ** 2 + z ** 2)
    D = oc.greens(np.array(X))
    Xs = [1000 * v for v in X]
    for a, b, c in TRLESS7:
        expr = sp.diff(f, x, a, y, b, z, c) if (a or b or c) else f
        v = float(expr.subs({x: Xs[0], y: Xs[1], z: Xs[2]}))
        assert D[oc.trless_index(a, b, c, 7)] == pytest.approx(v, rel=1e-12, abs=1e-300)


def test_greens_full_components_against_sympy(oc):
    """Components with two or more z indices come from the trace rule of the
    storage (cosmictiger/tensor.hpp:104-110); they must equal the derivatives."""
    sp = pytest.importorskip("sympy")
    x, y, z = sp.symbols("x y z", real=True)
    f = -1 / sp.sqrt(x ** 2 + y ** 2 + z ** 2)
    X = np.array([0.031, -0.017, 0.022])
    D = oc.greens(X)
    Xs = [1000 * v for v in X]
    full = lambda a, b, c: oc.lib().oc_full(D.ctypes.data_as(oc._D), 7, a, b, c)
    for a, b, c in [(0, 0, 3), (1, 0, 2), (0, 1, 4), (2, 1, 3), (0, 0, 6), (1, 1, 2)]:
        v = float(sp.diff(f, x, a, y, b, z, c).subs({x: Xs[0], y: Xs[1], z: Xs[2]}))
        assert full(a, b, c) == pytest.approx(v, rel=1e-11)
    assert abs(full(2, 0, 0) + full(0, 2, 0) + full(0, 0, 2)) <= 1e-12 * abs(full(2, 0, 0))


def test_gsoft_closed_form(oc):
    # Ferrers n=4: phi(1) = f(1) = 1 (matches the point mass at r = h), f = -2 P'(q^2)
    h = 0.37
    f1, p1 = oc.gsoft(h * h, h)
    assert f1 == pytest.approx(1 / h ** 3, rel=1e-14)
    assert p1 == pytest.approx(1 / h, rel=1e-14)
    assert oc.self_phi() == 693.0 / 256.0
    for q in [0.1, 0.5, 0.9]:
        e = 1e-6
        fp, pp = oc.gsoft((q + e) ** 2 * h * h, h)
        fm, pm = oc.gsoft((q - e) ** 2 * h * h, h)
        f, _ = oc.gsoft(q * q * h * h, h)
        # force per r = -(1/r) dphi/dr
        dphidr = (pp - pm) / (2 * e * h)
        assert f == pytest.approx(-dphidr / (q * h), rel=1e-6)
    # density of the softened potential integrates to unit mass: r^2 f(r) -> 1 at r = h
    # and -Laplacian(phi)/(4 pi) = rho(r) >= 0 inside
    rs = np.linspace(1e-3, 0.999, 200) * h
    rho = []
    for r in rs:
        e = 1e-5 * h
        fa, _ = oc.gsoft((r + e) ** 2, h)
        fb, _ = oc.gsoft((r - e) ** 2, h)
        # enclosed mass m(r) = r^3 f(r); rho = m'(r) / (4 pi r^2)
        dm = ((r + e) ** 3 * fa - (r - e) ** 3 * fb) / (2 * e)
        rho.append(dm / (4 * np.pi * r * r))
    rho = np.array(rho)
    assert np.all(rho >= -1e-9)
    mass = np.trapezoid(4 * np.pi * rs ** 2 * rho, rs)
    assert mass == pytest.approx(1.0, abs=2e-3)


def test_ewald_table_counts(oc):
    real, four = oc.ewald_tables()
    assert len(real) == 178
    assert len(four) == 92
    assert not np.any(np.all(real == 0, axis=1))
    n2 = np.sum(four ** 2, axis=1)
    assert n2.min() >= 1 and n2.max() <= 8
    # the real-image rule: |max(|n| - 1/2, 0)| < 2.6
    d = np.linalg.norm(np.maximum(np.abs(real) - 0.5, 0), axis=1)
    assert d.max() < 2.6
    # sorted by |n|^2 descending
    assert np.all(np.diff(np.sum(real ** 2, axis=1)) <= 0)


def test_cubic_lattice_constant(oc):
    # lim_{r->0} psi(r) - 1/r for the simple cubic lattice with a neutralising
    # background: -2.837297479480620 (Madelung-type constant, alpha-independent)
    assert oc.ewald_psi0() == pytest.approx(-2.8372974794806196, abs=1e-10)


@pytest.mark.parametrize("X", [(0.1, 0.2, -0.3), (0.45, -0.01, 0.2), (1e-3, 2e-3, -1e-3), (0.0, 0.0, 0.0)])
def test_ewald_greens_alpha_independent(oc, X):
    """D_0 and D_1 of the alpha=2 Ewald correction equal the independent alpha=9
    evaluation of the zero-mean periodic potential: D = -(psi - 1/r)."""
    X = np.array(X)
    D = oc.ewald_greens(X)
    r = np.linalg.norm(X)
    if r == 0:
        assert D[0] == pytest.approx(-oc.ewald_psi0() * 1e-3, rel=1e-9)
        assert np.all(np.abs(D[[1, 2, 28]]) < 1e-15)
        return
    psi, g = oc.ewald_psi(X)
    assert D[0] == pytest.approx(-(psi - 1 / r) * 1e-3, rel=1e-9)
    grad = -(g + X / r ** 3) * 1e-6
    assert D[1] == pytest.approx(grad[0], rel=1e-8, abs=1e-14)
    assert D[2] == pytest.approx(grad[1], rel=1e-8, abs=1e-14)
    assert D[28] == pytest.approx(grad[2], rel=1e-8, abs=1e-14)


def test_ewald_greens_higher_derivatives(oc):
    """Degree-2 derivatives by finite differences of the degree-1 ones, and the
    degree-2 trace equal to that of the periodic potential: Laplacian
    (psi - 1/r) = 4 pi (uniform background), so trace D_2 = -4 pi 1e-9."""
    X = np.array([0.12, -0.07, 0.21])
    D = oc.ewald_greens(X)
    e = 1e-5
    for (i, name) in enumerate([(2, 0, 0), (0, 2, 0)]):
        Xp = X.copy()
        Xp[i] += e
        Xm = X.copy()
        Xm[i] -= e
        d1 = (oc.ewald_greens(Xp)[1 + i] - oc.ewald_greens(Xm)[1 + i]) / (2 * e * 1000)
        assert D[oc.trless_index(*name, 7)] == pytest.approx(d1, rel=1e-6)
    tr = D[oc.trless_index(2, 0, 0, 7)] + D[oc.trless_index(0, 2, 0, 7)] + D[oc.trless_index(0, 0, 2, 7)]
    assert tr == pytest.approx(-4 * np.pi * 1e-9, rel=1e-6)


def _cluster(rng, c, r, n):
    return c + rng.uniform(-1, 1, (n, 3)) * r / np.sqrt(3)


def test_p2m_m2l_l2p_far_field(oc):
    """The multipole chain reproduces the direct 1/r sum to expansion accuracy."""
    rng = np.random.default_rng(1)
    cs = np.array([0.1, 0.2, 0.3])
    ct = cs + np.array([0.05, -0.03, 0.04])
    src = _cluster(rng, cs, 0.005, 20)
    snk = _cluster(rng, ct, 0.005, 5)
    M = sum(oc.p2m(s - cs) for s in src)
    L = oc.m2l(M, oc.greens(ct - cs))
    for t in snk:
        L4 = oc.l2p(L, t - ct)
        d = t - src
        r = np.linalg.norm(d, axis=1)
        assert 1e3 * L4[0] == pytest.approx(-np.sum(1 / r), rel=1e-8)
        g = -1e6 * L4[1:]
        gd = -np.sum(d / r[:, None] ** 3, axis=0)
        assert np.linalg.norm(g - gd) <= 1e-6 * np.linalg.norm(gd)


def test_m2l_error_decreases_with_distance(oc):
    rng = np.random.default_rng(2)
    cs = np.array([0.3, 0.3, 0.3])
    src = _cluster(rng, cs, 0.02, 30)
    M = sum(oc.p2m(s - cs) for s in src)
    errs = []
    for dist in [0.08, 0.16, 0.32]:
        ct = cs + np.array([dist, 0, 0])
        L = oc.m2l(M, oc.greens(ct - cs))
        L4 = oc.l2p(L, np.zeros(3))
        exact = -np.sum(1 / np.linalg.norm(ct - src, axis=1))
        errs.append(abs(1e3 * L4[0] - exact) / abs(exact))
    # order-6 truncation: halving the ratio cuts the error by ~2^6
    assert errs[1] < errs[0] / 20 and errs[2] < errs[1] / 20


def test_m2m_translation_identity(oc):
    rng = np.random.default_rng(3)
    c1 = np.array([0.4, 0.5, 0.6])
    c2 = c1 + np.array([0.002, -0.001, 0.003])
    src = _cluster(rng, c1, 0.004, 25)
    M1 = sum(oc.p2m(s - c1) for s in src)
    M2 = sum(oc.p2m(s - c2) for s in src)
    assert np.allclose(oc.m2m(M1, c1 - c2), M2, rtol=1e-10, atol=1e-12 * np.max(np.abs(M2)))


def test_l2l_translation_identity(oc):
    rng = np.random.default_rng(4)
    L = oc.m2l(oc.p2m(np.array([1e-3, -2e-3, 5e-4])), oc.greens(np.array([0.05, 0.02, -0.04])))
    d = np.array([1e-3, 2e-3, -1e-3])
    Lb = oc.l2l(L, d)
    for _ in range(4):
        y = rng.uniform(-2e-3, 2e-3, 3)
        assert np.allclose(oc.l2p(Lb, y), oc.l2p(L, y + d), rtol=1e-12, atol=1e-15 * np.max(np.abs(L)))


def test_p2m_traceless_above_degree_two(oc):
    M = oc.p2m(np.array([3e-3, -1e-3, 2e-3]))
    full = lambda a, b, c: oc.lib().oc_full(M.ctypes.data_as(oc._D), 6, a, b, c)
    y = -1000 * np.array([3e-3, -1e-3, 2e-3])
    # degree 2 keeps the trace: full quadrupole of y
    assert full(2, 0, 0) == pytest.approx(y[0] ** 2)
    assert full(0, 0, 2) == pytest.approx(y[2] ** 2)
    # degree 3: traceless projection, e.g. (3,0,0) = y0^3 - 3/5 r^2 y0
    r2 = np.sum(y ** 2)
    assert full(3, 0, 0) == pytest.approx(y[0] ** 3 - 0.6 * r2 * y[0], rel=1e-12)
    for n in [(1, 0, 0), (0, 1, 1), (1, 1, 1), (2, 0, 1)]:
        if sum(n) + 2 >= 6:
            continue
        tr = full(n[0] + 2, n[1], n[2]) + full(n[0], n[1] + 2, n[2]) + full(n[0], n[1], n[2] + 2)
        assert abs(tr) <= 1e-10 * r2 ** ((sum(n) + 2) / 2)


def test_fixed32_distance_and_conversion(oc):
    assert oc.distance(0, 0) == 0.0
    assert oc.distance(1, 0) == pytest.approx(2.0 ** -32)
    assert oc.distance(0, 1) == pytest.approx(-(2.0 ** -32))
    assert oc.distance(0, 2 ** 32 - 1) == pytest.approx(2.0 ** -32)  # wraps across 0/1
    assert oc.distance(2 ** 31, 0) == -0.5  # int32 wrap: -2^31
    assert oc.distance(0, 2 ** 31) == -0.5
    assert oc.distance(2 ** 31 - 1, 0) == pytest.approx(0.5, abs=2e-10)
    assert oc.fixed32_from_double(0.5) == 2 ** 31
    assert oc.fixed32_from_double(0.999999999999) == 2 ** 32 - 1  # truncation
    assert oc.fixed32_from_double(1e-12) == 0


def test_drift_closed_form(oc):
    x = np.array([0, 2 ** 31, 2 ** 32 - 1, 100], dtype=np.uint32)
    y = x.copy()
    z = x.copy()
    vel = np.array([[0.25, -0.25, 0], [0.25, 0, 0], [1e-9, 0, 0], [0, 0, 0]], np.float32)
    rung = np.array([0, 0, 0, 1], np.int8)
    ox, oy, oz = oc.drift(x, y, z, vel, rung, 0, 1.0, 1.0)
    assert ox[0] == 2 ** 30  # 0 + 0.25
    assert oy[0] == 3 * 2 ** 30  # 0 - 0.25 wraps to 0.75
    assert ox[1] == 2 ** 31 + 2 ** 30
    # 1 - 2^-32 + 1e-9 wraps past 1
    expect = int(np.floor((((2 ** 32 - 1) / 2 ** 32 + float(np.float32(1e-9))) - 1.0) * 2 ** 32))
    assert ox[2] == expect
    assert ox[3] == 100  # other rung untouched
    # a = 2 halves the displacement (c1 = dt / a)
    ox2, _, _ = oc.drift(x, y, z, vel, rung, 0, 2.0, 1.0)
    assert ox2[0] == 2 ** 29
