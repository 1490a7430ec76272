"""Interpolation of the oracle (ULagrangeInterpolation / NULagrangeInterpolation).

Neville's scheme of order n reproduces polynomials of degree <= n exactly, so
cubic polynomials sampled on uniform (U...) and non-uniform (NU...) grids pin
d1..d3.  The reference's d3 defect (NULagrangeInterpolation.cu:236, outer
abscissae taken from the table start) is checked in compat mode against the
value it must then produce."""
import numpy as np
import pytest


def test_u_d1_exact_for_cubics(oracle):
    rng = np.random.default_rng(1)
    x0, h = -1.3, 0.37
    xs = x0 + h * np.arange(20)
    for _ in range(20):
        c = rng.normal(size=4)
        f = c[0] + c[1] * xs + c[2] * xs ** 2 + c[3] * xs ** 3
        for x in rng.uniform(xs[0] - 0.5, xs[-1] + 0.5, 10):
            want = c[0] + c[1] * x + c[2] * x ** 2 + c[3] * x ** 3
            got = oracle.u_d1(f, x0, h, 3, x)
            assert got == pytest.approx(want, rel=1e-9, abs=1e-9)


def test_u_d2_exact_for_bicubics(oracle):
    import ctypes
    rng = np.random.default_rng(2)
    n0, n1 = 11, 9
    x0 = np.array([0.5, -2.0])
    h = np.array([0.25, 0.6])
    g0 = x0[0] + h[0] * np.arange(n0)
    g1 = x0[1] + h[1] * np.arange(n1)
    c = rng.normal(size=(4, 4))
    F = sum(c[i, j] * g0[:, None] ** i * g1[None, :] ** j for i in range(4) for j in range(4))
    F = np.ascontiguousarray(F)
    L = oracle.lib()
    for _ in range(50):
        x = np.array([rng.uniform(g0[0], g0[-1]), rng.uniform(g1[0], g1[-1])])
        want = sum(c[i, j] * x[0] ** i * x[1] ** j for i in range(4) for j in range(4))
        got = L.oracle_u_d2(oracle.dptr(F), n0, n1, oracle.dptr(x0), oracle.dptr(h), 3, oracle.dptr(x))
        assert got == pytest.approx(want, rel=1e-9, abs=1e-9)


@pytest.fixture(scope="module")
def poly_scene(tmp_path_factory, M):
    from cudamonsel_amd import scenes, tables
    d = tmp_path_factory.mktemp("poly")
    paths, P = tables.polynomial_tables(str(d))
    out = {}
    for compat in (True, False):
        sc = scenes.Scene()
        ch = scenes._chamber(sc)
        msm = scenes.si_tabulated(sc, paths)
        shp = scenes._half_space(0.0)
        reg = M.Region(ch, msm, shp)
        sc.hold(shp, reg)
        scenes._finish(sc, 1000.0, 0.5, 0.0)
        sc.monte.setRefCompat(M.MONSEL_COMPAT_REFERENCE if compat else M.MONSEL_COMPAT_FIXED)
        out[compat] = sc
    return out, P


def _scene(oracle, sc):
    cs = sc.monte.compileScene()
    return oracle.OracleScene(cs), cs


def test_nu_d1_d2_d3_exact_for_polynomials(oracle, poly_scene):
    scs, P = poly_scene
    osc, cs = _scene(oracle, scs[False])
    rng = np.random.default_rng(3)
    n = 200
    x = np.stack([rng.uniform(P["g0"][0], P["g0"][-1], n), rng.uniform(P["g1"][0], P["g1"][-1], n),
                  rng.uniform(P["g2"][0], P["g2"][-1], n)], axis=1)
    # tables are added in slot order: 0 IIMFP(1d), 1 RDE(2d), 2 theta(3d), 3 SE(2d)
    got1 = osc.nu_table(0, 1, x[:, :1], 1)
    assert np.allclose(got1, P["p1"](x[:, 0]), rtol=1e-10, atol=1e-10)
    got2 = osc.nu_table(1, 3, x[:, :2], 2)
    assert np.allclose(got2, P["p2"](x[:, 0], x[:, 1]), rtol=1e-9, atol=1e-9)
    got3 = osc.nu_table(2, 3, x, 3)
    assert np.allclose(got3, P["p3"](x[:, 0], x[:, 1], x[:, 2]), rtol=1e-9, atol=1e-9)


def _locate(xs, order, x):
    """First index of the order+1 window (NULagrangeInterpolation.cu:5-68)."""
    n = len(xs)
    lo, up = -1, n
    while up - lo > 1:
        mid = (up + lo) >> 1
        if x > xs[mid]:
            lo = mid
        else:
            up = mid
    if lo < 0:
        return 0
    if up > n - 1:
        return n - 1 - order
    first, last = lo, up
    while last - first < order and first > 0 and last < n - 1:
        if (xs[lo] - xs[first - 1]) + xs[up] - xs[last + 1] >= 0:
            last += 1
        else:
            first -= 1
    if last >= n - 1:
        return n - 1 - order
    return max(first, 0)


def _lagrange(xs, ys, x):
    tot = 0.0
    for i in range(len(xs)):
        w = 1.0
        for j in range(len(xs)):
            if j != i:
                w *= (x - xs[j]) / (xs[i] - xs[j])
        tot += w * ys[i]
    return tot


def test_nu_d3_compat_uses_table_start_abscissae(oracle, poly_scene):
    scs, P = poly_scene
    osc, cs = _scene(oracle, scs[True])
    rng = np.random.default_rng(4)
    g0 = P["g0"]
    for _ in range(50):
        x = np.array([rng.uniform(g0[0], g0[-1]), rng.uniform(P["g1"][0], P["g1"][-1]),
                      rng.uniform(P["g2"][0], P["g2"][-1])])
        i0 = _locate(g0, 3, x[0])
        ys = [P["p3"](g0[i0 + i], x[1], x[2]) for i in range(4)]
        want = _lagrange(g0[:4], ys, x[0])
        got = osc.nu_table(2, 3, x[None, :], 3)[0]
        assert got == pytest.approx(want, rel=1e-8, abs=1e-8)
        if i0 > 0:
            assert abs(got - P["p3"](*x)) > 1e-6 * (1 + abs(got))
