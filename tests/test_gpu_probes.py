"""Per-function parity of the FP32 device code with the FP64 oracle, through
the C-ABI probes (monsel_probe_*), on identical inputs and injected uniforms.

Tolerances (FP32 transport against FP64): rates and CSD energy changes 1e-4
relative; scattering cos(theta) 5e-5 absolute; boundary crossings 1e-3 nm
along the segment and the same next region for >= 99.5 % of random
segments (the rest are FP32 rounding at near-tangent or near-corner hits)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def pmma(oracle):
    from cudamonsel_amd import scenes
    from cudamonsel_amd.abi import probe_for
    sc = scenes.flat("PMMA", beam_eV=500.0)
    pr, cs = probe_for(sc)
    return sc, cs, pr, oracle.OracleScene(cs)


@pytest.fixture(scope="module")
def si(oracle):
    from cudamonsel_amd import scenes
    from cudamonsel_amd.abi import probe_for
    sc = scenes.flat("Si", beam_eV=1000.0)
    pr, cs = probe_for(sc)
    return sc, cs, pr, oracle.OracleScene(cs)


def test_rates(pmma, si):
    # through the screened-Rutherford branch of the NIST Mott elements (>= 20 keV)
    E = np.geomspace(0.06, 50000.0, 500).astype(np.float32).astype(np.float64)
    for sc, cs, pr, osc in (pmma, si):
        g = pr.rates(1, E, 3).astype(np.float64)
        o = osc.rates(1, E, 3)
        scale = np.maximum(np.abs(o), 1e-30)
        assert np.all((np.abs(g - o) / scale < 1e-4) | (np.abs(g - o) < 1e-9)), np.max(np.abs(g - o) / scale)


def test_csd(pmma, si):
    rng = np.random.default_rng(1)
    E = np.geomspace(0.6, 19000.0, 500).astype(np.float32).astype(np.float64)
    L = rng.uniform(0.001, 20.0, E.size).astype(np.float32).astype(np.float64)
    for sc, cs, pr, osc in (pmma, si):
        g = pr.csd(1, L, E).astype(np.float64)
        o = osc.csd(1, L, E)
        assert np.all(np.abs(g - o) <= 1e-4 * np.abs(o) + 2e-6 * E), np.max(np.abs(g - o) / np.maximum(np.abs(o), 1e-9))


@pytest.mark.parametrize("slot", [0, 1, 2])
def test_elastic_angles(pmma, slot):
    sc, cs, pr, osc = pmma
    rng = np.random.default_rng(2 + slot)
    n = 20000
    E = np.geomspace(1.0, 19500.0, n).astype(np.float32).astype(np.float64)
    u = rng.random((n, 2)).astype(np.float32).astype(np.float64)
    g = pr.elastic_angle(1, slot, E, u).astype(np.float64)
    o = osc.elastic_angle(1, slot, E, u)
    assert np.max(np.abs(np.cos(g) - np.cos(o))) < 5e-5


@pytest.mark.parametrize("slot", [0, 1, 2])
def test_elastic_angles_screened_rutherford(pmma, slot):
    """NISTMottRS above its tables (>= 20 keV, NISTMottRS.cu:151-153)."""
    sc, cs, pr, osc = pmma
    rng = np.random.default_rng(12 + slot)
    n = 20000
    E = np.geomspace(20100.0, 50000.0, n).astype(np.float32).astype(np.float64)
    u = rng.random((n, 2)).astype(np.float32).astype(np.float64)
    g = pr.elastic_angle(1, slot, E, u).astype(np.float64)
    o = osc.elastic_angle(1, slot, E, u)
    assert np.max(np.abs(np.cos(g) - np.cos(o))) < 5e-5


@pytest.mark.parametrize("compat", [True, False], ids=["compat", "fixed"])
def test_find_end_of_step(oracle, compat):
    from cudamonsel_amd import scenes
    from cudamonsel_amd.abi import probe_for
    sc = scenes.line_on_substrate(h_nm=50, w_nm=20, theta_deg=3, r_nm=5, compat=compat)
    pr, cs = probe_for(sc)
    osc = oracle.OracleScene(cs)
    rng = np.random.default_rng(10)
    n = 20000
    p0 = np.stack([rng.uniform(-25, 25, n), rng.uniform(-30, 30, n), rng.uniform(-60, 10, n)], 1)
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1)[:, None]
    L = rng.uniform(0.5, 40, n)
    p1 = p0 + d * L[:, None]
    p0 = p0.astype(np.float32).astype(np.float64)
    p1 = p1.astype(np.float32).astype(np.float64)
    reg = osc.locate(p0)
    assert np.mean(pr.locate(p0) == reg) > 0.999
    agree = 0
    tot = 0
    for r in np.unique(reg):
        if r < 0:
            continue
        m = reg == r
        ou, onx, _ = osc.find_end_of_step(r, p0[m], p1[m])
        gu, gnx, _ = pr.find_end_of_step(r, p0[m], p1[m])
        same = (onx == gnx)
        hit = (ou <= 1.0)
        # crossing distance apart, for segments that cross (u is infinite otherwise)
        apart = np.zeros(hit.shape)
        apart[hit] = np.abs(ou[hit] - gu[hit].astype(np.float64)) * L[m][hit]
        ok = same & (~hit | (apart < 1e-3))
        agree += ok.sum()
        tot += m.sum()
    # measured: 20,000 of 20,000 in both modes; the allowance is for FP32
    # rounding of grazing crossings against the FP64 oracle
    print(f"findEndOfStep agreement {agree}/{tot}")
    assert tot - agree <= 10, (agree, tot)


@pytest.mark.parametrize("geometry", ["line-compat", "line-fixed", "three-lines", "holes", "step"])
def test_safety_bounds_the_next_boundary(oracle, geometry):
    """The transport kernel skips findEndOfStep for a step shorter than the
    safety bound (monsel_probe_safety).  Pins the bound: no segment from a
    point shorter than its bound reaches a boundary, on the device or in the
    FP64 oracle; in the substrate away from every line the bound is the depth
    (the one surface there); outside a region's own shape it is 0 (the
    containment test opening the next step decides there); and the bound is
    not trivially small."""
    from cudamonsel_amd import scenes
    from cudamonsel_amd.abi import probe_for
    sc = {"line-compat": lambda: scenes.line_on_substrate(compat=True),
          "line-fixed": lambda: scenes.line_on_substrate(compat=False),
          "three-lines": lambda: scenes.random_lines(seed=101),
          "holes": lambda: scenes.contact_holes(radius_nm=12.0, depth_nm=60.0),
          "step": lambda: scenes.step_edge("Si")}[geometry]()
    pr, cs = probe_for(sc)
    osc = oracle.OracleScene(cs)
    rng = np.random.default_rng(7)
    n = 20000
    p0 = np.stack([rng.uniform(-90, 90, n), rng.uniform(-40, 40, n), rng.uniform(-100, 40, n)], 1)
    p0 = p0.astype(np.float32).astype(np.float64)
    reg = osc.locate(p0)
    saf = np.full(n, np.nan)
    for r in np.unique(reg):
        if r >= 0:
            saf[reg == r] = pr.safety(r, p0[reg == r])
    ok = (reg >= 0) & np.isfinite(saf)
    assert ok.mean() > 0.99
    assert np.all(saf[ok] >= 0.0)
    if geometry != "step":
        # points below z = 0 beyond the lines and holes: the bound is the
        # distance to the planes z = 0 (and z = 60 nm, the layer's bottom)
        far = ok & (reg == reg[np.argmax(p0[:, 2])]) & (np.abs(p0[:, 0]) > 80) & (p0[:, 2] > 0.5)
        z = p0[far, 2]
        expect = np.minimum(z, 60.0 - z) if geometry == "holes" else z
        assert far.sum() > 100 and np.allclose(saf[far], expect, rtol=1e-5, atol=1e-4)
    sub = reg[np.argmax(p0[:, 2])]
    vac = ok & (reg == reg[np.argmin(p0[:, 2])]) & (reg != sub) & (p0[:, 2] < -1.0)
    assert vac.sum() > 100 and np.all(pr.safety(sub, p0[vac]) == 0.0)
    k = 4
    sel = ok & (saf > 2e-3) & (saf < 1e6)
    P = np.repeat(p0[sel], k, 0)
    R = np.repeat(reg[sel], k)
    L = np.repeat(0.999 * saf[sel] - 1e-3, k)
    d = rng.normal(size=(P.shape[0], 3))
    d /= np.linalg.norm(d, axis=1)[:, None]
    P1 = (P + d * L[:, None]).astype(np.float32).astype(np.float64)
    for r in np.unique(R):
        m = R == r
        gu, gnx, _ = pr.find_end_of_step(r, P[m], P1[m])
        ou, onx, _ = osc.find_end_of_step(r, P[m], P1[m])
        assert np.all(gu > 1.0) and np.all(gnx == r), (geometry, r, int(np.sum(gu <= 1.0)))
        assert np.all(ou > 1.0), (geometry, r, int(np.sum(ou <= 1.0)))
    assert np.median(saf[ok]) > 1.0


@pytest.mark.parametrize("scale0", [1.0, 1.602e-16], ids=["unit", "joule"])
def test_nu_tables(oracle, tmp_path, M, scale0):
    """Tabulated-inelastic lookups; the joule case puts the first axis in the
    units the inelastic tables use (energies ~1e-16 J)."""
    from cudamonsel_amd import scenes, tables
    from cudamonsel_amd.abi import probe_for
    paths, P = tables.polynomial_tables(str(tmp_path), scale0=scale0)
    for compat in (True, False):
        sc = scenes.Scene()
        ch = scenes._chamber(sc)
        msm = scenes.si_tabulated(sc, paths)
        shp = scenes._half_space(0.0)
        reg = M.Region(ch, msm, shp)
        sc.hold(shp, reg)
        scenes._finish(sc, 1000.0, 0.5, 0.0)
        sc.monte.setRefCompat(M.MONSEL_COMPAT_REFERENCE if compat else M.MONSEL_COMPAT_FIXED)
        pr, cs = probe_for(sc)
        osc = oracle.OracleScene(cs)
        rng = np.random.default_rng(5)
        n = 500
        x = np.stack([rng.uniform(P["g0"][0], P["g0"][-1], n), rng.uniform(P["g1"][0], P["g1"][-1], n),
                      rng.uniform(P["g2"][0], P["g2"][-1], n)], 1).astype(np.float32).astype(np.float64)
        for table, order, dim in ((0, 1, 1), (1, 3, 2), (2, 3, 3)):
            g = pr.nu_table(table, order, x[:, :dim], dim).astype(np.float64)
            o = osc.nu_table(table, order, x[:, :dim], dim)
            scale = np.max(np.abs(o)) + 1.0
            assert np.max(np.abs(g - o)) / scale < 1e-4


def _unit(rng, n):
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1)[:, None]
    return d.astype(np.float32).astype(np.float64)


def _compare_scatter(g, o, rows_ok=0.995, sdir_tol=1e-3):
    """Per-row agreement of two scatter-probe results; rows whose outcome
    differs (FP32 vs FP64 at a choice boundary) are allowed up to 1-rows_ok."""
    same = (g["has"] == o["has"]) & (g["trapped"] == o["trapped"])
    ddir = np.max(np.abs(g["dir"] - o["dir"]), axis=1)
    dE = np.abs(g["E"] - o["E"]) <= 1e-5 * np.abs(o["E"]) + 1e-6
    dsE = np.abs(g["sE"] - o["sE"]) <= 1e-4 * np.abs(o["sE"]) + 1e-5
    dsd = np.max(np.abs(g["sdir"] - o["sdir"]), axis=1) <= sdir_tol
    ok = same & (ddir < 2e-4) & dE & dsE & (dsd | ~o["has"])
    assert ok.mean() >= rows_ok, (ok.mean(), np.max(ddir))


@pytest.mark.parametrize("mech", [0, 1, 2], ids=["elastic", "fitted", "polaron"])
def test_scatter_pmma(pmma, mech):
    sc, cs, pr, osc = pmma
    rng = np.random.default_rng(20 + mech)
    n = 20000
    E = np.geomspace(2.0, 19000.0, n).astype(np.float32).astype(np.float64)
    prev = (E * rng.uniform(1.0, 1.05, n)).astype(np.float32).astype(np.float64)
    d = _unit(rng, n)
    u = rng.random((n, 6)).astype(np.float32).astype(np.float64)
    _compare_scatter(pr.scatter(1, mech, prev, E, d, u), osc.scatter(1, mech, prev, E, d, u))


def test_scatter_si_phonon(si):
    sc, cs, pr, osc = si
    rng = np.random.default_rng(30)
    n = 20000
    E = np.geomspace(0.064, 50.0, n).astype(np.float32).astype(np.float64)
    d = _unit(rng, n)
    u = rng.random((n, 6)).astype(np.float32).astype(np.float64)
    _compare_scatter(pr.scatter(1, 2, E, E, d, u), osc.scatter(1, 2, E, E, d, u))


@pytest.mark.parametrize("method,dispersion", [(1, False), (2, False), (3, False), (1, True), (3, True)],
                         ids=["1", "2", "3", "1-E0-dispersion", "3-E0-dispersion"])
def test_scatter_tabulated(oracle, M, tmp_path, method, dispersion):
    """TabulatedInelasticSM::scatter, with setE0fromDispersion(true) in the
    last two cases (the core-level choice then compares the core energies
    with the plasmon energy of TabulatedInelasticSM.cu:103-151)."""
    from cudamonsel_amd import scenes, tables
    from cudamonsel_amd.abi import probe_for
    paths = tables.synthetic_inelastic_tables(str(tmp_path))
    sc = scenes.Scene()
    ch = scenes._chamber(sc)
    msm = scenes.si_tabulated(sc, paths, method_se=method, e0_from_dispersion=dispersion)
    shp = scenes._half_space(0.0)
    reg = M.Region(ch, msm, shp)
    sc.hold(shp, reg)
    scenes._finish(sc, 1000.0, 0.5, 0.0)
    pr, cs = probe_for(sc)
    osc = oracle.OracleScene(cs)
    rng = np.random.default_rng(40 + method)
    n = 20000
    E = np.geomspace(5.0, 5000.0, n).astype(np.float32).astype(np.float64)
    d = _unit(rng, n)
    u = rng.random((n, 6)).astype(np.float32).astype(np.float64)
    _compare_scatter(pr.scatter(1, 1, E, E, d, u), osc.scatter(1, 1, E, E, d, u), rows_ok=0.99, sdir_tol=2e-3)


@pytest.mark.parametrize("compat,lam", [(False, None), (True, None), (False, 0.0), (False, 0.3e-9)],
                         ids=["fixed", "compat", "sharp-qm", "smooth-qm"])
def test_barrier(oracle, M, compat, lam):
    from cudamonsel_amd import scenes
    from cudamonsel_amd.abi import probe_for
    sc = scenes.flat("PMMA", beam_eV=500.0, compat=compat, barrier_lambda_m=lam)
    pr, cs = probe_for(sc)
    osc = oracle.OracleScene(cs)
    rng = np.random.default_rng(50)
    n = 20000
    d = _unit(rng, n)
    nb = _unit(rng, n)
    E = rng.uniform(0.05, 5.0, n).astype(np.float32).astype(np.float64)
    u = rng.random(n).astype(np.float32).astype(np.float64)
    g = pr.barrier(1, 0, d, nb, E, u)
    o = osc.barrier(1, 0, d, nb, E, u)
    same = (g["region"] == o["region"]) & (g["sign"] == o["sign"])
    close = (np.max(np.abs(g["dir"] - o["dir"]), axis=1) < 2e-4) & (np.abs(g["E"] - o["E"]) < 1e-5 * E + 1e-6)
    assert np.mean(same & close) >= 0.999, np.mean(same & close)


@pytest.mark.parametrize("factory", ["NISTMottRS.Factory100", "NISTMottRS.Factory100Lin",
                                     "BrowningEmpiricalCrossSection.Factory",
                                     "ScreenedRutherfordScatteringAngle.Factory"])
def test_elastic_factories(oracle, M, factory):
    """PMMA's elastic mechanism built with each of the reference's other
    SelectableElasticSM factories (NISTMottRS.cu:16-60: the Mott tables with
    Browning or linear extrapolation below 100 eV instead of 50 eV; Browning
    or screened Rutherford alone): mechanism rates from 0.06 eV to 50 keV and
    every element's angles from 1 eV to 50 keV against the oracle."""
    from cudamonsel_amd import scenes
    from cudamonsel_amd.abi import probe_for
    cls, attr = factory.split(".")
    f = getattr(getattr(M, cls), attr)
    sc = scenes.Scene()
    ch = scenes._chamber(sc)
    msm = scenes.pmma(sc, elastic_factory=f)
    shape = scenes._half_space(0.0)
    reg = M.Region(ch, msm, shape)
    sc.hold(shape, reg)
    scenes._finish(sc, 500.0, 0.5, 0.0)
    pr, cs = probe_for(sc)
    osc = oracle.OracleScene(cs)
    E = np.geomspace(0.06, 50000.0, 500).astype(np.float32).astype(np.float64)
    g = pr.rates(1, E, 3).astype(np.float64)
    o = osc.rates(1, E, 3)
    scale = np.maximum(np.abs(o), 1e-30)
    assert np.all((np.abs(g - o) / scale < 1e-4) | (np.abs(g - o) < 1e-9)), np.max(np.abs(g - o) / scale)
    rng = np.random.default_rng(21)
    n = 20000
    E = np.geomspace(1.0, 50000.0, n).astype(np.float32).astype(np.float64)
    for slot in range(3):
        u = rng.random((n, 2)).astype(np.float32).astype(np.float64)
        ga = pr.elastic_angle(1, slot, E, u).astype(np.float64)
        oa = osc.elastic_angle(1, slot, E, u)
        assert np.max(np.abs(np.cos(ga) - np.cos(oa))) < 5e-5, (slot, np.max(np.abs(np.cos(ga) - np.cos(oa))))
