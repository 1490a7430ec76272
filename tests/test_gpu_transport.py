"""Statistical parity of GPU transport with the CPU oracle on identical scenes,
job parameters and seeds (FP32 device transport against FP64 oracle).

Gates: per-pixel and all-pixel BSE/SE yields within 3 sigma (4 sigma for
per-pixel maxima over many pixels, plus a chi-square test of the pixel
z-scores), exit-energy and exit-angle histograms by two-sample KS at
p > 0.01 for each configuration (the three histograms of a configuration
share that level: each must pass p > 0.01/3, Bonferroni).  Workloads are reduced forms of the benchmark configurations: bulk
Si at 1 keV with 200 energy bins over 0-1 keV (the bulk-Si configuration),
a PMMA line on Si at 0.5 keV (the line-scan configuration) in both
reference-compat and fixed modes, PMMA at 0.5 keV, the tabulated-inelastic
Si path on synthetic tables, a contact hole, Cu, and bulk Si at 25 keV
(the screened-Rutherford elastic branch)."""
import numpy as np
import pytest

import stats_util as su

pytestmark = pytest.mark.gpu


def run_both(oracle, M, sc, E_eV, centers, per, nbins, seed=11, cutoff=50.0):
    sc.monte.setSeed(seed)
    r = sc.monte.runPixels(centers, per, nbins, cutoff)
    cs = sc.monte.compileScene()
    osc = oracle.OracleScene(cs)
    o = osc.run(centers, per, M.ToSI.eV(E_eV), sc.gun.getWidth(), seed=seed, nbins=nbins,
                se_cutoff_eV=cutoff, threads=0)
    assert r.stats.queue_overflow == 0
    assert r.stats.watchdog_drops == 0 and o["stats"]["watchdog_drops"] == 0
    return r, o


def check_yields(r, o, per_pixel_max=4.0):
    g = su.from_gpu(r)
    c = su.from_oracle(o)
    zs, zb = su.total_z(g, c)
    assert abs(zs) < 3.0 and abs(zb) < 3.0, (zs, zb)
    if len(g[0]) > 1:
        pzs, pzb = su.z_scores(g, c)
        assert np.max(np.abs(pzs)) < per_pixel_max and np.max(np.abs(pzb)) < per_pixel_max
        assert su.chi2_pixels(pzs) > 1e-3 and su.chi2_pixels(pzb) > 1e-3
    return g, c


def check_histograms(r, o):
    names = ("backEnergyHist", "elevationHist", "azimuthHist")
    for name in names:
        hg = getattr(r, name)
        ho = o[name]
        D, p = su.ks_binned(hg, ho)
        assert p > 0.01 / len(names), (name, D, p)


def test_flat_pmma_500eV(oracle, M):
    from cudamonsel_amd import scenes
    sc = scenes.flat("PMMA", beam_eV=500.0)
    centers = np.array([[0.0, 0.0, sc.beam_z]] * 16)
    r, o = run_both(oracle, M, sc, 500.0, centers, 1000, 50)
    check_yields(r, o)
    check_histograms(r, o)
    # same number of steps per primary within 2 %
    assert r.stats.steps / o["stats"]["steps"] == pytest.approx(1.0, rel=0.02)


def test_bulk_si_1keV_200_bins(oracle, M):
    """Bulk Si, 1 keV, sigma 0.5 nm, 20000 primaries, 200 bins over 0-1 keV."""
    from cudamonsel_amd import scenes
    sc = scenes.flat("Si", beam_eV=1000.0)
    centers = np.array([[0.0, 0.0, sc.beam_z]] * 16)
    r, o = run_both(oracle, M, sc, 1000.0, centers, 1250, 200)
    check_yields(r, o)
    check_histograms(r, o)


def test_bulk_si_25keV(oracle, M):
    """Bulk Si at 25 keV, 4000 primaries: above 20 keV the elastic
    cross-sections and angles are the screened Rutherford ones
    (ScreenedRutherfordScatteringAngle.cu:24-39), so this runs that branch end
    to end (about 20,000 steps per primary, most of them secondaries)."""
    from cudamonsel_amd import scenes
    sc = scenes.flat("Si", beam_eV=25000.0)
    centers = np.array([[0.0, 0.0, sc.beam_z]] * 16)
    r, o = run_both(oracle, M, sc, 25000.0, centers, 250, 50)
    check_yields(r, o)
    check_histograms(r, o)
    assert r.stats.steps / o["stats"]["steps"] == pytest.approx(1.0, rel=0.02)


@pytest.mark.parametrize("compat", [True, False], ids=["compat", "fixed"])
def test_line_scan_500eV(oracle, M, compat):
    """PMMA line (h 50, w 20 nm, 3 deg, r 5 nm) on Si; pixels across the line."""
    from cudamonsel_amd import scenes
    sc = scenes.line_on_substrate(compat=compat)
    xs = np.arange(-20.0, 20.0, 2.5)
    centers = scenes.scan_centers(xs, [0.0], sc.beam_z)
    r, o = run_both(oracle, M, sc, 500.0, centers, 800, 50)
    check_yields(r, o)


def test_tabulated_si_synthetic_tables(oracle, M, tmp_path):
    from cudamonsel_amd import scenes, tables
    paths = tables.synthetic_inelastic_tables(str(tmp_path))
    sc = scenes.Scene()
    ch = scenes._chamber(sc)
    msm = scenes.si_tabulated(sc, paths)
    shp = scenes._half_space(0.0)
    reg = M.Region(ch, msm, shp)
    sc.hold(shp, reg)
    scenes._finish(sc, 1000.0, 0.5, 0.0)
    centers = np.array([[0.0, 0.0, sc.beam_z]] * 16)
    r, o = run_both(oracle, M, sc, 1000.0, centers, 500, 100)
    check_yields(r, o)
    check_histograms(r, o)


def test_contact_hole(oracle, M):
    from cudamonsel_amd import scenes
    sc = scenes.contact_holes(radius_nm=15.0, depth_nm=60.0)
    centers = scenes.scan_centers(np.linspace(-24, 24, 12), [0.0], sc.beam_z)
    r, o = run_both(oracle, M, sc, 500.0, centers, 600, 50)
    check_yields(r, o)


def test_cu_2keV(oracle, M):
    from cudamonsel_amd import scenes
    sc = scenes.flat("Cu", beam_eV=2000.0)
    centers = np.array([[0.0, 0.0, sc.beam_z]] * 16)
    r, o = run_both(oracle, M, sc, 2000.0, centers, 500, 200)
    check_yields(r, o)
    check_histograms(r, o)


def test_deterministic_and_independent_of_job_mix(M):
    """Integer tallies are identical for the same seed, whether the job runs
    alone or queued with other jobs; FP64 energy sums agree to 1e-12."""
    import ctypes
    from cudamonsel_amd import abi, scenes
    sc = scenes.line_on_substrate()
    centers = scenes.scan_centers(np.linspace(-10, 10, 8), [0.0], sc.beam_z)
    sc.monte.setSeed(77)
    a = sc.monte.runPixels(centers, 300, 50, 50.0)
    sc.monte.setSeed(77)
    b = sc.monte.runPixels(centers, 300, 50, 50.0)
    for k in ("back", "fwd", "se", "backEnergyHist", "elevationHist"):
        assert np.array_equal(getattr(a, k), getattr(b, k)), k
    assert np.allclose(a.backEnergy, b.backEnergy, rtol=1e-12)
    # the same job submitted between two other jobs through the C-ABI
    cs = sc.monte.compileScene()
    dev = M.Device.get(0)
    sid = dev.upload(cs)
    L = M.gpu_lib()
    ctx = ctypes.c_void_p(dev.ctx_address())
    c = np.ascontiguousarray(centers, dtype=np.float64)
    other = np.ascontiguousarray(scenes.scan_centers([30.0, -30.0], [5.0], sc.beam_z))

    def desc(cen, per, seed):
        return abi.JobDesc(M.ToSI.eV(500.0), 0.5e-9, cen.shape[0], per,
                           cen.ctypes.data_as(abi.c_double_p), seed, 50, 0, 50.0, 0)
    jobs = [desc(other, 700, 5), desc(c, 300, 77), desc(other, 400, 6)]
    ids = []
    for j in jobs:
        jid = ctypes.c_uint64()
        assert L.monsel_job_submit(ctx, ctypes.c_uint64(sid), ctypes.byref(j), ctypes.byref(jid)) == 0
        ids.append(jid)
    n = c.shape[0]
    back = np.zeros(n, np.uint64)
    se = np.zeros(n, np.uint64)
    t = abi.Tallies()
    t.back = back.ctypes.data_as(abi.c_u64_p)
    t.se = se.ctypes.data_as(abi.c_u64_p)
    for jid in ids:
        assert L.monsel_job_wait(ctx, jid, ctypes.c_int64(-1)) == 0
    assert L.monsel_job_read(ctx, ids[1], ctypes.byref(t)) == 0
    for jid in ids:
        assert L.monsel_job_release(ctx, jid) == 0
    assert np.array_equal(back, a.back) and np.array_equal(se, a.se)


def test_no_stuck_electrons_at_line_corners(M):
    """Regression: in seeded three-line scenes FP32 electrons in the vacuum
    stopped at the same grazing crossing near a line's top corner step after
    step until the watchdog dropped them (several per million primaries, each
    holding a launch for seconds).  A 64 x 64 raster, 307 primaries/pixel at
    377 eV on random_lines(seed=103) must finish without watchdog drops."""
    from cudamonsel_amd import scenes, serving
    from cudamonsel_amd.abi import Engine
    sc = scenes.random_lines(seed=103)
    cs = sc.monte.compileScene()
    cen = serving.raster_centers(64, sc.beam_z)
    with Engine() as eng:
        sid = eng.upload(cs)
        for seed in (1, 2):
            r = eng.result(eng.submit(sid, cen, 307, 377.0, 0.5e-9, seed, max_steps=200000), timeout_s=60)
            assert r["stats"]["watchdog_drops"] == 0 and r["stats"]["queue_overflow"] == 0


def test_step_watchdog_against_oracle(oracle, M):
    """The per-electron step limit (monsel_job_desc.max_steps) at 40 steps:
    an electron is dropped by the step that takes it past the limit, so no
    electron takes more than 41 steps; the GPU and the oracle drop the same
    fraction of electrons and keep the same yields within their statistics."""
    import math
    from cudamonsel_amd import scenes
    from cudamonsel_amd.abi import Engine
    sc = scenes.flat("PMMA", beam_eV=500.0)
    cen = np.array([[0.0, 0.0, sc.beam_z]] * 16)
    per, lim = 500, 40
    with Engine() as eng:
        r = eng.result(eng.submit(eng.upload(sc.monte.compileScene()), cen, per, 500.0, 0.5e-9, 13, 50,
                                  max_steps=lim))
    osc = oracle.OracleScene(sc.monte.compileScene())
    o = osc.run(cen, per, M.ToSI.eV(500.0), 0.5e-9, seed=13, nbins=50, threads=0, max_steps=lim)
    g, c = r["stats"], o["stats"]
    assert g["watchdog_drops"] > 0 and c["watchdog_drops"] > 0
    assert g["steps"] <= (lim + 1) * g["electrons"]
    # drop fractions of the electrons that took a step (binomial)
    pg, po = g["watchdog_drops"] / g["electrons"], c["watchdog_drops"] / c["electrons"]
    s = math.sqrt(pg * (1 - pg) / g["electrons"] + po * (1 - po) / c["electrons"])
    assert abs(pg - po) < 4 * s, (pg, po, s)
    # SE and BSE yields (BSE = back-escapes that are not SE), with each run's
    # variance from its per-primary second moments
    n = 16 * per

    def yields(back, se, se_sq, bse_sq):
        ys, yb = float(np.sum(se)) / n, float(np.sum(back) - np.sum(se)) / n
        return ys, yb, (float(np.sum(se_sq)) / n - ys * ys) / n, (float(np.sum(bse_sq)) / n - yb * yb) / n

    gs, gb, gvs, gvb = yields(r["back"], r["se"], r["se_sq"], r["bse_sq"])
    os_, ob, ovs, ovb = yields(o["back"], o["se"], o["seSq"], o["bseSq"])
    assert abs(gs - os_) < 3 * math.sqrt(gvs + ovs), (gs, os_)
    assert abs(gb - ob) < 3 * math.sqrt(gvb + ovb), (gb, ob)


def test_min_energy_for_tracking(oracle, M):
    """MONSEL_MaterialScatterModel::setMinEforTracking at 20 eV on PMMA:
    electrons below it end, and secondaries born below it are not tracked
    (MonteCarloSS.cu:165-174).  GPU against the oracle, and fewer SEs than
    without the limit."""
    from cudamonsel_amd import scenes
    sc = scenes.Scene()
    ch = scenes._chamber(sc)
    msm = scenes.pmma(sc, min_e_eV=20.0)
    shape = scenes._half_space(0.0)
    reg = M.Region(ch, msm, shape)
    sc.hold(shape, reg)
    scenes._finish(sc, 500.0, 0.5, 0.0)
    sc.monte.setRefCompat(M.MONSEL_COMPAT_REFERENCE)
    centers = np.array([[0.0, 0.0, sc.beam_z]] * 16)
    r, o = run_both(oracle, M, sc, 500.0, centers, 1000, 50)
    g, _ = check_yields(r, o)
    check_histograms(r, o)
    free = scenes.flat("PMMA", beam_eV=500.0)
    free.monte.setSeed(11)
    f = free.monte.runPixels(centers, 1000, 50, 50.0)
    assert np.sum(r.se) < 0.9 * np.sum(f.se)
