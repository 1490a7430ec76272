"""The benchmark configurations at their full sizes, GPU against the CPU
oracle on identical scenes, job parameters and seeds, and against
themselves where the oracle would take minutes.

- Line scan: one PMMA line (h 50 nm, bottom width 20 nm, 3 deg, top corner
  radius 5 nm) on Si at 0.5 keV, 128 pixels at 1 nm pitch x 2,000 primaries,
  in both reference-compat and fixed modes: per-pixel 3 sigma (4 sigma for
  the maximum over 128 pixels), chi-square of the pixel z-scores, KS on
  exit energy, elevation and azimuth.
- SEM image: three seeded random lines (seed 3), 0.5 keV, 300 primaries per
  pixel: a 32 x 32 sub-raster over the middle line against the oracle, and
  the full 256 x 256 image through conservation properties (every primary
  counted once, no watchdog drops or queue overflow, every escape in every
  histogram and in exactly one pixel's counts).
- Batch: all 512 single-pixel yield jobs of the seeded batch workload run
  together equal the same jobs run one at a time bit for bit; the first 64
  jobs agree with the oracle job by job.
- Serving: a shortened Poisson stream served with continuous batching gives
  every request the tallies it has when run alone; its request geometries
  (random line arrays, contact holes) at 0.3, 1 and 2 keV against the oracle.

Reference: LinesOnLayers.cu:351-380 (scan loop, yields), MonteCarloSS.cu:
222-231 (runMultipleTrajectories), BackscatterStats.cu:140-181 (tallies)."""
import numpy as np
import pytest

import stats_util as su

pytestmark = pytest.mark.gpu

KEYS = ("back", "fwd", "se", "se_sq", "bse_sq", "back_energy", "back_energy_hist", "fwd_energy_hist",
        "elevation_hist", "azimuth_hist")


def _same(a, b):
    for k in KEYS:
        assert np.array_equal(a[k], b[k]), k
    for k in ("primaries", "steps", "electrons", "escapes", "watchdog_drops", "queue_overflow"):
        assert a["stats"][k] == b["stats"][k], k


def _oracle_run(oracle, M, sc, centers, per, E_eV, seed, nbins, cutoff=50.0):
    osc = oracle.OracleScene(sc.monte.compileScene())
    return osc.run(centers, per, M.ToSI.eV(E_eV), sc.gun.getWidth(), seed=seed, nbins=nbins,
                   se_cutoff_eV=cutoff, threads=0)


def _gpu_run(M, sc, centers, per, seed, nbins, cutoff=50.0):
    sc.monte.setSeed(seed)
    return sc.monte.runPixels(centers, per, nbins, cutoff)


def _check_pixels(g, c, per_pixel_max):
    zs, zb = su.total_z(g, c)
    assert abs(zs) < 3.0 and abs(zb) < 3.0, (zs, zb)
    pzs, pzb = su.z_scores(g, c)
    assert np.max(np.abs(pzs)) < per_pixel_max and np.max(np.abs(pzb)) < per_pixel_max, \
        (np.max(np.abs(pzs)), np.max(np.abs(pzb)))
    assert su.chi2_pixels(pzs) > 1e-3 and su.chi2_pixels(pzb) > 1e-3, (su.chi2_pixels(pzs), su.chi2_pixels(pzb))


def _check_hists(r, o):
    names = ("backEnergyHist", "elevationHist", "azimuthHist")
    for name in names:
        D, p = su.ks_binned(getattr(r, name), o[name])
        assert p > 0.01 / len(names), (name, D, p)


@pytest.mark.parametrize("compat", [True, False], ids=["compat", "fixed"])
def test_line_scan_full(oracle, M, compat):
    """128 pixels x 2,000 primaries (256,000) at 0.5 keV."""
    from cudamonsel_amd import scenes
    sc = scenes.line_on_substrate(compat=compat)
    centers = scenes.scan_centers(np.arange(-64.0, 64.0, 1.0), [0.0], sc.beam_z)
    r = _gpu_run(M, sc, centers, 2000, 2, 50)
    o = _oracle_run(oracle, M, sc, centers, 2000, 500.0, 2, 50)
    assert r.stats.primaries == 256000 and o["stats"]["primaries"] == 256000
    assert r.stats.queue_overflow == 0 and r.stats.watchdog_drops == 0 and o["stats"]["watchdog_drops"] == 0
    # 256 per-pixel z-scores: the maximum of |z| stays below 4 with 98 % probability
    _check_pixels(su.from_gpu(r), su.from_oracle(o), 4.0)
    _check_hists(r, o)
    assert r.stats.steps / o["stats"]["steps"] == pytest.approx(1.0, rel=0.01)


@pytest.mark.parametrize("geo_seed", [3, 11])
def test_image_subraster_against_oracle(oracle, M, geo_seed):
    """32 x 32 pixels (1 nm pitch) of the 256 x 256 image over the middle line
    of random_lines(seed) -- seed 3 is the configuration's; seed 11 draws
    other widths, heights and sidewall angles -- 300 primaries per pixel at
    0.5 keV."""
    from cudamonsel_amd import scenes
    sc = scenes.random_lines(seed=geo_seed)
    centers = scenes.scan_centers(np.arange(-16.0, 16.0, 1.0), np.arange(-16.0, 16.0, 1.0), sc.beam_z)
    r = _gpu_run(M, sc, centers, 300, 3, 50)
    o = _oracle_run(oracle, M, sc, centers, 300, 500.0, 3, 50)
    assert r.stats.queue_overflow == 0 and r.stats.watchdog_drops == 0 and o["stats"]["watchdog_drops"] == 0
    # 2,048 per-pixel z-scores: max |z| < 4.5 with 99 % probability
    _check_pixels(su.from_gpu(r), su.from_oracle(o), 4.5)
    _check_hists(r, o)


@pytest.mark.parametrize("geometry", ["lines", "holes"])
@pytest.mark.parametrize("E_eV", [300.0, 1000.0, 2000.0])
def test_request_geometries_against_oracle(oracle, M, geometry, E_eV):
    """The serving stream's request geometries at the ends and middle of its
    0.3-2 keV range: a seeded random-line array (seed 104) and a contact
    hole (radius 14 nm, depth 70 nm), a 4 x 4 raster at 4 nm pitch across
    the middle line or the hole's edge, 500 primaries per pixel."""
    from cudamonsel_amd import scenes
    if geometry == "lines":
        sc = scenes.random_lines(seed=104, beam_eV=E_eV)
    else:
        sc = scenes.contact_holes(radius_nm=14.0, depth_nm=70.0, beam_eV=E_eV)
    centers = scenes.scan_centers(np.arange(-6.0, 10.0, 4.0) + (14.0 if geometry == "holes" else 0.0),
                                  np.arange(-6.0, 10.0, 4.0), sc.beam_z)
    nbins = int(E_eV / 10.0)       # 10 eV bins, as LinesOnLayers.cu:359 sizes them
    r = _gpu_run(M, sc, centers, 500, 41, nbins)
    o = _oracle_run(oracle, M, sc, centers, 500, E_eV, 41, nbins)
    assert r.stats.queue_overflow == 0 and r.stats.watchdog_drops == 0 and o["stats"]["watchdog_drops"] == 0
    assert o["se"].sum() > 0
    _check_pixels(su.from_gpu(r), su.from_oracle(o), 4.0)
    _check_hists(r, o)


def test_image_full_conservation(M):
    """The full 256 x 256 x 300 image (19,660,800 primaries), on the engine's
    default settings (262,144 primaries per launch, geometry gathering chosen
    by the scene) and again with every primary admitted at once and no
    gathering: the same tallies."""
    from cudamonsel_amd import scenes, serving
    from cudamonsel_amd.abi import Engine
    sc = scenes.random_lines(seed=3)
    cen = serving.raster_centers(256, sc.beam_z)
    cs = sc.monte.compileScene()
    with Engine() as eng:
        sid = eng.upload(cs)
        r = eng.result(eng.submit(sid, cen, 300, 500.0, 0.5e-9, 3), timeout_s=100)
    with Engine(primaries_per_launch=1 << 30, geometry_batch=1) as eng:
        sid = eng.upload(cs)
        other = eng.result(eng.submit(sid, cen, 300, 500.0, 0.5e-9, 3), timeout_s=100)
    assert r["stats"]["transport_launches"] > other["stats"]["transport_launches"]
    _same(r, other)
    st = r["stats"]
    assert st["primaries"] == 256 * 256 * 300
    assert st["watchdog_drops"] == 0 and st["queue_overflow"] == 0
    esc = int(st["escapes"])
    assert esc > st["primaries"]   # > 1 electron out per primary at 0.5 keV on PMMA
    assert int(r["back"].sum() + r["fwd"].sum()) == esc
    assert int(r["elevation_hist"].sum()) == esc and int(r["azimuth_hist"].sum()) == esc
    assert int(r["back_energy_hist"].sum() + r["fwd_energy_hist"].sum()) == esc
    assert int(r["back_energy_hist"].sum()) == int(r["back"].sum())
    # the SE count of a pixel is part of its back-escapes; sums of squares
    # are at least the square of the mean count
    assert np.all(r["se"] <= r["back"])
    assert np.all(r["se_sq"] * 300 >= r["se"].astype(np.float64) ** 2 * (1 - 1e-12))
    # every pixel was illuminated: at 300 primaries no pixel lacks escapes
    assert np.all(r["back"] > 0)
    # energy: no back-escape carries more than the beam energy
    assert np.all(r["back_energy"] <= r["back"] * 500.0 * (1 + 1e-6))


def _batch_scene(cache, job):
    from cudamonsel_amd import serving
    return serving._yield_scene(cache, job)


def test_batch_512_jobs_batched_equal_alone(M):
    """All 512 jobs of batch_workload(512, seed=4): 7.6 M primaries."""
    from cudamonsel_amd import serving
    from cudamonsel_amd.abi import Engine
    jobs = serving.batch_workload(512, seed=4)
    with Engine() as eng:
        batched, _ = serving.run_batch(eng, jobs)
        alone, _ = serving.run_batch(eng, jobs, plain=True)
    for j, a, b in zip(jobs, batched, alone):
        assert a["stats"]["primaries"] == j.primaries
        assert a["stats"]["queue_overflow"] == 0 and a["stats"]["watchdog_drops"] == 0
        _same(a, b)


def test_batch_subset_against_oracle(oracle, M):
    """The first 64 jobs of the batch workload, GPU (batched, every job at its
    full size) against the oracle job by job: 128 z-scores (SE and BSE
    yields), max |z| < 4 and chi-square of the z-scores.  The oracle runs
    up to 8,000 primaries of each job, as 16 pixels at the job's beam
    centre (its thread pool is pixel-granular)."""
    from cudamonsel_amd import serving
    from cudamonsel_amd.abi import Engine
    jobs = serving.batch_workload(512, seed=4)[:64]
    with Engine() as eng:
        batched, _ = serving.run_batch(eng, jobs)
        cache = serving.SceneCache(eng)
        zs_all, zb_all = [], []
        for j, g in zip(jobs, batched):
            sc, cs, sid = _batch_scene(cache, j)
            per = max(1, min(j.primaries, 8000) // 16)
            o = _oracle_run(oracle, M, sc, [[0.0, 0.0, sc.beam_z]] * 16, per, j.beam_eV, j.seed, 50)
            gy = su.pixel_yields(g["back"], g["se"], g["se_sq"], g["bse_sq"], j.primaries)
            oy = su.pixel_yields([o["back"].sum()], [o["se"].sum()], [o["seSq"].sum()], [o["bseSq"].sum()],
                                 16 * per)
            zs, zb = su.z_scores(gy, oy)
            zs_all.append(float(zs[0]))
            zb_all.append(float(zb[0]))
    zs_all, zb_all = np.array(zs_all), np.array(zb_all)
    assert np.max(np.abs(zs_all)) < 4.0 and np.max(np.abs(zb_all)) < 4.0, (zs_all, zb_all)
    assert su.chi2_pixels(zs_all) > 1e-3 and su.chi2_pixels(zb_all) > 1e-3


def test_served_requests_equal_alone(M):
    """A Poisson stream (2 requests/s over 6 s, replayed 4x faster, so that
    several requests are in flight at once) of 32- and 64-pixel images with
    50-200 primaries/pixel at 0.3-2 keV on line arrays and contact holes:
    each request served in the mix has the tallies it has alone."""
    from cudamonsel_amd import serving
    from cudamonsel_amd.abi import Engine
    reqs = serving.request_stream(rate_per_s=2.0, duration_s=6.0, seed=5, sides=(32, 64), per_pixel=(50, 200))
    assert len(reqs) >= 6
    with Engine() as eng:
        out = serving.serve(eng, reqs, time_scale=4.0, keep_results=True)
        assert out["max_in_flight"] >= 2
        cache = serving.SceneCache(eng)
        for r in reqs:
            sc, cs, sid = serving._request_scene(cache, r)
            cen = serving.raster_centers(r.side, sc.beam_z)
            alone = eng.result(eng.submit(sid, cen, r.per_pixel, r.beam_eV, 0.5e-9, 5000 + r.rid))
            _same(alone, out["results"][r.rid])
