"""The C-ABI's edge cases and error behaviour (include/monsel_gpu.h): every
malformed job descriptor is refused with MONSEL_E_INVALID before it is
queued, unknown ids give MONSEL_E_NOTFOUND, a running job reads as
MONSEL_E_NOTREADY and a zero timeout as MONSEL_E_TIMEOUT, scenes beyond the
kernels' limits (convex parts per shape, planes per multi-plane shape) are
refused at upload with MONSEL_E_UNSUPPORTED, and the
smallest and the largest histogram shapes give consistent tallies (the
histograms go through LDS when they fit and global atomics when they do
not)."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

OK, INVALID, UNSUPPORTED, NOTFOUND, TIMEOUT, NOTREADY = 0, -1, -2, -5, -6, -7


def _desc(centers, per=100, E_eV=500.0, sigma=0.5e-9, seed=1, nbins=50, log_detected=0):
    from cudamonsel_amd import abi
    c = np.ascontiguousarray(centers, dtype=np.float64).reshape(-1, 3)
    jd = abi.JobDesc(E_eV * 1.60217653e-19, sigma, c.shape[0], per, c.ctypes.data_as(abi.c_double_p), seed, nbins,
                     0, 50.0, 0, log_detected)
    return jd, c


def _submit(eng, sid, jd):
    jid = ctypes.c_uint64()
    rc = eng.L.monsel_job_submit(eng.ctx, ctypes.c_uint64(sid), ctypes.byref(jd), ctypes.byref(jid))
    return rc, jid.value


@pytest.fixture(scope="module")
def flat_pmma(M):
    from cudamonsel_amd import scenes
    from cudamonsel_amd.abi import Engine
    sc = scenes.flat("PMMA", beam_eV=500.0)
    eng = Engine()
    sid = eng.upload(sc.monte.compileScene())
    yield eng, sid, sc
    eng.close()


def test_malformed_job_descriptors_are_refused(flat_pmma):
    from cudamonsel_amd import abi
    eng, sid, sc = flat_pmma
    cen = [[0.0, 0.0, sc.beam_z]]
    cases = []
    jd, keep = _desc(cen)
    jd.n_pixels = 0
    cases.append(("no pixels", jd, keep))
    jd, keep = _desc(cen)
    jd.pixel_center = ctypes.cast(None, abi.c_double_p)
    cases.append(("NULL centres", jd, keep))
    jd, keep = _desc(cen, per=0)
    cases.append(("no primaries", jd, keep))
    jd, keep = _desc(cen, E_eV=0.0)
    cases.append(("zero energy", jd, keep))
    jd, keep = _desc(cen, sigma=-1e-9)
    cases.append(("negative width", jd, keep))
    jd, keep = _desc(np.repeat(cen, 3, axis=0), per=(1 << 31) - 1)
    cases.append(("over 2^32 primaries", jd, keep))
    jd, keep = _desc(cen, nbins=(1 << 20) + 1)
    cases.append(("too many bins", jd, keep))
    jd, keep = _desc(cen, log_detected=-1)
    cases.append(("negative log size", jd, keep))
    for name, jd, keep in cases:
        rc, _ = _submit(eng, sid, jd)
        assert rc == INVALID, name
    # the largest accepted primary count is 2^32 - 1 over all pixels
    jd, keep = _desc(np.repeat(cen, 2, axis=0), per=(1 << 31) - 1)
    assert jd.n_pixels * jd.primaries_per_pixel == (1 << 32) - 2
    jd.primary_offset = 2
    rc, _ = _submit(eng, sid, jd)
    assert rc == INVALID
    # nothing refused reached the queue: a valid job still runs at once
    r = eng.result(eng.submit(sid, cen, 10, 500.0, 0.5e-9, 3))
    assert r["stats"]["primaries"] == 10


def test_unknown_ids(flat_pmma):
    eng, sid, sc = flat_pmma
    jd, keep = _desc([[0.0, 0.0, sc.beam_z]])
    rc, _ = _submit(eng, sid + 1000, jd)
    assert rc == NOTFOUND
    bad = ctypes.c_uint64(1 << 60)
    done = ctypes.c_int()
    assert eng.L.monsel_job_poll(eng.ctx, bad, ctypes.byref(done)) == NOTFOUND
    assert eng.L.monsel_job_wait(eng.ctx, bad, ctypes.c_int64(0)) == NOTFOUND
    from cudamonsel_amd import abi
    t = abi.Tallies()
    assert eng.L.monsel_job_read(eng.ctx, bad, ctypes.byref(t)) == NOTFOUND
    assert eng.L.monsel_job_release(eng.ctx, bad) == NOTFOUND


def test_running_job_reads_not_ready(flat_pmma):
    """A 2.56 M-primary job is read and waited on with a zero timeout right
    after its submission (it takes about 0.1 s on the device)."""
    from cudamonsel_amd import abi, scenes
    eng, sid, sc = flat_pmma
    cen = scenes.scan_centers(np.arange(-64.0, 64.0, 1.0), [0.0], sc.beam_z)
    jid = eng.submit(sid, cen, 20000, 500.0, 0.5e-9, 5)
    t = abi.Tallies()
    rc_read = eng.L.monsel_job_read(eng.ctx, ctypes.c_uint64(jid), ctypes.byref(t))
    rc_wait = eng.L.monsel_job_wait(eng.ctx, ctypes.c_uint64(jid), ctypes.c_int64(0))
    done = ctypes.c_int(1)
    assert eng.L.monsel_job_poll(eng.ctx, ctypes.c_uint64(jid), ctypes.byref(done)) == OK
    assert rc_read == NOTREADY and rc_wait == TIMEOUT and done.value == 0
    r = eng.result(jid)
    assert r["stats"]["primaries"] == 128 * 20000 and r["stats"]["queue_overflow"] == 0
    # a completed job cannot be cancelled into failure
    jid = eng.submit(sid, cen[:1], 10, 500.0, 0.5e-9, 6)
    assert eng.L.monsel_job_wait(eng.ctx, ctypes.c_uint64(jid), ctypes.c_int64(-1)) == OK
    assert eng.L.monsel_job_cancel(eng.ctx, ctypes.c_uint64(jid)) == OK
    assert eng.result(jid)["stats"]["primaries"] == 10


@pytest.mark.parametrize("nbins", [1, 50, 20000])
def test_histogram_shapes(flat_pmma, nbins):
    """One primary up to 20,000 energy bins (histograms beyond the LDS
    budget go to global atomics): every back-scattered electron is in one
    energy bin, and every escape in one elevation and one azimuth bin."""
    eng, sid, sc = flat_pmma
    cen = [[0.0, 0.0, sc.beam_z]]
    one = eng.result(eng.submit(sid, cen, 1, 500.0, 0.5e-9, 7, nbins=nbins))
    assert one["stats"]["primaries"] == 1 and one["stats"]["steps"] > 0
    r = eng.result(eng.submit(sid, cen, 3000, 500.0, 0.5e-9, 8, nbins=nbins))
    assert len(r["back_energy_hist"]) == nbins + 2 and len(r["fwd_energy_hist"]) == nbins + 2
    back = int(r["back"][0])
    assert int(np.sum(r["back_energy_hist"])) == back
    assert int(np.sum(r["elevation_hist"])) == r["stats"]["escapes"]
    assert int(np.sum(r["azimuth_hist"])) == r["stats"]["escapes"]
    assert back > 0


def _upload_status(eng, sc):
    cs = sc.monte.compileScene()   # owns the descriptor's arrays during the upload
    sid = ctypes.c_uint64()
    return eng.L.monsel_scene_upload(eng.ctx, ctypes.c_void_p(cs.desc_address()), ctypes.byref(sid))


def test_multiplane_shape_limit(M, oracle):
    """A 17-sided prism (a multi-plane shape of 17 planes; the kernels take
    16) is refused at upload with MONSEL_E_UNSUPPORTED; the 16-sided one,
    the largest primitive the kernels take (a PMMA prism of 20 nm inradius
    and 50 nm height under the beam), agrees with the oracle: BSE and SE
    yields within 3 sigma, exit histograms by KS."""
    import stats_util as su
    from cudamonsel_amd import scenes
    from cudamonsel_amd.abi import Engine, last_error
    NM = 1e-9

    def prism(sides):
        sc = scenes.Scene()
        ch = scenes._chamber(sc)
        msm = scenes.MATERIALS["PMMA"](sc)
        s = M.NormalMultiPlaneShape()
        for k in range(sides - 2):
            a = 2.0 * np.pi * k / (sides - 2)
            s.addPlane([np.cos(a), np.sin(a), 0.0], [20 * NM * np.cos(a), 20 * NM * np.sin(a), 0.0])
        s.addPlane([0.0, 0.0, -1.0], [0.0, 0.0, 0.0])        # z >= 0, under the beam
        s.addPlane([0.0, 0.0, 1.0], [0.0, 0.0, 50 * NM])     # z <= 50 nm
        assert s.getNumPlanes() == sides
        reg = M.Region(ch, msm, s)
        sc.hold(s, reg)
        scenes._finish(sc, 500.0, 0.5, 0.0)
        return sc

    with Engine() as eng:
        assert _upload_status(eng, prism(17)) == UNSUPPORTED, last_error()
        assert "16 planes" in last_error()
    sc = prism(16)
    centers = np.array([[x * NM, 0.0, sc.beam_z] for x in np.linspace(-24.0, 24.0, 16)])
    sc.monte.setSeed(12)
    r = sc.monte.runPixels(centers, 500, 50, 50.0)
    o = oracle.OracleScene(sc.monte.compileScene()).run(centers, 500, M.ToSI.eV(500.0), 0.5e-9, seed=12, nbins=50,
                                                        se_cutoff_eV=50.0, threads=0)
    assert r.stats.queue_overflow == 0 and r.stats.watchdog_drops == 0 and o["stats"]["watchdog_drops"] == 0
    zs, zb = su.total_z(su.from_gpu(r), su.from_oracle(o))
    assert abs(zs) < 3 and abs(zb) < 3, (zs, zb)
    for name in ("backEnergyHist", "elevationHist", "azimuthHist"):
        D, p = su.ks_binned(getattr(r, name), o[name])
        assert p > 0.01 / 3, (name, D, p)


def test_scene_beyond_kernel_limits_is_refused(M):
    """A region shape of nine convex parts (the kernels take eight) is
    refused at upload with MONSEL_E_UNSUPPORTED; the context still runs
    other scenes."""
    from cudamonsel_amd import scenes
    from cudamonsel_amd.abi import Engine
    sc = scenes.Scene()
    ch = scenes._chamber(sc)
    msm = scenes.MATERIALS["PMMA"](sc)
    NM = 1e-9
    shape = M.NormalCylindricalShape([0.0, 0.0, 0.0], [0.0, 0.0, 10 * NM], 2 * NM)
    parts = [shape]
    for k in range(1, 9):
        c = M.NormalCylindricalShape([k * 5 * NM, 0.0, 0.0], [k * 5 * NM, 0.0, 10 * NM], 2 * NM)
        parts.append(c)
        shape = M.NormalUnionShape(shape, c)
        parts.append(shape)
    reg = M.Region(ch, msm, shape)
    sc.hold(*parts, reg)
    scenes._finish(sc, 500.0, 0.5, 0.0)
    from cudamonsel_amd.abi import last_error
    cs = sc.monte.compileScene()   # owns the descriptor's arrays during the upload
    with Engine() as eng:
        sid = ctypes.c_uint64()
        rc = eng.L.monsel_scene_upload(eng.ctx, ctypes.c_void_p(cs.desc_address()), ctypes.byref(sid))
        assert rc == UNSUPPORTED, last_error()
        ok = scenes.flat("Si")
        r = eng.result(eng.submit(eng.upload(ok.monte.compileScene()), [[0.0, 0.0, ok.beam_z]], 100, 1000.0,
                                  0.5e-9, 9))
        assert r["stats"]["primaries"] == 100


def test_destroy_with_jobs_in_flight(M):
    """monsel_destroy with jobs still queued and running: the dispatcher
    finishes them before it stops, destroy returns MONSEL_OK, and a new
    context on the device works."""
    from cudamonsel_amd import scenes
    from cudamonsel_amd.abi import Engine
    sc = scenes.flat("PMMA", beam_eV=500.0)
    cs = sc.monte.compileScene()
    cen = scenes.scan_centers(np.arange(-8.0, 8.0, 1.0), [0.0], sc.beam_z)
    eng = Engine()
    sid = eng.upload(cs)
    for k in range(4):
        eng.submit(sid, cen, 5000, 500.0, 0.5e-9, 60 + k)
    ctx = eng.ctx
    eng.ctx = ctypes.c_void_p()
    assert eng.L.monsel_destroy(ctx) == OK
    with Engine() as again:
        r = again.result(again.submit(again.upload(cs), cen[:1], 100, 500.0, 0.5e-9, 70))
        assert r["stats"]["primaries"] == 100


def test_scene_released_while_its_jobs_wait(M):
    """A scene released right after jobs were submitted on it (some of them
    still queued behind a large job) stays alive for those jobs: they complete
    with the results they have when the scene is kept; new submissions on the
    released id are refused."""
    from cudamonsel_amd import scenes
    from cudamonsel_amd.abi import Engine
    big = scenes.flat("PMMA", beam_eV=500.0)
    sc = scenes.flat("Si", beam_eV=1000.0)
    cen_big = scenes.scan_centers(np.arange(-64.0, 64.0, 1.0), [0.0], big.beam_z)
    cen = [[0.0, 0.0, sc.beam_z]]
    with Engine(max_jobs_in_flight=1) as eng:
        ref = eng.result(eng.submit(eng.upload(sc.monte.compileScene()), cen, 2000, 1000.0, 0.5e-9, 81))
        first = eng.submit(eng.upload(big.monte.compileScene()), cen_big, 4000, 500.0, 0.5e-9, 80)
        sid = eng.upload(sc.monte.compileScene())
        ids = [eng.submit(sid, cen, 2000, 1000.0, 0.5e-9, 81) for _ in range(2)]
        assert eng.L.monsel_scene_release(eng.ctx, ctypes.c_uint64(sid)) == OK
        jd, keep = _desc(cen)
        rc, _ = _submit(eng, sid, jd)
        assert rc == NOTFOUND
        for i in ids:
            r = eng.result(i, timeout_s=60)
            assert np.array_equal(r["back"], ref["back"]) and np.array_equal(r["se"], ref["se"])
        eng.result(first, timeout_s=60)
