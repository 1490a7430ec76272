"""monsel_run_static: the allocation-free variant of monsel_run.  Through the
C-ABI, two calls in a row -- the second on a smaller job and another scene,
so that the static array holds the first call's contents -- each give the
tallies and statistics monsel_run gives for the same job, bit for bit; a
job whose tallies do not fit the array is rejected before any launch."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

KEYS = ("back", "fwd", "se", "se_sq", "bse_sq", "back_energy", "back_energy_hist", "fwd_energy_hist",
        "elevation_hist", "azimuth_hist")


def _same(a, b):
    for k in KEYS:
        assert np.array_equal(a[k], b[k]), k
    for k in ("primaries", "steps", "electrons", "escapes", "watchdog_drops", "queue_overflow"):
        assert a["stats"][k] == b["stats"][k], k


def test_run_static_matches_run(M):
    from cudamonsel_amd import scenes
    from cudamonsel_amd.abi import Engine
    line = scenes.line_on_substrate()
    si = scenes.flat("Si", beam_eV=1000.0)
    big = (scenes.scan_centers(np.arange(-32.0, 32.0, 2.0), [0.0], line.beam_z), 1500, 500.0, 0.5e-9, 61, 50)
    small = (np.array([[0.0, 0.0, si.beam_z]] * 3), 700, 1000.0, 0.5e-9, 62, 200)
    with Engine() as eng:
        s_line = eng.upload(line.monte.compileScene())
        s_si = eng.upload(si.monte.compileScene())
        ref_big = eng.run(s_line, *big)
        ref_small = eng.run(s_si, *small)
        got_big = eng.run(s_line, *big, static=True)
        got_small = eng.run(s_si, *small, static=True)
        # and once more in the other order: nothing of a previous call leaks
        again_big = eng.run(s_line, *big, static=True)
    for ref, got in ((ref_big, got_big), (ref_small, got_small), (ref_big, again_big)):
        assert got["stats"]["queue_overflow"] == 0 and got["stats"]["watchdog_drops"] == 0
        assert got["stats"]["steps"] > 0
        _same(ref, got)
    assert got_big["stats"]["primaries"] == 32 * 1500 and got_small["stats"]["primaries"] == 3 * 700


def test_run_static_rejects_jobs_beyond_its_buffer(M):
    """16 B per primary: 10 M primaries need more than the 128 MiB array."""
    import ctypes
    from cudamonsel_amd import abi, scenes
    sc = scenes.flat("PMMA", beam_eV=500.0)
    with abi.Engine() as eng:
        sid = eng.upload(sc.monte.compileScene())
        c = np.ascontiguousarray([[0.0, 0.0, sc.beam_z]] * 10, dtype=np.float64)
        jd = abi.JobDesc(500.0 * 1.60217653e-19, 0.5e-9, 10, 1_000_000, c.ctypes.data_as(abi.c_double_p), 1, 50, 0,
                         50.0, 0)
        rc = eng.L.monsel_run_static(eng.ctx, ctypes.c_uint64(sid), ctypes.byref(jd), None, None)
        assert rc == -9, rc
        assert "static buffer" in M.last_error()
        # the context runs on
        r = eng.run(sid, c[:2], 100, 500.0, 0.5e-9, 3, static=True)
        assert r["stats"]["primaries"] == 200
