"""The detected-electron log (BackscatterStats::setLogDetected,
BackscatterStats.cu:166-168): one record per electron that reaches the
detector, whose contents reproduce the job's own tallies exactly -- every
record's elevation, azimuth and energy fall in the histogram bins the job
counted (reference-compat binning, the oracle's Histogram semantics), every
back-escape is in its primary's pixel -- plus record count = escapes, the
points lie on the chamber sphere, truncation at the requested capacity, and
the same records in a repeat run."""
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _bins(oracle, vals, lo, hi, n):
    return np.array([oracle.hist_bin(float(v), lo, hi, n, True) for v in vals], dtype=np.int64) + 1


def test_detected_log_reproduces_the_tallies(oracle, M):
    from cudamonsel_amd import scenes
    from cudamonsel_amd.abi import Engine
    sc = scenes.line_on_substrate()
    cen = scenes.scan_centers(np.linspace(-12, 12, 16), [0.0], sc.beam_z)
    per, nb = 300, 50
    with Engine() as eng:
        sid = eng.upload(sc.monte.compileScene())
        r = eng.result(eng.submit(sid, cen, per, 500.0, 0.5e-9, 9, nb, log_detected=64 * 16 * per))
        again = eng.result(eng.submit(sid, cen, per, 500.0, 0.5e-9, 9, nb, log_detected=64 * 16 * per))
        small = eng.result(eng.submit(sid, cen, per, 500.0, 0.5e-9, 9, nb, log_detected=10))
    d = r["detected"]
    esc = r["stats"]["escapes"]
    assert r["n_detected"] == esc == len(d) > 16 * per
    p = d["position"]
    R = math.sqrt(float(np.max(np.sum(p * p, axis=1))))
    assert abs(R - M.ChamberRadius) < 1e-6 * M.ChamberRadius
    onwall = np.abs(np.sqrt(np.sum(p * p, axis=1)) - M.ChamberRadius) < 1e-6 * M.ChamberRadius
    assert onwall.mean() > 0.999
    elev = math.pi / 2 - np.arctan2(p[:, 2], np.sqrt(p[:, 0] ** 2 + p[:, 1] ** 2))
    azim = np.arctan2(p[:, 1], p[:, 0])
    azim = np.where(azim < 0, 2 * math.pi + azim, azim)
    assert np.array_equal(np.bincount(_bins(oracle, elev, 0.0, math.pi, 180), minlength=182), r["elevation_hist"])
    assert np.array_equal(np.bincount(_bins(oracle, azim, 0.0, 2 * math.pi, 360), minlength=362), r["azimuth_hist"])
    back = elev >= math.pi / 2
    eb = _bins(oracle, d["energy_ev"], 0.0, 500.0, nb)
    assert np.array_equal(np.bincount(eb[back], minlength=nb + 2), r["back_energy_hist"])
    assert np.array_equal(np.bincount(eb[~back], minlength=nb + 2), r["fwd_energy_hist"])
    pix = d["primary"] // per
    assert np.array_equal(np.bincount(pix[back], minlength=16), r["back"])
    assert np.array_equal(np.bincount(pix[~back], minlength=16), r["fwd"])
    assert np.all(d["step"] > 0) and np.all(d["energy_ev"] <= 500.0)
    assert np.all((d["theta"] >= 0) & (d["theta"] <= math.pi))
    # the same records in a repeat run (any order)
    key = lambda a: np.sort(a, order=["primary", "electron", "step"])
    assert np.array_equal(key(d), key(again["detected"]))
    # truncation: all are counted, the first 10 kept
    assert small["n_detected"] == esc and len(small["detected"]) == 10


def test_backscatter_stats_log(M):
    """The host API: BackscatterStats::setLogDetected / getLog."""
    from cudamonsel_amd import scenes
    sc = scenes.flat("PMMA", beam_eV=500.0)
    bs = M.BackscatterStats(sc.monte, 50)
    bs.setLogDetected(True)
    sc.monte.addActionListener(bs)
    sc.monte.setSeed(4)
    sc.monte.runMultipleTrajectories(2000)
    log = bs.getLog()
    assert log.shape[1] == 9 and len(log) == sc.monte.lastStats().escapes and not bs.logTruncated()
    back = bs.backscatterEnergyHistogram().totalCounts()
    elev = math.pi / 2 - np.arctan2(log[:, 6], np.hypot(log[:, 4], log[:, 5]))
    assert int(np.sum(elev >= math.pi / 2)) == back
