"""The C++ LinesOnLayers-style driver (cudamonsel_amd/host/lines_on_layers.cpp),
written against the host API alone and built with g++: it runs the
reference's scan loop (LinesOnLayers.cu:350-380, one runMultipleTrajectories
call per landing position with a BackscatterStats listener) and writes the
reference's five-column text.  Its rows equal, digit for digit, the same
calls made through the Python bindings of the same host library on the
Python-built scene (scenes.lines_on_layers) with the same seed, so the C++
object graph compiles to the same scene; and its flat-region rows agree
with the oracle on that scene."""
import os
import subprocess

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "cudamonsel_amd", "lines_on_layers")
TRAJ, NX, SEED = 400, 12, 77


@pytest.fixture(scope="module")
def driver_rows(tmp_path_factory):
    from cudamonsel_amd import outputs
    assert os.path.exists(EXE), "build the package first (python -m cudamonsel_amd.build)"
    out = tmp_path_factory.mktemp("lol") / "output.txt"
    p = subprocess.run([EXE, "--traj", str(TRAJ), "--xmax", str(NX), "--seed", str(SEED), "--out", str(out)],
                       capture_output=True, text=True, timeout=300)
    assert p.returncode == 0, p.stderr
    t = outputs.read_scan_text(str(out))
    assert len(t["x_nm"]) == NX and np.all(t["beamE"] == 500.0) and np.all(t["y_nm"] == 0.0)
    assert t["elapsed_s"] is not None and any("Trajectories at each landing position: 400" in h for h in t["header"])
    # the printed rows and the file's rows are the same
    printed = [l for l in p.stdout.splitlines() if l.strip()]
    assert len(printed) == NX
    return t, printed


def test_driver_rows_equal_the_bindings(M, driver_rows):
    from cudamonsel_amd import scenes
    t, printed = driver_rows
    # the reference's grid starts at 300.5 nm left of the line's foot, 5 nm apart
    assert np.array_equal(t["x_nm"], 40.0 - 300.5 + 5.0 * np.arange(NX))
    sc = scenes.lines_on_layers()
    sc.monte.setSeed(SEED)
    h = 120.0 * 1e-9
    for k, x_nm in enumerate(t["x_nm"]):
        sc.gun.setCenter([x_nm * 1e-9, 0.0, -h - 20.0 * 1e-9])   # the driver's expression for the beam's start
        back = M.BackscatterStats(sc.monte, 50)
        sc.monte.addActionListener(back)
        sc.monte.runMultipleTrajectories(TRAJ)
        hist = back.backscatterEnergyHistogram()
        se = sum(hist.counts(j) for j in range(int(50.0 / (500.0 / hist.binCount())))) / TRAJ
        bse = back.backscatterFraction() - se
        sc.monte.removeActionListener(back)
        row = " ".join("%f" % v for v in (500.0, x_nm, 0.0, bse, se))
        assert printed[k] == row, (k, printed[k], row)


def test_driver_rows_agree_with_the_oracle(M, oracle, driver_rows):
    """Every row lies 260 nm or more left of the line's foot (flat PMMA over
    Si): the driver's per-row yields, pooled, against the oracle's on the
    same scene and landing positions, within 3 sigma."""
    from cudamonsel_amd import scenes
    t, _ = driver_rows
    sc = scenes.lines_on_layers()
    centers = np.array([[x * 1e-9, 0.0, -120.0e-9 - 20.0e-9] for x in t["x_nm"]])
    o = oracle.OracleScene(sc.monte.compileScene()).run(centers, TRAJ, M.ToSI.eV(500.0), 0.5e-9, seed=SEED,
                                                        nbins=50, se_cutoff_eV=50.0, threads=0)
    n = TRAJ * NX
    # SE as the driver counts it: the back-scatter histogram's bins below
    # 50 eV (slot 0 is the underflow)
    o_se = float(np.sum(np.asarray(o["backEnergyHist"])[1:6])) / n
    o_bse = float(np.sum(o["back"])) / n - o_se
    g_se, g_bse = float(np.mean(t["se"])), float(np.mean(t["bse"]))
    # per-primary variances from the oracle's sums of squares
    var_se = float(np.sum(o["seSq"])) / n - o_se ** 2
    var_bse = max(o_bse * (1.0 - o_bse), 1e-12)
    z_se = (g_se - o_se) / np.sqrt(2.0 * var_se / n)
    z_bse = (g_bse - o_bse) / np.sqrt(2.0 * var_bse / n)
    assert abs(z_se) < 3 and abs(z_bse) < 3, (g_se, o_se, z_se, g_bse, o_bse, z_bse)


def test_batched_driver_rows_equal_run_pixels(M, tmp_path):
    """--batched runs the same landing positions as one runPixels job; its
    rows equal runPixels through the bindings digit for digit."""
    from cudamonsel_amd import outputs, scenes
    out = tmp_path / "batched.txt"
    p = subprocess.run([EXE, "--batched", "--traj", str(TRAJ), "--xmax", str(NX), "--seed", str(SEED), "--out",
                        str(out)], capture_output=True, text=True, timeout=300)
    assert p.returncode == 0, p.stderr
    t = outputs.read_scan_text(str(out))
    printed = [l for l in p.stdout.splitlines() if l.strip()]
    assert len(printed) == NX == len(t["x_nm"])
    sc = scenes.lines_on_layers()
    sc.monte.setSeed(SEED)
    centers = [[x * 1e-9, 0.0, -120.0 * 1e-9 - 20.0 * 1e-9] for x in t["x_nm"]]
    r = sc.monte.runPixels(centers, TRAJ, 50, 50.0)
    for k, x_nm in enumerate(t["x_nm"]):
        se = int(r.se[k]) / TRAJ
        bse = (int(r.back[k]) - int(r.se[k])) / TRAJ
        row = " ".join("%f" % v for v in (500.0, x_nm, 0.0, bse, se))
        assert printed[k] == row, (k, printed[k], row)
