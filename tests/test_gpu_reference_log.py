"""The GPU path against the reference's own committed run log
(CudaMONSEL/CudaMONSEL/outputs/output_release.txt, per-x-column means in
tests/golden/reference_run_log.json): the LinesOnLayers geometry (PMMA line
h 120 nm, w 80 nm, 3 deg, r 20 nm on an 80 nm PMMA layer over Si), 0.5 keV,
reference-compat mode, the log's 80 x columns at the log's 29,600 primaries
per column.

- BSE yield per column: chi-square against the log without any scaling.
- SE yield: the shape per column after one common scale factor, and that
  factor itself.  The CPU oracle on this geometry gives 1.0109 +- 0.0007
  times the log's SE yield uniformly across the profile (flat top, corner
  peak, sidewall, foot, far field), with the BSE profile matching (chi-square
  p = 0.37); no modelled deviation explains the offset (reference-mode
  toggles, 15-bit or MSVC uniforms, element order), so the log's unknown
  revision is taken to differ in level, and the test pins the level at the
  oracle's measured factor within 0.006.
- The SE peak (the right top corner) and the dip just beyond the line's
  foot fall in the log's columns."""
import json
import os

import numpy as np
import pytest
from scipy import stats

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(__file__), "golden", "reference_run_log.json")
ORACLE_SE_SCALE = 1.0109


def test_line_profile_matches_reference_log(M):
    from cudamonsel_amd import scenes
    cols = json.load(open(GOLD))["line_profile_x_nm"]["columns"]
    xs = sorted(cols, key=float)
    per = cols[xs[0]][2] * 100       # rows x 100 primaries per row
    log_bse = np.array([cols[x][0] for x in xs])
    log_se = np.array([cols[x][1] for x in xs])
    sc = scenes.lines_on_layers()
    cen = np.array([[float(x) * 1e-9, 0.0, sc.beam_z] for x in xs])
    sc.monte.setSeed(2024)
    r = sc.monte.runPixels(cen, per, 50, 50.0)
    assert r.stats.queue_overflow == 0 and r.stats.watchdog_drops == 0
    se = np.asarray(r.se, np.float64) / per
    bse = (np.asarray(r.back, np.float64) - np.asarray(r.se, np.float64)) / per
    se_err = np.sqrt((np.asarray(r.seSq) / per - se ** 2) / per)
    # BSE: no scaling; binomial errors of both runs
    eb = np.sqrt(bse * (1 - bse) / per + log_bse * (1 - log_bse) / per)
    zb = (bse - log_bse) / eb
    assert stats.chi2.sf(np.sum(zb ** 2), len(zb)) > 1e-3 and np.max(np.abs(zb)) < 4.5, np.max(np.abs(zb))
    # SE: the log's per-column error taken equal to this run's
    err = np.sqrt(2.0) * se_err / log_se
    w = 1.0 / err ** 2
    ratio = se / log_se
    k = np.sum(w * ratio) / np.sum(w)
    assert abs(k - ORACLE_SE_SCALE) < 0.006, k
    z = (ratio / k - 1.0) / err
    assert stats.chi2.sf(np.sum(z ** 2), len(z) - 1) > 1e-4 and np.max(np.abs(z)) < 4.5, np.max(np.abs(z))
    # features: the SE peak at the right top corner, the dip past the foot
    xv = np.array([float(x) for x in xs])
    assert abs(xv[np.argmax(se)] - xv[np.argmax(log_se)]) <= 1.0
    far = xv > 40.0
    assert abs(xv[far][np.argmin(se[far])] - xv[far][np.argmin(log_se[far])]) <= 1.0
