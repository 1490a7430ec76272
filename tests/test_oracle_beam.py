"""GaussianBeam::createElectron (GaussianBeam.cu:52-65) in the oracle: the
start point is center + r (cos th, sin th, 0) with r = width sqrt(-2 ln U1)
and th = 2 pi U2, so for independent uniforms the radial offset follows the
Rayleigh distribution of scale `width` (P(r <= x) = 1 - exp(-x^2 / 2 w^2))
and the azimuth is uniform on [0, 2 pi).  A dropped factor (w sqrt(-ln U)),
a wrong scale or a swapped cos/sin fails one of the checks below."""
import numpy as np
from scipy import stats


def _scene(M):
    from cudamonsel_amd import scenes
    sc = scenes.flat("PMMA", beam_eV=500.0)
    return sc, sc.monte.compileScene()


def test_beam_spot_is_rayleigh_and_isotropic(oracle, M):
    sc, cs = _scene(M)
    osc = oracle.OracleScene(cs)
    w = 0.5e-9
    center = np.array([[3.0e-9, -2.0e-9, sc.beam_z]])
    n = 100000
    p, reg = osc.primaries(center, n, w, seed=9, n=n)
    d = p - center
    assert np.all(d[:, 2] == 0.0)
    r = np.hypot(d[:, 0], d[:, 1])
    assert stats.kstest(r / w, stats.rayleigh.cdf).pvalue > 1e-3
    th = np.mod(np.arctan2(d[:, 1], d[:, 0]), 2 * np.pi)
    assert stats.kstest(th / (2 * np.pi), "uniform").pvalue > 1e-3
    # second moment: <r^2> = 2 w^2
    assert abs(np.mean(r * r) / (2 * w * w) - 1.0) < 5 * np.sqrt(1.0 / n)
    # the beam starts above the sample, in the chamber's vacuum region
    assert np.all(reg == 0)


def test_beam_spot_pixels_and_streams(oracle, M):
    """Primary k of a job belongs to pixel k // per_pixel, and its offset
    depends only on (seed, primary index): a job split by primary_offset
    draws the same start points."""
    sc, cs = _scene(M)
    osc = oracle.OracleScene(cs)
    centers = np.array([[x * 1e-9, 0.0, sc.beam_z] for x in (-10.0, 0.0, 10.0)])
    p, _ = osc.primaries(centers, 4, 0.5e-9, seed=3, n=12)
    for k in range(12):
        assert np.linalg.norm(p[k, :2] - centers[k // 4, :2]) < 5e-9
    # primaries 4..11 of the same stream, as a second job starting at pixel 1
    q, _ = osc.primaries(centers[1:], 4, 0.5e-9, seed=3, n=8, primary_offset=4)
    assert np.array_equal(p[4:], q)
    # zero width: every primary at its pixel centre
    z, _ = osc.primaries(centers, 4, 0.0, seed=3, n=12)
    assert np.array_equal(z, np.repeat(centers, 4, 0))
