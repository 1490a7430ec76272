"""TabulatedInelasticSM on the dielectric-model tables (dielectric.py), GPU
against the oracle: bulk Si at 1 keV with 20,000 primaries and 200 energy
bins over 0-1 keV (the reference's single-pixel Si configuration, with
model tables in place of its unshipped files), and bulk Cu at 2 keV, in
fixed mode: in reference mode the 3-D angle lookup extrapolates from the
table's first energies (the NULagrangeInterpolation d3 defect,
NULagrangeInterpolation.cu:236) and scatters inelastically into nearly
isotropic angles.  Yields within 3 sigma, exit-energy and angle histograms
by KS; the Si yields lie where measurements for Si at 1 keV do
(backscatter 0.1-0.35 counting every electron above 50 eV, secondary yield
0.3-1.2; the measured data scatter widely)."""
import numpy as np
import pytest

import stats_util as su

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("material,E_eV,per,nbins", [("Si-dielectric", 1000.0, 1250, 200),
                                                      ("Cu-dielectric", 2000.0, 500, 200)])
def test_dielectric_tables_gpu_vs_oracle(oracle, M, monkeypatch, tmp_path, material, E_eV, per, nbins):
    monkeypatch.setenv("MONSEL_TABLE_CACHE", str(tmp_path))
    from cudamonsel_amd import scenes
    sc = scenes.flat(material, beam_eV=E_eV, compat=False)
    centers = np.array([[0.0, 0.0, sc.beam_z]] * 16)
    sc.monte.setSeed(21)
    r = sc.monte.runPixels(centers, per, nbins, 50.0)
    osc = oracle.OracleScene(sc.monte.compileScene())
    o = osc.run(centers, per, M.ToSI.eV(E_eV), 0.5e-9, seed=21, nbins=nbins, se_cutoff_eV=50.0, threads=0)
    assert r.stats.queue_overflow == 0 and r.stats.watchdog_drops == 0 and o["stats"]["watchdog_drops"] == 0
    g, c = su.from_gpu(r), su.from_oracle(o)
    zs, zb = su.total_z(g, c)
    assert abs(zs) < 3 and abs(zb) < 3, (zs, zb)
    for name in ("backEnergyHist", "elevationHist", "azimuthHist"):
        D, p = su.ks_binned(getattr(r, name), o[name])
        assert p > 0.01 / 3, (name, D, p)
    if material == "Si-dielectric":
        se, bse = g[0].mean(), g[1].mean()
        assert 0.1 < bse < 0.35 and 0.3 < se < 1.2, (se, bse)
