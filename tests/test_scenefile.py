"""Scene/job files and result files (CPU): JSON scenes build the same scene
as the Python builders (the oracle, which reads the compiled scene, gives
identical physics and geometry for both); the five-column text reads back
what was written and parses the reference's own run log; raster maps."""
import json

import numpy as np
import pytest


def test_reference_output_parses():
    """The head of the reference's run log (outputs/output_release.txt)."""
    import os
    from cudamonsel_amd import outputs
    path = os.path.join(os.path.dirname(__file__), "golden", "reference_output_excerpt.txt")
    t = outputs.read_scan_text(path)
    assert t["beamE"].tolist() == [500.0] * 5
    assert t["x_nm"].tolist() == [-60.5, -55.5, -50.5, -45.5, -40.5]
    assert t["y_nm"].tolist() == [0.0] * 5
    assert t["bse"].tolist() == [0.1, 0.1, 0.16, 0.13, 0.09]
    assert t["se"].tolist() == [2.37, 2.19, 2.31, 2.08, 2.31]
    assert "# Trajectories at each landing position: 100" in t["header"]
    assert t["elapsed_s"] == pytest.approx(6426.023773)


def test_text_round_trip(tmp_path):
    from cudamonsel_amd import outputs
    rng = np.random.default_rng(1)
    n, per = 40, 1500
    se = rng.integers(2000, 4000, n).astype(np.uint64)
    back = se + rng.integers(100, 300, n).astype(np.uint64)
    cen = np.stack([np.linspace(-20e-9, 19e-9, n), np.full(n, 3e-9), np.full(n, -70e-9)], 1)
    p = str(tmp_path / "scan.txt")
    outputs.write_scan_text(p, 500.0, cen, back, se, per, header={"Line height (nm)": "50.000000"}, sigma_nm=0.5,
                            elapsed_s=1.25)
    t = outputs.read_scan_text(p)
    bse, sey = outputs.scan_yields(back, se, per)
    assert np.allclose(t["se"], sey, atol=5e-7) and np.allclose(t["bse"], bse, atol=5e-7)
    assert np.allclose(t["x_nm"], cen[:, 0] * 1e9, atol=5e-7) and np.allclose(t["y_nm"], 3.0)
    assert t["elapsed_s"] == 1.25
    # the reference's header order
    text = open(p).read().splitlines()
    assert text[0] == "# Trajectories at each landing position: 1500"
    assert text[1] == "Line height (nm): 50.000000"
    assert text[2] == "Beam landing energies (eV): 500.000000"
    assert text[3] == "Beam size (standard deviation, in nm): 0.500000"
    assert text[4] == "" and text[5] == outputs.COLUMNS
    # SE yield in single precision, BSE = back fraction - SE yield (LinesOnLayers.cu:370-371)
    assert sey[0] == float(np.float32(se[0]) / np.float32(per))
    assert bse[0] == back[0] / per - sey[0]


def test_maps_and_tallies(tmp_path):
    from cudamonsel_amd import outputs
    nx, ny, per = 6, 4, 100
    se = np.arange(nx * ny, dtype=np.uint64) * 10
    back = se + 7
    b, s = outputs.raster_maps(back, se, per, nx, ny)
    assert b.shape == (ny, nx) and b.dtype == np.float32
    assert s[1, 2] == np.float32((1 * nx + 2) * 10 / per)
    pb, ps = outputs.save_maps(str(tmp_path / "img"), back, se, per, nx, ny)
    assert np.array_equal(np.load(pb), b) and np.array_equal(np.load(ps), s)
    r = {k: np.arange(5, dtype=np.uint64) for k in outputs.TALLY_KEYS}
    outputs.save_tallies(str(tmp_path / "t.npz"), r, np.zeros((5, 3)), per, 500.0)
    z = outputs.load_tallies(str(tmp_path / "t.npz"))
    assert all(np.array_equal(z[k], r[k]) for k in outputs.TALLY_KEYS) and int(z["per_pixel"]) == per


@pytest.mark.parametrize("compat", [True, False], ids=["compat", "fixed"])
def test_scene_file_builds_the_builder_scene(oracle, tmp_path, compat):
    """line_scan_spec (through a JSON file) and scenes.line_on_substrate give
    the same regions, rates and boundary crossings."""
    from cudamonsel_amd import scenefile, scenes
    p = tmp_path / "scene.json"
    p.write_text(json.dumps(scenefile.line_scan_spec(compat=compat)))
    a = scenefile.load_scene(str(p))
    b = scenes.line_on_substrate(compat=compat)
    assert a.beam_z == b.beam_z
    ca, cb = a.monte.compileScene(), b.monte.compileScene()
    oa, ob = oracle.OracleScene(ca), oracle.OracleScene(cb)
    rng = np.random.default_rng(3)
    n = 4000
    p0 = np.stack([rng.uniform(-30, 30, n), rng.uniform(-30, 30, n), rng.uniform(-70, 20, n)], 1)
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1)[:, None]
    p1 = p0 + d * rng.uniform(0.5, 30, n)[:, None]
    ra, rb = oa.locate(p0), ob.locate(p0)
    ma = np.array([ca.msm_index_of_region(int(r)) if r >= 0 else -1 for r in ra])
    mb = np.array([cb.msm_index_of_region(int(r)) if r >= 0 else -1 for r in rb])
    E = np.geomspace(1.0, 5000.0, 64)
    for m_a, m_b in set(zip(ma.tolist(), mb.tolist())):
        if m_a < 0 or m_b < 0:
            assert m_a == m_b
            continue
        assert np.array_equal(oa.rates(m_a, E, 3), ob.rates(m_b, E, 3))
        assert np.array_equal(oa.csd(m_a, np.full(E.size, 0.7), E), ob.csd(m_b, np.full(E.size, 0.7), E))
    for r_a in np.unique(ra):
        if r_a < 0:
            continue
        sel = ra == r_a
        r_b = int(rb[sel][0])
        assert np.all(rb[sel] == r_b)
        ua, na, nra = oa.find_end_of_step(int(r_a), p0[sel], p1[sel])
        ub, nb, nrb = ob.find_end_of_step(r_b, p0[sel], p1[sel])
        assert np.array_equal(ua, ub) and np.array_equal(nra, nrb)


def test_scene_file_errors():
    from cudamonsel_amd import scenefile
    spec = scenefile.line_scan_spec()
    spec["shapes"]["line"]["kind"] = "torus"
    with pytest.raises(ValueError, match="torus"):
        scenefile.load_scene(spec)
    spec = scenefile.line_scan_spec()
    spec["models"]["PMMA"]["mechanisms"].append({"kind": "Magic"})
    with pytest.raises(ValueError, match="Magic"):
        scenefile.load_scene(spec)
