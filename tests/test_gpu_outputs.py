"""A JSON job (scene file + raster + outputs) run through the C-ABI writes
the five-column text, the BSE/SE maps and the tallies; the tallies equal
those of the same job on the builder's scene bit for bit, and the files
hold the yields the reference's driver would print (LinesOnLayers.cu:366-377)."""
import json

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_json_job_outputs(tmp_path):
    from cudamonsel_amd import outputs, scenefile, scenes
    from cudamonsel_amd.abi import Engine
    (tmp_path / "scene.json").write_text(json.dumps(scenefile.line_scan_spec()))
    job = {"scene": "scene.json", "raster": {"x_nm": [-8, 8, 1], "y_nm": [0, 4, 1]}, "primaries_per_pixel": 200,
           "seed": 7, "energy_bins": 50, "outputs": {"text": "scan.txt", "maps": "img", "tallies": "t.npz"}}
    (tmp_path / "job.json").write_text(json.dumps(job))
    with Engine() as eng:
        r, cen = scenefile.run_job(str(tmp_path / "job.json"), engine=eng)
        sc = scenes.line_on_substrate()
        sid = eng.upload(sc.monte.compileScene())
        direct = eng.result(eng.submit(sid, cen, 200, 500.0, 0.5e-9, 7, 50, 50.0))
    assert cen.shape == (64, 3)
    z = outputs.load_tallies(str(tmp_path / "t.npz"))
    for k in outputs.TALLY_KEYS:
        assert np.array_equal(z[k], direct[k]), k
    t = outputs.read_scan_text(str(tmp_path / "scan.txt"))
    bse, se = outputs.scan_yields(direct["back"], direct["se"], 200)
    assert np.allclose(t["se"], se, atol=5e-7) and np.allclose(t["bse"], bse, atol=5e-7)
    assert np.allclose(t["x_nm"], np.tile(np.arange(-8, 8), 4)) and np.allclose(t["y_nm"], np.repeat(np.arange(4), 16))
    b = np.load(str(tmp_path / "img_bse.npy"))
    s = np.load(str(tmp_path / "img_se.npy"))
    assert b.shape == (4, 16) and np.allclose(s.ravel(), se.astype(np.float32))
