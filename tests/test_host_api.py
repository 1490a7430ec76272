"""Host API (the reference-mirror classes) and the C-ABI library, on the CPU.

- weight fractions and SEmaterial energetics against closed forms from the
  reference's definitions (Composition.cu:320-373, SEmaterial.cu:201, 392-399);
- MONSEL_MaterialScatterModel's default tracking cutoff
  (MONSEL_MaterialScatterModel.cu:25-36);
- invalid constructor arguments raise instead of leaving undefined state;
- libmonsel_gpu.so exports every function include/monsel_gpu.h declares, and
  the scene compiler runs without a device (monsel_scene_footprint)."""
import ctypes
import math
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_pmma_weight_fractions(M):
    A = {e: getattr(M.Elements, e).getAtomicWeight() for e in ("C", "O", "H")}
    comp = M.Composition()
    comp.defineByMoleFraction([M.Elements.C, M.Elements.O, M.Elements.H], [5, 2, 8])
    tot = 5 * A["C"] + 2 * A["O"] + 8 * A["H"]
    assert comp.weightFraction(M.Elements.C, True) == pytest.approx(5 * A["C"] / tot, rel=1e-14)
    assert comp.weightFraction(M.Elements.O, True) == pytest.approx(2 * A["O"] / tot, rel=1e-14)
    assert comp.weightFraction(M.Elements.H, True) == pytest.approx(8 * A["H"] / tot, rel=1e-14)
    # AtomicWeights.csv values held as float (Element.cu:606)
    assert A["C"] == float(__import__("numpy").float32(12.0107))


def test_semat_energetics(M):
    ev = M.ToSI.eV(1.0)
    mat = M.SEmaterial()
    mat.setWorkfunction(5.5 * ev)
    assert mat.getEnergyCBbottom() == pytest.approx(-5.5 * ev)  # clamped below -W
    mat.setEnergyCBbottom(-0.5 * ev)
    assert mat.getEFermi() == pytest.approx(-5.0 * ev)


def test_msm_tracking_cutoff(M):
    from cudamonsel_amd import scenes
    sc = scenes.Scene()
    pm = scenes.pmma(sc)
    assert pm.getMinEforTracking() == pytest.approx(M.ToSI.eV(0.5))
    vac = scenes.vacuum(sc)
    assert vac.getMinEforTracking() == -math.inf


def test_invalid_construction_raises(M):
    mat = M.SEmaterial([M.Elements.Si], [1.0], 2330.0, "Si")
    mat.setWorkfunction(M.ToSI.eV(4.85))
    with pytest.raises(M.MonselError):
        M.JoyLuoNieminenCSD(mat, M.ToSI.eV(5.0))  # breakE must exceed W + 1 eV
    with pytest.raises(M.MonselError):
        M.NormalCylindricalShape([0, 0, 0], [0, 0, 0], 1e-9)
    with pytest.raises(M.MonselError):
        M.GanachaudMokraniPolaronTrapSM(-1.0, 1.0)


def header_functions():
    txt = open(os.path.join(ROOT, "include", "monsel_gpu.h")).read()
    return sorted(set(re.findall(r"^\s*(?:int|const char\*)\s+(monsel_\w+)\(", txt, re.M)))


def test_library_exports_header(M):
    lib = ctypes.CDLL(M.GPU_LIB)
    names = header_functions()
    assert len(names) >= 20
    for n in names:
        assert hasattr(lib, n), n
    version = int(re.search(r"#define MONSEL_ABI_VERSION (\d+)u", open(os.path.join(ROOT, "include", "monsel_gpu.h")).read()).group(1))
    assert lib.monsel_abi_version() == version


def test_scene_footprint_without_device(M):
    from cudamonsel_amd import scenes
    sc = scenes.line_on_substrate()
    cs = sc.monte.compileScene()
    lib = ctypes.CDLL(M.GPU_LIB)
    nbytes = ctypes.c_size_t(0)
    rc = lib.monsel_scene_footprint(ctypes.c_void_p(cs.desc_address()), ctypes.byref(nbytes))
    assert rc == 0 and 0 < nbytes.value < 64 << 20
    # regions: chamber, substrate, line; compat line = 4 distinct convex parts
    assert cs.n_regions == 3


def test_unsupported_scene_is_rejected(M):
    """A region shape with more than 8 convex parts is refused at upload."""
    sc_mod = __import__("cudamonsel_amd.scenes", fromlist=["x"])
    sc = sc_mod.Scene()
    ch = sc_mod._chamber(sc)
    msm = sc_mod.pmma(sc)
    shapes = []
    for i in range(9):
        s = M.NormalMultiPlaneShape()
        s.addPlane([0.0, 0.0, -1.0], [0.0, 0.0, i * 1e-9])
        shapes.append(s)
    u = shapes[0]
    for s in shapes[1:]:
        u = M.NormalUnionShape(u, s)
        shapes.append(u)
    reg = M.Region(ch, msm, u)
    sc.hold(reg, *shapes)
    sc_mod._finish(sc, 500.0, 0.5, 0.0)
    cs = sc.monte.compileScene()
    lib = ctypes.CDLL(M.GPU_LIB)
    lib.monsel_last_error.restype = ctypes.c_char_p
    nbytes = ctypes.c_size_t(0)
    rc = lib.monsel_scene_footprint(ctypes.c_void_p(cs.desc_address()), ctypes.byref(nbytes))
    assert rc == -2 and b"convex" in lib.monsel_last_error()


def test_job_without_device_fails_loudly(M):
    """No CPU fallback: without a device a job raises."""
    from cudamonsel_amd import scenes
    import numpy as np
    lib = ctypes.CDLL(M.GPU_LIB)
    ctx = ctypes.c_void_p()
    if lib.monsel_create(0, None, ctypes.byref(ctx)) == 0:
        lib.monsel_destroy(ctx)
        pytest.skip("a GPU is present")
    sc = scenes.flat("PMMA")
    with pytest.raises(M.MonselError):
        sc.monte.runPixels(np.array([[0.0, 0.0, sc.beam_z]]), 10)


def test_gaussian_beam_create_electron(M):
    """GaussianBeam::createElectron (GaussianBeam.cu:52-65): at the beam
    energy, along +z, in the plane of the centre, at a radius that follows
    Rayleigh(width) and an angle uniform in [0, 2 pi) (KS); MonteCarloSS
    takes it as its initial electron, as LinesOnLayers.cu:329 passes it."""
    import numpy as np
    from scipy import stats
    w = 2e-9
    g = M.GaussianBeam(w, M.ToSI.eV(500.0), [1e-9, -3e-9, -5e-8])
    es = [g.createElectron() for _ in range(4000)]
    p = np.array([e.getPosition() for e in es])
    assert all(e.getEnergy() == M.ToSI.eV(500.0) and e.getTheta() == 0.0 and e.getPhi() == 0.0 for e in es)
    assert np.all(p[:, 2] == -5e-8)
    dx, dy = p[:, 0] - 1e-9, p[:, 1] + 3e-9
    r = np.hypot(dx, dy)
    assert stats.kstest(r, stats.rayleigh(scale=w).cdf).pvalue > 1e-3
    ang = np.mod(np.arctan2(dy, dx), 2 * np.pi)
    assert stats.kstest(ang, stats.uniform(0, 2 * np.pi).cdf).pvalue > 1e-3
    assert len({e.getIdent() for e in es}) == len(es)
    from cudamonsel_amd import scenes
    sc = scenes.flat("Si")
    mc = M.MonteCarloSS(sc.gun, sc.chamber, g.createElectron())
    assert mc.getSeed() is not None
