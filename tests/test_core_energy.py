"""SEmaterial::setEstimatedCoreEnergy (SEmaterial.cu:347-366): core energies
from the occupied shells' edge energies (NIST-Chantler 2005, the FFastEdgeDB
rows shipped with the reference) down to the first shell at or below the
cutoff.  Pinned by the table rows themselves (Si: K 1838.9, L1 148.7, L2 = L3
99.2, M1 11.36 eV; C: K 283.8, L1 19.51; O: K 532.0, L1 23.7, L2 7.1;
H: K 13.6; Cu: K 8978.9, L1 1096.1, L2 951.0, L3 931.1, M1 119.8, M2 = M3
73.6, M4 1.6 eV) and by the loop's semantics: Si's estimate is the set of
core levels LinesOnLayers.cu:184 writes by hand (99.2, 99.8, 149.7, 1839
eV) within the two tables' differences."""
import pytest


def _estimate(M, els, fr, cutoff_eV=None):
    c = M.Composition()
    c.defineByMoleFraction(els, fr)
    m = M.SEmaterial(c, 1000.0)
    if cutoff_eV is None:
        m.setEstimatedCoreEnergy()
    else:
        m.setEstimatedCoreEnergy(M.ToSI.eV(cutoff_eV))
    return [round(M.FromSI.eV(e), 6) for e in m.getCoreEnergyArray()]


def test_estimated_core_energies(M):
    E = M.Elements
    assert _estimate(M, [E.Si], [1.0]) == [99.2, 148.7, 1838.9]
    assert _estimate(M, [E.C, E.O, E.H], [5, 2, 8]) == [23.7, 283.8, 532.0]
    assert _estimate(M, [E.H], [1.0]) == []
    assert _estimate(M, [E.Cu], [1.0]) == [73.6, 119.8, 931.1, 951.0, 1096.1, 8978.9]
    # the cutoff: shells at or below it end the walk
    assert _estimate(M, [E.Si], [1.0], 100.0) == [148.7, 1838.9]
    assert _estimate(M, [E.Si], [1.0], 148.7) == [1838.9]
    assert _estimate(M, [E.Si], [1.0], 5.0) == [5.08305, 11.3572, 99.2, 148.7, 1838.9]
    # close to LinesOnLayers' hand-written Si core levels
    est = _estimate(M, [E.Si], [1.0])
    for a, b in zip(est, [99.2, 149.7, 1839.0]):
        assert abs(a - b) / b < 0.01


def test_scene_file_estimated_core(M):
    from cudamonsel_amd import scenefile
    spec = scenefile.line_scan_spec()
    spec["materials"]["Silicon"].pop("core_energies_eV")
    spec["materials"]["Silicon"]["estimated_core_energy_eV"] = 20.0
    sc = scenefile.load_scene(spec)
    assert sc.monte.compileScene() is not None
