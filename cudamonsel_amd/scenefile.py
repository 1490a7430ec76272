"""JSON scene and job files.

A scene file declares the reference's object graph by name -- materials
(SEmaterial), material scatter models (MONSEL_MaterialScatterModel with
their mechanisms, barrier and slowing-down algorithm), shapes and the region
tree -- and `load_scene` builds it with the same classes and constructor
arguments that LinesOnLayers.cu:80-282 uses, so a scene file reproduces a
driver's hand-written scene.  Lengths are in nm, energies in eV, densities
in kg/m^3; the chamber (a vacuum sphere, LinesOnLayers.cu:230-243) is
implicit and named "chamber".

  {"materials": {"PMMA": {"elements": ["C", "O", "H"], "mole_fractions": [5, 2, 8],
                          "density": 1190, "workfunction_eV": 5.5, "bandgap_eV": 5.0,
                          "cb_bottom_eV": -0.5}},
   "models": {"PMMA": {"material": "PMMA", "barrier": {"lambda_nm": null},
                       "csd": {"kind": "JoyLuoNieminenCSD", "breakE_eV": 45},
                       "mechanisms": [{"kind": "SelectableElasticSM", "factory": "NISTMottRS"},
                                      {"kind": "FittedInelSM", "energySEgen_eV": 65.4},
                                      {"kind": "GanachaudMokraniPolaronTrapSM",
                                       "prefactor_per_m": 2e7, "extinction_per_eV": 0.25}]}},
   "shapes": {"sub": {"kind": "NormalMultiPlaneShape",
                      "planes": [{"normal": [0, 0, -1], "point_nm": [0, 0, 0]}]}},
   "regions": [{"name": "substrate", "parent": "chamber", "model": "PMMA", "shape": "sub"}],
   "beam": {"energy_eV": 500, "sigma_nm": 0.5, "center_nm": [0, 0, -20]},
   "ref_compat": "reference"}

A material may give its core levels as "core_energies_eV" or ask for
SEmaterial::setEstimatedCoreEnergy with "estimated_core_energy_eV" (the
cutoff).  Shape kinds: NormalMultiPlaneShape (planes), NormalCylindricalShape (end0_nm,
end1_nm, radius_nm), line (NShapes::createLine: topz_nm, width_nm,
length_nm, thetal_deg, thetar_deg, radl_nm, radr_nm, reference_compat,
translate_nm), step_edge (NShapes::createStepEdge: z_nm, step_nm),
NormalIntersectionShape and NormalUnionShape (two shape names).  Mechanism
kinds: SelectableElasticSM (factory NISTMottRS, NISTMottRS100,
NISTMottRS100Lin, Browning or ScreenedRutherford), FittedInelSM,
GanachaudMokraniPolaronTrapSM, GanachaudMokraniPhononInelasticSM,
TabulatedInelasticSM (method_se, tables, energy_offset_eV,
e0_from_dispersion, min_egen_se_eV, rate_mult).

A job file names a scene file (or holds one under "scene") and the job:
pixels as a raster {"x_nm": [start, stop, step], "y_nm": [...]} or a list
"centers_nm", primaries per pixel, seed, energy bins, SE cutoff, and the
outputs to write ("text", "maps", "tallies").
"""
import json
import math
import os
import time

import numpy as np

from . import _monsel as M
from . import outputs, scenes

NM = 1e-9


def _material(spec):
    els = [getattr(M.Elements, e) for e in spec["elements"]]
    comp = M.Composition()
    if "mole_fractions" in spec:
        comp.defineByMoleFraction(els, [float(v) for v in spec["mole_fractions"]])
    else:
        comp.defineByWeightFraction(els, [float(v) for v in spec["weight_fractions"]])
    mat = M.SEmaterial(comp, float(spec["density"]))
    if "name" in spec:
        mat.setName(spec["name"])
    if "workfunction_eV" in spec:
        mat.setWorkfunction(M.ToSI.eV(spec["workfunction_eV"]))
    if "bandgap_eV" in spec:
        mat.setBandgap(M.ToSI.eV(spec["bandgap_eV"]))
    if "cb_bottom_eV" in spec:
        mat.setEnergyCBbottom(M.ToSI.eV(spec["cb_bottom_eV"]))
    if "core_energies_eV" in spec:
        mat.setCoreEnergy([M.ToSI.eV(v) for v in spec["core_energies_eV"]])
    if "estimated_core_energy_eV" in spec:   # SEmaterial::setEstimatedCoreEnergy(cutoff)
        mat.setEstimatedCoreEnergy(M.ToSI.eV(float(spec["estimated_core_energy_eV"])))
    return comp, mat


_FACTORIES = {
    "NISTMottRS": lambda: M.NISTMottRS.Factory,
    "NISTMottRS100": lambda: M.NISTMottRS.Factory100,
    "NISTMottRS100Lin": lambda: M.NISTMottRS.Factory100Lin,
    "Browning": lambda: M.BrowningEmpiricalCrossSection.Factory,
    "ScreenedRutherford": lambda: M.ScreenedRutherfordScatteringAngle.Factory,
}


def _csd(spec, mat):
    if spec is None:
        return None
    if spec["kind"] == "JoyLuoNieminenCSD":
        return M.JoyLuoNieminenCSD(mat, M.ToSI.eV(spec["breakE_eV"]))
    if spec["kind"] == "ZeroCSD":
        return M.ZeroCSD()
    raise ValueError(f"unknown slowing-down algorithm {spec['kind']!r}")


def _mechanism(spec, mat, csd, base_dir, held):
    k = spec["kind"]
    if k == "SelectableElasticSM":
        return M.SelectableElasticSM(mat, _FACTORIES[spec.get("factory", "NISTMottRS")]())
    if k == "FittedInelSM":
        c = _csd(spec["csd"], mat) if "csd" in spec else csd
        if c is None:
            raise ValueError("FittedInelSM needs a slowing-down algorithm (its own 'csd' or the model's)")
        held.append(c)
        return M.FittedInelSM(mat, M.ToSI.eV(spec["energySEgen_eV"]), c)
    if k == "GanachaudMokraniPolaronTrapSM":
        return M.GanachaudMokraniPolaronTrapSM(float(spec["prefactor_per_m"]),
                                               1.0 / M.ToSI.eV(1.0 / float(spec["extinction_per_eV"])))
    if k == "GanachaudMokraniPhononInelasticSM":
        return M.GanachaudMokraniPhononInelasticSM(float(spec["ratemult"]), M.ToSI.eV(spec["phononE_eV"]),
                                                   float(spec["temperature_K"]), float(spec["eps0"]),
                                                   float(spec["epsInfinity"]))
    if k == "TabulatedInelasticSM":
        tables = [t if os.path.isabs(t) else os.path.join(base_dir, t) for t in spec["tables"]]
        m = M.TabulatedInelasticSM(mat, int(spec.get("method_se", 3)), tables,
                                   M.ToSI.eV(spec.get("energy_offset_eV", 0.0)))
        if spec.get("e0_from_dispersion"):
            m.setE0fromDispersion(True)
        if "min_egen_se_eV" in spec:
            m.setMinEgenSE(M.ToSI.eV(spec["min_egen_se_eV"]))
        if "rate_mult" in spec:
            m.setRateMult(float(spec["rate_mult"]))
        return m
    raise ValueError(f"unknown scatter mechanism {k!r}")


def _model(spec, mats, base_dir, sc):
    comp, mat = mats[spec["material"]]
    lam = (spec.get("barrier") or {}).get("lambda_nm")
    bar = M.ExpQMBarrierSM(mat) if lam is None else M.ExpQMBarrierSM(mat, float(lam) * NM)
    msm = M.MONSEL_MaterialScatterModel(mat, bar)
    held = [bar, msm]
    csd = _csd(spec.get("csd"), mat)
    if csd is not None:
        held.append(csd)
    for ms in spec.get("mechanisms", []):
        m = _mechanism(ms, mat, csd, base_dir, held)
        held.append(m)
        msm.addScatterMechanism(m)
    if csd is not None:
        msm.setCSD(csd)
    if spec.get("minEforTracking_eV") is not None:
        msm.setMinEforTracking(M.ToSI.eV(spec["minEforTracking_eV"]))
    sc.hold(*held)
    return msm


def _shape(name, specs, built, sc):
    if name in built:
        return built[name]
    s = specs[name]
    k = s["kind"]
    if k == "NormalMultiPlaneShape":
        shp = M.NormalMultiPlaneShape()
        for p in s["planes"]:
            shp.addPlane([float(v) for v in p["normal"]], [float(v) * NM for v in p["point_nm"]])
    elif k == "NormalCylindricalShape":
        shp = M.NormalCylindricalShape([v * NM for v in s["end0_nm"]], [v * NM for v in s["end1_nm"]],
                                       float(s["radius_nm"]) * NM)
    elif k == "line":
        shp = M.NShapes.createLine(s["topz_nm"] * NM, s["width_nm"] * NM, s["length_nm"] * NM,
                                   s["thetal_deg"] * math.pi / 180.0, s["thetar_deg"] * math.pi / 180.0,
                                   s["radl_nm"] * NM, s["radr_nm"] * NM, bool(s.get("reference_compat", True)))
    elif k == "step_edge":
        shp = M.NShapes.createStepEdge(s.get("z_nm", 0.0) * NM, s["step_nm"] * NM)
    elif k in ("NormalIntersectionShape", "NormalUnionShape"):
        a, b = (_shape(n, specs, built, sc) for n in s["shapes"])
        shp = (M.NormalIntersectionShape if k == "NormalIntersectionShape" else M.NormalUnionShape)(a, b)
    else:
        raise ValueError(f"unknown shape kind {k!r}")
    if "translate_nm" in s and k != "step_edge":
        M.NShapes.translate(shp, [v * NM for v in s["translate_nm"]])
    sc.hold(shp)
    built[name] = shp
    return shp


def load_scene(spec, base_dir="."):
    """Builds a scenes.Scene from a scene spec (a dict, or the path of a JSON
    file); sc.regions maps region names to their Region objects."""
    if isinstance(spec, str):
        base_dir = os.path.dirname(os.path.abspath(spec))
        spec = json.load(open(spec))
    sc = scenes.Scene()
    ch = scenes._chamber(sc)
    mats = {}
    for name, ms in spec.get("materials", {}).items():
        comp, mat = _material(dict(ms, name=ms.get("name", name)))
        sc.hold(comp, mat)
        mats[name] = (comp, mat)
    models = {"vacuum": None}
    for name, ms in spec.get("models", {}).items():
        models[name] = _model(ms, mats, base_dir, sc)
    built = {}
    regions = {"chamber": ch}
    for rs in spec["regions"]:
        parent = regions[rs.get("parent", "chamber")]
        model = models[rs["model"]] if rs["model"] != "vacuum" else scenes.vacuum(sc)
        reg = M.Region(parent, model, _shape(rs["shape"], spec["shapes"], built, sc))
        sc.hold(reg)
        regions[rs["name"]] = reg
        sc.regions[rs["name"]] = reg
    b = spec["beam"]
    c = [v * NM for v in b["center_nm"]]
    sc.gun = M.GaussianBeam(float(b["sigma_nm"]) * NM, M.ToSI.eV(b["energy_eV"]), c)
    sc.monte = M.MonteCarloSS(sc.gun, ch)
    sc.beam_z = c[2]
    sc.monte.setRefCompat(M.MONSEL_COMPAT_REFERENCE if spec.get("ref_compat", "reference") == "reference"
                          else M.MONSEL_COMPAT_FIXED)
    sc.spec = spec
    return sc


def job_centers(job, beam_z):
    """Pixel centres (m) of a job spec: a raster (x fastest, as
    LinesOnLayers.cu:351-357 loops) or an explicit list."""
    if "centers_nm" in job:
        return np.asarray(job["centers_nm"], np.float64) * NM
    xs = np.arange(*job["raster"]["x_nm"])
    ys = np.arange(*job["raster"]["y_nm"]) if "y_nm" in job["raster"] else np.array([0.0])
    return scenes.scan_centers(xs, ys, beam_z)


def run_job(job, base_dir=".", engine=None):
    """Runs a job spec (a dict, or the path of a JSON file) through the
    C-ABI and writes the outputs it names.  Returns (result, centres)."""
    from .abi import Engine
    if isinstance(job, str):
        base_dir = os.path.dirname(os.path.abspath(job))
        job = json.load(open(job))
    scene = job["scene"]
    sc = load_scene(scene if isinstance(scene, dict) else os.path.join(base_dir, scene), base_dir)
    cen = job_centers(job, sc.beam_z)
    E = float(job.get("beam_eV", sc.spec["beam"]["energy_eV"]))
    per = int(job["primaries_per_pixel"])
    nbins = int(job.get("energy_bins", max(1, int(E / 10.0))))
    own = engine is None
    eng = Engine() if own else engine
    try:
        sid = eng.upload(sc.monte.compileScene())
        t0 = time.perf_counter()
        r = eng.result(eng.submit(sid, cen, per, E, float(sc.spec["beam"]["sigma_nm"]) * NM, int(job.get("seed", 1)),
                                  nbins, float(job.get("se_cutoff_eV", 50.0))))
        dt = time.perf_counter() - t0
    finally:
        if own:
            eng.close()
    out = job.get("outputs", {})
    path = lambda p: p if os.path.isabs(p) else os.path.join(base_dir, p)
    if "text" in out:
        outputs.write_scan_text(path(out["text"]), E, cen, r["back"], r["se"], per,
                                sigma_nm=float(sc.spec["beam"]["sigma_nm"]), elapsed_s=dt)
    if "maps" in out:
        nx = len(np.arange(*job["raster"]["x_nm"]))
        outputs.save_maps(path(out["maps"]), r["back"], r["se"], per, nx, cen.shape[0] // nx)
    if "tallies" in out:
        outputs.save_tallies(path(out["tallies"]), r, cen, per, E)
    return r, cen


def line_scan_spec(h_nm=50.0, w_nm=20.0, theta_deg=3.0, r_nm=5.0, length_nm=1000.0, beam_eV=500.0, sigma_nm=0.5,
                   compat=True):
    """The line-scan scene (one PMMA line on Si, scenes.line_on_substrate)
    as a scene spec."""
    pmma = {"elements": ["C", "O", "H"], "mole_fractions": [5, 2, 8], "density": 1190.0, "workfunction_eV": 5.5,
            "bandgap_eV": 5.0, "cb_bottom_eV": -0.5}
    si = {"elements": ["Si"], "mole_fractions": [1.0], "density": 2330.0, "workfunction_eV": 4.85,
          "bandgap_eV": 1.1, "cb_bottom_eV": -3.75, "core_energies_eV": [99.2, 99.8, 149.7, 1839.0]}
    si_fit = scenes.SI_FITTED
    return {
        "materials": {"PMMA": pmma, "Silicon": si},
        "models": {
            "PMMA": {"material": "PMMA", "csd": {"kind": "JoyLuoNieminenCSD", "breakE_eV": 45.0},
                     "mechanisms": [{"kind": "SelectableElasticSM", "factory": "NISTMottRS"},
                                    {"kind": "FittedInelSM", "energySEgen_eV": 65.4},
                                    {"kind": "GanachaudMokraniPolaronTrapSM", "prefactor_per_m": 2.0e7,
                                     "extinction_per_eV": 0.25}]},
            "Si": {"material": "Silicon", "csd": {"kind": "JoyLuoNieminenCSD", "breakE_eV": si_fit["breakE_eV"]},
                   "mechanisms": [{"kind": "SelectableElasticSM", "factory": "NISTMottRS"},
                                  {"kind": "FittedInelSM", "energySEgen_eV": si_fit["energySEgen_eV"]},
                                  {"kind": "GanachaudMokraniPhononInelasticSM", "ratemult": 3.0,
                                   "phononE_eV": 0.063, "temperature_K": 300.0, "eps0": 11.7,
                                   "epsInfinity": 1.0}]},
        },
        "shapes": {
            "substrate": {"kind": "NormalMultiPlaneShape", "planes": [{"normal": [0, 0, -1], "point_nm": [0, 0, 0]}]},
            "line": {"kind": "line", "topz_nm": -h_nm, "width_nm": w_nm, "length_nm": length_nm,
                     "thetal_deg": theta_deg, "thetar_deg": theta_deg, "radl_nm": r_nm, "radr_nm": r_nm,
                     "reference_compat": compat},
        },
        "regions": [{"name": "substrate", "parent": "chamber", "model": "Si", "shape": "substrate"},
                    {"name": "line0", "parent": "chamber", "model": "PMMA", "shape": "line"}],
        "beam": {"energy_eV": beam_eV, "sigma_nm": sigma_nm, "center_nm": [0.0, 0.0, -(h_nm + 20.0)]},
        "ref_compat": "reference" if compat else "fixed",
    }


def main(argv=None):
    """python -m cudamonsel_amd.scenefile JOB.json [JOB.json ...]: runs each
    job file on GPU 0 and writes the outputs it names."""
    import sys
    paths = sys.argv[1:] if argv is None else argv
    if not paths:
        print(main.__doc__, file=sys.stderr)
        return 2
    for p in paths:
        r, cen = run_job(p)
        print(f"{p}: {cen.shape[0]} pixels, {r['stats']['primaries']} primaries, "
              f"{r['stats']['steps']} electron steps", flush=True)
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
