"""Scene builders for the benchmark and test workloads.

Materials and region trees are built with the reference's classes and
constructor arguments, as nanoscalemetrology/JMONSELTests/LinesOnLayers.cu:80-282
does.  All lengths in metres, energies in joules (reference units); the
helpers take nm / eV arguments and convert.

Si and Cu: the reference reads Si's TabulatedInelasticSM tables from files
that are not shipped (LinesOnLayers.cu:187-193) and defines no Cu model.  The
"Si" and "Cu" models here use the reference's FittedInelSM +
JoyLuoNieminenCSD mechanisms with the parameters below (a stand-in, stated in
every report); "Si-dielectric" and "Cu-dielectric" build the reference's
TabulatedInelasticSM form on tables computed from a dielectric model
(dielectric.py); si_tabulated() takes any four table files.
"""
import math

import numpy as np

from . import _monsel as M

NM = 1e-9
PI = math.pi

# stand-in fitted-inelastic parameters for materials without shipped tables;
# not calibrated against measured yields (in fixed mode, bulk Si at 1 keV
# gives SE 0.64 with them, as the dielectric-model tables do; bulk Cu at
# 2 keV gives SE 0.37, the tables 0.74)
SI_FITTED = dict(energySEgen_eV=63.0, breakE_eV=45.0)
CU_FITTED = dict(energySEgen_eV=36.0, breakE_eV=45.0)


class Scene:
    """Keeps every object of a scene alive and owns its MonteCarloSS."""

    def __init__(self):
        self.keep = []
        self.chamber = None
        self.gun = None
        self.monte = None
        self.regions = {}

    def hold(self, *objs):
        self.keep.extend(objs)
        return objs[0] if len(objs) == 1 else objs


def vacuum(sc):
    """LinesOnLayers.cu:80-84"""
    vac = M.SEmaterial()
    vac.setName("SE vacuum")
    bar = M.ExpQMBarrierSM(vac)
    msm = M.MONSEL_MaterialScatterModel(vac, bar)
    sc.hold(vac, bar, msm)
    return msm


def pmma(sc, min_e_eV=None, barrier_lambda_m=None, elastic_factory=None):
    """PMMA exactly as LinesOnLayers.cu:86-119; barrier_lambda_m selects the
    quantum-mechanical ExpQMBarrierSM(mat, lambda) instead of the classical
    one; elastic_factory another SelectableElasticSM factory than
    NISTMottRS.Factory."""
    breakE = M.ToSI.eV(45.0)
    density = 1190.0
    workfun, bandgap = 5.5, 5.0
    EFermi = -bandgap
    potU = -workfun - EFermi
    comp = M.Composition()
    comp.defineByMoleFraction([M.Elements.C, M.Elements.O, M.Elements.H], [5, 2, 8])
    mat = M.SEmaterial(comp, density)
    mat.setName("PMMA")
    mat.setWorkfunction(M.ToSI.eV(workfun))
    mat.setBandgap(M.ToSI.eV(bandgap))
    mat.setEnergyCBbottom(M.ToSI.eV(potU))
    el = M.SelectableElasticSM(mat, elastic_factory or M.NISTMottRS.Factory)
    csd = M.JoyLuoNieminenCSD(mat, breakE)
    fit = M.FittedInelSM(mat, M.ToSI.eV(65.4), csd)
    pol = M.GanachaudMokraniPolaronTrapSM(2.0e7, 1.0 / M.ToSI.eV(4.0))
    bar = M.ExpQMBarrierSM(mat) if barrier_lambda_m is None else M.ExpQMBarrierSM(mat, barrier_lambda_m)
    msm = M.MONSEL_MaterialScatterModel(mat, bar)
    msm.addScatterMechanism(el)
    msm.addScatterMechanism(fit)
    msm.addScatterMechanism(pol)
    msm.setCSD(csd)
    if min_e_eV is not None:
        msm.setMinEforTracking(M.ToSI.eV(min_e_eV))
    sc.hold(comp, mat, el, csd, fit, pol, bar, msm)
    return msm


def _si_material():
    """LinesOnLayers.cu:172-185"""
    density, workfun, bandgap = 2330.0, 4.85, 1.1
    EFermi = -bandgap
    potU = -workfun - EFermi
    mat = M.SEmaterial([M.Elements.Si], [1.0], density, "Silicon")
    mat.setWorkfunction(M.ToSI.eV(workfun))
    mat.setEnergyCBbottom(M.ToSI.eV(potU))
    mat.setBandgap(M.ToSI.eV(bandgap))
    mat.setCoreEnergy([M.ToSI.eV(99.2), M.ToSI.eV(99.8), M.ToSI.eV(149.7), M.ToSI.eV(1839.0)])
    return mat


def si(sc, phonon=True):
    """Si with NIST Mott elastic, the fitted-inelastic stand-in and the
    phonon mechanism of LinesOnLayers.cu:198."""
    mat = _si_material()
    el = M.SelectableElasticSM(mat, M.NISTMottRS.Factory)
    csd = M.JoyLuoNieminenCSD(mat, M.ToSI.eV(SI_FITTED["breakE_eV"]))
    fit = M.FittedInelSM(mat, M.ToSI.eV(SI_FITTED["energySEgen_eV"]), csd)
    bar = M.ExpQMBarrierSM(mat)
    msm = M.MONSEL_MaterialScatterModel(mat, bar)
    msm.addScatterMechanism(el)
    msm.addScatterMechanism(fit)
    objs = [mat, el, csd, fit, bar, msm]
    if phonon:
        ph = M.GanachaudMokraniPhononInelasticSM(3.0, M.ToSI.eV(0.063), 300.0, 11.7, 1.0)
        msm.addScatterMechanism(ph)
        objs.append(ph)
    msm.setCSD(csd)
    sc.hold(*objs)
    return msm


def si_tabulated(sc, tables, energy_offset_eV=13.54, method_se=3, e0_from_dispersion=False):
    """Si as LinesOnLayers.cu:195-212 with TabulatedInelasticSM on the given
    four table files (IIMFP, reduced deltaE, theta, SE E0)."""
    mat = _si_material()
    el = M.SelectableElasticSM(mat, M.NISTMottRS.Factory)
    tab = M.TabulatedInelasticSM(mat, method_se, list(tables), M.ToSI.eV(energy_offset_eV))
    if e0_from_dispersion:
        tab.setE0fromDispersion(True)
    ph = M.GanachaudMokraniPhononInelasticSM(3.0, M.ToSI.eV(0.063), 300.0, 11.7, 1.0)
    bar = M.ExpQMBarrierSM(mat)
    msm = M.MONSEL_MaterialScatterModel(mat, bar)
    msm.addScatterMechanism(el)
    msm.addScatterMechanism(tab)
    msm.addScatterMechanism(ph)
    sc.hold(mat, el, tab, ph, bar, msm)
    return msm


def cu(sc):
    """Cu: NIST Mott elastic + fitted-inelastic stand-in (no reference model)."""
    density, workfun = 8960.0, 4.65
    mat = M.SEmaterial([M.Elements.Cu], [1.0], density, "Copper")
    mat.setWorkfunction(M.ToSI.eV(workfun))
    mat.setEnergyCBbottom(M.ToSI.eV(-workfun - 7.0))
    el = M.SelectableElasticSM(mat, M.NISTMottRS.Factory)
    csd = M.JoyLuoNieminenCSD(mat, M.ToSI.eV(CU_FITTED["breakE_eV"]))
    fit = M.FittedInelSM(mat, M.ToSI.eV(CU_FITTED["energySEgen_eV"]), csd)
    bar = M.ExpQMBarrierSM(mat)
    msm = M.MONSEL_MaterialScatterModel(mat, bar)
    msm.addScatterMechanism(el)
    msm.addScatterMechanism(fit)
    msm.setCSD(csd)
    sc.hold(mat, el, csd, fit, bar, msm)
    return msm


def si_dielectric(sc, method_se=3):
    """Si as LinesOnLayers.cu:195-212 -- NIST Mott elastic, TabulatedInelasticSM
    (method 3, energy offset 13.54 eV) and the phonon mechanism -- on tables
    computed from a dielectric model (dielectric.py) in place of the
    reference's unshipped files."""
    from . import dielectric
    paths, offset = dielectric.make_tables("Si")
    return si_tabulated(sc, paths, energy_offset_eV=offset, method_se=method_se)


def cu_dielectric(sc, method_se=3):
    """Cu: NIST Mott elastic and TabulatedInelasticSM on dielectric-model
    tables (dielectric.py)."""
    from . import dielectric
    paths, offset = dielectric.make_tables("Cu")
    density, workfun = 8960.0, 4.65
    mat = M.SEmaterial([M.Elements.Cu], [1.0], density, "Copper")
    mat.setWorkfunction(M.ToSI.eV(workfun))
    mat.setEnergyCBbottom(M.ToSI.eV(-workfun - 7.0))
    mat.setEstimatedCoreEnergy()
    el = M.SelectableElasticSM(mat, M.NISTMottRS.Factory)
    tab = M.TabulatedInelasticSM(mat, method_se, list(paths), M.ToSI.eV(offset))
    bar = M.ExpQMBarrierSM(mat)
    msm = M.MONSEL_MaterialScatterModel(mat, bar)
    msm.addScatterMechanism(el)
    msm.addScatterMechanism(tab)
    sc.hold(mat, el, tab, bar, msm)
    return msm


MATERIALS = {"PMMA": pmma, "Si": si, "Cu": cu, "Si-dielectric": si_dielectric, "Cu-dielectric": cu_dielectric}


def _chamber(sc):
    """LinesOnLayers.cu:230-243"""
    null = M.NullMaterialScatterModel()
    sphere = M.Sphere([0.0, 0.0, 0.0], M.ChamberRadius)
    chamber = M.Region(None, null, sphere)
    vac = vacuum(sc)
    chamber.updateMaterial(null, vac)
    sc.hold(null, sphere, chamber)
    sc.chamber = chamber
    return chamber


def _half_space(z_nm=0.0):
    """NormalMultiPlaneShape with one plane, normal (0,0,-1): z >= z_nm
    (LinesOnLayers.cu:245-250)."""
    s = M.NormalMultiPlaneShape()
    s.addPlane([0.0, 0.0, -1.0], [0.0, 0.0, z_nm * NM])
    return s


def _finish(sc, beam_eV, sigma_nm, top_nm):
    center = [0.0, 0.0, -(top_nm + 20.0) * NM]
    gun = M.GaussianBeam(sigma_nm * NM, M.ToSI.eV(beam_eV), center)
    sc.gun = gun
    sc.monte = M.MonteCarloSS(gun, sc.chamber)
    sc.beam_z = center[2]
    return sc


def flat(material="Si", beam_eV=1000.0, sigma_nm=0.5, compat=True, barrier_lambda_m=None):
    """Bulk half-space z >= 0 of one material (barrier_lambda_m: PMMA only)."""
    sc = Scene()
    ch = _chamber(sc)
    msm = MATERIALS[material](sc) if barrier_lambda_m is None else pmma(sc, barrier_lambda_m=barrier_lambda_m)
    shape = _half_space(0.0)
    reg = M.Region(ch, msm, shape)
    sc.hold(shape, reg)
    sc.regions["substrate"] = reg
    _finish(sc, beam_eV, sigma_nm, 0.0)
    sc.monte.setRefCompat(M.MONSEL_COMPAT_REFERENCE if compat else M.MONSEL_COMPAT_FIXED)
    return sc


def line_on_substrate(h_nm=50.0, w_nm=20.0, theta_deg=3.0, r_nm=5.0, length_nm=1000.0, line="PMMA",
                      substrate="Si", beam_eV=500.0, sigma_nm=0.5, compat=True, xcenters_nm=(0.0,)):
    """Trapezoidal lines (NShapes::createLine) on a substrate half-space.
    compat=True builds the line as the reference does (NShapes.cu:140)."""
    sc = Scene()
    ch = _chamber(sc)
    sub_msm = MATERIALS[substrate](sc)
    sub = _half_space(0.0)
    sub_reg = M.Region(ch, sub_msm, sub)
    sc.hold(sub, sub_reg)
    sc.regions["substrate"] = sub_reg
    line_msm = MATERIALS[line](sc)
    th = theta_deg * PI / 180.0 if np.isscalar(theta_deg) else None
    for i, xc in enumerate(xcenters_nm):
        hh = h_nm[i] if not np.isscalar(h_nm) else h_nm
        ww = w_nm[i] if not np.isscalar(w_nm) else w_nm
        tt = (theta_deg[i] * PI / 180.0) if th is None else th
        shp = M.NShapes.createLine(-hh * NM, ww * NM, length_nm * NM, tt, tt, r_nm * NM, r_nm * NM, compat)
        if xc != 0.0:
            M.NShapes.translate(shp, [xc * NM, 0.0, 0.0])
        reg = M.Region(ch, line_msm, shp)
        sc.hold(reg)
        sc.regions[f"line{i}"] = reg
    top = h_nm if np.isscalar(h_nm) else max(h_nm)
    _finish(sc, beam_eV, sigma_nm, top)
    sc.monte.setRefCompat(M.MONSEL_COMPAT_REFERENCE if compat else M.MONSEL_COMPAT_FIXED)
    return sc


def random_lines(seed=3, n_lines=3, beam_eV=500.0, compat=True, r_nm=5.0):
    """Three parallel PMMA lines on Si with widths U[10,30], pitch U[40,60],
    heights U[30,80] nm and sidewall angles U[0,8] deg drawn from a seed;
    the middle line is centred at x = 0."""
    rng = np.random.default_rng(seed)
    widths = rng.uniform(10, 30, n_lines)
    pitch = rng.uniform(40, 60)
    heights = rng.uniform(30, 80, n_lines)
    angles = rng.uniform(0, 8, n_lines)
    xc = [(i - (n_lines - 1) / 2.0) * pitch for i in range(n_lines)]
    # the rounded top corner must fit inside each line
    r = min(r_nm, 0.45 * float(min(widths)))
    sc = line_on_substrate(h_nm=list(heights), w_nm=list(widths), theta_deg=list(angles), r_nm=r,
                           beam_eV=beam_eV, compat=compat, xcenters_nm=xc)
    sc.params = dict(widths=widths, pitch=pitch, heights=heights, angles=angles, xcenters=xc, radius=r)
    return sc


def lines_on_layers(h_nm=120.0, w_nm=80.0, theta_deg=3.0, r_nm=20.0, length_nm=1000.0, arc_nm=80.0,
                    beam_eV=500.0, sigma_nm=0.5, compat=True):
    """The geometry of LinesOnLayers::run (LinesOnLayers.cu:245-282): one
    PMMA line (createLine, h 120 nm, bottom width 80 nm, 3 deg, corner radius
    20 nm) standing on a PMMA layer ("ARC", the line region a child of the
    chamber) over Si.  The driver's third layer plane is placed at the second
    layer's position (:260), so Si starts right below the ARC and the glassy
    carbon layer between them is shadowed; the beam starts 20 nm above the
    line top (:356)."""
    sc = Scene()
    ch = _chamber(sc)
    arc = pmma(sc)
    arc_shape = _half_space(0.0)
    arc_reg = M.Region(ch, arc, arc_shape)
    si_msm = si(sc)
    si_shape = _half_space(arc_nm)
    si_reg = M.Region(arc_reg, si_msm, si_shape)
    sc.hold(arc_shape, arc_reg, si_shape, si_reg)
    line_msm = pmma(sc)
    th = theta_deg * PI / 180.0
    shp = M.NShapes.createLine(-h_nm * NM, w_nm * NM, length_nm * NM, th, th, r_nm * NM, r_nm * NM, compat)
    reg = M.Region(ch, line_msm, shp)
    sc.hold(reg)
    sc.regions.update(arc=arc_reg, substrate=si_reg, line0=reg)
    _finish(sc, beam_eV, sigma_nm, h_nm)
    sc.monte.setRefCompat(M.MONSEL_COMPAT_REFERENCE if compat else M.MONSEL_COMPAT_FIXED)
    return sc


def step_edge(material="Si", step_nm=20.0, beam_eV=1000.0, sigma_nm=0.5, compat=True):
    """Substrate z >= 0 joined with a step of height step_nm for x <= 0."""
    sc = Scene()
    ch = _chamber(sc)
    msm = MATERIALS[material](sc)
    shp = M.NShapes.createStepEdge(0.0, step_nm * NM)
    reg = M.Region(ch, msm, shp)
    sc.hold(reg)
    sc.regions["substrate"] = reg
    _finish(sc, beam_eV, sigma_nm, step_nm)
    sc.monte.setRefCompat(M.MONSEL_COMPAT_REFERENCE if compat else M.MONSEL_COMPAT_FIXED)
    return sc


def contact_holes(radius_nm=15.0, depth_nm=60.0, pitch_nm=None, beam_eV=500.0, compat=True, n=1):
    """A PMMA layer of thickness depth_nm on Si with n x n vacuum cylinder
    holes through it (region children with the vacuum model)."""
    sc = Scene()
    ch = _chamber(sc)
    si_msm = si(sc)
    sub = _half_space(depth_nm)
    sub_reg = M.Region(ch, si_msm, sub)
    sc.hold(sub, sub_reg)
    layer = M.NormalMultiPlaneShape()
    layer.addPlane([0.0, 0.0, -1.0], [0.0, 0.0, 0.0])
    layer.addPlane([0.0, 0.0, 1.0], [0.0, 0.0, depth_nm * NM])
    pm = pmma(sc)
    layer_reg = M.Region(ch, pm, layer)
    sc.hold(layer, layer_reg)
    vac = vacuum(sc)
    pitch = pitch_nm or 4.0 * radius_nm
    for i in range(n):
        for j in range(n):
            x = (i - (n - 1) / 2.0) * pitch
            y = (j - (n - 1) / 2.0) * pitch
            cyl = M.NormalCylindricalShape([x * NM, y * NM, 0.0], [x * NM, y * NM, depth_nm * NM], radius_nm * NM)
            reg = M.Region(layer_reg, vac, cyl)
            sc.hold(cyl, reg)
    sc.regions["layer"] = layer_reg
    sc.regions["substrate"] = sub_reg
    _finish(sc, beam_eV, 0.5, 0.0)
    sc.monte.setRefCompat(M.MONSEL_COMPAT_REFERENCE if compat else M.MONSEL_COMPAT_FIXED)
    return sc


def scan_centers(xs_nm, ys_nm, z_m):
    """Beam centres (m) for a raster, row-major in y then x (LinesOnLayers.cu:351-357)."""
    xs = np.asarray(xs_nm, dtype=np.float64) * NM
    ys = np.asarray(ys_nm, dtype=np.float64) * NM
    X, Y = np.meshgrid(xs, ys)
    return np.stack([X.ravel(), Y.ravel(), np.full(X.size, z_m)], axis=1)
