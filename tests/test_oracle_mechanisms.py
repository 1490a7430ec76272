"""Pins the oracle's scatter mechanisms and barrier crossing (the CPU
restatement the GPU kernels are checked against) with closed forms and
invariants that do not re-use the restated formulas:

- ExpQMBarrierSM (ExpQMBarrierSM.cu:98-219): the classical threshold
  E cos^2(alpha) > dU, conservation of the tangential momentum sqrt(E) sin(alpha)
  on transmission, mirror reflection, the compat-mode defect of line 111
  (normal := direction), and the textbook step-potential transmission
  probability 4 k1 k2 / (k1 + k2)^2 for the sharp quantum barrier;
- FittedInelSM (FittedInelSM.cu:13-18): isotropic SE at E_SEgen + E_F;
- GanachaudMokraniPolaronTrapSM: trapping;
- GanachaudMokraniPhononInelasticSM (:19-44): the sampled cos(theta) follows
  the 1/q^2 density 1/(a - b cos(theta)), a = 2 - x, b = 2 sqrt(1 - x), x = hw/E
  (two-sample KS against its inverse CDF), and E loses hw;
- SelectableElasticSM (:49-67): element choice in proportion to
  sigma_i * n_i (PMMA C5O2H8: n = 5, 2, 8), deflection = the element's
  scattering angle;
- TabulatedInelasticSM (:209-434): with designed constant tables, the PE
  loses dE = c (E + offset), is deflected by theta0, and the SE has energy
  dE + E0 - offset at pi/2 - theta0 from the PE's original direction."""
import numpy as np
import pytest
from scipy import stats


def _unit(rng, n):
    d = rng.normal(size=(n, 3))
    return d / np.linalg.norm(d, axis=1)[:, None]


def _scene(M, compat, barrier_lambda=None):
    from cudamonsel_amd import scenes
    return scenes.flat("PMMA", beam_eV=500.0, compat=compat, barrier_lambda_m=barrier_lambda)


def _oracle_scene(oracle, sc):
    cs = sc.monte.compileScene()
    return oracle.OracleScene(cs), cs


# PMMA conduction-band bottom: -workfunction - E_F = -5.5 + 5.0 eV
# (LinesOnLayers.cu:86-95); vacuum 0 -> dU leaving PMMA = +0.5 eV
DU_PMMA_TO_VACUUM = 0.5


def test_barrier_fixed_classical(oracle, M):
    sc = _scene(M, compat=False)
    osc, cs = _oracle_scene(oracle, sc)
    rng = np.random.default_rng(1)
    n = 4000
    d = _unit(rng, n)
    nb = _unit(rng, n)
    E = rng.uniform(0.05, 3.0, n)
    r = osc.barrier(1, 0, d, nb, E, np.full(n, 0.5))
    cosa = np.sum(d * nb, axis=1)
    back = cosa <= 0
    assert np.all(r["sign"][back] == -1) and np.all(r["region"][back] == 1)
    assert np.allclose(r["dir"][back], d[back], atol=1e-12) and np.allclose(r["E"][back], E[back])
    trans = (~back) & (E * cosa ** 2 > DU_PMMA_TO_VACUUM)
    refl = (~back) & ~trans
    assert trans.sum() > 200 and refl.sum() > 200
    # reflection: mirror image, energy kept
    mirror = d - 2 * cosa[:, None] * nb
    assert np.allclose(r["dir"][refl], mirror[refl], atol=1e-9)
    assert np.allclose(r["E"][refl], E[refl], rtol=1e-12)
    assert np.all(r["region"][refl] == 1) and np.all(r["sign"][refl] == -1)
    # transmission: E - dU, tangential momentum conserved, normal momentum
    # sqrt(E cos^2 - dU)
    Et = E[trans] - DU_PMMA_TO_VACUUM
    assert np.allclose(r["E"][trans], Et, rtol=1e-9, atol=1e-12)
    nf = r["dir"][trans]
    n0 = d[trans]
    nn = nb[trans]
    ptan_in = np.sqrt(E[trans])[:, None] * (n0 - cosa[trans][:, None] * nn)
    ptan_out = np.sqrt(Et)[:, None] * (nf - np.sum(nf * nn, axis=1)[:, None] * nn)
    assert np.allclose(ptan_in, ptan_out, atol=1e-9)
    assert np.allclose(np.sqrt(Et) * np.sum(nf * nn, axis=1),
                       np.sqrt(E[trans] * cosa[trans] ** 2 - DU_PMMA_TO_VACUUM), atol=1e-9)
    assert np.all(r["region"][trans] == 0) and np.all(r["sign"][trans] == 1)


def test_barrier_compat_uses_direction_as_normal(oracle, M):
    """ExpQMBarrierSM.cu:111: the normal is replaced by the direction, so the
    crossing is at normal incidence whatever the surface."""
    sc = _scene(M, compat=True)
    osc, cs = _oracle_scene(oracle, sc)
    rng = np.random.default_rng(2)
    n = 2000
    d = _unit(rng, n)
    nb = _unit(rng, n)
    E = rng.uniform(0.05, 2.0, n)
    r = osc.barrier(1, 0, d, nb, E, np.full(n, 0.5))
    trans = E > DU_PMMA_TO_VACUUM
    assert np.allclose(r["dir"][trans], d[trans], atol=1e-9)
    assert np.allclose(r["E"][trans], E[trans] - DU_PMMA_TO_VACUUM, atol=1e-12)
    assert np.allclose(r["dir"][~trans], -d[~trans], atol=1e-9)
    assert np.all(r["region"][trans] == 0) and np.all(r["region"][~trans] == 1)


def test_barrier_sharp_quantum_transmission(oracle, M):
    """lambda = 0: transmission iff u < 4 k1 k2 / (k1 + k2)^2 (the step
    potential), k1 = sqrt(E cos^2 a), k2 = sqrt(E cos^2 a - dU)."""
    sc = _scene(M, compat=False, barrier_lambda=0.0)
    osc, cs = _oracle_scene(oracle, sc)
    rng = np.random.default_rng(3)
    n = 1000
    nb = np.tile([0.0, 0.0, -1.0], (n, 1))
    ct = rng.uniform(0.3, 1.0, n)
    st = np.sqrt(1 - ct ** 2)
    d = np.stack([st, np.zeros(n), -ct], 1)
    E = rng.uniform(0.6, 3.0, n) / ct ** 2
    k1 = np.sqrt(E * ct ** 2)
    k2 = np.sqrt(E * ct ** 2 - DU_PMMA_TO_VACUUM)
    T = 4 * k1 * k2 / (k1 + k2) ** 2
    below = osc.barrier(1, 0, d, nb, E, T * (1 - 1e-7))
    above = osc.barrier(1, 0, d, nb, E, np.minimum(T * (1 + 1e-7), 1 - 1e-12))
    assert np.all(below["region"] == 0)
    assert np.all(above["region"][T < 1 - 1e-6] == 1)


def test_fitted_inelastic_and_polaron(oracle, M):
    from cudamonsel_amd import scenes
    sc = scenes.flat("PMMA", beam_eV=500.0)
    osc, cs = _oracle_scene(oracle, sc)
    rng = np.random.default_rng(4)
    n = 500
    d = _unit(rng, n)
    u = rng.random((n, 6))
    E = rng.uniform(80, 500, n)
    r = osc.scatter(1, 1, E, E, d, u)
    assert np.all(r["has"]) and np.all(r["draws"] == 2)
    # E_SEgen + E_F = 65.4 - 5.0 eV (LinesOnLayers.cu:108-109, SEmaterial.cu:392-399)
    assert np.allclose(r["sE"], 60.4, rtol=1e-12)
    assert np.allclose(r["dir"], d) and np.allclose(r["E"], E)
    # theta = acos(1 - 2 u1), phi = 2 pi u0
    assert np.allclose(r["sdir"][:, 2], 1 - 2 * u[:, 1], atol=1e-12)
    phi = np.mod(np.arctan2(r["sdir"][:, 1], r["sdir"][:, 0]), 2 * np.pi)
    ok = np.sin(np.arccos(np.clip(1 - 2 * u[:, 1], -1, 1))) > 1e-6
    dphi = np.angle(np.exp(1j * (phi - 2 * np.pi * u[:, 0])))
    assert np.max(np.abs(dphi[ok])) < 1e-9
    p = osc.scatter(1, 2, E, E, d, u)
    assert np.all(p["trapped"]) and np.all(p["E"] == 0) and not np.any(p["has"])


def test_phonon_angle_follows_inverse_q2_density(oracle, M):
    from cudamonsel_amd import scenes
    sc = scenes.flat("Si", beam_eV=1000.0)
    osc, cs = _oracle_scene(oracle, sc)
    rng = np.random.default_rng(5)
    hw = 0.063
    for E0 in (0.07, 0.2, 0.6, 3.0):  # x = hw/E spans both sampler branches
        n = 20000
        d = np.tile([0.0, 0.0, 1.0], (n, 1))
        u = rng.random((n, 6))
        E = np.full(n, E0)
        r = osc.scatter(1, 2, E, E, d, u)
        assert np.allclose(r["E"], E0 - hw, rtol=1e-12)
        c = r["dir"][:, 2]
        x = hw / E0
        a, b = 2 - x, 2 * np.sqrt(1 - x)
        # inverse CDF of the density 1/(a - b c) on [-1, 1]
        v = rng.random(n)
        ref = (a - (a + b) * ((a - b) / (a + b)) ** v) / b
        assert stats.ks_2samp(c, ref).pvalue > 1e-3, E0


def test_elastic_element_choice_and_deflection(oracle, M):
    from cudamonsel_amd import scenes
    sc = scenes.flat("PMMA", beam_eV=500.0)
    osc, cs = _oracle_scene(oracle, sc)
    n = 3000
    E = 300.0
    sig = np.array([osc.elastic_sigma(1, k, [E])[0] for k in range(3)])
    w = sig * np.array([5.0, 2.0, 8.0])
    cum = np.cumsum(w) / w.sum()
    rng = np.random.default_rng(6)
    u = rng.random((n, 6))
    d = _unit(rng, n)
    r = osc.scatter(1, 0, np.full(n, E), np.full(n, E - 1.0), d, u)
    assert np.allclose(np.linalg.norm(r["dir"], axis=1), 1.0)
    defl = np.arccos(np.clip(np.sum(r["dir"] * d, axis=1), -1, 1))
    elem = np.searchsorted(cum, u[:, 0])
    ang = np.empty(n)
    for k in range(3):
        m = elem == k
        ang[m] = osc.elastic_angle(1, k, np.full(m.sum(), E), u[m][:, 1:3])
    assert np.allclose(defl, ang, atol=1e-6)
    assert np.allclose(r["E"], E - 1.0, rtol=1e-12) and not np.any(r["has"])


def test_tabulated_with_designed_tables(oracle, M, tmp_path):
    from cudamonsel_amd import scenes, tables
    EV = tables.EV
    c, theta0, e0 = 0.1, 0.3, 5.0
    eg = np.geomspace(20.0, 5000.0, 12) * EV
    u = np.linspace(0, 1, 6)
    paths = [str(tmp_path / f"d_{k}.csv") for k in ("iimfp", "rde", "theta", "see0")]
    tables.write_nu_table(paths[0], [eg], np.full(eg.size, 1e9))
    tables.write_nu_table(paths[1], [eg, u], np.full((eg.size, u.size), c))
    tables.write_nu_table(paths[2], [eg, u, u], np.full((eg.size, u.size, u.size), theta0))
    tables.write_nu_table(paths[3], [eg, u], np.full((eg.size, u.size), e0 * EV))
    sc = scenes.Scene()
    ch = scenes._chamber(sc)
    msm = scenes.si_tabulated(sc, paths, energy_offset_eV=13.54, method_se=2)
    shp = scenes._half_space(0.0)
    reg = M.Region(ch, msm, shp)
    sc.hold(shp, reg)
    scenes._finish(sc, 1000.0, 0.5, 0.0)
    sc.monte.setRefCompat(M.MONSEL_COMPAT_FIXED)
    osc, cs = _oracle_scene(oracle, sc)
    rng = np.random.default_rng(7)
    n = 400
    E = rng.uniform(60.0, 700.0, n)  # dE = 0.1 (E + 13.54) stays below the 99.2 eV core level
    d = _unit(rng, n)
    uu = rng.random((n, 6))
    r = osc.scatter(1, 1, E, E, d, uu)
    dE = c * (E + 13.54)
    assert np.allclose(r["E"], E - dE, rtol=1e-9)
    assert np.allclose(np.arccos(np.clip(np.sum(r["dir"] * d, axis=1), -1, 1)), theta0, atol=1e-7)
    # an SE is kept when its energy above the band bottom, E_SE + CB with
    # CB = -4.85 + 1.1 eV (LinesOnLayers.cu:172-185), reaches minEgenSE (0)
    esE = dE + e0 - 13.54
    keep = esE - 3.75 >= 0.0
    assert keep.sum() > 300 and (~keep).sum() > 10
    assert np.array_equal(r["has"], keep)
    assert np.allclose(r["sE"][keep], esE[keep], rtol=1e-9)
    sang = np.arccos(np.clip(np.sum(r["sdir"][keep] * d[keep], axis=1), -1, 1))
    assert np.allclose(sang, np.pi / 2 - theta0, atol=1e-7)
