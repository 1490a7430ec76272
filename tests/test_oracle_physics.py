"""Scattering physics of the oracle, pinned by the shipped NIST SRD-64 tables,
published closed forms and the reference's own known-answer tests.

- NISTMottRS sigma at the table's energy nodes equals a0^2 * Spwem[j]
  (NISTMottRS.cu:114) and the angle at (node, U = k/200) equals
  acos(1 - 2 X1[j][k]^2) (NISTMottRS.cu:147-149);
- the Browning angle sampler follows the screened-Rutherford distribution
  dsigma/dOmega ~ (1 - mu + 2 alpha)^-2 mixed with isotropic scattering in
  the ratio r/(r+1) (Browning et al.), tested with a KS test;
- the Joy-Luo stopping power equals the Bethe form 2 pi e^4 N_A rho w Z /
  ((4 pi eps0)^2 A E) ln(1.166 E/J + beta), with J from the reference's
  MeanIonizationPotentialTest known answers (Wilson41: Si 0.161 keV, ...);
- the Nieminen branch is the exact solution of dE/ds = -gamma E^2.5 (semigroup
  property) and joins the Joy-Luo branch continuously at breakE."""
import math
import os
import struct

import numpy as np
import pytest
from scipy import stats

A0_NM = 0.05291772083
EV = 1.60217653e-19


def load_mott(Z):
    from cudamonsel_amd import DATA_DIR
    buf = open(os.path.join(DATA_DIR, "nistmott.bin"), "rb").read()
    count = struct.unpack_from("<I", buf, 8)[0]
    off = 12
    for _ in range(count):
        z = struct.unpack_from("<i", buf, off)[0]
        off += 4
        sp = np.frombuffer(buf, "<f8", 61, off)
        off += 61 * 8
        x1 = np.frombuffer(buf, "<f8", 61 * 201, off).reshape(61, 201)
        off += 61 * 201 * 8
        if z == Z:
            return sp, x1
    raise KeyError(Z)


@pytest.fixture(scope="module")
def pmma(oracle):
    from cudamonsel_amd import scenes
    sc = scenes.flat("PMMA", beam_eV=500.0)
    cs = sc.monte.compileScene()
    return sc, cs, oracle.OracleScene(cs)


@pytest.fixture(scope="module")
def si(oracle):
    from cudamonsel_amd import scenes
    sc = scenes.flat("Si", beam_eV=1000.0)
    cs = sc.monte.compileScene()
    return sc, cs, oracle.OracleScene(cs)


def node_energies_eV():
    return 50.0 * 400.0 ** (np.arange(61) / 60.0)


@pytest.mark.parametrize("slot,Z", [(0, 6), (1, 8), (2, 1)])
def test_mott_sigma_at_nodes(pmma, slot, Z):
    _, _, osc = pmma
    sp, _ = load_mott(Z)
    E = node_energies_eV()[:60]
    got = osc.elastic_sigma(1, slot, E)
    assert np.allclose(got, A0_NM ** 2 * sp[:60], rtol=1e-10, atol=0)


@pytest.mark.parametrize("slot,Z", [(0, 6), (2, 1)])
def test_mott_angle_at_nodes(pmma, slot, Z):
    _, _, osc = pmma
    _, x1 = load_mott(Z)
    E = node_energies_eV()
    rng = np.random.default_rng(Z)
    js = rng.integers(0, 60, 40)
    ks = rng.integers(1, 200, 40)
    u = np.stack([ks / 200.0, np.zeros(40)], axis=1)
    got = osc.elastic_angle(1, slot, E[js], u)
    q = x1[js, ks]
    assert np.allclose(np.cos(got), 1 - 2 * q * q, rtol=0, atol=1e-9)


def test_elastic_sigma_continuous_at_extrapolation_energy(pmma):
    _, _, osc = pmma
    E = np.array([50.0 * (1 - 1e-12), 50.0])
    for slot in range(3):
        s = osc.elastic_sigma(1, slot, E)
        assert s[0] == pytest.approx(s[1], rel=1e-9)


@pytest.mark.parametrize("E_eV,slot,Z", [(20.0, 0, 6), (5.0, 1, 8), (35.0, 2, 1)])
def test_browning_angle_distribution(pmma, E_eV, slot, Z):
    _, _, osc = pmma
    n = 40000
    rng = np.random.default_rng(11)
    u = rng.random((n, 2))
    th = osc.elastic_angle(1, slot, np.full(n, E_eV), u)
    e = E_eV * 1e-3
    r = 300.0 * e / Z + Z ** 3 / (3e5 * e)
    pr = r / (r + 1)
    a = 7.0e-3 / e

    def cdf(mu):
        rut = 1.0 - (1.0 - mu) * (1.0 + a) / (1.0 + 2.0 * a - mu)
        iso = (1.0 + mu) / 2.0
        return pr * rut + (1.0 - pr) * iso

    p = stats.kstest(np.cos(th), cdf).pvalue
    assert p > 1e-3


def bethe_K():
    e, eps0, NA = 1.60217653e-19, 8.8541878176203898505365630317108e-12, 6.0221415e23
    return 2 * math.pi * e ** 4 / (4 * math.pi * eps0) ** 2 * NA * 1000.0


def test_wilson41_known_answers():
    # MeanIonizationPotentialTest.cu:72-77: Wilson41 Si 0.161, Al 0.149, Ca 0.230, Ti 0.253 keV (+-0.001)
    for Z, want in ((14, 0.161), (13, 0.149), (20, 0.230), (22, 0.253)):
        assert abs(11.5 * Z * 1e-3 - want) <= 0.001 + 1e-12


def stopping_power(E_eV, Z, w, A, rho, J_eV, bh_eV):
    """Bethe form with Joy-Luo's 1.166 and the barrier-shifted beta (eV/nm)."""
    K = bethe_K()
    E = E_eV * EV
    tot = 0.0
    for z, wi, ai, j in zip(Z, w, A, J_eV):
        recipJ = 1.166 / (j * EV)
        beta = 1.0 - recipJ * (bh_eV + 1.0) * EV
        tot += K * rho * wi * z / ai * math.log(recipJ * E + beta)
    return tot / E / EV * 1e-9


def test_joy_luo_stopping_power_pmma(pmma, M):
    _, _, osc = pmma
    comp = [M.Elements.C, M.Elements.O, M.Elements.H]
    Z = [6, 8, 1]
    A = [el.getAtomicWeight() for el in comp]
    x = np.array([5.0, 2.0, 8.0])
    w = x * np.array(A) / np.dot(x, A)
    J = [11.5 * z for z in Z]
    for E in (60.0, 200.0, 500.0, 2000.0, 10000.0):
        s = -osc.csd(1, [1e-3], [E])[0] / 1e-3
        assert s == pytest.approx(stopping_power(E, Z, w, A, 1190.0, J, 0.5), rel=2e-3)


def test_joy_luo_stopping_power_si_uses_sternheimer(si):
    _, cs, osc = si
    J = [9.76 * 14 + 58.8 * 14 ** -0.19]
    for E in (100.0, 1000.0, 5000.0):
        s = -osc.csd(1, [1e-3], [E])[0] / 1e-3
        assert s == pytest.approx(stopping_power(E, [14], [1.0], [28.0855], 2330.0, J, 3.75), rel=2e-3)


def test_nieminen_is_exact_ode_solution(pmma):
    _, _, osc = pmma
    for E in (2.0, 10.0, 30.0, 44.0):
        for s1, s2 in ((0.3, 0.5), (2.0, 7.0)):
            d1 = osc.csd(1, [s1], [E])[0]
            d2 = osc.csd(1, [s2], [E + d1])[0]
            d12 = osc.csd(1, [s1 + s2], [E])[0]
            if E + d1 > 0.5:
                assert d1 + d2 == pytest.approx(d12, rel=1e-9, abs=1e-12)


def test_nieminen_joins_joy_luo_at_break_energy(pmma):
    _, _, osc = pmma
    h = 1e-6
    below = -osc.csd(1, [h], [45.0 * (1 - 1e-9)])[0] / h
    above = -osc.csd(1, [h], [45.0 * (1 + 1e-9)])[0] / h
    assert below == pytest.approx(above, rel=1e-5)


def test_csd_halving_for_large_losses(pmma):
    _, _, osc = pmma
    # 1 keV over 50 nm loses far more than 10%: the step is halved recursively
    E, L = 1000.0, 50.0
    whole = osc.csd(1, [L], [E])[0]
    h1 = osc.csd(1, [L / 2], [E])[0]
    h2 = osc.csd(1, [L / 2], [E + h1])[0]
    assert whole == pytest.approx(h1 + h2, rel=1e-12)
    assert whole < 0 and -whole < E


def test_fitted_polaron_rates(pmma):
    _, _, osc = pmma
    E = np.array([5.0, 40.0, 58.0, 58.1, 60.0, 61.0, 100.0, 500.0])
    r = osc.rates(1, E, 3)
    # FittedInelSM.cu:20-25: stopping power over a 0.1 nm step / energySEgen
    for i, e in enumerate(E):
        if e <= 65.4 - 5.0:
            assert r[i, 1] == 0.0
        else:
            sp = -osc.csd(1, [0.1], [e])[0] / 0.1
            assert r[i, 1] == pytest.approx(sp / 65.4, rel=1e-12)
    # GanachaudMokraniPolaronTrapSM.cu:21-27: 2e7 /m * exp(-E / 4 eV) below the cutoff
    cutoff = -math.log(10.0 / 2e7) * 4.0
    for i, e in enumerate(E):
        want = 0.02 * math.exp(-e / 4.0) if e <= cutoff else 0.0
        assert r[i, 2] == pytest.approx(want, rel=1e-9, abs=1e-300)
    assert np.allclose(r[:, 3], r[:, 0] + r[:, 1] + r[:, 2], rtol=1e-12)


def test_phonon_rate_closed_form(si):
    _, _, osc = si
    E = np.array([0.05, 0.07, 0.5, 1.0, 5.0])
    r = osc.rates(1, E, 3)
    kT = 1.3806505e-23 * 300.0 / EV
    occ = 0.5 * (1 + 1 / (math.exp(0.063 / kT) - 1))
    pref = 3.0 * occ * (11.7 - 1.0) / 11.7 / 1.0 / A0_NM
    for i, e in enumerate(E):
        x = 0.063 / e
        if e < 0.063 or x >= 1:
            want = 0.0
        elif x < 0.1:
            want = pref * x * math.log(4 / x)
        else:
            t = math.sqrt(1 - x)
            want = pref * x * math.log((1 + t) / (1 - t))
        assert r[i, 2] == pytest.approx(want, rel=1e-9, abs=1e-300)
