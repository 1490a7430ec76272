"""Oracle functions pinned by what the physics fixes, independently of how the
oracle writes them:

- TabulatedInelasticSM::computeE0fromDispersion (TabulatedInelasticSM.cu:103-151)
  returns the plasmon energy E0 of the dispersion relation
      dE^2 = E0^2 + (4/3) E_F(E0) Eq + Eq^2,
  E_F(E0) the Fermi energy of the free-electron gas whose plasma energy is E0
  (E_F = (hbar^2/2m)(3 pi^2 n)^(2/3), n = eps0 m E0^2/(hbar^2 e^2)), i.e.
  omega^2 = omega_p^2 + v_F^2 q^2 / 3 + (hbar q^2/2m)^2: checked from CODATA
  constants, on both branches of the cubic's solution.
- ScreenedRutherfordScatteringAngle (ScreenedRutherfordScatteringAngle.cu:24-39):
  the total cross-section is the integral of the screened Rutherford
  dsigma/dOmega = Z(Z+1) (e^2/4 pi eps0)^2 / (4 E^2 (1 - cos t + 2 a)^2)
  with Joy's screening a = 3.4 eV Z^(2/3) / E, i.e.
  pi Z(Z+1) (e^2/4 pi eps0)^2 / (4 E^2 a (1 + a)); sampled angles follow the
  normalised form of the same law (KS).  The oracle uses it at >= 20 keV
  (NISTMottRS.cu:116-118, 151-153).
- ExpQMBarrierSM with lambda > 0 (ExpQMBarrierSM.cu:21-48): the transmission
  equals that of the smooth step U(x) = dU / (1 + exp(-2 x / lambda)) found by
  integrating the Schroedinger equation numerically; it tends to the sharp
  step's 4 k1 k2 / (k1 + k2)^2 as lambda -> 0 and to 1 for wide barriers."""
import numpy as np
import pytest
from scipy import stats
from scipy.integrate import solve_ivp

J = 1.60217653e-19                 # C (the reference's PhysicalConstants)
HBAR = 1.054571817e-34             # CODATA 2018
ME = 9.1093837015e-31
QE = 1.602176634e-19
EPS0 = 8.8541878128e-12


def fermi_energy_of_plasma(E0):
    n = EPS0 * ME * E0 ** 2 / (HBAR ** 2 * QE ** 2)
    return HBAR ** 2 / (2 * ME) * (3 * np.pi ** 2 * n) ** (2.0 / 3.0)


def test_e0_from_dispersion_solves_the_penn_dispersion(oracle):
    L = oracle.lib()
    rng = np.random.default_rng(3)
    seen_trig = seen_cardano = False
    for _ in range(400):
        dE = rng.uniform(2.0, 300.0) * J
        Eq = rng.uniform(1e-4, 0.999) * dE
        E0 = L.oracle_e0_from_dispersion(Eq, dE)
        assert 0 < E0 <= dE
        rhs = E0 ** 2 + (4.0 / 3.0) * fermi_energy_of_plasma(E0) * Eq + Eq ** 2
        assert rhs / dE ** 2 == pytest.approx(1.0, rel=2e-6)
        # which branch of the cubic's solution was taken (c3 > 0: three real roots)
        x, y = Eq / dE, 3.77300614251479e17 * Eq
        if x * x * (27 + 4 * y) > 27:
            seen_trig = True
        else:
            seen_cardano = True
    assert seen_trig and seen_cardano
    # q -> 0: no dispersion
    assert L.oracle_e0_from_dispersion(0.0, 17 * J) == 17 * J
    assert L.oracle_e0_from_dispersion(1e-12 * J, 17 * J) == pytest.approx(17 * J, rel=1e-9)


def _rutherford_scene(M):
    from cudamonsel_amd import scenes
    sc = scenes.flat("PMMA", beam_eV=500.0)
    return sc, sc.monte.compileScene()


def test_screened_rutherford_cross_section(oracle, M):
    sc, cs = _rutherford_scene(M)
    osc = oracle.OracleScene(cs)
    k = QE ** 2 / (4 * np.pi * EPS0)
    E = np.geomspace(20.01e3, 1.0e6, 50)   # eV; the NIST Mott element switches to it at 20 keV
    for slot, Z in enumerate((6, 8, 1)):   # PMMA's elements in the material's order (C5O2H8)
        sig = osc.elastic_sigma(1, slot, E) * 1e-18   # nm^2 -> m^2
        EJ = E * J
        a = 5.44968e-19 * Z ** (2.0 / 3.0) / EJ
        want = np.pi * Z * (Z + 1) * k ** 2 / (4 * EJ ** 2 * a * (1 + a))
        assert np.allclose(sig, want, rtol=2e-4), (Z, np.max(np.abs(sig / want - 1)))
        # Joy's screening constant: 3.4e-3 keV
        assert 5.44968e-19 / J == pytest.approx(3.4015, rel=1e-4)


@pytest.mark.parametrize("E_keV", [20.5, 50.0, 300.0])
def test_screened_rutherford_angles(oracle, M, E_keV):
    sc, cs = _rutherford_scene(M)
    osc = oracle.OracleScene(cs)
    rng = np.random.default_rng(int(E_keV))
    n = 40000
    for slot, Z in enumerate((6, 8, 1)):
        E = np.full(n, E_keV * 1e3)
        th = osc.elastic_angle(1, slot, E, rng.random((n, 2)))
        t = 1.0 - np.cos(th)
        a = 5.44968e-19 * Z ** (2.0 / 3.0) / (E_keV * 1e3 * J)
        # P(1 - cos <= t) from the normalised density ~ 1/(t + 2a)^2 on [0, 2]
        cdf = lambda x: x * (1 + a) / (2 * a + x)
        assert stats.kstest(t, cdf).pvalue > 1e-3, (Z, E_keV)


def _smooth_step_T(lam, E, dU):
    """Transmission of U(x) = dU/(1 + exp(-2x/lam)) for an electron of
    energy E, by integrating the Schroedinger equation from the transmitted
    plane wave on the right to the incident one on the left."""
    a = lam / 2.0
    k1 = np.sqrt(2 * ME * E) / HBAR
    k2 = np.sqrt(2 * ME * (E - dU)) / HBAR
    X = 40 * a

    def f(x, y):
        psi = y[0] + 1j * y[1]
        d2 = 2 * ME * (dU / (1 + np.exp(-x / a)) - E) / HBAR ** 2 * psi
        return [y[2], y[3], d2.real, d2.imag]

    p0 = np.exp(1j * k2 * X)
    d0 = 1j * k2 * p0
    sol = solve_ivp(f, [X, -X], [p0.real, p0.imag, d0.real, d0.imag], rtol=1e-10, atol=1e-14, method="DOP853")
    psi = sol.y[0, -1] + 1j * sol.y[1, -1]
    dpsi = sol.y[2, -1] + 1j * sol.y[3, -1]
    A = (psi + dpsi / (1j * k1)) / 2 * np.exp(1j * k1 * X)
    return (k2 / k1) / abs(A) ** 2


@pytest.mark.parametrize("lam_nm", [0.01, 0.03, 0.1])
def test_smooth_barrier_transmission(oracle, lam_nm):
    L = oracle.lib()
    for E_eV, dU_eV in ((5.0, 4.9), (5.0, 3.0), (2.0, 0.5), (10.0, 9.0), (0.6, 0.5)):
        t = L.oracle_barrier_T(lam_nm * 1e-9, E_eV * J, dU_eV * J)
        want = _smooth_step_T(lam_nm * 1e-9, E_eV * J, dU_eV * J)
        assert t == pytest.approx(want, abs=2e-6), (lam_nm, E_eV, dU_eV)


def test_barrier_transmission_limits(oracle):
    L = oracle.lib()
    for E_eV, dU_eV in ((5.0, 4.9), (2.0, 0.5), (0.6, 0.5)):
        k1, k2 = np.sqrt(E_eV), np.sqrt(E_eV - dU_eV)
        sharp = 4 * k1 * k2 / (k1 + k2) ** 2
        assert L.oracle_barrier_T(0.0, E_eV * J, dU_eV * J) == pytest.approx(sharp, rel=1e-12)
        assert L.oracle_barrier_T(1e-16, E_eV * J, dU_eV * J) == pytest.approx(sharp, rel=1e-6)
        assert L.oracle_barrier_T(5e-9, E_eV * J, dU_eV * J) == pytest.approx(1.0, abs=1e-9)
    # below the step: reflected
    assert L.oracle_barrier_T(0.1e-9, 0.4 * J, 0.5 * J) == 0.0
