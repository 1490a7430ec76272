"""Inelastic scattering tables for TabulatedInelasticSM from a dielectric model.

The reference's driver reads four tables per material (LinesOnLayers.cu:187-196;
IIMFPFullPennInterpSiSI.csv and the like) that are not shipped with it.  This
module computes tables of the same meaning, layout and units from a model
energy-loss function (ELF) and writes them in the NUTableInterpolation format
(tables.write_nu_table), so that TabulatedInelasticSM runs on physical Si and
Cu data:

  IIMFP     (E)           inverse inelastic mean free path, 1/m
  RedDeltaE (E, U)        energy loss / E at cumulative probability U
  Theta     (E, r, U)     polar scattering angle (rad) for the reduced loss
                          r = (dE - gap) / (E - E_F - 2 gap)
                          (TabulatedInelasticSM.cu:254-262) at cumulative
                          probability U of the momentum transfer
  SEE0      (dE, U)       initial energy (J, above the band bottom) of the
                          electron a loss dE excites, at probability U

Energies E are kinetic energies above the bottom of the valence band, the
convention of the reference's energyOffset (E_band = E_CB + offset,
TabulatedInelasticSM.cu:211-212); abscissae are in J.

Model (all in atomic units inside):
- ELF at q = 0: one Drude-Lorentz oscillator per valence band and a power-law
  edge per inner shell,
      Im[-1/eps(0, w)] = A w0 g w / ((w0^2 - w^2)^2 + (g w)^2),  w > gap,
      Im[-1/eps(0, w)] = A (E_edge / w)^3,                         w >= E_edge,
  each scaled so that the Bethe sum rule holds for its electrons,
      (2 / pi) integral w Im[-1/eps(0, w)] dw = w_p^2 = 4 pi n_shell.
- Dispersion: Penn's single-pole approximation (D. R. Penn, Phys. Rev. B 35,
  482 (1987); as used by Z.-J. Ding and R. Shimizu, Scanning 18, 92 (1996)):
      Im[-1/eps(q, w)] = (w0 / w) Im[-1/eps(0, w0)],  w0^2 = w^2 - (q^2 / 2)^2.
- Differential inverse mean free path of an electron of energy E
  (non-relativistic first Born approximation):
      d(1/lambda)/dw = 1 / (pi E) integral_{q-}^{q+} dq / q Im[-1/eps(q, w)],
      q+- = sqrt(2E) +- sqrt(2(E - w)),   w <= E - E_F.
- The secondary's initial energy E0 below the Fermi level follows the joint
  density of free-electron states sqrt(E0) sqrt(E0 + dE) on
  [max(0, E_F - dE), E_F] (Ding and Shimizu 1996).

Parameters: Si -- 4 valence electrons per atom (plasma energy 16.6 eV for
2330 kg/m^3), oscillator at the bulk plasmon 16.7 eV with width 3.7 eV
(H. Raether, Excitation of Plasmons and Interband Transitions by Electrons,
Springer 1980), gap 1.1 eV, L1 / L2,3 / K shells (2 / 6 / 2 electrons) at the
edge energies of the Chantler table (148.7, 99.2, 1838.9 eV); its IMFP
follows TPP-2M within 4 % from 0.2 to 10 keV with no parameter fitted.
Cu -- 11 valence electrons (3d10 4s1) in a broad oscillator at 35 eV (width
40 eV) and the 8 M-shell electrons in one at 300 eV (width 400 eV), whose
positions and widths were chosen so that the IMFP follows TPP-2M within 6 %
from 0.1 to 10 keV, then L2,3 / L1 / K shells (6 / 2 / 2 electrons) at
931.1, 1096.1, 8978.9 eV; its Fermi level is that of one conduction
electron per atom (7.0 eV), at the bottom of whose band the energies start
(no energy offset).  These are models, not the reference's
tables: transport on them is not comparable with the reference's Si results
(which need its unshipped files) -- the IMFP is checked against the TPP-2M
predictive formula (S. Tanuma, C. J. Powell, D. R. Penn, Surf. Interface
Anal. 21, 165 (1994)) instead.
"""
import hashlib
import json
import os

import numpy as np

from . import tables

HARTREE_EV = 27.211386245988
BOHR_M = 5.29177210903e-11
EV_J = 1.60217653e-19
N_A = 6.02214076e23

MATERIALS = {
    "Si": dict(density=2330.0, molar_mass=28.0855, gap_eV=1.1, energy_offset_eV=13.54,
               valence=[dict(electrons=4.0, w0_eV=16.7, gamma_eV=3.7)],
               shells=[dict(electrons=6.0, edge_eV=99.2), dict(electrons=2.0, edge_eV=148.7),
                       dict(electrons=2.0, edge_eV=1838.9)]),
    "Cu": dict(density=8960.0, molar_mass=63.546, gap_eV=0.0, energy_offset_eV=0.0, fermi_electrons=1.0,
               valence=[dict(electrons=11.0, w0_eV=35.0, gamma_eV=40.0),
                        dict(electrons=8.0, w0_eV=300.0, gamma_eV=400.0)],
               shells=[dict(electrons=6.0, edge_eV=931.1), dict(electrons=2.0, edge_eV=1096.1),
                       dict(electrons=2.0, edge_eV=8978.9)]),
}


def atom_density(mat):
    """atoms / bohr^3"""
    n_m3 = mat["density"] * 1e3 / mat["molar_mass"] * N_A
    return n_m3 * BOHR_M ** 3


class ELF:
    """Im[-1/eps(0, w)] of a material model (w in Hartree)."""

    def __init__(self, mat):
        self.mat = mat
        n_at = atom_density(mat)
        self.gap = mat["gap_eV"] / HARTREE_EV
        w = np.geomspace(1e-4, 1e4, 400001)
        self.terms = []
        for osc in mat["valence"]:
            w0, g = osc["w0_eV"] / HARTREE_EV, osc["gamma_eV"] / HARTREE_EV
            shape = lambda x, w0=w0, g=g: np.where(x > self.gap, w0 * g * x / ((w0 ** 2 - x ** 2) ** 2 + (g * x) ** 2),
                                                   0.0)
            self.terms.append(self._normalised(shape, osc["electrons"] * n_at, w))
        for sh in mat["shells"]:
            e = sh["edge_eV"] / HARTREE_EV
            shape = lambda x, e=e: np.where(x >= e, (e / np.maximum(x, 1e-300)) ** 3, 0.0)
            self.terms.append(self._normalised(shape, sh["electrons"] * n_at, w))

    @staticmethod
    def _normalised(shape, n, w):
        wp2 = 4.0 * np.pi * n
        integral = np.trapezoid(w * shape(w), w)
        scale = (np.pi / 2.0) * wp2 / integral
        return lambda x: scale * shape(x)

    def __call__(self, w):
        w = np.asarray(w, np.float64)
        return sum(t(w) for t in self.terms)

    def sum_rule_electrons(self, w_max_eV=1e6):
        """Electrons per atom recovered by the Bethe sum rule up to w_max."""
        w = np.geomspace(1e-4, w_max_eV / HARTREE_EV, 400001)
        wp2 = (2.0 / np.pi) * np.trapezoid(w * self(w), w)
        return wp2 / (4.0 * np.pi * atom_density(self.mat))


def diimfp(elf, E, w, e_fermi, nq=160):
    """d(1/lambda)/dw (1/bohr per Hartree) for energies E (Hartree, above the
    band bottom) and losses w (Hartree); shape (len(E), len(w)).  Zero where
    the loss leaves the electron below the Fermi level."""
    E = np.atleast_1d(E)[:, None, None]
    w = np.atleast_1d(w)[None, :, None]
    ok = (w < E - e_fermi) & (w > 0)
    ww = np.where(ok, w, 0.5 * E)
    qm = np.sqrt(2 * E) - np.sqrt(2 * (E - ww))
    qp = np.sqrt(2 * E) + np.sqrt(2 * (E - ww))
    qmax = np.minimum(qp, np.sqrt(2 * ww) * (1.0 - 1e-12))   # w0 real: q^2 / 2 < w
    t = np.linspace(0.0, 1.0, nq)[None, None, :]
    lo, hi = np.log(qm), np.log(np.maximum(qmax, qm))
    lq = lo + (hi - lo) * t
    q = np.exp(lq)
    w0 = np.sqrt(np.maximum(ww ** 2 - (q * q / 2) ** 2, 0.0))
    integrand = (w0 / ww) * elf(w0)                           # dq / q = d ln q
    val = np.trapezoid(integrand, lq, axis=2) / (np.pi * E[:, :, 0])
    return np.where(ok[:, :, 0], val, 0.0)


def _inverse_cdf(x, pdf, u):
    """x at cumulative probabilities u of a density sampled on grid x."""
    c = np.concatenate([[0.0], np.cumsum(0.5 * (pdf[1:] + pdf[:-1]) * np.diff(x))])
    if c[-1] <= 0:
        return np.full(u.shape, x[0])
    c /= c[-1]
    return np.interp(u, c, x)


def make_tables(material="Si", directory=None, n_e=64, n_u=41, n_r=36, n_w=400, e_max_eV=40000.0):
    """Computes the four tables of `material` and writes them to `directory`
    (default: a cache directory keyed by the parameters).  Returns (paths in
    TabulatedInelasticSM order, energy offset in eV)."""
    mat = MATERIALS[material]
    key = hashlib.sha1(json.dumps([mat, n_e, n_u, n_r, n_w, e_max_eV, 3], sort_keys=True).encode()).hexdigest()[:12]
    if directory is None:
        directory = os.path.join(os.environ.get("MONSEL_TABLE_CACHE", os.path.join(os.path.expanduser("~"), ".cache",
                                                                                  "cudamonsel_amd")),
                                 f"dielectric_{material}_{key}")
    names = [os.path.join(directory, f"{material}_{k}.csv") for k in ("IIMFP", "RedDeltaE", "Theta", "SEE0")]
    if all(os.path.exists(p) for p in names):
        return names, mat["energy_offset_eV"]
    os.makedirs(directory, exist_ok=True)
    elf = ELF(mat)
    gap = mat["gap_eV"] / HARTREE_EV
    ef = fermi_energy_eV(material) / HARTREE_EV               # above the band bottom
    e_min = ef + 2 * gap + 1.0 / HARTREE_EV
    E = np.geomspace(e_min, e_max_eV / HARTREE_EV, n_e)
    U = np.linspace(0.0, 1.0, n_u)
    # reduced losses: dense where the valence losses lie (r of a 20 eV loss
    # is 2e-3 at 10 keV), sparse towards binary-collision losses (r = 1)
    R = np.concatenate([[0.0], np.geomspace(1e-5, 1.0, n_r - 1)])
    iimfp = np.zeros(n_e)
    red = np.zeros((n_e, n_u))
    theta = np.zeros((n_e, n_r, n_u))
    w_lo = max(gap, 1e-3 / HARTREE_EV)
    for i, e in enumerate(E):
        wmax = e - ef
        w = np.geomspace(w_lo, wmax * (1 - 1e-9), n_w)
        d = diimfp(elf, e, w, ef)[0]
        iimfp[i] = np.trapezoid(d, w) / BOHR_M
        red[i] = _inverse_cdf(w, d, U) / e
        for j, r in enumerate(R):
            # the loss at r = 0 is the gap itself, where the ELF vanishes:
            # its angles are those of the smallest loss the ELF allows
            dw = gap + max(r, 1e-6) * (e - ef - 2 * gap)
            dw = min(max(dw, w_lo), wmax * (1 - 1e-9))
            qm = np.sqrt(2 * e) - np.sqrt(2 * (e - dw))
            qp = np.sqrt(2 * e) + np.sqrt(2 * (e - dw))
            qmax = min(qp, np.sqrt(2 * dw) * (1 - 1e-12))
            if qmax <= qm:
                theta[i, j] = 0.0
                continue
            q = np.geomspace(qm, qmax, 400)
            w0 = np.sqrt(np.maximum(dw ** 2 - (q * q / 2) ** 2, 0.0))
            pdf = (w0 / dw) * elf(w0) / q
            qs = _inverse_cdf(q, pdf, U)
            k, k2 = np.sqrt(2 * e), np.sqrt(2 * (e - dw))
            ct = np.clip((k * k + k2 * k2 - qs * qs) / (2 * k * k2), -1.0, 1.0)
            theta[i, j] = np.arccos(ct)
    # secondaries' initial energies below the Fermi level
    dE = np.geomspace(max(gap, 0.1 / HARTREE_EV), e_max_eV / HARTREE_EV, 48)
    see0 = np.zeros((dE.size, n_u))
    for i, de in enumerate(dE):
        e0 = np.linspace(max(0.0, ef - de), ef, 400)
        see0[i] = _inverse_cdf(e0, np.sqrt(e0) * np.sqrt(e0 + de), U)
    J = HARTREE_EV * EV_J
    tables.write_nu_table(names[0], [E * J], iimfp)
    tables.write_nu_table(names[1], [E * J, U], red)
    tables.write_nu_table(names[2], [E * J, R, U], theta)
    tables.write_nu_table(names[3], [dE * J, U], see0 * J)
    return names, mat["energy_offset_eV"]


def fermi_energy_eV(material):
    """Free-electron Fermi energy of the valence electrons (eV), measured
    from the band bottom."""
    mat = MATERIALS[material]
    n = mat.get("fermi_electrons", sum(o["electrons"] for o in mat["valence"])) * atom_density(mat)
    return 0.5 * (3 * np.pi ** 2 * n) ** (2.0 / 3.0) * HARTREE_EV


def imfp_nm(material, energies_eV):
    """Inelastic mean free path (nm) of the model at kinetic energies above
    the band bottom."""
    mat = MATERIALS[material]
    elf = ELF(mat)
    ef = fermi_energy_eV(material) / HARTREE_EV
    gap = mat["gap_eV"] / HARTREE_EV
    out = []
    for e in np.atleast_1d(energies_eV) / HARTREE_EV:
        w = np.geomspace(max(gap, 1e-3 / HARTREE_EV), (e - ef) * (1 - 1e-9), 600)
        out.append(1.0 / (np.trapezoid(diimfp(elf, e, w, ef)[0], w) / BOHR_M) * 1e9)
    return np.array(out)


def tpp2m_nm(E_eV, Nv, density_g_cm3, M, Eg_eV):
    """The TPP-2M predictive IMFP formula (Tanuma, Powell and Penn, Surf.
    Interface Anal. 21, 165 (1994)), nm."""
    E = np.asarray(E_eV, np.float64)
    Ep = 28.8 * np.sqrt(Nv * density_g_cm3 / M)
    beta = -0.10 + 0.944 / np.sqrt(Ep ** 2 + Eg_eV ** 2) + 0.069 * density_g_cm3 ** 0.1
    gamma = 0.191 / np.sqrt(density_g_cm3)
    U = Nv * density_g_cm3 / M
    C = 1.97 - 0.91 * U
    D = 53.4 - 20.8 * U
    lam_A = E / (Ep ** 2 * (beta * np.log(gamma * E) - C / E + D / E ** 2))
    return lam_A / 10.0
