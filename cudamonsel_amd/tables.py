"""Writers for NUTableInterpolation table files and a synthetic table set.

File format (NUTableInterpolation::ReadTable,
nanoscalemetrology/JMONSEL/NUTableInterpolation.cu:117-228): whitespace
separated numbers -- the dimension N, then for each input variable its
point count and monotonic abscissae, then the tabulated values with the last
variable varying fastest.

The reference's Si tables (IIMFPFullPennInterpSiSI.csv, ...) are not shipped
with it.  synthetic_inelastic_tables() writes a four-table set with the same
layout and units (SI: J, 1/m, rad) from simple closed forms, so that the
TabulatedInelasticSM path runs; it is not a dielectric-function model and
its yields are not physical predictions.
"""
import math
import os

import numpy as np

EV = 1.60217653e-19


def write_nu_table(path, axes, values):
    axes = [np.asarray(a, dtype=np.float64) for a in axes]
    values = np.asarray(values, dtype=np.float64)
    assert values.shape == tuple(len(a) for a in axes)
    with open(path, "w") as f:
        f.write(f"{len(axes)}\n")
        for a in axes:
            f.write(f"{len(a)}\n")
            f.write(" ".join(repr(float(v)) for v in a) + "\n")
        flat = values.ravel(order="C")
        for i in range(0, flat.size, 8):
            f.write(" ".join(repr(float(v)) for v in flat[i:i + 8]) + "\n")


def synthetic_inelastic_tables(directory, e_fermi_eV=12.44, gap_eV=1.1, name="synthSi"):
    """Write IIMFP (1-D), reduced deltaE (2-D), theta (3-D) and SE E0 (2-D)
    tables; returns the four paths in TabulatedInelasticSM order.  Energies
    are kinetic energies above the bottom of the scattering band, as the
    reference's energyOffset convention assumes."""
    os.makedirs(directory, exist_ok=True)
    e_ev = np.geomspace(e_fermi_eV + 2.5 * gap_eV + 0.5, 40000.0, 64)
    u = np.linspace(0.0, 1.0, 41)
    r = np.linspace(0.0, 1.0, 11)
    # inverse inelastic mean free path: universal-curve-like shape (1/m)
    a = 0.27
    lam_nm = 538.0 * a / (e_ev - e_fermi_eV) ** 2 + 0.41 * a ** 1.5 * np.sqrt(e_ev)
    onset = 1.0 - np.exp(-(e_ev - e_fermi_eV - 2 * gap_eV) / 4.0)
    iimfp = onset / (lam_nm * 1e-9)
    # energy loss: gap + (max - gap) * U^4, stored as deltaE / E
    dmax = e_ev - e_fermi_eV - gap_eV
    de = gap_eV + np.outer(dmax - gap_eV, u ** 4)
    rde = de / e_ev[:, None]
    # scattering angle: 0.5 sqrt(reduced energy) * U (rad)
    th = 0.5 * np.sqrt(r)[None, :, None] * u[None, None, :] * np.ones((e_ev.size, 1, 1))
    # SE initial energy below the Fermi level: uniform in [0, E_F] (J)
    dE = np.geomspace(gap_eV, 40000.0, 48)
    e0 = np.outer(np.ones(dE.size), u) * e_fermi_eV * EV
    paths = [os.path.join(directory, f"{name}_{k}.csv") for k in ("IIMFP", "RedDeltaE", "Theta", "SEE0")]
    write_nu_table(paths[0], [e_ev * EV], iimfp)
    write_nu_table(paths[1], [e_ev * EV, u], rde)
    write_nu_table(paths[2], [e_ev * EV, r, u], th)
    write_nu_table(paths[3], [dE * EV, u], e0)
    return paths


def polynomial_tables(directory, seed=0, scale0=1.0):
    """Tables whose values are cubic polynomials on non-uniform grids (for
    interpolation-exactness tests): returns (paths, coefficient dict).  The
    first axis is stored multiplied by scale0 (e.g. an energy grid in J), the
    values are those of the unscaled polynomial."""
    rng = np.random.default_rng(seed)
    os.makedirs(directory, exist_ok=True)
    g0 = np.sort(rng.uniform(1.0, 9.0, 12))
    g1 = np.sort(rng.uniform(-2.0, 3.0, 9))
    g2 = np.sort(rng.uniform(0.0, 1.0, 8))
    c = rng.normal(size=(4, 4, 4))

    def p3(x, y, z):
        return sum(c[i, j, k] * x ** i * y ** j * z ** k for i in range(4) for j in range(4) for k in range(4))

    def p2(x, y):
        return sum(c[i, j, 0] * x ** i * y ** j for i in range(4) for j in range(4))

    def p1(x):
        return c[0, 0, 0] + c[1, 0, 0] * x

    X, Y, Z = np.meshgrid(g0, g1, g2, indexing="ij")
    X2, Y2 = np.meshgrid(g0, g1, indexing="ij")
    paths = [os.path.join(directory, f"poly_{k}.csv") for k in ("1d", "2d", "3d", "2dse")]
    s0 = g0 * scale0
    write_nu_table(paths[0], [s0], p1(g0))
    write_nu_table(paths[1], [s0, g1], p2(X2, Y2))
    write_nu_table(paths[2], [s0, g1, g2], p3(X, Y, Z))
    write_nu_table(paths[3], [s0, g1], p2(X2, Y2))
    return paths, dict(g0=s0, g1=g1, g2=g2, p1=p1, p2=p2, p3=p3, scale0=scale0)
