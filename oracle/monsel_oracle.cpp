/*
 * TEST INFRASTRUCTURE -- the CPU oracle of the electron-transport path.
 *
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
 * load or call this library, and only as the checker / the CPU baseline; the
 * product path (libmonsel_gpu.so) never links or calls it.
 *
 * A from-scratch FP64 restatement, in SI units, of the reference algorithms
 * (m1bcodes/CudaMONSEL, paths relative to CudaMONSEL/CudaMONSEL/gov/nist/):
 *   microanalysis/NISTMonte/{MonteCarloSS, Electron, GaussianBeam, RegionBase,
 *     Sphere, MultiPlaneShape, CylindricalShape, SumShape, BackscatterStats}.cu
 *   nanoscalemetrology/JMONSEL/{MONSEL_MaterialScatterModel, SelectableElasticSM,
 *     NISTMottRS, FittedInelSM, JoyLuoNieminenCSD, GanachaudMokrani*,
 *     TabulatedInelasticSM, NUTableInterpolation, ExpQMBarrierSM,
 *     NormalMultiPlaneShape, NormalCylindricalShape, NormalIntersectionShape,
 *     NormalUnionShape}.cu
 *   nanoscalemetrology/JMONSELutils/{ULagrangeInterpolation,NULagrangeInterpolation}.cu
 *   microanalysis/EPQLibrary/{BrowningEmpiricalCrossSection,
 *     ScreenedRutherfordScatteringAngle, MeanIonizationPotential}.cu
 *   microanalysis/Utility/Histogram.cu, Amphibian/Algorithm.cu
 * Each function below cites the lines it follows.  The scene comes in as the
 * same monsel_scene_desc the GPU engine consumes; every derived constant is
 * recomputed here from the raw constructor parameters.
 *
 * Deliberate differences from the reference (shared with the GPU engine):
 *   - random numbers come from per-electron Philox streams (oracle_rng.h)
 *     instead of the global rand(); uniforms lie in the open interval (0, 1);
 *   - the RefCompat flags of monsel_gpu.h select the reference's deterministic
 *     defects (barrier normal, histogram bin shift, d3 abscissae) or their fix;
 *   - the reference's undefined behaviour (uninitialised cylinder axis in
 *     NormalCylindricalShape, uninitialised delta in
 *     NormalMultiPlaneShape::contains(p0,p1)) is not reproduced: the intended
 *     values are used;
 *   - an electron exceeding job.max_steps steps is dropped and counted.
 */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <atomic>
#include <string>
#include <thread>
#include <vector>

#include "../include/monsel_gpu.h"
#include "oracle_rng.h"

namespace oracle {

/* PhysicalConstants.cu / ToSI.cu / FromSI.cu */
const double PI = 3.14159265358979323846;
const double ECHARGE = 1.60217653e-19;
const double AVOGADRO = 6.0221415e23;
const double BOHR = 5.291772083e-11;
const double KBOLTZ = 1.3806505e-23;
const double HBAR = 1.05457266e-34;
const double EMASS = 9.1093826e-31;
const double TOSI_EV = 1.60217653e-19;
const double TOSI_KEV = 1.60217653e-19 * 1.0e3;
const double FROMSI_KEV = 1.0 / (1.60217653e-19 * 1.0e3);
const double FROMSI_EV = 1.0 / 1.60217653e-19;
/* MonteCarloSS.cu:32-33 */
const double CHAMBER_RADIUS_DEFAULT = 0.1;
const double SMALL_DISP = 1.0e-15;
/* NormalIntersectionShape.cu:72, NormalUnionShape.cu:105 */
const double EXTRAU = 1.0e-10;

thread_local std::string g_err;
int g_rng15 = 0;
thread_local uint32_t g_lcg = 1;

/* ======================= interpolation ================================== */

/* ULagrangeInterpolation.cu:5-33 (Neville on a unit-spaced grid). */
static double u_neville(const double* f, int offset, int order, double x, double* dy_out) {
  int ns = (int)::round(x);
  if (ns < 0) ns = 0;
  else if (ns > order) ns = order;
  double c[16] = {}, d[16] = {};
  for (int i = 0; i <= order; ++i) c[i] = d[i] = f[offset + i];
  double y = c[ns--];
  double dy = 0;
  for (int m = 1; m <= order; m++) {
    for (int i = 0; i <= order - m; i++) {
      double ho = i - x;
      double hp = i + m - x;
      double w = c[i + 1] - d[i];
      d[i] = -hp * w / m;
      c[i] = -ho * w / m;
    }
    dy = (2 * ns < (order - 1 - m)) ? c[ns + 1] : d[ns--];
    y += dy;
  }
  if (dy_out) *dy_out = dy;
  return y;
}

/* ULagrangeInterpolation.cu:35-48 */
static double u_d1(const double* f, int len, double x0, double xinc, int order, double x) {
  const double reducedx = (x - x0) / xinc;
  int index0 = (int)reducedx - order / 2;
  if (index0 < 0) index0 = 0;
  else if (index0 > len - order - 1) index0 = len - order - 1;
  return u_neville(f, index0, order, reducedx - index0, nullptr);
}

/* ULagrangeInterpolation.cu:55-78; f row-major [flen0][flen1]. */
static double u_d2(const double* f, int flen0, int flen1, const double x0[2], const double xinc[2], int order,
                   const double x[2]) {
  const double reducedx1 = (x[0] - x0[0]) / xinc[0];
  int index0 = (int)reducedx1 - order / 2;
  if (index0 < 0) index0 = 0;
  else if (index0 > flen0 - order - 1) index0 = flen0 - order - 1;
  double y[16];
  for (int i = 0; i <= order; i++) y[i] = u_d1(f + (size_t)(index0 + i) * flen1, flen1, x0[1], xinc[1], order, x[1]);
  return u_d1(y, order + 1, x0[0] + index0 * xinc[0], xinc[0], order, x[0]);
}

/* NULagrangeInterpolation.cu:5-68 -> (first index, nearest index). */
static void nu_locate(double x, const double* xs, int n, int order, int* first, int* nearest) {
  int lowlim = -1, uplim = n, maxindex = n - 1, nindex;
  bool ascending = xs[maxindex] > xs[0];
  while ((uplim - lowlim) > 1) {
    int mid = (uplim + lowlim) >> 1;
    if ((x > xs[mid]) == ascending) lowlim = mid;
    else uplim = mid;
  }
  if (lowlim < 0) { *first = 0; *nearest = 0; return; }
  if (uplim > maxindex) { *first = maxindex - order; *nearest = maxindex; return; }
  if (ascending) nindex = ((x - xs[lowlim]) <= (xs[uplim] - x)) ? lowlim : uplim;
  else nindex = ((xs[lowlim] - x) <= (x - xs[uplim])) ? lowlim : uplim;
  int firstInd = lowlim, lastInd = uplim;
  while (((lastInd - firstInd) < order) && (firstInd > 0) && (lastInd < maxindex)) {
    if (((((xs[lowlim] - xs[firstInd - 1]) + xs[uplim]) - xs[lastInd + 1]) >= 0) == ascending) lastInd += 1;
    else firstInd -= 1;
  }
  if (lastInd >= maxindex) { *first = maxindex - order; *nearest = nindex; return; }
  if (firstInd <= 0) { *first = 0; *nearest = nindex; return; }
  *first = firstInd;
  *nearest = nindex;
}

/* NULagrangeInterpolation.cu:85-112 */
static double nu_neville(const double* f, const double* xs, int offset, int nearest, int order, double x) {
  int ns = nearest - offset;
  double c[16] = {}, d[16] = {};
  for (int i = 0; i <= order; ++i) c[i] = d[i] = f[offset + i];
  double y = c[ns--];
  for (int m = 1; m <= order; m++) {
    for (int i = 0; i <= order - m; i++) {
      double ho = xs[i + offset] - x;
      double hp = xs[i + m + offset] - x;
      double w = c[i + 1] - d[i];
      double den = w / (ho - hp);
      d[i] = hp * den;
      c[i] = ho * den;
    }
    double dy = ((2 * ns) < (order - 1 - m)) ? c[ns + 1] : d[ns--];
    y += dy;
  }
  return y;
}

/* NULagrangeInterpolation.cu:131-147 */
static double nu_d1(const double* f, const double* xs, int n, int order, double x) {
  int first, nearest;
  nu_locate(x, xs, n, order, &first, &nearest);
  return nu_neville(f, xs, first, nearest, order, x);
}

struct NUTable {
  int dim = 0;
  int n[4] = {0, 0, 0, 0};
  std::vector<double> x[4];
  std::vector<double> v;
  double domain[4][2];
  double range[2];
  size_t stride(int d) const {
    size_t s = 1;
    for (int k = d + 1; k < dim; ++k) s *= (size_t)n[k];
    return s;
  }
};

/* NULagrangeInterpolation.cu:169-192; f points at an [n_a][n_b] block. */
static double nu_d2(const double* f, const double* xa, int na, const double* xb, int nb, int order, const double xv[2]) {
  int index0, nearest;
  nu_locate(xv[0], xa, na, order, &index0, &nearest);
  double y[16];
  for (int i = 0; i <= order; i++) y[i] = nu_d1(f + (size_t)(index0 + i) * nb, xb, nb, order, xv[1]);
  return nu_d1(y, xa + index0, order + 1, order, xv[0]);
}

/* NULagrangeInterpolation.cu:215-240.  With compat_d3 the outer abscissae are
 * taken from the start of the grid, as line 236 does. */
static double nu_d3(const NUTable& t, const double* f, int order, const double xv[3], bool compat_d3) {
  int index0, nearest;
  nu_locate(xv[0], t.x[0].data(), t.n[0], order, &index0, &nearest);
  double y[16];
  size_t s0 = (size_t)t.n[1] * t.n[2];
  for (int i = 0; i <= order; i++)
    y[i] = nu_d2(f + (index0 + i) * s0, t.x[1].data(), t.n[1], t.x[2].data(), t.n[2], order, xv + 1);
  const double* xt = compat_d3 ? t.x[0].data() : t.x[0].data() + index0;
  return nu_d1(y, xt, order + 1, order, xv[0]);
}

/* NULagrangeInterpolation.cu:263-306 */
static double nu_d4(const NUTable& t, int order, const double xv[4], bool compat_d3) {
  int index0, nearest;
  nu_locate(xv[0], t.x[0].data(), t.n[0], order, &index0, &nearest);
  double y[16];
  size_t s0 = (size_t)t.n[1] * t.n[2] * t.n[3];
  NUTable sub;
  sub.dim = 3;
  for (int k = 0; k < 3; ++k) { sub.n[k] = t.n[k + 1]; sub.x[k] = t.x[k + 1]; }
  for (int i = 0; i <= order; i++) y[i] = nu_d3(sub, t.v.data() + (index0 + i) * s0, order, xv + 1, compat_d3);
  return nu_d1(y, t.x[0].data() + index0, order + 1, order, xv[0]);
}

/* NUTableInterpolation.cu:98-115 */
static double nu_interpolate(const NUTable& t, const double* xv, int order, bool compat_d3) {
  switch (t.dim) {
    case 1: return nu_d1(t.v.data(), t.x[0].data(), t.n[0], order, xv[0]);
    case 2: return nu_d2(t.v.data(), t.x[0].data(), t.n[0], t.x[1].data(), t.n[1], order, xv);
    case 3: return nu_d3(t, t.v.data(), order, xv, compat_d3);
    case 4: return nu_d4(t, order, xv, compat_d3);
  }
  return NAN;
}

/* ======================= elastic cross-sections ========================== */

/* BrowningEmpiricalCrossSection.cu:11-45 */
struct Browning {
  double z, zp17, zp2, zp3;
  void init(int Z) {
    z = Z;
    zp17 = ::pow((double)Z, 1.7);
    zp2 = ::pow((double)Z, 2.0);
    zp3 = ::pow((double)Z, 3.0);
  }
  double total(double energy) const {
    const double e = energy * FROMSI_KEV;
    const double re = ::sqrt(e);
    return 3.0e-22 * zp17 / (e + 0.005 * zp17 * re + 0.0007 * zp2 / re);
  }
  double angle(double energy, double r1, double r2) const {
    const double e = energy * FROMSI_KEV;
    const double r = (300.0 * e / z) + (zp3 / (3.0e5 * e));
    if (r1 <= r / (r + 1)) {
      const double alpha = 7.0e-3 / e;
      return ::acos(1.0 - ((2.0 * alpha * r2) / (alpha - r2 + 1)));
    }
    return ::acos(1 - 2.0 * r2);
  }
};

/* ScreenedRutherfordScatteringAngle.cu:24-39 */
struct Rutherford {
  double z;
  void init(int Z) { z = Z; }
  double total(double energy) const {
    double zp = ::pow(z, 1.0 / 3.0);
    return (7.670843088080456e-38 * zp * (1.0 + z)) / ((energy + ((5.44967975966321e-19 * zp * zp))));
  }
  double angle(double energy, double r) const {
    double alpha = (5.44968e-19 * ::pow(z, 2.0 / 3.0)) / energy;
    return ::acos(1 - 2.0 * alpha * r / (1 + alpha - r));
  }
};

/* NISTMottRS.cu:16-26 */
static const double MAX_NISTMOTT = 20.0 * TOSI_KEV;
static const double MIN_NISTMOTT = 0.050 * TOSI_KEV;
static double DL50() { return ::log(MIN_NISTMOTT); }
static double PARAM() { return (::log(MAX_NISTMOTT) - DL50()) / 60.0; }

/* One element's randomized scatter: NISTMottRS (NISTMottRS.cu:76-154) or a
 * plain Browning / screened-Rutherford model. */
struct ElementScatter {
  int kind = MONSEL_MECH_NISTMOTT_ELASTIC;
  int Z = 0;
  int method = 1;
  const double* spwem = nullptr;
  const double* x1 = nullptr;
  Browning browning;
  Rutherford rutherford;
  double extrapolateBelowEnergy = 0, MottXSatMinEnergy = 0, sfBrowning = 0;

  double mott_table_sigma(double energy) const {
    return BOHR * BOHR * u_d1(spwem, MONSEL_MOTT_NE, DL50(), PARAM(), 3, ::log(energy));
  }
  void init_mott(int Zin, int meth, const double* sp, const double* x) {
    kind = MONSEL_MECH_NISTMOTT_ELASTIC;
    Z = Zin;
    method = meth;
    spwem = sp;
    x1 = x;
    browning.init(Z);
    rutherford.init(Z);
    extrapolateBelowEnergy = method == 1 ? 50.0 * TOSI_EV : 100.0 * TOSI_EV;
    MottXSatMinEnergy = total(extrapolateBelowEnergy);
    sfBrowning = MottXSatMinEnergy / browning.total(extrapolateBelowEnergy);
  }
  /* NISTMottRS.cu:103-119 */
  double total(double energy) const {
    if (kind == MONSEL_MECH_BROWNING_ELASTIC) return browning.total(energy);
    if (kind == MONSEL_MECH_RUTHERFORD_ELASTIC) return rutherford.total(energy);
    if (energy < extrapolateBelowEnergy) {
      if (method == 3) return MottXSatMinEnergy * energy / extrapolateBelowEnergy;
      return sfBrowning * browning.total(energy);
    } else if (energy < MAX_NISTMOTT) {
      return mott_table_sigma(energy);
    }
    return rutherford.total(energy);
  }
  /* NISTMottRS.cu:121-154 */
  double angle(double energy, Stream& rng) const {
    if (kind == MONSEL_MECH_BROWNING_ELASTIC) {
      double r1 = rng.next(), r2 = rng.next();
      return browning.angle(energy, r1, r2);
    }
    if (kind == MONSEL_MECH_RUTHERFORD_ELASTIC) return rutherford.angle(energy, rng.next());
    if (energy < extrapolateBelowEnergy) {
      double r1 = rng.next(), r2 = rng.next();
      return browning.angle(energy, r1, r2);
    } else if (energy < MAX_NISTMOTT) {
      return mott_angle(energy, rng.next());
    }
    return rutherford.angle(energy, rng.next());
  }
  double mott_angle(double energy, double u) const {
    const double x0[] = {DL50(), 0.};
    const double xinc[] = {PARAM(), 0.005};
    const double x[] = {::log(energy), u};
    const double q = u_d2(x1, MONSEL_MOTT_NE, MONSEL_MOTT_NQ, x0, xinc, 3, x);
    const double com = 1.0 - (2.0 * q * q);
    return com > -1.0 ? (com < 1.0 ? ::acos(com) : 0.0) : PI;
  }
};

/* ======================= materials and mechanisms ======================== */

struct Material {
  int nce = 0;
  int Z[MONSEL_MAX_ELEMENTS];
  double w[MONSEL_MAX_ELEMENTS], A[MONSEL_MAX_ELEMENTS];
  double density = 0, workfunction = 0, cb = 0, bandgap = 0;
  std::vector<double> core;
  double eFermi() const { return -cb - workfunction; } /* SEmaterial.cu:201 */
};

/* JoyLuoNieminenCSD.cu:42-121 */
struct CSD {
  int nce = 0;
  double breakE = 0, bh = 0, bhplus1eV = 0, gamma = 0, minE = 0;
  double recipJ[MONSEL_MAX_ELEMENTS], coef[MONSEL_MAX_ELEMENTS], beta[MONSEL_MAX_ELEMENTS];
  bool zero = true;
  void init(const Material& m, double brk) {
    zero = false;
    breakE = brk;
    bh = -m.cb;
    nce = m.nce;
    if (nce == 0) return;
    bhplus1eV = bh + TOSI_EV;
    if (breakE < bhplus1eV) breakE = bhplus1eV;
    for (int i = 0; i < nce; ++i) {
      double z = m.Z[i];
      /* MeanIonizationPotential.cu:85-89 (Wilson41), :43-47 (Sternheimer64) */
      double J = m.Z[i] < 13 ? TOSI_EV * (11.5 * z) : TOSI_EV * (9.76 * z + 58.8 * ::pow(z, -0.19));
      recipJ[i] = 1.166 / J;
      beta[i] = 1. - (recipJ[i] * bhplus1eV);
      coef[i] = 2.01507E-28 * m.density * m.w[i] * m.Z[i] / m.A[i];
    }
    minE = bh;
    gamma = 0.;
    for (int i = 0; i < nce; i++) gamma += coef[i] * ::log((recipJ[i] * breakE) + beta[i]);
    gamma /= ::pow(breakE, 3.5);
  }
  double compute(double len, double kE, int depth = 0) const {
    if (zero) return 0.;
    if ((nce == 0) || (kE < minE) || (kE <= 0.)) return 0.;
    if (kE <= breakE) return (kE / ::pow(1. + (1.5 * gamma * len * kE * ::sqrt(kE)), 2. / 3.)) - kE;
    double loss = 0.;
    for (int i = 0; i < nce; i++) loss += coef[i] * ::log((recipJ[i] * kE) + beta[i]);
    loss *= len / kE;
    if (loss <= (0.1 * kE) || depth > 200) return -loss;
    loss = compute(len / 2., kE, depth + 1);
    return loss + compute(len / 2., kE + loss, depth + 1);
  }
};

struct Electron;

struct Mech {
  int kind = 0;
  /* SelectableElasticSM.cu:69-89 */
  std::vector<ElementScatter> rse;
  std::vector<double> scalefactor;
  double densityNa = 0;
  /* FittedInelSM.cu:8-33 */
  double energySEgen = 0, eFermi = 0;
  CSD sdAlg;
  /* GanachaudMokraniPolaronTrapSM.cu:8-13 */
  double prefactor = 0, extinction = 0, cutoff = 0;
  /* GanachaudMokraniPhononInelasticSM.cu:8-17 */
  double phononE = 0, phPrefactor = 0;
  /* TabulatedInelasticSM.cu:23-66, 461-498 */
  int methodSE = 1;
  double energyOffset = 0, minEgenSE = 0, rateMult = 1;
  bool E0fromDispersion = false;
  const NUTable *tIIMFP = nullptr, *tRDE = nullptr, *tTheta = nullptr, *tSEE0 = nullptr;
  double energyCBbottom = 0, bandgap = 0, energyGap = 0, offsetFermiEnergy = 0, bEref = 0;
  std::vector<double> coreEnergies;
  double tableEiDomain[2], tableIIMFPEiDomain[2], energyRangeSE0[2];
};

struct MSM {
  bool isNull = false;
  int material = 0;
  std::vector<Mech> mech;
  CSD csd;
  int barrierKind = 0;
  double u0 = 0, lambda = 0, lambdaFactor = 0;
  double minE = 0;
};

struct Plane {
  double n[3], c[3];
};

struct Shape {
  int kind = 0;
  int a = -1, b = -1;
  std::vector<Plane> planes;
  /* NormalCylindricalShape.cu:8-16 (intended member values) */
  double end0[3], axis[3], naxis[3], radius2 = 0, len2 = 0, len = 0;
  double delta[3];
  /* Sphere */
  double center[3], radius = 0;
};

struct Region {
  int parent = -1, shape = -1, msm = -1;
  std::vector<int> children;
};

struct Scene {
  uint32_t compat = MONSEL_COMPAT_REFERENCE;
  double chamberRadius = CHAMBER_RADIUS_DEFAULT;
  std::vector<Material> materials;
  std::vector<MSM> msms;
  std::vector<Shape> shapes;
  std::vector<Region> regions;
  std::vector<std::vector<double>> mottSpwem, mottX1;
  std::vector<int> mottZ;
  std::vector<NUTable> nu;
  bool barrierCompat() const { return compat & MONSEL_COMPAT_BARRIER_NORMAL; }
  bool histCompat() const { return compat & MONSEL_COMPAT_HISTOGRAM_BIN; }
  bool d3Compat() const { return compat & MONSEL_COMPAT_D3_ABSCISSAE; }
};

/* ======================= electron ======================================== */

/* Electron.cu:15-143 */
struct Electron {
  double pos[3], prevPos[3];
  double theta = 0, phi = 0, E = 0, prevE = 0;
  int region = -1, prevRegion = -1;
  long steps = 0;
  bool complete = false;
  Stream rng;
  uint32_t nchildren = 0;

  void init(const double p[3], double th, double ph, double kE) {
    for (int i = 0; i < 3; ++i) pos[i] = prevPos[i] = p[i];
    region = prevRegion = -1;
    E = prevE = kE;
    theta = th;
    phi = ph;
    steps = 0;
    complete = false;
    nchildren = 0;
  }
  void candidatePoint(double dS, double out[3]) const {
    double st = ::sin(theta);
    out[0] = pos[0] + dS * ::cos(phi) * st;
    out[1] = pos[1] + dS * ::sin(phi) * st;
    out[2] = pos[2] + dS * ::cos(theta);
  }
  void updateDirection(double dTheta, double dPhi) {
    double ct = ::cos(theta), st = ::sin(theta);
    double cp = ::cos(phi), sp = ::sin(phi);
    double ca = ::cos(dTheta), sa = ::sin(dTheta);
    double cb = ::cos(dPhi);
    double xx = cb * ct * sa + ca * st;
    double yy = sa * ::sin(dPhi);
    double dx = cp * xx - sp * yy;
    double dy = cp * yy + sp * xx;
    double dz = ca * ct - cb * sa * st;
    theta = ::atan2(::sqrt(dx * dx + dy * dy), dz);
    phi = ::atan2(dy, dx);
  }
  void move(const double p[3], double dE) {
    for (int i = 0; i < 3; ++i) { prevPos[i] = pos[i]; pos[i] = p[i]; }
    prevE = E;
    E += dE;
    ++steps;
  }
  void setRegion(int r) { prevRegion = region; region = r; }
};

/* TabulatedInelasticSM.cu:68-85 */
static void update_direction_pair(double theta, double phi, double dTheta, double dPhi, double* ot, double* op) {
  double ct = ::cos(theta), st = ::sin(theta);
  double cp = ::cos(phi), sp = ::sin(phi);
  double ca = ::cos(dTheta), sa = ::sin(dTheta);
  double cb = ::cos(dPhi);
  double xx = (cb * ct * sa) + (ca * st);
  double yy = sa * ::sin(dPhi);
  double dx = (cp * xx) - (sp * yy);
  double dy = (cp * yy) + (sp * xx);
  double dz = (ca * ct) - (cb * sa * st);
  *ot = ::atan2(::sqrt((dx * dx) + (dy * dy)), dz);
  *op = ::atan2(dy, dx);
}

/* ======================= geometry ======================================== */

struct NV {
  double n[3];
  double u;
};
static NV no_intersection() { return NV{{0., 0., 0.}, INFINITY}; }

/* NormalMultiPlaneShape.cu:26-38 */
static bool contains_tie_break(const double* n) {
  if (n[0] < 0.) return false;
  if (n[0] == 0.) {
    if (n[1] < 0.) return false;
    if (n[1] == 0.)
      if (n[2] < 0.) return false;
  }
  return true;
}

struct Geometry {
  const Scene& sc;
  explicit Geometry(const Scene& s) : sc(s) {}

  /* Shape::contains(pos) for each shape kind. */
  bool contains(int s, const double* p) const {
    const Shape& sh = sc.shapes[s];
    switch (sh.kind) {
      case MONSEL_SHAPE_PLANES: /* NormalMultiPlaneShape.cu:230-240, MultiPlaneShape.cu:23-27 */
        for (const Plane& pl : sh.planes)
          if (!(((p[0] - pl.c[0]) * pl.n[0] + (p[1] - pl.c[1]) * pl.n[1] + (p[2] - pl.c[2]) * pl.n[2]) <= 0.0))
            return false;
        return true;
      case MONSEL_SHAPE_CYLINDER: { /* CylindricalShape.cu:41-45 */
        double u = (sh.delta[0] * (p[0] - sh.end0[0]) + sh.delta[1] * (p[1] - sh.end0[1]) +
                    sh.delta[2] * (p[2] - sh.end0[2])) / sh.len2;
        double d2 = 0;
        for (int i = 0; i < 3; ++i) {
          double t = p[i] - (sh.end0[i] + u * sh.delta[i]);
          d2 += t * t;
        }
        return (u >= 0) && (u <= 1.0) && (d2 <= sh.radius2);
      }
      case MONSEL_SHAPE_INTERSECTION: /* NormalIntersectionShape.cu:9-12 */
        return contains(sh.a, p) && contains(sh.b, p);
      case MONSEL_SHAPE_UNION: /* SumShape.cu:38-44 */
        return contains(sh.a, p) || contains(sh.b, p);
      case MONSEL_SHAPE_SPHERE: { /* Sphere.cu:11-14 */
        double s2 = 0;
        for (int i = 0; i < 3; ++i) s2 += (p[i] - sh.center[i]) * (p[i] - sh.center[i]);
        return s2 <= sh.radius * sh.radius;
      }
    }
    return false;
  }

  /* NormalShape::contains(pos0, pos1): inside just after pos0 along the step. */
  bool contains2(int s, const double* p0, const double* p1) const {
    const Shape& sh = sc.shapes[s];
    switch (sh.kind) {
      case MONSEL_SHAPE_PLANES: { /* NormalMultiPlaneShape.cu:178-204 */
        double delta[3] = {p1[0] - p0[0], p1[1] - p0[1], p1[2] - p0[2]};
        for (const Plane& pl : sh.planes) {
          double p0cdotn = (p0[0] - pl.c[0]) * pl.n[0] + (p0[1] - pl.c[1]) * pl.n[1] + (p0[2] - pl.c[2]) * pl.n[2];
          if (p0cdotn > 0.) return false;
          if (p0cdotn == 0.) {
            double dn = delta[0] * pl.n[0] + delta[1] * pl.n[1] + delta[2] * pl.n[2];
            if (dn > 0.) return false;
            if ((dn == 0.) && !contains_tie_break(pl.n)) return false;
          }
        }
        return true;
      }
      case MONSEL_SHAPE_CYLINDER: { /* NormalCylindricalShape.cu:24-98 */
        const double p0c[] = {p0[0] - sh.end0[0], p0[1] - sh.end0[1], p0[2] - sh.end0[2]};
        const double* na = sh.naxis;
        const double p0cdotn = p0c[0] * na[0] + p0c[1] * na[1] + p0c[2] * na[2];
        const double p0cSq = p0c[0] * p0c[0] + p0c[1] * p0c[1] + p0c[2] * p0c[2];
        const double r2 = p0cSq - p0cdotn * p0cdotn;
        const bool on_end0 = (p0cdotn == 0) && (r2 <= sh.radius2);
        const bool on_end1 = (p0cdotn == sh.len) && (r2 <= sh.radius2);
        const bool on_cyl = (r2 == sh.radius2) && (p0cdotn >= 0) && (p0cdotn <= sh.len);
        const double delta[] = {p1[0] - p0[0], p1[1] - p0[1], p1[2] - p0[2]};
        const double dd = delta[0] * na[0] + delta[1] * na[1] + delta[2] * na[2];
        auto tie = [&]() {
          if (na[0] == 0) {
            if (na[1] == 0) return na[2] > 0;
            return na[1] > 0;
          }
          return na[0] > 0;
        };
        if (on_end0) return dd == 0 ? tie() : dd > 0;
        if (on_end1) return dd == 0. ? tie() : dd < 0;
        if (on_cyl) {
          const double nv[] = {p0c[0] - p0cdotn * na[0], p0c[1] - p0cdotn * na[1], p0c[2] - p0cdotn * na[2]};
          return (nv[0] * delta[0] + nv[1] * delta[1] + nv[2] * delta[2]) < 0;
        }
        return (r2 < sh.radius2) && (p0cdotn > 0) && (p0cdotn < sh.len);
      }
      case MONSEL_SHAPE_INTERSECTION: /* NormalIntersectionShape.cu:14-17 */
        return contains2(sh.a, p0, p1) && contains2(sh.b, p0, p1);
      case MONSEL_SHAPE_UNION: /* NormalUnionShape.cu:17-23 */
        return contains2(sh.a, p0, p1) || contains2(sh.b, p0, p1);
      case MONSEL_SHAPE_SPHERE:
        return contains(s, p0);
    }
    return false;
  }

  /* NormalMultiPlaneShape.cu:40-171 */
  NV planes_first(const Shape& sh, const double* pos0, const double* pos1) const {
    double umin = -INFINITY, umax = INFINITY, u;
    int minindex = -1, maxindex = -1;
    NV result = no_intersection();
    const double delta[] = {pos1[0] - pos0[0], pos1[1] - pos0[1], pos1[2] - pos0[2]};
    for (size_t i = 0; i < sh.planes.size(); i++) {
      const Plane& pl = sh.planes[i];
      double numerator = (pos0[0] - pl.c[0]) * pl.n[0] + (pos0[1] - pl.c[1]) * pl.n[1] + (pos0[2] - pl.c[2]) * pl.n[2];
      double denominator = delta[0] * pl.n[0] + delta[1] * pl.n[1] + delta[2] * pl.n[2];
      if (denominator == 0) {
        if ((numerator < 0) || ((numerator == 0) && contains_tie_break(pl.n))) continue;
        return result;
      }
      u = -numerator / denominator;
      if (denominator > 0) {
        if (u < umax) {
          if ((u < 0) || (u <= umin)) return result;
          umax = u;
          maxindex = (int)i;
        }
      } else if (u > umin) {
        if ((u > 1) || (u >= umax)) return result;
        umin = u;
        minindex = (int)i;
      }
    }
    if (umin > 0) {
      result.u = umin;
      for (int k = 0; k < 3; ++k) result.n[k] = sh.planes[minindex].n[k];
      return result;
    }
    if ((umax <= 1) && (umax > 0.)) {
      result.u = umax;
      for (int k = 0; k < 3; ++k) result.n[k] = sh.planes[maxindex].n[k];
      return result;
    }
    return result;
  }

  /* NormalCylindricalShape.cu:111-322 */
  NV cylinder_first(const Shape& sh, const double* pos0, const double* pos1) const {
    NV nv = no_intersection();
    const double* na = sh.naxis;
    const double p0c[] = {pos0[0] - sh.end0[0], pos0[1] - sh.end0[1], pos0[2] - sh.end0[2]};
    const double p0cSquared = p0c[0] * p0c[0] + p0c[1] * p0c[1] + p0c[2] * p0c[2];
    const double delta[] = {pos1[0] - pos0[0], pos1[1] - pos0[1], pos1[2] - pos0[2]};
    const double p0cdotn = p0c[0] * na[0] + p0c[1] * na[1] + p0c[2] * na[2];
    const double deltadotn = delta[0] * na[0] + delta[1] * na[1] + delta[2] * na[2];
    const double deltadotnSquared = deltadotn * deltadotn;
    const double p0cdotdelta = delta[0] * p0c[0] + delta[1] * p0c[1] + delta[2] * p0c[2];
    const double deltaSquared = delta[0] * delta[0] + delta[1] * delta[1] + delta[2] * delta[2];
    int inum = 0;
    double u = 0., u1 = 0., u2 = 0., projection = 0., savedprojection = 0.;
    if (deltadotn != 0.) {
      u1 = -p0cdotn / deltadotn;
      if ((u1 > 0.) && (u1 <= 1.) && ((p0cSquared + (2 * u1 * p0cdotdelta) + (u1 * u1 * deltaSquared)) <= sh.radius2)) {
        nv.u = u1;
        inum = 1;
      }
      u2 = u1 + (sh.len / deltadotn);
      if ((u2 > 0.) && (u2 <= 1.) && (u2 < nv.u) &&
          (((p0cSquared + (2 * u2 * p0cdotdelta) + (u2 * u2 * deltaSquared) + sh.len2) -
            (2 * sh.len * (p0cdotn + (u2 * deltadotn)))) <= sh.radius2)) {
        nv.u = u2;
        inum = 2;
      }
    }
    if (!(((u1 < 0) && (u2 < 0)) || ((u1 > 1) && (u2 > 1)))) {
      const double r0Squared = p0cSquared - (p0cdotn * p0cdotn);
      const double r1Squared = (r0Squared + (2. * (p0cdotdelta - (p0cdotn * deltadotn))) + deltaSquared) - deltadotnSquared;
      if ((r0Squared > sh.radius2) || (r1Squared > sh.radius2)) {
        const double a = deltaSquared - deltadotnSquared;
        const double minusbover2 = (p0cdotn * deltadotn) - p0cdotdelta;
        double term = (minusbover2 * minusbover2) - (a * (r0Squared - sh.radius2));
        if (term >= 0) {
          term = ::sqrt(term);
          u = (minusbover2 + term) / a;
          if ((u > 0) && (u <= 1) && (u < nv.u)) {
            projection = p0cdotn + (u * deltadotn);
            if ((projection >= 0) && (projection <= sh.len)) {
              nv.u = u;
              inum = 3;
              savedprojection = projection;
            }
          }
          u = (minusbover2 - term) / a;
          if ((u > 0) && (u <= 1) && (u < nv.u)) {
            projection = p0cdotn + (u * deltadotn);
            if ((projection >= 0) && (projection <= sh.len)) {
              nv.u = u;
              inum = 4;
              savedprojection = projection;
            }
          }
        }
      }
    }
    switch (inum) {
      case 1:
        for (int k = 0; k < 3; ++k) nv.n[k] = -na[k];
        return nv;
      case 2:
        for (int k = 0; k < 3; ++k) nv.n[k] = na[k];
        return nv;
      case 3:
      case 4: {
        double v[3], m = 0;
        for (int k = 0; k < 3; ++k) {
          v[k] = (p0c[k] + (nv.u * delta[k])) - (savedprojection * na[k]);
          m += v[k] * v[k];
        }
        m = ::sqrt(m);
        for (int k = 0; k < 3; ++k) nv.n[k] = v[k] / m;
        return nv;
      }
    }
    return nv;
  }

  /* Sphere.cu:21-45 */
  NV sphere_first(const Shape& sh, const double* pos0, const double* pos1) const {
    NV r = no_intersection();
    double d[3], m[3];
    for (int k = 0; k < 3; ++k) { d[k] = pos1[k] - pos0[k]; m[k] = pos0[k] - sh.center[k]; }
    const double ma2 = -2.0 * (d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
    const double b = 2.0 * (m[0] * d[0] + m[1] * d[1] + m[2] * d[2]);
    const double c2 = 2.0 * ((m[0] * m[0] + m[1] * m[1] + m[2] * m[2]) - sh.radius * sh.radius);
    const double f = b * b + ma2 * c2;
    if (f >= 0) {
      double up = (b + ::sqrt(f)) / ma2;
      double un = (b - ::sqrt(f)) / ma2;
      if (up < 0.0) up = INFINITY;
      if (un < 0.0) un = INFINITY;
      r.u = ::fmin(up, un);
      if (r.u != INFINITY) {
        double len = 0;
        for (int k = 0; k < 3; ++k) { r.n[k] = m[k] + r.u * d[k]; len += r.n[k] * r.n[k]; }
        len = ::sqrt(len);
        for (int k = 0; k < 3; ++k) r.n[k] /= len;
      }
    }
    return r;
  }

  static void advance(const double* pos0, const double* delta, double u, double* out) {
    for (int k = 0; k < 3; ++k) out[k] = pos0[k] + (u * delta[k]);
  }
  static double dotd(const double* a, const NV& v) { return a[0] * v.n[0] + a[1] * v.n[1] + a[2] * v.n[2]; }

  /* NormalIntersectionShape.cu:19-239 */
  NV intersection_first(const Shape& sh, const double* pos0, const double* pos1) const {
    const NV nointersection = no_intersection();
    int adepth, bdepth;
    double u;
    const double delta[] = {pos1[0] - pos0[0], pos1[1] - pos0[1], pos1[2] - pos0[2]};
    NV nva = first_normal(sh.a, pos0, pos1);
    if (nva.u <= 1.) adepth = dotd(delta, nva) > 0 ? 1 : 0;
    else {
      if (!contains2(sh.a, pos0, pos1)) return nointersection;
      adepth = 1;
    }
    NV nvb = first_normal(sh.b, pos0, pos1);
    if (nvb.u <= 1.) bdepth = dotd(delta, nvb) > 0 ? 1 : 0;
    else {
      if (!contains2(sh.b, pos0, pos1)) return nointersection;
      if (adepth == 1) return nva;
      bdepth = 1;
    }
    int cdepth = adepth + bdepth;
    for (int guard = 0; guard < 100000; ++guard) {
      if (nva.u < nvb.u) {
        if (bdepth == 1) return nva;
        cdepth = cdepth ^ 1;
        u = nva.u + EXTRAU;
        double tmp[3];
        advance(pos0, delta, u, tmp);
        nva = first_normal(sh.a, tmp, pos1);
        if (nva.u < INFINITY) nva.u = (nva.u * (1. - u)) + u;
        if (nva.u > 1) {
          if (cdepth == 0) return nointersection;
          return nvb;
        }
        adepth = adepth ^ 1;
      } else if (nva.u > nvb.u) {
        if (adepth == 1) return nvb;
        cdepth = cdepth ^ 1;
        u = nvb.u + EXTRAU;
        double tmp[3];
        advance(pos0, delta, u, tmp);
        nvb = first_normal(sh.b, tmp, pos1);
        if (nvb.u < INFINITY) nvb.u = (nvb.u * (1. - u)) + u;
        if (nvb.u > 1) {
          if (cdepth == 0) return nointersection;
          return nva;
        }
        bdepth = bdepth ^ 1;
      } else {
        int depthchange = (((adepth ^ 1) - adepth) + (bdepth ^ 1)) - bdepth;
        if (depthchange == 0) {
          u = nva.u + EXTRAU;
          double tmp1[3];
          advance(pos0, delta, u, tmp1);
          nva = first_normal(sh.a, tmp1, pos1);
          if (nva.u < INFINITY) nva.u = (nva.u * (1. - u)) + u;
          u = nvb.u + EXTRAU;
          double tmp2[3];
          advance(pos0, delta, u, tmp2);
          nvb = first_normal(sh.b, tmp2, pos1);
          if (nvb.u < INFINITY) nvb.u = (nvb.u * (1. - u)) + u;
          if (nva.u > 1) {
            if (adepth == 0) return nvb;
            return nointersection;
          }
          if (nvb.u > 1) {
            if (bdepth == 0) return nva;
            return nointersection;
          }
          adepth = adepth ^ 1;
          bdepth = bdepth ^ 1;
        } else {
          NV r;
          for (int k = 0; k < 3; ++k) r.n[k] = (nva.n[k] + nvb.n[k]) / 2.;
          r.u = nva.u;
          return r;
        }
      }
    }
    return nointersection;
  }

  /* NormalUnionShape.cu:30-271 */
  NV union_first(const Shape& sh, const double* pos0, const double* pos1) const {
    const NV nointersection = no_intersection();
    int adepth, bdepth;
    double u;
    const double delta[] = {pos1[0] - pos0[0], pos1[1] - pos0[1], pos1[2] - pos0[2]};
    NV nva = first_normal(sh.a, pos0, pos1);
    if (nva.u <= 1.) adepth = dotd(delta, nva) > 0 ? 1 : 0;
    else {
      if (contains2(sh.a, pos0, pos1)) return nointersection;
      adepth = 0;
    }
    NV nvb = first_normal(sh.b, pos0, pos1);
    if (nvb.u <= 1.) bdepth = dotd(delta, nvb) > 0 ? 1 : 0;
    else {
      if (contains2(sh.b, pos0, pos1)) return nointersection;
      return nva;
    }
    int cdepth = adepth + bdepth;
    for (int guard = 0; guard < 100000; ++guard) {
      if (nva.u < nvb.u) {
        if (adepth == cdepth) return nva;
        cdepth = (cdepth == 1 ? 2 : 1);
        u = nva.u + EXTRAU;
        double tmp[3];
        advance(pos0, delta, u, tmp);
        nva = first_normal(sh.a, tmp, pos1);
        if (nva.u < INFINITY) nva.u = (nva.u * (1. - u)) + u;
        if (nva.u > 1) {
          if (cdepth == bdepth) return nvb;
          return nointersection;
        }
        adepth = adepth ^ 1;
      } else if (nva.u > nvb.u) {
        if (bdepth == cdepth) return nvb;
        cdepth = (cdepth == 1 ? 2 : 1);
        u = nvb.u + EXTRAU;
        double tmp[3];
        advance(pos0, delta, u, tmp);
        nvb = first_normal(sh.b, tmp, pos1);
        if (nvb.u < INFINITY) nvb.u = (nvb.u * (1. - u)) + u;
        if (nvb.u > 1) {
          if (cdepth == adepth) return nva;
          return nointersection;
        }
        bdepth = bdepth ^ 1;
      } else {
        const int depthchange = (((adepth ^ 1) - adepth) + (bdepth ^ 1)) - bdepth;
        if (depthchange == 0) {
          u = nva.u + EXTRAU;
          double tmp0[3];
          advance(pos0, delta, u, tmp0);
          nva = first_normal(sh.a, tmp0, pos1);
          if (nva.u < INFINITY) nva.u = (nva.u * (1. - u)) + u;
          u = nvb.u + EXTRAU;
          double tmp1[3];
          advance(pos0, delta, u, tmp1);
          nvb = first_normal(sh.b, tmp1, pos1);
          if (nvb.u < INFINITY) nvb.u = (nvb.u * (1. - u)) + u;
          if (nva.u > 1) {
            if (cdepth != bdepth) return nvb;
            return nointersection;
          }
          if (nvb.u > 1) {
            if (cdepth != adepth) return nva;
            return nointersection;
          }
          adepth = adepth ^ 1;
          bdepth = bdepth ^ 1;
        } else {
          NV r;
          for (int k = 0; k < 3; ++k) r.n[k] = (nva.n[k] + nvb.n[k]) / 2.;
          r.u = nva.u;
          return r;
        }
      }
    }
    return nointersection;
  }

  NV first_normal(int s, const double* p0, const double* p1) const {
    const Shape& sh = sc.shapes[s];
    switch (sh.kind) {
      case MONSEL_SHAPE_PLANES: return planes_first(sh, p0, p1);
      case MONSEL_SHAPE_CYLINDER: return cylinder_first(sh, p0, p1);
      case MONSEL_SHAPE_INTERSECTION: return intersection_first(sh, p0, p1);
      case MONSEL_SHAPE_UNION: return union_first(sh, p0, p1);
      case MONSEL_SHAPE_SPHERE: return sphere_first(sh, p0, p1);
    }
    return no_intersection();
  }

  /* RegionBase.cu:50-60 */
  int containing_sub_region(int r, const double* pos) const {
    const Region& rg = sc.regions[r];
    if (contains(rg.shape, pos)) {
      for (int c : rg.children) {
        int csr = containing_sub_region(c, pos);
        if (csr >= 0) return csr;
      }
      return r;
    }
    return -1;
  }

  /* RegionBase.cu:141-147 */
  bool is_containing_region(int r, int target) const {
    for (int c : sc.regions[r].children)
      if (c == target || is_containing_region(c, target)) return true;
    return false;
  }

  /* RegionBase.cu:74-119.  shapeNormal[] receives the normal each top-level
   * shape query returned (the NormalShape::getPreviousNormal side channel). */
  int find_end_of_step(int r, const double* p0, double* p1, std::vector<NV>& shapeNormal, double* tOut) const {
    const Region& rg = sc.regions[r];
    int base = r;
    NV own = first_normal(rg.shape, p0, p1);
    shapeNormal[rg.shape] = own;
    double t = own.u;
    if ((t <= 1.0) && rg.parent >= 0) base = rg.parent;
    int res = r;
    for (int c : rg.children) {
      int cs = sc.regions[c].shape;
      NV cand = first_normal(cs, p0, p1);
      shapeNormal[cs] = cand;
      if ((cand.u <= 1.0) && (cand.u < t)) {
        t = cand.u;
        base = c;
      }
    }
    if (tOut) *tOut = t;
    if (t <= 1.0) {
      double delta[3] = {p1[0] - p0[0], p1[1] - p0[1], p1[2] - p0[2]};
      for (int k = 0; k < 3; ++k) p1[k] = p0[k] + (t * delta[k]);
      double m = ::sqrt(delta[0] * delta[0] + delta[1] * delta[1] + delta[2] * delta[2]);
      double over[3];
      for (int k = 0; k < 3; ++k) over[k] = p1[k] + SMALL_DISP * (delta[k] / m);
      while (base >= 0) {
        res = containing_sub_region(base, over);
        if (res >= 0) return res;
        base = sc.regions[base].parent;
      }
      return -1;
    }
    return res;
  }
};

/* ======================= scatter mechanisms ============================== */

struct Ctx {
  const Scene& sc;
  Geometry geo;
  std::vector<NV> shapeNormal;
  explicit Ctx(const Scene& s) : sc(s), geo(s), shapeNormal(s.shapes.size(), no_intersection()) {}
};

/* SelectableElasticSM.cu:25-40: cumulative scaled cross-sections at kE. */
static double elastic_cumulative(const Mech& m, double kE, double* cum) {
  double total = 0.;
  for (size_t i = 0; i < m.rse.size(); i++) {
    total += m.rse[i].total(kE) * m.scalefactor[i];
    cum[i] = total;
  }
  return total;
}

/* ScatterMechanism::scatterRate implementations. */
static double mech_rate(const Scene& sc, const Mech& m, double E) {
  switch (m.kind) {
    case MONSEL_MECH_NISTMOTT_ELASTIC:
    case MONSEL_MECH_BROWNING_ELASTIC:
    case MONSEL_MECH_RUTHERFORD_ELASTIC: { /* SelectableElasticSM.cu:42-47 */
      double cum[MONSEL_MAX_ELEMENTS];
      return elastic_cumulative(m, E, cum) * m.densityNa;
    }
    case MONSEL_MECH_FITTED_INELASTIC: /* FittedInelSM.cu:20-25 */
      if (E <= (m.energySEgen + m.eFermi)) return 0.;
      return (-m.sdAlg.compute(1.e-10, E) * 1.e10) / m.energySEgen;
    case MONSEL_MECH_POLARON_TRAP: /* GanachaudMokraniPolaronTrapSM.cu:21-27 */
      if (E > m.cutoff) return 0.;
      return m.prefactor * ::exp(-m.extinction * E);
    case MONSEL_MECH_PHONON: { /* GanachaudMokraniPhononInelasticSM.cu:46-72 */
      if (E < m.phononE) return 0.;
      double x = m.phononE / E;
      if (x < 0.1) return m.phPrefactor * x * ::log(4. / x);
      if (x >= 1.) return 0.;
      double temp = ::sqrt(1. - x);
      return m.phPrefactor * x * ::log((1. + temp) / (1. - temp));
    }
    case MONSEL_MECH_TABULATED_INELASTIC: { /* TabulatedInelasticSM.cu:436-459 */
      double kEa = E + m.energyOffset;
      if (kEa < m.tableIIMFPEiDomain[0]) return 0.;
      return m.rateMult * nu_interpolate(*m.tIIMFP, &kEa, 1, sc.d3Compat());
    }
  }
  return 0.;
}

/* TabulatedInelasticSM.cu:87-101 */
static void sim_ESEf(const Mech& m, double Eq, double deltaE, double r, double* ESEf, double* theta) {
  double q = ::sqrt(Eq);
  double kz = (deltaE - Eq) / 2. / q;
  double kzf = kz + q;
  double Ezq = kzf * kzf;
  double minE = (m.offsetFermiEnergy + m.bandgap) - Ezq;
  if (minE < 0.) minE = 0.;
  double maxE = m.offsetFermiEnergy - (kz * kz);
  double Exy = (minE * (1. - r)) + (maxE * r);
  *ESEf = Exy + Ezq;
  *theta = ::acos(kzf / ::sqrt(*ESEf));
}

/* TabulatedInelasticSM.cu:103-151 */
static double compute_E0_from_dispersion(double Eq, double deltaE) {
  if (Eq == 0.) return deltaE;
  double x = Eq / deltaE;
  double y = 3.77300614251479e17 * Eq;
  double x13 = ::pow(x, 1 / 3.);
  double x2 = x * x;
  double c1 = x2 * x2 * (27. + (18. * y) + (2 * y * y));
  double c2 = x2 * (27. + (4. * y));
  double c3 = c2 - 27.;
  double c4 = x2 * (3. + y);
  double c6 = 1 - x2;
  if (c3 > 0.) {
    double tanPart = ::atan2(3. * c6 * ::sqrt(3. * c3 * c6), (18. * c4) - c1 - 27.);
    double trigPart = ::cos(tanPart / 3.);
    double prefactor = 2. * x * ::sqrt(y * (y - (c6 * (y + 6.))));
    trigPart *= prefactor;
    return deltaE * ::sqrt(((3. - c4) + trigPart) / 3.);
  }
  double c5 = ::pow(-c1 + (18. * c4) + (3. * (-9. + (c6 * ::sqrt(-(3. * c3 * c6))))), 1. / 3.);
  double term1 = -2. * c4;
  double x23 = x13 * x13;
  double x43 = x * x13;
  double y13 = ::pow(y, 1. / 3.);
  double y23 = y13 * y13;
  double x43y23 = x43 * y23;
  double two13 = ::pow(2., 1. / 3.);
  double term2 = -12. * two13 * x43y23 * c6;
  double term3 = 2. * two13 * x43y23 * x2 * y;
  double term23 = (term2 + term3) / c5;
  double term4 = two13 * two13 * x23 * y13 * c5;
  return deltaE * ::sqrt((6 + term1 + term23 + term4) / 6.);
}

/* TabulatedInelasticSM.cu:153-207 (default branching ratios). */
static double pick_BE(const Mech& m, double Eq, double deltaE) {
  if ((m.coreEnergies.size() > 0) && (deltaE <= m.coreEnergies[0])) return 0.;
  double energyForIonization = m.E0fromDispersion ? compute_E0_from_dispersion(Eq, deltaE) : deltaE;
  size_t i;
  for (i = 0; (i < m.coreEnergies.size()) && (m.coreEnergies[i] <= energyForIonization); i++)
    ;
  if (i == 0) return 0.;
  return m.coreEnergies[i - 1];
}

struct Spawn {
  bool has = false;
  double theta = 0, phi = 0, E = 0;
  /* relative rotation applied after construction (Electron::updateDirection) */
  bool rotate = false;
  double dTheta = 0, dPhi = 0;
};

/* TabulatedInelasticSM.cu:209-434 */
static Spawn tabulated_scatter(const Scene& sc, const Mech& m, Electron& pe) {
  Spawn se;
  double kE0 = pe.E;
  double kE = kE0 + m.energyOffset;
  if (kE < m.tableEiDomain[0]) return se;
  double theta = 0., phi = 0.;
  double randoms[] = {pe.rng.next(), pe.rng.next(), pe.rng.next(), pe.rng.next()};
  double in[3] = {kE, randoms[0], 0};
  double deltaE = kE * nu_interpolate(*m.tRDE, in, 3, sc.d3Compat());
  if ((deltaE < m.energyGap) && (deltaE > (0.95 * m.energyGap))) deltaE = m.energyGap;
  double theta0PE = pe.theta, phi0PE = pe.phi;
  if (deltaE >= m.bandgap) {
    in[1] = (deltaE - m.energyGap) / (kE - m.offsetFermiEnergy - (2. * m.energyGap));
    if (in[1] > 1.) in[1] = 1.;
    else if (in[1] < 0.) in[1] = 0.;
    in[2] = randoms[1];
    theta = nu_interpolate(*m.tTheta, in, 3, sc.d3Compat());
    phi = 2. * PI * randoms[2];
    pe.updateDirection(theta, phi);
  }
  pe.E = kE0 - deltaE;
  if (deltaE < m.bandgap) return se;
  double Eq = (2. * kE) - deltaE - (2. * ::sqrt(kE * (kE - deltaE)) * ::cos(theta));
  double energySE, be;
  switch (m.methodSE) {
    case 1:
      energySE = (deltaE + m.bEref) - pick_BE(m, Eq, deltaE);
      if ((energySE + m.energyCBbottom) < m.minEgenSE) return se;
      se.has = true;
      se.theta = theta0PE; se.phi = phi0PE; se.E = energySE;
      se.rotate = true; se.dTheta = PI / 2. - theta; se.dPhi = phi + PI;
      return se;
    case 2:
      be = pick_BE(m, Eq, deltaE);
      if (be > 0.) energySE = (deltaE + m.bEref) - be;
      else {
        double in2[2] = {deltaE, randoms[3]};
        double energy0SE = nu_interpolate(*m.tSEE0, in2, 3, sc.d3Compat());
        if (energy0SE < m.energyRangeSE0[0]) energy0SE = m.energyRangeSE0[0];
        else if (energy0SE > m.energyRangeSE0[1]) energy0SE = m.energyRangeSE0[1];
        energySE = (deltaE + energy0SE) - m.energyOffset;
      }
      if ((energySE + m.energyCBbottom) < m.minEgenSE) return se;
      se.has = true;
      se.theta = theta0PE; se.phi = phi0PE; se.E = energySE;
      se.rotate = true; se.dTheta = PI / 2. - theta; se.dPhi = phi + PI;
      return se;
    case 3:
      be = pick_BE(m, Eq, deltaE);
      if (be > 0.) {
        energySE = (deltaE + m.bEref) - be;
        if ((energySE + m.energyCBbottom) < m.minEgenSE) return se;
        double ts = ::acos(1. - (2. * pe.rng.next()));
        double ps = 2. * PI * pe.rng.next();
        se.has = true; se.theta = ts; se.phi = ps; se.E = energySE;
        return se;
      } else {
        double root = 2. * ::sqrt(m.offsetFermiEnergy * (m.offsetFermiEnergy + deltaE));
        double sum = (2. * m.offsetFermiEnergy) + deltaE;
        double Eqmin = sum - root, Eqmax = sum + root;
        if ((Eqmin <= Eq) && (Eq <= Eqmax)) {
          double ESEf, th;
          sim_ESEf(m, Eq, deltaE, randoms[3], &ESEf, &th);
          energySE = ESEf - m.energyOffset;
          if ((energySE + m.energyCBbottom) < m.minEgenSE) return se;
          double thetaSE = PI / 2. - theta, phiSE = phi + PI;
          double nt, np;
          update_direction_pair(thetaSE, phiSE, th, 2. * PI * pe.rng.next(), &nt, &np);
          se.has = true; se.theta = theta0PE; se.phi = phi0PE; se.E = energySE;
          se.rotate = true; se.dTheta = nt; se.dPhi = np;
          return se;
        } else {
          double in2[2] = {deltaE, randoms[3]};
          double energy0SE = nu_interpolate(*m.tSEE0, in2, 3, sc.d3Compat());
          if (energy0SE < m.energyRangeSE0[0]) energy0SE = m.energyRangeSE0[0];
          else if (energy0SE > m.energyRangeSE0[1]) energy0SE = m.energyRangeSE0[1];
          energySE = (deltaE + energy0SE) - m.energyOffset;
          if ((energySE + m.energyCBbottom) < m.minEgenSE) return se;
          double ts = ::acos(1. - (2. * pe.rng.next()));
          double ps = 2 * PI * pe.rng.next();
          se.has = true; se.theta = ts; se.phi = ps; se.E = energySE;
          return se;
        }
      }
  }
  return se;
}

/* ScatterMechanism::scatter implementations; the returned Spawn describes
 * the secondary (new Electron(*pe, theta, phi, E) [+ updateDirection]). */
static Spawn mech_scatter(const Scene& sc, const Mech& m, Electron& pe) {
  Spawn se;
  switch (m.kind) {
    case MONSEL_MECH_NISTMOTT_ELASTIC:
    case MONSEL_MECH_BROWNING_ELASTIC:
    case MONSEL_MECH_RUTHERFORD_ELASTIC: { /* SelectableElasticSM.cu:49-67 */
      double kE = pe.prevE;
      double cum[MONSEL_MAX_ELEMENTS];
      double total = elastic_cumulative(m, kE, cum);
      double r = pe.rng.next() * total;
      int index = 0;
      while (cum[index] < r) index++;
      double alpha = m.rse[index].angle(kE, pe.rng);
      double beta = 2 * PI * pe.rng.next();
      pe.updateDirection(alpha, beta);
      return se;
    }
    case MONSEL_MECH_FITTED_INELASTIC: { /* FittedInelSM.cu:13-18 */
      double phi = 2 * PI * pe.rng.next();
      double theta = ::acos(1. - (2. * pe.rng.next()));
      se.has = true; se.theta = theta; se.phi = phi; se.E = m.energySEgen + m.eFermi;
      return se;
    }
    case MONSEL_MECH_POLARON_TRAP: /* GanachaudMokraniPolaronTrapSM.cu:15-20 */
      pe.E = 0.;
      pe.complete = true;
      return se;
    case MONSEL_MECH_PHONON: { /* GanachaudMokraniPhononInelasticSM.cu:19-44 */
      double kE0 = pe.E;
      if (kE0 < m.phononE) return se;
      double x = m.phononE / kE0;
      double r0 = pe.rng.next(), r1 = pe.rng.next();
      double costheta;
      if (x < 0.1) costheta = 1 + (((x * x) - (::pow(16., r0) * ::pow(x, 2. - (2. * r0)))) / 8.);
      else {
        double root = ::sqrt(1. - x);
        double temp = ::pow(((-2. * (1 + root)) + x) / ((-2. * (1 - root)) + x), r0);
        costheta = temp + (((x - 2.) * (temp - 1)) / 2. / root);
      }
      double phi = 2. * PI * r1;
      pe.updateDirection(::acos(costheta), phi);
      pe.E = kE0 - m.phononE;
      pe.prevE = kE0;
      return se;
    }
    case MONSEL_MECH_TABULATED_INELASTIC:
      return tabulated_scatter(sc, m, pe);
  }
  return se;
}

/* ExpQMBarrierSM.cu:30-48 */
static double sharp_barrier_T(double rootPerpE, double rootDiff) {
  double ksum = rootPerpE + rootDiff;
  return (4. * rootPerpE * rootDiff) / (ksum * ksum);
}
static double general_barrier_T(double lambdaFactor, double rootPerpE, double rootDiff) {
  double k1 = lambdaFactor * rootPerpE;
  if (k1 > 50.) return 1.;
  double k2 = lambdaFactor * rootDiff;
  double sinhPlus = ::sinh(k1 + k2), sinhMinus = ::sinh(k1 - k2);
  double ratio = sinhMinus / sinhPlus;
  return 1. - (ratio * ratio);
}

/* ExpQMBarrierSM.cu:98-219: the crossing itself, given the outward barrier
 * normal nb (ExpQMBarrierSM.cu:77-96 resolves it from the struck shape).
 * Returns +1 when the electron moved (or was nudged) along +nb, -1 along -nb. */
static int barrier_core(const Scene& sc, const MSM& msm, Electron& pe, int nextRegion, const double nbIn[3]) {
  int cur = pe.region;
  const Material& curMat = sc.materials[sc.msms[sc.regions[cur].msm].material];
  const Material& nextMat = sc.materials[sc.msms[sc.regions[nextRegion].msm].material];
  double deltaU = -curMat.cb + nextMat.cb;
  double nb[3] = {nbIn[0], nbIn[1], nbIn[2]};
  double st0 = ::sin(pe.theta);
  double n0[3] = {st0 * ::cos(pe.phi), st0 * ::sin(pe.phi), ::cos(pe.theta)};
  if (sc.barrierCompat()) {
    for (int i = 0; i < 3; ++i) nb[i] = n0[i]; /* ExpQMBarrierSM.cu:111 */
  } else {
    double m = ::sqrt(nb[0] * nb[0] + nb[1] * nb[1] + nb[2] * nb[2]);
    if (m > 0) for (int i = 0; i < 3; ++i) nb[i] /= m;
    else for (int i = 0; i < 3; ++i) nb[i] = n0[i];
  }
  double cosalpha = n0[0] * nb[0] + n0[1] * nb[1] + n0[2] * nb[2];
  if (cosalpha <= 0.) {
    for (int i = 0; i < 3; ++i) pe.pos[i] -= SMALL_DISP * nb[i];
    return -1;
  }
  if (deltaU == 0.) {
    for (int i = 0; i < 3; ++i) pe.pos[i] += SMALL_DISP * nb[i];
    pe.setRegion(nextRegion);
    return 1;
  }
  double kE0 = pe.E;
  double perpE = kE0 <= 0. ? 0. : cosalpha * cosalpha * kE0;
  double rootPerpE = 0., rootDiff = 0.;
  bool transmits;
  if ((perpE == 0.) || (perpE <= deltaU)) transmits = false;
  else {
    rootPerpE = ::sqrt(perpE);
    rootDiff = ::sqrt(perpE - deltaU);
    if (msm.barrierKind == MONSEL_BARRIER_EXPQM_CLASSICAL) transmits = true;
    else {
      double tp = msm.lambda == 0. ? sharp_barrier_T(rootPerpE, rootDiff)
                                   : general_barrier_T(msm.lambdaFactor, rootPerpE, rootDiff);
      transmits = pe.rng.next() < tp;
    }
  }
  double nf[3];
  if (transmits) {
    double factor = cosalpha * ((rootDiff / rootPerpE) - 1.);
    for (int i = 0; i < 3; i++) nf[i] = n0[i] + (factor * nb[i]);
    nf[2] /= ::sqrt((nf[0] * nf[0]) + (nf[1] * nf[1]) + (nf[2] * nf[2]));
    pe.E = kE0 - deltaU;
    pe.setRegion(nextRegion);
    for (int i = 0; i < 3; ++i) pe.pos[i] += SMALL_DISP * nb[i];
  } else {
    double two = 2. * cosalpha;
    for (int i = 0; i < 3; i++) nf[i] = n0[i] - (nb[i] * two);
    for (int i = 0; i < 3; ++i) pe.pos[i] -= SMALL_DISP * nb[i];
  }
  pe.theta = ::acos(nf[2]);
  pe.phi = ::atan2(nf[1], nf[0]);
  return transmits ? 1 : -1;
}

/* ExpQMBarrierSM.cu:50-219 (and NullMaterialScatterModel.cu:14-20). */
static void barrier_scatter(Ctx& cx, const MSM& msm, Electron& pe, int nextRegion) {
  const Scene& sc = cx.sc;
  if (msm.isNull || msm.barrierKind == MONSEL_BARRIER_NONE) {
    pe.setRegion(nextRegion);
    return;
  }
  int cur = pe.region;
  if (cur == nextRegion) return;
  /* ExpQMBarrierSM.cu:77-96: outward normal of the struck shape, negated
   * when the electron enters a child region */
  double nb[3] = {0, 0, 0};
  if (cx.geo.is_containing_region(cur, nextRegion)) {
    int struck = nextRegion;
    while (sc.regions[struck].parent != cur) struck = sc.regions[struck].parent;
    const NV& v = cx.shapeNormal[sc.regions[struck].shape];
    for (int i = 0; i < 3; i++) nb[i] = -v.n[i];
  } else {
    const NV& v = cx.shapeNormal[sc.regions[cur].shape];
    for (int i = 0; i < 3; i++) nb[i] = v.n[i];
  }
  barrier_core(sc, msm, pe, nextRegion, nb);
}

/* MONSEL_MaterialScatterModel.cu:44-62: cumulative mechanism rates. */
static double msm_rates(const Scene& sc, const MSM& msm, double E, double* cum) {
  double total = 0.;
  for (size_t i = 0; i < msm.mech.size(); ++i) {
    total += mech_rate(sc, msm.mech[i], E);
    cum[i] = total;
  }
  return total;
}

/* ======================= histograms (Histogram.cu, Algorithm.cu) ========= */

/* Histogram::bin with bin edges min + i*delta (Histogram.cu:25-42, 73-79).
 * compat: reference insertion-point semantics of Algorithm::binarySearch
 * (values between edges one bin high; out-of-range values clamped to the
 * over-range slot instead of writing past the array). */
static int hist_bin(double val, double min, double max, int nbins, bool compat) {
  const double delta = (max - min) / nbins;
  int l = 0, r = nbins;
  while (l <= r) {
    int mid = l + (r - l) / 2;
    double edge = mid == 0 ? min : min + mid * delta;
    if (edge == val) return mid;
    if (edge < val) l = mid + 1;
    else r = mid - 1;
  }
  if (compat) return l > nbins ? nbins : l;
  return l - 1;
}

/* ======================= transport ======================================= */

struct PixelTally {
  uint64_t back = 0, fwd = 0, se = 0;
  double se_sq = 0, bse_sq = 0, back_energy = 0;
};

struct JobHist {
  std::vector<uint64_t> backE, fwdE, elev, azim;
  void init(int nb) {
    backE.assign(nb + 2, 0);
    fwdE.assign(nb + 2, 0);
    elev.assign(182, 0);
    azim.assign(362, 0);
  }
};

struct Counters {
  uint64_t electrons = 0, steps = 0, escapes = 0, watchdog = 0;
};

struct JobParams {
  double E0, width;
  int nbins;
  double E0eV;
  int maxSEbin;
  long maxSteps;
  uint64_t seed;
};

/* GaussianBeam::createElectron (GaussianBeam.cu:52-65): the start point
 * center + r (cos th, sin th, 0), r = width sqrt(-2 ln U1), th = 2 pi U2,
 * drawn from the electron's stream. */
static void gaussian_beam_start(const JobParams& jp, const double* center, Stream& rng, double* ipos) {
  ipos[0] = center[0];
  ipos[1] = center[1];
  ipos[2] = center[2];
  double r = ::sqrt(-2. * ::log(rng.next())) * jp.width;
  double th = 2.0 * PI * (rng.next());
  ipos[0] += r * ::cos(th);
  ipos[1] += r * ::sin(th);
}

/* MonteCarloSS.cu:99-231 for one primary; returns (nSE, nBack). */
static void run_trajectory(Ctx& cx, const JobParams& jp, const double* center, uint32_t primary, PixelTally& pt,
                           JobHist& hist, Counters& cnt) {
  const Scene& sc = cx.sc;
  std::vector<Electron> stack;
  Electron cur;
  cur.rng.init(jp.seed, 0, primary);
  double ipos[3];
  gaussian_beam_start(jp, center, cur.rng, ipos);
  cur.init(ipos, 0., 0., jp.E0);
  cur.setRegion(cx.geo.containing_sub_region(0, ipos));
  cnt.electrons++;
  long nSE = 0, nBack = 0;
  for (;;) {
    /* allElectronsComplete (MonteCarloSS.cu:194-207) */
    while (cur.complete && !stack.empty()) {
      cur = stack.back();
      stack.pop_back();
    }
    if (cur.complete) break;
    /* takeStep (MonteCarloSS.cu:107-163) */
    cnt.steps++;
    int region = cur.region;
    if (region < 0 || !cx.geo.contains(sc.regions[region].shape, cur.pos)) {
      region = cx.geo.containing_sub_region(0, cur.pos);
      cur.setRegion(region);
      if (region < 0) { cur.complete = true; continue; }
    }
    const MSM& msm = sc.msms[sc.regions[region].msm];
    double cum[MONSEL_MAX_MECH];
    double total = 0;
    double freePath;
    if (msm.isNull) freePath = 1.0; /* NullMaterialScatterModel.cu:32-35 */
    else {
      total = msm_rates(sc, msm, cur.E, cum);
      double maxFreePath = 2. * sc.chamberRadius;
      if (total != 0.) {
        double fp = -::log(cur.rng.next()) / total;
        freePath = fp > maxFreePath ? maxFreePath : fp;
      } else {
        if (cur.steps > 1000000) cur.complete = true;
        freePath = maxFreePath;
      }
    }
    double pos0[3] = {cur.pos[0], cur.pos[1], cur.pos[2]};
    double pos1[3];
    cur.candidatePoint(freePath, pos1);
    int next = cx.geo.find_end_of_step(region, pos0, pos1, cx.shapeNormal, nullptr);
    double d = ::sqrt((pos1[0] - pos0[0]) * (pos1[0] - pos0[0]) + (pos1[1] - pos0[1]) * (pos1[1] - pos0[1]) +
                      (pos1[2] - pos0[2]) * (pos1[2] - pos0[2]));
    double dE = msm.isNull ? 0. : msm.csd.compute(d, cur.E);
    cur.move(pos1, dE);
    if (cur.steps > jp.maxSteps) {
      cur.complete = true;
      cnt.watchdog++;
      continue;
    }
    bool tc = (cur.E < msm.minE) || cur.complete;
    cur.complete = tc;
    if (tc) continue;
    if (next == region) {
      /* MONSEL_MaterialScatterModel.cu:93-134 */
      Spawn sp;
      if (!msm.isNull && total != 0.) {
        double rr = cur.rng.next() * total;
        int index = 0;
        while (cum[index] < rr) index++;
        sp = mech_scatter(sc, msm.mech[index], cur);
      }
      cur.complete = (cur.E < msm.minE) || cur.complete;
      if (sp.has) {
        Electron se;
        se.init(cur.pos, sp.theta, sp.phi, sp.E);
        se.region = se.prevRegion = cur.region;
        if (sp.rotate) se.updateDirection(sp.dTheta, sp.dPhi);
        se.rng.init(jp.seed, child_lineage(cur.rng.lineage, cur.nchildren++), primary);
        /* trackSecondaryElectron (MonteCarloSS.cu:165-174) */
        double minE = sc.msms[sc.regions[se.region].msm].minE;
        if (se.E > minE) {
          stack.push_back(cur);
          cur = se;
          cnt.electrons++;
        }
      }
    } else if (next >= 0) {
      barrier_scatter(cx, msm, cur, next);
      double np[3];
      cur.candidatePoint(SMALL_DISP, np);
      for (int k = 0; k < 3; ++k) cur.pos[k] = np[k];
      if (cur.region < 0 || !cx.geo.contains(sc.regions[cur.region].shape, cur.pos))
        cur.setRegion(cx.geo.containing_sub_region(0, cur.pos));
    } else {
      /* BackscatterEvent -> BackscatterStats.cu:147-170 */
      const double* pos = cur.pos;
      const double elevation = (PI / 2) - ::atan2(pos[2], ::sqrt((pos[0] * pos[0]) + (pos[1] * pos[1])));
      double azimuth = ::atan2(pos[1], pos[0]);
      if (azimuth < 0.0) azimuth = (2.0 * PI) + azimuth;
      bool compat = sc.histCompat();
      hist.elev[hist_bin(elevation, 0.0, PI, 180, compat) + 1]++;
      hist.azim[hist_bin(azimuth, 0.0, 2.0 * PI, 360, compat) + 1]++;
      const double kEeV = cur.E * FROMSI_EV;
      int bin = hist_bin(kEeV, 0.0, jp.E0eV, jp.nbins, compat);
      if (elevation < (PI / 2.0)) {
        hist.fwdE[bin + 1]++;
        pt.fwd++;
      } else {
        hist.backE[bin + 1]++;
        pt.back++;
        pt.back_energy += kEeV;
        nBack++;
        if (bin >= 0 && bin < jp.maxSEbin) { pt.se++; nSE++; }
      }
      cnt.escapes++;
      cur.region = -1;
      cur.complete = true;
    }
  }
  pt.se_sq += (double)nSE * (double)nSE;
  pt.bse_sq += (double)(nBack - nSE) * (double)(nBack - nSE);
}

/* ======================= scene import ==================================== */

static bool fail(const std::string& m) {
  g_err = m;
  return false;
}

static bool import_scene(const monsel_scene_desc* d, Scene& sc) {
  if (!d || d->abi_version != MONSEL_ABI_VERSION) return fail("bad scene descriptor / abi version");
  sc.compat = d->compat_flags;
  sc.chamberRadius = d->chamber_radius > 0 ? d->chamber_radius : CHAMBER_RADIUS_DEFAULT;
  for (int i = 0; i < d->n_mott; ++i) {
    sc.mottZ.push_back(d->mott[i].Z);
    sc.mottSpwem.emplace_back(d->mott[i].spwem, d->mott[i].spwem + MONSEL_MOTT_NE);
    sc.mottX1.emplace_back(d->mott[i].x1, d->mott[i].x1 + MONSEL_MOTT_NE * MONSEL_MOTT_NQ);
  }
  for (int i = 0; i < d->n_nu_tables; ++i) {
    const monsel_nu_table& t = d->nu_tables[i];
    NUTable u;
    u.dim = t.dim;
    size_t tot = 1;
    for (int k = 0; k < t.dim; ++k) {
      u.n[k] = t.n[k];
      u.x[k].assign(t.x[k], t.x[k] + t.n[k]);
      tot *= (size_t)t.n[k];
      if (u.x[k][0] < u.x[k][t.n[k] - 1]) { u.domain[k][0] = u.x[k][0]; u.domain[k][1] = u.x[k][t.n[k] - 1]; }
      else { u.domain[k][1] = u.x[k][0]; u.domain[k][0] = u.x[k][t.n[k] - 1]; }
    }
    u.v.assign(t.values, t.values + tot);
    /* NUTableInterpolation.cu:169-172: if / else-if range update */
    u.range[0] = INFINITY;
    u.range[1] = -INFINITY;
    for (double v : u.v) {
      if (v < u.range[0]) u.range[0] = v;
      else if (v > u.range[1]) u.range[1] = v;
    }
    sc.nu.push_back(std::move(u));
  }
  for (int i = 0; i < d->n_materials; ++i) {
    const monsel_material_desc& m = d->materials[i];
    Material mm;
    mm.nce = m.n_elements;
    for (int k = 0; k < m.n_elements; ++k) { mm.Z[k] = m.Z[k]; mm.w[k] = m.weight_fraction[k]; mm.A[k] = m.atomic_weight[k]; }
    mm.density = m.density;
    mm.workfunction = m.workfunction;
    mm.cb = m.energy_cb_bottom;
    mm.bandgap = m.bandgap;
    mm.core.assign(m.core_energy, m.core_energy + m.n_core);
    sc.materials.push_back(mm);
  }
  for (int i = 0; i < d->n_msm; ++i) {
    const monsel_msm_desc& md = d->msms[i];
    MSM msm;
    msm.isNull = md.is_null != 0;
    msm.material = md.material;
    const Material& mat = sc.materials[md.material];
    msm.minE = md.min_e_for_tracking;
    if (md.csd_kind == MONSEL_CSD_JOY_LUO_NIEMINEN) msm.csd.init(mat, md.csd_break_e);
    msm.barrierKind = md.barrier_kind;
    msm.u0 = mat.cb;
    msm.lambda = md.barrier_lambda;
    msm.lambdaFactor = (PI * md.barrier_lambda * ::sqrt(EMASS / 2.)) / HBAR;
    for (int k = 0; k < md.n_mech; ++k) {
      const monsel_mech_desc& e = md.mech[k];
      Mech m;
      m.kind = e.kind;
      switch (e.kind) {
        case MONSEL_MECH_NISTMOTT_ELASTIC:
        case MONSEL_MECH_BROWNING_ELASTIC:
        case MONSEL_MECH_RUTHERFORD_ELASTIC:
          m.densityNa = mat.density * AVOGADRO;
          for (int j = 0; j < mat.nce; ++j) {
            ElementScatter es;
            es.kind = e.kind;
            es.Z = mat.Z[j];
            es.browning.init(mat.Z[j]);
            es.rutherford.init(mat.Z[j]);
            if (e.kind == MONSEL_MECH_NISTMOTT_ELASTIC) {
              int t = -1;
              for (size_t q = 0; q < sc.mottZ.size(); ++q) if (sc.mottZ[q] == mat.Z[j]) t = (int)q;
              if (t < 0) return fail("missing NIST Mott table for Z=" + std::to_string(mat.Z[j]));
              es.init_mott(mat.Z[j], e.method < 1 ? 1 : e.method, sc.mottSpwem[t].data(), sc.mottX1[t].data());
            }
            m.rse.push_back(es);
            m.scalefactor.push_back((1000. * mat.w[j]) / mat.A[j]);
          }
          break;
        case MONSEL_MECH_FITTED_INELASTIC:
          m.energySEgen = e.p[0];
          m.eFermi = mat.eFermi();
          if ((int)e.p[2] == MONSEL_CSD_JOY_LUO_NIEMINEN) m.sdAlg.init(mat, e.p[1]);
          break;
        case MONSEL_MECH_POLARON_TRAP:
          m.prefactor = e.p[0];
          m.extinction = e.p[1];
          m.cutoff = -::log(10. / m.prefactor) / m.extinction;
          break;
        case MONSEL_MECH_PHONON: {
          double ratemult = e.p[0], phE = e.p[1], T = e.p[2], eps0 = e.p[3], epsInf = e.p[4];
          double occ = 0.5 * (1. + (1. / (::exp(phE / (KBOLTZ * T)) - 1.)));
          double epsRatio = (eps0 - epsInf) / eps0 / epsInf;
          m.phononE = phE;
          m.phPrefactor = (ratemult * occ * epsRatio) / BOHR;
          break;
        }
        case MONSEL_MECH_TABULATED_INELASTIC: {
          m.methodSE = (e.method != 2 && e.method != 3) ? 1 : e.method;
          m.energyOffset = e.p[0];
          m.minEgenSE = e.p[1];
          m.rateMult = e.p[2];
          m.E0fromDispersion = e.p[3] != 0;
          for (int q = 0; q < 3; ++q)
            if (e.table[q] < 0 || e.table[q] >= (int)sc.nu.size()) return fail("tabulated mechanism table index");
          m.tIIMFP = &sc.nu[e.table[0]];
          m.tRDE = &sc.nu[e.table[1]];
          m.tTheta = &sc.nu[e.table[2]];
          if (m.methodSE >= 2) {
            if (e.table[3] < 0 || e.table[3] >= (int)sc.nu.size()) return fail("tabulated SE table index");
            m.tSEE0 = &sc.nu[e.table[3]];
            m.energyRangeSE0[0] = m.tSEE0->range[0];
            m.energyRangeSE0[1] = m.tSEE0->range[1];
          }
          m.energyCBbottom = mat.cb;
          m.bandgap = mat.bandgap;
          m.energyGap = mat.bandgap;
          m.offsetFermiEnergy = mat.eFermi() + m.energyOffset;
          m.coreEnergies = mat.core;
          m.bEref = mat.bandgap > 0. ? -mat.bandgap : mat.eFermi();
          m.tableEiDomain[0] = m.tRDE->domain[0][0];
          m.tableEiDomain[1] = m.tRDE->domain[0][1];
          if (m.tTheta->domain[0][0] > m.tableEiDomain[0]) m.tableEiDomain[0] = m.tTheta->domain[0][0];
          if (m.tTheta->domain[0][1] < m.tableEiDomain[1]) m.tableEiDomain[1] = m.tTheta->domain[0][1];
          m.tableIIMFPEiDomain[0] = m.tIIMFP->domain[0][0];
          m.tableIIMFPEiDomain[1] = m.tIIMFP->domain[0][1];
          break;
        }
        default:
          return fail("unknown mechanism kind " + std::to_string(e.kind));
      }
      msm.mech.push_back(std::move(m));
    }
    sc.msms.push_back(std::move(msm));
  }
  for (int i = 0; i < d->n_shapes; ++i) {
    const monsel_shape_desc& s = d->shapes[i];
    Shape sh;
    sh.kind = s.kind;
    sh.a = s.a;
    sh.b = s.b;
    if (s.kind == MONSEL_SHAPE_PLANES) {
      for (int k = 0; k < s.n_planes; ++k) {
        const monsel_plane_desc& p = d->planes[s.first_plane + k];
        Plane pl;
        for (int q = 0; q < 3; ++q) { pl.n[q] = p.normal[q]; pl.c[q] = p.point[q]; }
        sh.planes.push_back(pl);
      }
    } else if (s.kind == MONSEL_SHAPE_CYLINDER) {
      for (int q = 0; q < 3; ++q) {
        sh.end0[q] = s.end0[q];
        sh.axis[q] = s.end1[q] - s.end0[q];
        sh.delta[q] = sh.axis[q];
      }
      sh.radius2 = s.radius * s.radius;
      sh.len2 = sh.axis[0] * sh.axis[0] + sh.axis[1] * sh.axis[1] + sh.axis[2] * sh.axis[2];
      sh.len = ::sqrt(sh.len2);
      for (int q = 0; q < 3; ++q) sh.naxis[q] = sh.axis[q] / sh.len;
    } else if (s.kind == MONSEL_SHAPE_SPHERE) {
      for (int q = 0; q < 3; ++q) sh.center[q] = s.center[q];
      sh.radius = s.radius;
    } else if (s.kind != MONSEL_SHAPE_INTERSECTION && s.kind != MONSEL_SHAPE_UNION) {
      return fail("unknown shape kind");
    }
    sc.shapes.push_back(sh);
  }
  for (int i = 0; i < d->n_regions; ++i) {
    Region r;
    r.parent = d->regions[i].parent;
    r.shape = d->regions[i].shape;
    r.msm = d->regions[i].msm;
    sc.regions.push_back(r);
  }
  for (int i = 0; i < d->n_regions; ++i)
    if (sc.regions[i].parent >= 0) sc.regions[sc.regions[i].parent].children.push_back(i);
  return true;
}

}  // namespace oracle

/* ======================= C API ============================================ */

using namespace oracle;

extern "C" {

const char* oracle_last_error(void) { return g_err.c_str(); }

void* oracle_scene_create(const monsel_scene_desc* d) {
  Scene* sc = new Scene();
  if (!import_scene(d, *sc)) {
    delete sc;
    return nullptr;
  }
  return sc;
}

void oracle_scene_destroy(void* s) { delete (Scene*)s; }

/* Runs a job on `threads` host threads (<= 0: hardware concurrency). */
int oracle_run_job(void* scene, const monsel_job_desc* job, monsel_tallies* out, monsel_job_stats* stats,
                   int threads) {
  const Scene& sc = *(const Scene*)scene;
  JobParams jp;
  jp.E0 = job->beam_energy;
  jp.width = job->beam_width;
  jp.nbins = job->n_energy_bins > 0 ? job->n_energy_bins : 1;
  jp.E0eV = job->beam_energy * FROMSI_EV;
  double perBin = jp.E0eV / jp.nbins;
  jp.maxSEbin = (int)(job->se_cutoff_ev / perBin);
  jp.maxSteps = job->max_steps > 0 ? job->max_steps : 1000000;
  jp.seed = job->seed;
  int np = job->n_pixels;
  std::vector<PixelTally> px(np);
  if (threads <= 0) {
    /* default: the host's cores, at most 16 (one GPU's share of a shared box) */
    threads = (int)std::thread::hardware_concurrency();
    if (threads > 16) threads = 16;
  }
  if (threads < 1) threads = 1;
  if (threads > np) threads = np > 0 ? np : 1;
  std::vector<JobHist> hists(threads);
  std::vector<Counters> cnts(threads);
  for (auto& h : hists) h.init(jp.nbins);
  std::atomic<int> nextPixel(0);
  auto worker = [&](int tid) {
    Ctx cx(sc);
    g_lcg = (uint32_t)(job->seed + 7919u * (uint32_t)tid);
    for (;;) {
      int p = nextPixel.fetch_add(1);
      if (p >= np) break;
      for (int i = 0; i < job->primaries_per_pixel; ++i) {
        uint32_t prim = (uint32_t)(job->primary_offset + (uint64_t)p * job->primaries_per_pixel + i);
        run_trajectory(cx, jp, job->pixel_center + 3 * p, prim, px[p], hists[tid], cnts[tid]);
      }
    }
  };
  std::vector<std::thread> pool;
  for (int t = 0; t < threads; ++t) pool.emplace_back(worker, t);
  for (auto& t : pool) t.join();
  JobHist tot;
  tot.init(jp.nbins);
  Counters c;
  for (int t = 0; t < threads; ++t) {
    for (size_t k = 0; k < tot.backE.size(); ++k) { tot.backE[k] += hists[t].backE[k]; tot.fwdE[k] += hists[t].fwdE[k]; }
    for (size_t k = 0; k < tot.elev.size(); ++k) tot.elev[k] += hists[t].elev[k];
    for (size_t k = 0; k < tot.azim.size(); ++k) tot.azim[k] += hists[t].azim[k];
    c.electrons += cnts[t].electrons;
    c.steps += cnts[t].steps;
    c.escapes += cnts[t].escapes;
    c.watchdog += cnts[t].watchdog;
  }
  if (out) {
    for (int p = 0; p < np; ++p) {
      if (out->back) out->back[p] = px[p].back;
      if (out->fwd) out->fwd[p] = px[p].fwd;
      if (out->se) out->se[p] = px[p].se;
      if (out->se_sq) out->se_sq[p] = px[p].se_sq;
      if (out->bse_sq) out->bse_sq[p] = px[p].bse_sq;
      if (out->back_energy) out->back_energy[p] = px[p].back_energy;
    }
    if (out->back_energy_hist) std::copy(tot.backE.begin(), tot.backE.end(), out->back_energy_hist);
    if (out->fwd_energy_hist) std::copy(tot.fwdE.begin(), tot.fwdE.end(), out->fwd_energy_hist);
    if (out->elevation_hist) std::copy(tot.elev.begin(), tot.elev.end(), out->elevation_hist);
    if (out->azimuth_hist) std::copy(tot.azim.begin(), tot.azim.end(), out->azimuth_hist);
  }
  if (stats) {
    memset(stats, 0, sizeof(*stats));
    stats->primaries = (uint64_t)np * job->primaries_per_pixel;
    stats->electrons = c.electrons;
    stats->steps = c.steps;
    stats->escapes = c.escapes;
    stats->watchdog_drops = c.watchdog;
  }
  return 0;
}

/* ---- probes (double precision; energies in eV, lengths in nm) ---------- */

/* Start points (m) and regions of primaries 0..n-1 of a job, as
 * run_trajectory creates them (GaussianBeam.cu:52-65). */
int oracle_probe_primaries(void* scene, const monsel_job_desc* job, int n, double* out) {
  if (!scene || !job || n < 0 || (n > 0 && !out) || job->primaries_per_pixel <= 0) return -1;
  Ctx cx(*(Scene*)scene);
  JobParams jp{};
  jp.E0 = job->beam_energy;
  jp.width = job->beam_width;
  jp.seed = job->seed;
  for (int i = 0; i < n; ++i) {
    const uint32_t primary = (uint32_t)(job->primary_offset + (uint64_t)i);
    const int pix = i / job->primaries_per_pixel;
    if (pix >= job->n_pixels) return -1;
    Stream rng;
    rng.init(jp.seed, 0, primary);
    double p[3];
    gaussian_beam_start(jp, job->pixel_center + 3 * (size_t)pix, rng, p);
    out[4 * i] = p[0];
    out[4 * i + 1] = p[1];
    out[4 * i + 2] = p[2];
    out[4 * i + 3] = cx.geo.containing_sub_region(0, p);
  }
  return 0;
}

/* TabulatedInelasticSM.cu:103-151 as the oracle evaluates it (J). */
double oracle_e0_from_dispersion(double Eq, double deltaE) { return compute_E0_from_dispersion(Eq, deltaE); }

/* ExpQMBarrierSM.cu:21-48: transmission probability of a barrier of width
 * lambda (m; 0 = the sharp step) for perpendicular energy perpE above a step
 * deltaU (J), as barrier_core computes it; 0 where the classical threshold
 * reflects. */
double oracle_barrier_T(double lambda, double perpE, double deltaU) {
  if (perpE <= 0. || perpE <= deltaU) return 0.;
  const double rootPerpE = ::sqrt(perpE), rootDiff = ::sqrt(perpE - deltaU);
  if (lambda == 0.) return sharp_barrier_T(rootPerpE, rootDiff);
  return general_barrier_T((PI * lambda * ::sqrt(EMASS / 2.)) / HBAR, rootPerpE, rootDiff);
}

int oracle_probe_rates(void* scene, int msm, int n, const double* e_ev, double* out) {
  const Scene& sc = *(const Scene*)scene;
  const MSM& m = sc.msms[msm];
  int nm = (int)m.mech.size();
  for (int i = 0; i < n; ++i) {
    double tot = 0;
    for (int k = 0; k < nm; ++k) {
      double r = mech_rate(sc, m.mech[k], e_ev[i] * TOSI_EV) * 1e-9;
      out[i * (nm + 1) + k] = r;
      tot += r;
    }
    out[i * (nm + 1) + nm] = tot;
  }
  return 0;
}

/* Elastic angle with injected uniforms u[2i], u[2i+1] (the first is used by
 * the NIST Mott table and screened Rutherford; Browning uses both). */
int oracle_probe_elastic_angle(void* scene, int msm, int element, int n, const double* e_ev, const double* u,
                               double* out) {
  const Scene& sc = *(const Scene*)scene;
  const MSM& m = sc.msms[msm];
  const Mech* el = nullptr;
  for (const Mech& q : m.mech)
    if (q.kind == MONSEL_MECH_NISTMOTT_ELASTIC || q.kind == MONSEL_MECH_BROWNING_ELASTIC ||
        q.kind == MONSEL_MECH_RUTHERFORD_ELASTIC)
      el = &q;
  if (!el || element < 0 || element >= (int)el->rse.size()) return MONSEL_E_INVALID;
  const ElementScatter& es = el->rse[element];
  for (int i = 0; i < n; ++i) {
    double E = e_ev[i] * TOSI_EV;
    double r1 = u[2 * i], r2 = u[2 * i + 1];
    if (es.kind == MONSEL_MECH_BROWNING_ELASTIC) out[i] = es.browning.angle(E, r1, r2);
    else if (es.kind == MONSEL_MECH_RUTHERFORD_ELASTIC) out[i] = es.rutherford.angle(E, r1);
    else if (E < es.extrapolateBelowEnergy) out[i] = es.browning.angle(E, r1, r2);
    else if (E < MAX_NISTMOTT) out[i] = es.mott_angle(E, r1);
    else out[i] = es.rutherford.angle(E, r1);
  }
  return 0;
}

/* Elastic total cross-section of element slot `element` (nm^2). */
int oracle_probe_elastic_sigma(void* scene, int msm, int element, int n, const double* e_ev, double* out) {
  const Scene& sc = *(const Scene*)scene;
  const MSM& m = sc.msms[msm];
  const Mech* el = nullptr;
  for (const Mech& q : m.mech)
    if (q.kind == MONSEL_MECH_NISTMOTT_ELASTIC || q.kind == MONSEL_MECH_BROWNING_ELASTIC ||
        q.kind == MONSEL_MECH_RUTHERFORD_ELASTIC)
      el = &q;
  if (!el || element < 0 || element >= (int)el->rse.size()) return MONSEL_E_INVALID;
  for (int i = 0; i < n; ++i) out[i] = el->rse[element].total(e_ev[i] * TOSI_EV) * 1e18;
  return 0;
}

int oracle_probe_csd(void* scene, int msm, int n, const double* len_nm, const double* e_ev, double* out) {
  const Scene& sc = *(const Scene*)scene;
  const MSM& m = sc.msms[msm];
  for (int i = 0; i < n; ++i) out[i] = m.csd.compute(len_nm[i] * 1e-9, e_ev[i] * TOSI_EV) * FROMSI_EV;
  return 0;
}

int oracle_probe_find_end_of_step(void* scene, int region, int n, const double* p0, const double* p1, double* out_u,
                                  int* out_next, double* out_normal) {
  const Scene& sc = *(const Scene*)scene;
  Ctx cx(sc);
  for (int i = 0; i < n; ++i) {
    double a[3], b[3];
    for (int k = 0; k < 3; ++k) { a[k] = p0[3 * i + k] * 1e-9; b[k] = p1[3 * i + k] * 1e-9; }
    /* The winning shape's normal: recompute the per-shape results. */
    double t;
    int next = cx.geo.find_end_of_step(region, a, b, cx.shapeNormal, &t);
    out_u[i] = t;
    out_next[i] = next;
    double nrm[3] = {0, 0, 0};
    if (t <= 1.0) {
      const Region& rg = sc.regions[region];
      const NV* win = &cx.shapeNormal[rg.shape];
      double tt = win->u;
      for (int c : rg.children) {
        const NV& cand = cx.shapeNormal[sc.regions[c].shape];
        if (cand.u <= 1.0 && cand.u < tt) { tt = cand.u; win = &cand; }
      }
      for (int k = 0; k < 3; ++k) nrm[k] = win->n[k];
    }
    for (int k = 0; k < 3; ++k) out_normal[3 * i + k] = nrm[k];
  }
  return 0;
}

static void dir_to_angles(const double* d, double* theta, double* phi) {
  *theta = ::atan2(::sqrt(d[0] * d[0] + d[1] * d[1]), d[2]);
  *phi = ::atan2(d[1], d[0]);
}
static void angles_to_dir(double theta, double phi, double* d) {
  double st = ::sin(theta);
  d[0] = st * ::cos(phi);
  d[1] = st * ::sin(phi);
  d[2] = ::cos(theta);
}

/* One ScatterMechanism::scatter of mechanism `mech` of MSM `msm` with the
 * electron at energy e_ev after a step from energy prev_ev, direction dir,
 * and the uniforms u[6i..6i+5] consumed in order.  Per sample, out[11i..]:
 * direction (3), energy (eV), secondary flag, secondary energy (eV),
 * secondary direction (3), trapped flag, uniforms consumed. */
int oracle_probe_scatter(void* scene, int msm, int mech, int n, const double* prev_ev, const double* e_ev,
                         const double* dir, const double* u, double* out) {
  const Scene& sc = *(const Scene*)scene;
  if (msm < 0 || msm >= (int)sc.msms.size()) return MONSEL_E_INVALID;
  const MSM& m = sc.msms[msm];
  if (mech < 0 || mech >= (int)m.mech.size()) return MONSEL_E_INVALID;
  for (int i = 0; i < n; ++i) {
    Electron pe;
    double p0[3] = {0, 0, 0}, th, ph;
    dir_to_angles(dir + 3 * i, &th, &ph);
    pe.init(p0, th, ph, e_ev[i] * TOSI_EV);
    pe.prevE = prev_ev[i] * TOSI_EV;
    pe.rng.script = u + 6 * (size_t)i;
    Spawn sp = mech_scatter(sc, m.mech[mech], pe);
    double* o = out + 11 * (size_t)i;
    angles_to_dir(pe.theta, pe.phi, o);
    o[3] = pe.E * FROMSI_EV;
    o[4] = sp.has ? 1. : 0.;
    o[5] = sp.has ? sp.E * FROMSI_EV : 0.;
    double sd[3] = {0, 0, 0};
    if (sp.has) {
      Electron se;
      se.init(p0, sp.theta, sp.phi, sp.E);
      if (sp.rotate) se.updateDirection(sp.dTheta, sp.dPhi);
      angles_to_dir(se.theta, se.phi, sd);
    }
    for (int k = 0; k < 3; ++k) o[6 + k] = sd[k];
    o[9] = pe.complete ? 1. : 0.;
    o[10] = (double)pe.rng.draws;
  }
  return 0;
}

/* One barrier crossing from region `from` into region `to` with the
 * outward barrier normal nb[3i..] already resolved (ExpQMBarrierSM.cu:77-96),
 * direction dir, energy e_ev and the transmission uniform u[i].  Per sample,
 * out[6i..]: direction (3), energy (eV), region after, nudge sign (+1 along
 * nb, -1 against). */
int oracle_probe_barrier(void* scene, int from, int to, int n, const double* dir, const double* nb,
                         const double* e_ev, const double* u, double* out) {
  const Scene& sc = *(const Scene*)scene;
  if (from < 0 || from >= (int)sc.regions.size() || to < 0 || to >= (int)sc.regions.size()) return MONSEL_E_INVALID;
  const MSM& m = sc.msms[sc.regions[from].msm];
  if (m.isNull || m.barrierKind == MONSEL_BARRIER_NONE || from == to) return MONSEL_E_INVALID;
  for (int i = 0; i < n; ++i) {
    Electron pe;
    double p0[3] = {0, 0, 0}, th, ph;
    dir_to_angles(dir + 3 * i, &th, &ph);
    pe.init(p0, th, ph, e_ev[i] * TOSI_EV);
    pe.region = from;
    pe.rng.script = u + i;
    int sign = barrier_core(sc, m, pe, to, nb + 3 * (size_t)i);
    double* o = out + 6 * (size_t)i;
    angles_to_dir(pe.theta, pe.phi, o);
    o[3] = pe.E * FROMSI_EV;
    o[4] = (double)pe.region;
    o[5] = (double)sign;
  }
  return 0;
}

int oracle_probe_locate(void* scene, int n, const double* pos, int* out) {
  const Scene& sc = *(const Scene*)scene;
  Geometry g(sc);
  for (int i = 0; i < n; ++i) {
    double p[3] = {pos[3 * i] * 1e-9, pos[3 * i + 1] * 1e-9, pos[3 * i + 2] * 1e-9};
    out[i] = g.containing_sub_region(0, p);
  }
  return 0;
}

/* Table values are returned in the table's own units (no conversion). */
int oracle_probe_nu_table(void* scene, int table, int order, int n, const double* x, double* out) {
  const Scene& sc = *(const Scene*)scene;
  if (table < 0 || table >= (int)sc.nu.size()) return MONSEL_E_INVALID;
  const NUTable& t = sc.nu[table];
  for (int i = 0; i < n; ++i) out[i] = nu_interpolate(t, x + (size_t)i * t.dim, order, sc.d3Compat());
  return 0;
}

/* Raw helpers for known-answer tests. */
void oracle_philox(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
  Philox4 c;
  for (int i = 0; i < 4; ++i) c.v[i] = ctr[i];
  Philox4 r = philox4x32_10(c, key[0], key[1]);
  for (int i = 0; i < 4; ++i) out[i] = r.v[i];
}

void oracle_stream(uint64_t seed, uint64_t lineage, uint32_t primary, int n, double* out) {
  Stream s;
  s.init(seed, lineage, primary);
  for (int i = 0; i < n; ++i) out[i] = s.next();
}

uint64_t oracle_child_lineage(uint64_t parent, uint32_t k) { return child_lineage(parent, k); }

/* MSVC rand()/RAND_MAX sequence after srand(seed) (oracle_rng.h). */
void oracle_lcg(uint32_t seed, int n, double* out) {
  g_lcg = seed;
  for (int i = 0; i < n; ++i) out[i] = msvc_rand01();
}

/* Stream mode (oracle_rng.h): 0 Philox, 1 quantised Philox, 2 MSVC rand(). */
void oracle_set_rng15(int on) { g_rng15 = on; }

double oracle_u_d1(const double* f, int len, double x0, double xinc, int order, double x) {
  return u_d1(f, len, x0, xinc, order, x);
}

double oracle_u_d2(const double* f, int n0, int n1, const double* x0, const double* xinc, int order, const double* x) {
  return u_d2(f, n0, n1, x0, xinc, order, x);
}

int oracle_hist_bin(double val, double min, double max, int nbins, int compat) {
  return hist_bin(val, min, max, nbins, compat != 0);
}

/* Runs one trajectory and reports per-step details of the primary only (for
 * debugging trajectories against the GPU engine); returns steps written. */
}
