/*
 * monsel_host.cpp -- implementation of the reference-mirror host API.
 * Serialisation of the object graph into monsel_scene_desc and job dispatch
 * through the C-ABI (include/monsel_gpu.h).  No transport runs here.
 */
#include "monsel_host.hpp"

#include <dlfcn.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <cmath>
#include <fstream>
#include <functional>
#include <mutex>
#include <sstream>

namespace monsel {

/* ---------------------------------------------------------------- data -- */

static std::string g_dataDir;
static std::mutex g_dataMu;

static std::string libraryDirectory() {
  Dl_info info;
  if (dladdr((void*)&libraryDirectory, &info) && info.dli_fname) {
    std::string p(info.dli_fname);
    size_t k = p.find_last_of('/');
    if (k != std::string::npos) return p.substr(0, k);
  }
  return ".";
}

void setDataDirectory(const std::string& dir) {
  std::lock_guard<std::mutex> g(g_dataMu);
  g_dataDir = dir;
}

std::string getDataDirectory() {
  std::lock_guard<std::mutex> g(g_dataMu);
  if (!g_dataDir.empty()) return g_dataDir;
  const char* env = getenv("MONSEL_DATA_DIR");
  if (env && *env) return env;
  return libraryDirectory() + "/data";
}

static const std::vector<double>& atomicWeights() {
  static std::vector<double> w;
  static std::once_flag once;
  std::call_once(once, [] {
    std::ifstream f(getDataDirectory() + "/atomic_weights.txt");
    std::string line;
    while (std::getline(f, line)) {
      if (line.empty() || line[0] == '/') continue;
      /* Element.cu:606 keeps the weights as float */
      w.push_back((double)std::stof(line));
    }
  });
  return w;
}

const Element& Element::byAtomicNumber(int Z) {
  static std::vector<std::unique_ptr<Element>> all;
  static std::once_flag once;
  std::call_once(once, [] {
    for (int z = 0; z <= 112; ++z) all.emplace_back(new Element(z));
  });
  if (Z < 1 || Z > 112) throw Error(MONSEL_E_INVALID, "atomic number out of range: " + std::to_string(Z));
  return *all[Z];
}

double Element::getAtomicWeight() const {
  const auto& w = atomicWeights();
  if (Z_ < 1 || Z_ > (int)w.size())
    throw Error(MONSEL_E_INVALID, "no atomic weight for Z=" + std::to_string(Z_) + " in " + getDataDirectory());
  return w[Z_ - 1];
}

namespace Elements {
const Element& H = Element::byAtomicNumber(1);
const Element& C = Element::byAtomicNumber(6);
const Element& N = Element::byAtomicNumber(7);
const Element& O = Element::byAtomicNumber(8);
const Element& Si = Element::byAtomicNumber(14);
const Element& Cu = Element::byAtomicNumber(29);
}  // namespace Elements

MottData loadMottTable(int Z) {
  static std::map<int, MottData> cache;
  static std::mutex mu;
  std::lock_guard<std::mutex> g(mu);
  auto it = cache.find(Z);
  if (it != cache.end()) return it->second;
  const int NE = MONSEL_MOTT_NE, NQ = MONSEL_MOTT_NQ;
  std::string dir = getDataDirectory();
  /* compact file written by data/make_data.py */
  std::ifstream f(dir + "/nistmott.bin", std::ios::binary);
  if (f) {
    char magic[4];
    uint32_t ver = 0, count = 0;
    f.read(magic, 4);
    f.read((char*)&ver, 4);
    f.read((char*)&count, 4);
    if (f && memcmp(magic, "NMOT", 4) == 0 && ver == 1) {
      for (uint32_t i = 0; i < count; ++i) {
        int32_t z;
        f.read((char*)&z, 4);
        auto sp = std::make_shared<std::vector<double>>(NE);
        auto x1 = std::make_shared<std::vector<double>>((size_t)NE * NQ);
        f.read((char*)sp->data(), sizeof(double) * NE);
        f.read((char*)x1->data(), sizeof(double) * NE * NQ);
        if (!f) break;
        cache[z] = MottData{sp, x1};
      }
    }
  }
  it = cache.find(Z);
  if (it != cache.end()) return it->second;
  /* fall back to an original SRD-64 file (NISTMottScatteringAngle.cu:44-66) */
  const char* xs = getenv("MONSEL_NISTXSEC_DIR");
  std::string d64 = std::string(xs ? xs : dir) + "/E" + (Z < 10 ? "0" : "") + std::to_string(Z) + ".D64";
  std::ifstream t(d64);
  if (!t) throw Error(MONSEL_E_INVALID, "no NIST Mott table for Z=" + std::to_string(Z));
  double v;
  t >> v;
  auto sp = std::make_shared<std::vector<double>>(NE);
  auto x1 = std::make_shared<std::vector<double>>((size_t)NE * NQ);
  for (int j = 0; j < NE; ++j) {
    t >> (*sp)[j];
    for (int k = 0; k < NQ; ++k) t >> (*x1)[(size_t)j * NQ + k];
  }
  if (!t) throw Error(MONSEL_E_INVALID, "truncated Mott table " + d64);
  cache[Z] = MottData{sp, x1};
  return cache[Z];
}

NUTableData readNUTable(const std::string& path) {
  std::ifstream f(path);
  if (!f) throw Error(MONSEL_E_INVALID, "cannot open table " + path);
  NUTableData t;
  double a;
  if (!(f >> a)) throw Error(MONSEL_E_INVALID, "empty table " + path);
  t.dim = (int)a;
  if (t.dim < 1 || t.dim > 4) throw Error(MONSEL_E_INVALID, "table dimension must be 1..4: " + path);
  size_t total = 1;
  for (int i = 0; i < t.dim; ++i) {
    f >> a;
    t.n[i] = (int)a;
    if (t.n[i] < 2) throw Error(MONSEL_E_INVALID, "table axis too short: " + path);
    t.x[i].resize(t.n[i]);
    for (int j = 0; j < t.n[i]; ++j) f >> t.x[i][j];
    total *= (size_t)t.n[i];
  }
  t.values.resize(total);
  for (size_t k = 0; k < total; ++k) f >> t.values[k];
  if (!f) throw Error(MONSEL_E_INVALID, "truncated table " + path);
  return t;
}

/* --------------------------------------------------------- composition -- */

/* Composition.cu:320-340 (mole fractions -> weight fractions, :355-373) */
void Composition::defineByMoleFraction(const Element* const elms[], int n, const double moleFracs[], int nf) {
  if (n != nf) throw Error(MONSEL_E_INVALID, "defineByMoleFraction: length mismatch");
  elements_.clear();
  wf_.clear();
  double mfSum = 0;
  for (int k = 0; k < nf; ++k) mfSum += moleFracs[k];
  std::vector<double> ap(n);
  double apSum = 0;
  for (int i = 0; i < n; ++i) {
    ap[i] = moleFracs[i] / mfSum;
    apSum += ap[i];
  }
  double totalWgt = 0;
  for (int i = 0; i < n; ++i) totalWgt += elms[i]->getAtomicWeight() * (ap[i] / apSum);
  for (int i = 0; i < n; ++i) {
    elements_.push_back(elms[i]->getAtomicNumber());
    wf_.push_back(elms[i]->getAtomicWeight() * ((ap[i] / apSum) / totalWgt));
  }
}

void Composition::defineByWeightFraction(const Element* const elms[], int n, const double wgtFracs[], int nf) {
  if (n != nf) throw Error(MONSEL_E_INVALID, "defineByWeightFraction: length mismatch");
  elements_.clear();
  wf_.clear();
  for (int i = 0; i < n; ++i) {
    elements_.push_back(elms[i]->getAtomicNumber());
    wf_.push_back(wgtFracs[i]);
  }
}

/* Composition.cu:204-212 */
double Composition::weightFraction(const Element& el, bool normalized) const {
  double norm = 0;
  for (double w : wf_) norm += w;
  for (size_t i = 0; i < elements_.size(); ++i)
    if (elements_[i] == el.getAtomicNumber()) return normalized ? (norm > 0 ? wf_[i] / norm : 0.) : wf_[i];
  return 0.;
}

SEmaterial::SEmaterial(const Element* const elms[], int n, const double weightFracs[], int nw, double density,
                       const char* name) {
  defineByWeightFraction(elms, n, weightFracs, nw);
  density_ = density;
  if (name) name_ = name;
}

void SEmaterial::setWorkfunction(double w) {
  workfunction_ = w;
  if (energyCBbottom_ > -w) energyCBbottom_ = -w;
}

/* Rows of whitespace-separated numbers, one per element (line Z-1). */
static std::vector<std::vector<double>> readRows(const std::string& file) {
  std::vector<std::vector<double>> rows;
  std::ifstream f(getDataDirectory() + "/" + file);
  if (!f) throw Error(MONSEL_E_INVALID, "cannot read " + getDataDirectory() + "/" + file);
  std::string line;
  while (std::getline(f, line)) {
    std::istringstream is(line);
    std::vector<double> r;
    double v;
    while (is >> v) r.push_back(v);
    rows.push_back(r);
  }
  return rows;
}

/* ChantlerEdgeEnergy (EdgeEnergy.cu:153-205): per element the non-zero edge
 * energies (eV) of shells K .. O5 then P1 .. P3. */
static const std::vector<std::vector<double>>& chantlerEdges() {
  static std::vector<std::vector<double>> rows;
  static std::once_flag once;
  std::call_once(once, [] { rows = readRows("edge_energies.txt"); });
  return rows;
}

/* AtomicShell::getGroundStateOccupancy (AtomicShell.cu:386-425). */
static const std::vector<std::vector<double>>& shellOccupancy() {
  static std::vector<std::vector<double>> rows;
  static std::once_flag once;
  std::call_once(once, [] { rows = readRows("electron_config.txt"); });
  return rows;
}

/* EdgeEnergy::SuperSet = Chantler 2005 for Z <= 92 (its compute returns 0
 * rather than throwing, so the NIST and DTSA fallbacks are never reached),
 * in J.  The shell numbering is AtomicShell's: K = 0, ..., O5 = 20,
 * O6 .. O9 = 21 .. 24 (no Chantler column), P1 = 25. */
static double chantlerEdgeEnergy(int Z, int shell) {
  const auto& E = chantlerEdges();
  int i = shell <= 20 ? shell : (shell >= 25 ? shell - 4 : -1);
  if (Z < 1 || Z > (int)E.size() || i < 0 || i >= (int)E[Z - 1].size()) return 0.0;
  return ToSI::eV(E[Z - 1][i]);
}

void SEmaterial::setEstimatedCoreEnergy() { setEstimatedCoreEnergy(ToSI::eV(20.)); }

void SEmaterial::setEstimatedCoreEnergy(double cutoff) {
  core_.clear();
  const auto& occ = shellOccupancy();
  for (int Z : elements_) {
    if (Z < 1 || Z > (int)occ.size()) throw Error(MONSEL_E_INVALID, "no shell occupancy for Z=" + std::to_string(Z));
    const auto& o = occ[Z - 1];
    /* SEmaterial.cu:357-364: occupied shells, innermost first, until one's
     * edge is at or below the cutoff.  The reference would go on past the
     * last occupied shell (an unoccupied shell gives NaN, which never meets
     * the cutoff); this stops there. */
    for (int shell = 0; shell < (int)o.size() && o[shell] > 0; ++shell) {
      const double e = chantlerEdgeEnergy(Z, shell);
      if (e <= cutoff) break;
      addCoreEnergy(e);
    }
  }
}

void SEmaterial::setCoreEnergy(const double coreEnergy[], int len) {
  core_.clear();
  for (int i = 0; i < len; ++i) {
    if (coreEnergy[i] < 0) throw Error(MONSEL_E_INVALID, "Core energies must be positive.");
    core_.insert(coreEnergy[i]);
  }
}

/* ---------------------------------------------------------- mechanisms -- */

JoyLuoNieminenCSD::JoyLuoNieminenCSD(const SEmaterial& mat, double breakE) : breakE_(breakE) {
  /* JoyLuoNieminenCSD.cu:8-16: the reference leaves breakE unset otherwise */
  if (!(breakE > mat.getWorkfunction() + ToSI::eV(1.)))
    throw Error(MONSEL_E_INVALID, "JoyLuoNieminenCSD: Supplied breakpoint energy is too small.");
}

static const RandomizedScatterFactory kNIST1(MONSEL_MECH_NISTMOTT_ELASTIC, 1);
static const RandomizedScatterFactory kNIST2(MONSEL_MECH_NISTMOTT_ELASTIC, 2);
static const RandomizedScatterFactory kNIST3(MONSEL_MECH_NISTMOTT_ELASTIC, 3);
static const RandomizedScatterFactory kBrowning(MONSEL_MECH_BROWNING_ELASTIC, 1);
static const RandomizedScatterFactory kRutherford(MONSEL_MECH_RUTHERFORD_ELASTIC, 1);
namespace NISTMottRS {
const RandomizedScatterFactory& Factory = kNIST1;
const RandomizedScatterFactory& Factory100 = kNIST2;
const RandomizedScatterFactory& Factory100Lin = kNIST3;
}  // namespace NISTMottRS
namespace BrowningEmpiricalCrossSection {
const RandomizedScatterFactory& Factory = kBrowning;
}
namespace ScreenedRutherfordScatteringAngle {
const RandomizedScatterFactory& Factory = kRutherford;
}

static monsel_mech_desc blankMech(int kind) {
  monsel_mech_desc d;
  memset(&d, 0, sizeof(d));
  d.kind = kind;
  for (int i = 0; i < 4; ++i) d.table[i] = -1;
  return d;
}

monsel_mech_desc SelectableElasticSM::describe() const {
  monsel_mech_desc d = blankMech(rsf_.kind());
  d.method = rsf_.method();
  return d;
}

FittedInelSM::FittedInelSM(const SEmaterial& mat, double energySEgen, const SlowingDownAlg& sdAlg)
    : energySEgen_(energySEgen), breakE_(sdAlg.kind() == MONSEL_CSD_JOY_LUO_NIEMINEN ? sdAlg.breakE() : 0.) {
  (void)mat;
}

monsel_mech_desc FittedInelSM::describe() const {
  monsel_mech_desc d = blankMech(MONSEL_MECH_FITTED_INELASTIC);
  d.p[0] = energySEgen_;
  d.p[1] = breakE_;
  d.p[2] = breakE_ > 0 ? MONSEL_CSD_JOY_LUO_NIEMINEN : MONSEL_CSD_ZERO;
  return d;
}

GanachaudMokraniPolaronTrapSM::GanachaudMokraniPolaronTrapSM(double prefactor, double extinction)
    : prefactor_(prefactor), extinction_(extinction) {
  if (prefactor < 0. || extinction < 0.)
    throw Error(MONSEL_E_INVALID, "GanachaudMokraniPolaronTrapSM: nonpositive prefactor or extinction");
}

monsel_mech_desc GanachaudMokraniPolaronTrapSM::describe() const {
  monsel_mech_desc d = blankMech(MONSEL_MECH_POLARON_TRAP);
  d.p[0] = prefactor_;
  d.p[1] = extinction_;
  return d;
}

GanachaudMokraniPhononInelasticSM::GanachaudMokraniPhononInelasticSM(double ratemultiplier, double phononE,
                                                                     double temperature, double eps0,
                                                                     double epsInfinity)
    : p_{ratemultiplier, phononE, temperature, eps0, epsInfinity} {
  if (ratemultiplier <= 0. || phononE <= 0. || (eps0 - epsInfinity) / eps0 / epsInfinity <= 0.)
    throw Error(MONSEL_E_INVALID, "GanachaudMokraniPhononInelasticSM: invalid parameters");
}

monsel_mech_desc GanachaudMokraniPhononInelasticSM::describe() const {
  monsel_mech_desc d = blankMech(MONSEL_MECH_PHONON);
  for (int i = 0; i < 5; ++i) d.p[i] = p_[i];
  return d;
}

TabulatedInelasticSM::TabulatedInelasticSM(const SEmaterial& mat, int methodSE, const std::string tables[])
    : TabulatedInelasticSM(mat, methodSE, tables, 0.) {}

TabulatedInelasticSM::TabulatedInelasticSM(const SEmaterial& mat, int methodSE, const std::string tables[],
                                           double energyOffset)
    : methodSE_((methodSE != 2) && (methodSE != 3) ? 1 : methodSE), energyOffset_(energyOffset) {
  (void)mat;
  int n = methodSE_ >= 2 ? 4 : 3;
  for (int i = 0; i < n; ++i) tables_.push_back(tables[i]);
}

monsel_mech_desc TabulatedInelasticSM::describe() const {
  monsel_mech_desc d = blankMech(MONSEL_MECH_TABULATED_INELASTIC);
  d.method = methodSE_;
  d.p[0] = energyOffset_;
  d.p[1] = minEgenSE_;
  d.p[2] = rateMult_;
  d.p[3] = e0FromDispersion_ ? 1. : 0.;
  return d;
}

/* ------------------------------------------------------------------ MSM -- */

MONSEL_MaterialScatterModel::MONSEL_MaterialScatterModel(const SEmaterial* mat, const BarrierScatterMechanism* bsm)
    : mat_(mat), barrier_(bsm) {
  /* MONSEL_MaterialScatterModel.cu:25-36 */
  if (mat->getElementCount() == 0) minE_ = -INFINITY;
  else {
    minE_ = -mat->getEnergyCBbottom();
    if (minE_ < 0.) minE_ = 0.;
  }
}

bool MONSEL_MaterialScatterModel::addScatterMechanism(const ScatterMechanism* mech) {
  if (std::find(mech_.begin(), mech_.end(), mech) != mech_.end()) return false;
  if ((int)mech_.size() >= MONSEL_MAX_MECH) throw Error(MONSEL_E_UNSUPPORTED, "too many scatter mechanisms");
  mech_.push_back(mech);
  return true;
}

bool MONSEL_MaterialScatterModel::removeScatterMechanism(const ScatterMechanism* mech) {
  auto it = std::find(mech_.begin(), mech_.end(), mech);
  if (it == mech_.end()) return false;
  mech_.erase(it);
  return true;
}

monsel_msm_desc MONSEL_MaterialScatterModel::describe(int materialIndex) const {
  monsel_msm_desc d;
  memset(&d, 0, sizeof(d));
  d.material = materialIndex;
  d.is_null = 0;
  d.n_mech = (int)mech_.size();
  for (size_t i = 0; i < mech_.size(); ++i) d.mech[i] = mech_[i]->describe();
  d.csd_kind = csd_ ? csd_->kind() : MONSEL_CSD_ZERO;
  d.csd_break_e = csd_ ? csd_->breakE() : 0.;
  d.barrier_kind = barrier_ ? barrier_->kind() : MONSEL_BARRIER_NONE;
  d.barrier_lambda = barrier_ ? barrier_->lambda() : 0.;
  d.min_e_for_tracking = minE_;
  return d;
}

NullMaterialScatterModel::NullMaterialScatterModel() : minE_(ToSI::eV(0.1)) {}

monsel_msm_desc NullMaterialScatterModel::describe(int materialIndex) const {
  monsel_msm_desc d;
  memset(&d, 0, sizeof(d));
  d.material = materialIndex;
  d.is_null = 1;
  d.barrier_kind = MONSEL_BARRIER_NONE;
  d.min_e_for_tracking = minE_;
  return d;
}

/* --------------------------------------------------------------- shapes -- */

Plane::Plane(const double n[], int nl, const double p[], int pl) {
  if (nl != 3 || pl != 3) throw Error(MONSEL_E_INVALID, "Plane: vectors must have length 3");
  normal = {n[0], n[1], n[2]};
  point = {p[0], p[1], p[2]};
}

void NormalMultiPlaneShape::addPlane(const double normal[], const double point[]) {
  planes_.push_back(Plane(normal, 3, point, 3));
}

void NormalMultiPlaneShape::translate(const double d[]) {
  for (auto& p : planes_)
    for (int k = 0; k < 3; ++k) p.point[k] += d[k];
}

NormalCylindricalShape::NormalCylindricalShape(const double e0[], const double e1[], double r)
    : end0{e0[0], e0[1], e0[2]}, end1{e1[0], e1[1], e1[2]}, radius(r) {
  double l2 = 0;
  for (int k = 0; k < 3; ++k) l2 += (e1[k] - e0[k]) * (e1[k] - e0[k]);
  if (r * r < 1.0e-30 || l2 < 1.0e-30)
    throw Error(MONSEL_E_INVALID, "NormalCylindricalShape: unrealistically small radius or length");
}

void NormalCylindricalShape::translate(const double d[]) {
  for (int k = 0; k < 3; ++k) {
    end0[k] += d[k];
    end1[k] += d[k];
  }
}

Sphere::Sphere(const double c[], double r) : center{c[0], c[1], c[2]}, radius(r) {}

namespace NShapes {

static std::vector<std::unique_ptr<Shape>>& pool() {
  static std::vector<std::unique_ptr<Shape>> p;
  return p;
}
static std::mutex g_poolMu;

template <class T, class... A>
static T* make(A&&... a) {
  T* s = new T(std::forward<A>(a)...);
  pool().emplace_back(s);
  return s;
}

/* NShapes.cu:55-149 */
NormalShape* createLine(double topz, double width, double length, double thetal, double thetar, double radl,
                        double radr, bool referenceCompat) {
  std::lock_guard<std::mutex> g(g_poolMu);
  if (radr < 0.) radr = 0.;
  if (radl < 0.) radl = 0.;
  auto* enclosure = make<NormalMultiPlaneShape>();
  double signz = !topz ? 0 : (topz > 0 ? 1 : -1);
  if (signz == 0.) signz = 1.;
  double n0[] = {0., 0., signz}, p0[] = {0., 0., topz};
  enclosure->addPlane(n0, p0);
  double n1[] = {0., 0., -signz}, p1[] = {0., 0., 0.};
  enclosure->addPlane(n1, p1);
  double n2[] = {0., 1., 0.}, p2[] = {0., length / 2., 0.};
  enclosure->addPlane(n2, p2);
  double n3[] = {0., -1., 0.}, p3[] = {0., -length / 2., 0.};
  enclosure->addPlane(n3, p3);

  auto* rightNMPS = make<NormalMultiPlaneShape>();
  NormalShape* rightSide = nullptr;
  double costhetar = ::cos(thetar), sinthetar = ::sin(thetar);
  double n4[] = {costhetar, 0., signz * sinthetar}, p4[] = {width / 2, 0., 0.};
  rightNMPS->addPlane(n4, p4);
  double root2 = ::sqrt(2.);
  double absz = signz * topz;
  if (radr > 0) {
    double rad = ::sqrt(1 - sinthetar);
    double nr[] = {rad / root2, 0., (signz * costhetar) / root2 / rad};
    double pr[] = {((width / 2.) - (radr / costhetar)) + (((radr - absz) * sinthetar) / costhetar), 0., topz};
    rightNMPS->addPlane(nr, pr);
    double xc = ((width / 2.) - (radr / ::cos(thetar))) + ((radr - absz) * ::tan(thetar));
    double zc = topz - (signz * radr);
    double e0[] = {xc, -length / 2., zc}, e1[] = {xc, length / 2., zc};
    auto* rcyl = make<NormalCylindricalShape>(e0, e1, radr);
    rightSide = make<NormalUnionShape>(*rightNMPS, *rcyl);
  } else {
    rightSide = rightNMPS;
  }

  auto* leftNMPS = make<NormalMultiPlaneShape>();
  NormalShape* leftSide = nullptr;
  double costhetal = ::cos(thetal), sinthetal = ::sin(thetal);
  double n6[] = {-costhetal, 0., signz * sinthetal}, p6[] = {-width / 2, 0., 0.};
  leftNMPS->addPlane(n6, p6);
  if (radl > 0.) {
    double rad = ::sqrt(1 - sinthetal);
    double n8[] = {-rad / root2, 0., (signz * costhetal) / root2 / rad};
    double p8[] = {((-width / 2.) + (radl / costhetal)) - (((radl - absz) * sinthetal) / costhetal), 0., topz};
    leftNMPS->addPlane(n8, p8);
    double xc = ((width / 2.) - (radl / ::cos(thetal))) + ((radl - absz) * ::tan(thetal));
    double zc = topz - (signz * radl);
    double e0[] = {-xc, -length / 2., zc}, e1[] = {-xc, length / 2., zc};
    auto* lcyl = make<NormalCylindricalShape>(e0, e1, radl);
    /* NShapes.cu:140 builds this union from rightNMPS */
    leftSide = make<NormalUnionShape>(referenceCompat ? *rightNMPS : *leftNMPS, *lcyl);
  } else {
    leftSide = leftNMPS;
  }
  auto* nts = make<NormalIntersectionShape>(*leftSide, *rightSide);
  return make<NormalIntersectionShape>(*nts, *enclosure);
}

NormalShape* createStepEdge(double x0, double h) {
  std::lock_guard<std::mutex> g(g_poolMu);
  auto* sub = make<NormalMultiPlaneShape>();
  double nz[] = {0., 0., -1.}, p0[] = {0., 0., 0.};
  sub->addPlane(nz, p0);
  auto* step = make<NormalMultiPlaneShape>();
  double nx[] = {1., 0., 0.}, px[] = {x0, 0., 0.};
  step->addPlane(nx, px);
  double nt[] = {0., 0., -1.}, pt[] = {0., 0., -h};
  step->addPlane(nt, pt);
  return make<NormalUnionShape>(*sub, *step);
}

void translate(NormalShape* s, const double d[]) {
  if (auto* m = dynamic_cast<NormalMultiPlaneShape*>(s)) m->translate(d);
  else if (auto* c = dynamic_cast<NormalCylindricalShape*>(s)) c->translate(d);
  else if (auto* i = dynamic_cast<NormalIntersectionShape*>(s)) {
    translate(i->a, d);
    translate(i->b, d);
  } else if (auto* u = dynamic_cast<NormalUnionShape*>(s)) {
    translate(u->a, d);
    translate(u->b, d);
  }
}

}  // namespace NShapes

/* -------------------------------------------------------------- regions -- */

Region::Region(Region* parent, IMaterialScatterModel* msm, Shape* shape) : parent_(parent), msm_(msm), shape_(shape) {
  if (parent_) parent_->children_.push_back(this);
}

void Region::updateMaterial(const IMaterialScatterModel& oldMat, IMaterialScatterModel& newMat) {
  if (msm_ == &oldMat) msm_ = &newMat;
  for (Region* r : children_) r->updateMaterial(oldMat, newMat);
}

void Region::removeSubRegion(Region& r) {
  auto it = std::find(children_.begin(), children_.end(), &r);
  if (it != children_.end()) children_.erase(it);
}

/* ----------------------------------------------------------------- beam -- */

static long g_last_electron = 0;

Electron::Electron(const double initialPos[], double theta, double phi, double kE)
    : pos_{initialPos[0], initialPos[1], initialPos[2]}, theta_(theta), phi_(phi), energy_(kE) {
  static std::mutex mu;
  std::lock_guard<std::mutex> lk(mu);
  ident_ = ++g_last_electron;
}

Electron* ElectronGun::createElectron() const {
  const std::array<double, 3> c = getCenter();
  return new Electron(c.data(), getBeamEnergy());
}

/* uniform in (0, 1] from a splitmix64 sequence */
static double unit_open0(uint64_t x) {
  x += 0x9E3779B97F4A7C15ull;
  x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
  x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
  x ^= x >> 31;
  return ((double)(x >> 11) + 1.0) * (1.0 / 9007199254740992.0);
}

Electron* GaussianBeam::createElectron() const {
  double p[3] = {center_[0], center_[1], center_[2]};
  const double r = std::sqrt(-2. * std::log(unit_open0(draws_++))) * width_;
  const double th = 2.0 * M_PI * unit_open0(draws_++);
  p[0] += r * std::cos(th);
  p[1] += r * std::sin(th);
  return new Electron(p, energy_);
}

GaussianBeam::GaussianBeam(double width) : width_(width), center_{0., 0., -0.99 * MonteCarloSS::ChamberRadius} {}

GaussianBeam::GaussianBeam(double width, double energy, const double c[])
    : width_(width), energy_(energy), center_{c[0], c[1], c[2]} {}

/* ----------------------------------------------------------- histograms -- */

Histogram::Histogram(double min, double max, int nBins) : min_(min), max_(max), nbins_(nBins) {
  if (max_ < min_) std::swap(min_, max_);
  if (nBins <= 0 || min_ == max_) throw Error(MONSEL_E_INVALID, "Histogram: invalid range or bin count");
  counts_.assign(nBins + 2, 0);
}

int Histogram::totalCounts() const {
  uint64_t s = 0;
  for (auto c : counts_) s += c;
  return (int)s;
}

double Histogram::minValue(int bin) const { return bin > -1 ? min_ + bin * ((max_ - min_) / nbins_) : -INFINITY; }
double Histogram::maxValue(int bin) const {
  return bin + 1 <= nbins_ ? min_ + (bin + 1) * ((max_ - min_) / nbins_) : INFINITY;
}
void Histogram::clear() { std::fill(counts_.begin(), counts_.end(), 0); }

BackscatterStats::BackscatterStats(const MonteCarloSS& mcss, int nEnergyBins)
    : nbins_(nEnergyBins),
      beamEnergyEV_(FromSI::eV(mcss.getBeamEnergy())),
      elev_(0.0, PhysicalConstants::PI, 180),
      azim_(0.0, 2.0 * PhysicalConstants::PI, 360),
      fwd_(0.0, beamEnergyEV_ > 0 ? beamEnergyEV_ : 1.0, nEnergyBins),
      back_(0.0, beamEnergyEV_ > 0 ? beamEnergyEV_ : 1.0, nEnergyBins) {}

double BackscatterStats::backscatterFraction() const { return (double)back_.totalCounts() / (double)eventCount_; }
double BackscatterStats::forwardscatterFraction() const { return (double)fwd_.totalCounts() / (double)eventCount_; }

void BackscatterStats::jobCompleted(const MonteCarloSS& mc, int n, const monsel_tallies& t) {
  beamEnergyEV_ = FromSI::eV(mc.getBeamEnergy());
  double e = beamEnergyEV_ > 0 ? beamEnergyEV_ : 1.0;
  fwd_ = Histogram(0.0, e, nbins_);
  back_ = Histogram(0.0, e, nbins_);
  std::copy(t.back_energy_hist, t.back_energy_hist + nbins_ + 2, back_.slots().begin());
  std::copy(t.fwd_energy_hist, t.fwd_energy_hist + nbins_ + 2, fwd_.slots().begin());
  std::copy(t.elevation_hist, t.elevation_hist + 182, elev_.slots().begin());
  std::copy(t.azimuth_hist, t.azimuth_hist + 362, azim_.slots().begin());
  eventCount_ = n;
}

void BackscatterStats::detectedLog(const std::vector<monsel_detected>& log, int64_t nDetected) {
  log_.clear();
  log_.reserve(log.size());
  for (const monsel_detected& d : log)
    log_.push_back(Datum{d.electron, d.primary, (long)d.step, d.energy_ev,
                         {d.position[0], d.position[1], d.position[2]}, d.theta, d.phi});
  logTruncated_ = nDetected > (int64_t)log.size();
}

/* ------------------------------------------------------- serialisation -- */

static void fnv(uint64_t& h, const void* p, size_t n) {
  const unsigned char* c = (const unsigned char*)p;
  for (size_t i = 0; i < n; ++i) {
    h ^= c[i];
    h *= 1099511628211ull;
  }
}

void CompiledScene::finalize(uint32_t compat, double chamberRadius) {
  memset(&desc, 0, sizeof(desc));
  desc.abi_version = MONSEL_ABI_VERSION;
  desc.compat_flags = compat;
  desc.chamber_radius = chamberRadius;
  desc.n_materials = (int)materials.size();
  desc.materials = materials.data();
  desc.n_msm = (int)msms.size();
  desc.msms = msms.data();
  desc.n_shapes = (int)shapes.size();
  desc.shapes = shapes.data();
  desc.n_planes = (int)planes.size();
  desc.planes = planes.data();
  desc.n_regions = (int)regions.size();
  desc.regions = regions.data();
  desc.n_mott = (int)mott.size();
  desc.mott = mott.data();
  desc.n_nu_tables = (int)nu.size();
  desc.nu_tables = nu.data();
  uint64_t h = 1469598103934665603ull;
  fnv(h, &compat, sizeof(compat));
  fnv(h, &chamberRadius, sizeof(chamberRadius));
  fnv(h, materials.data(), materials.size() * sizeof(materials[0]));
  fnv(h, msms.data(), msms.size() * sizeof(msms[0]));
  fnv(h, shapes.data(), shapes.size() * sizeof(shapes[0]));
  fnv(h, planes.data(), planes.size() * sizeof(planes[0]));
  fnv(h, regions.data(), regions.size() * sizeof(regions[0]));
  for (auto& m : mott) fnv(h, &m.Z, sizeof(m.Z));
  for (auto& t : nu) {
    fnv(h, &t.dim, sizeof(t.dim));
    fnv(h, t.n, sizeof(t.n));
    size_t tot = 1;
    for (int k = 0; k < t.dim; ++k) {
      fnv(h, t.x[k], sizeof(double) * t.n[k]);
      tot *= (size_t)t.n[k];
    }
    fnv(h, t.values, sizeof(double) * tot);
  }
  hash = h;
}

namespace {
struct Builder {
  CompiledScene& cs;
  std::map<const SEmaterial*, int> mat;
  std::map<const IMaterialScatterModel*, int> msm;
  std::map<const Shape*, int> shape;
  std::map<std::string, int> tables;
  std::set<int> mottZ;

  explicit Builder(CompiledScene& c) : cs(c) {}

  int addMaterial(const SEmaterial& m) {
    auto it = mat.find(&m);
    if (it != mat.end()) return it->second;
    monsel_material_desc d;
    memset(&d, 0, sizeof(d));
    const auto& zs = m.elementZ();
    if ((int)zs.size() > MONSEL_MAX_ELEMENTS) throw Error(MONSEL_E_UNSUPPORTED, "too many elements in material");
    d.n_elements = (int)zs.size();
    for (size_t i = 0; i < zs.size(); ++i) {
      const Element& el = Element::byAtomicNumber(zs[i]);
      d.Z[i] = zs[i];
      d.weight_fraction[i] = m.weightFraction(el, true);
      d.atomic_weight[i] = el.getAtomicWeight();
    }
    d.density = m.getDensity();
    d.workfunction = m.getWorkfunction();
    d.energy_cb_bottom = m.getEnergyCBbottom();
    d.bandgap = m.getBandgap();
    const auto& core = m.getCoreEnergyArray();
    if ((int)core.size() > MONSEL_MAX_CORE) throw Error(MONSEL_E_UNSUPPORTED, "too many core levels");
    d.n_core = (int)core.size();
    int k = 0;
    for (double c : core) d.core_energy[k++] = c;
    int idx = (int)cs.materials.size();
    cs.materials.push_back(d);
    mat[&m] = idx;
    return idx;
  }

  int addTable(const std::string& path) {
    auto it = tables.find(path);
    if (it != tables.end()) return it->second;
    NUTableData t = readNUTable(path);
    monsel_nu_table d;
    memset(&d, 0, sizeof(d));
    d.dim = t.dim;
    for (int k = 0; k < t.dim; ++k) {
      d.n[k] = t.n[k];
      auto v = std::make_shared<const std::vector<double>>(t.x[k]);
      cs.keep.push_back(v);
      d.x[k] = v->data();
    }
    auto vals = std::make_shared<const std::vector<double>>(t.values);
    cs.keep.push_back(vals);
    d.values = vals->data();
    int idx = (int)cs.nu.size();
    cs.nu.push_back(d);
    tables[path] = idx;
    return idx;
  }

  int addMSM(const IMaterialScatterModel* m) {
    auto it = msm.find(m);
    if (it != msm.end()) return it->second;
    int mi = addMaterial(m->getMaterial());
    monsel_msm_desc d = m->describe(mi);
    auto mechs = m->mechanisms();
    for (size_t k = 0; k < mechs.size(); ++k) {
      auto files = mechs[k]->tableFiles();
      for (size_t q = 0; q < files.size() && q < 4; ++q) d.mech[k].table[q] = addTable(files[q]);
      if (d.mech[k].kind == MONSEL_MECH_NISTMOTT_ELASTIC)
        for (int z : m->getMaterial().elementZ()) mottZ.insert(z);
    }
    int idx = (int)cs.msms.size();
    cs.msms.push_back(d);
    msm[m] = idx;
    return idx;
  }

  int addShape(const Shape* s) {
    auto it = shape.find(s);
    if (it != shape.end()) return it->second;
    monsel_shape_desc d;
    memset(&d, 0, sizeof(d));
    d.kind = s->kind();
    d.a = d.b = -1;
    if (auto* mp = dynamic_cast<const NormalMultiPlaneShape*>(s)) {
      d.first_plane = (int)cs.planes.size();
      d.n_planes = mp->getNumPlanes();
      for (const Plane& p : mp->planes()) {
        monsel_plane_desc pd;
        for (int k = 0; k < 3; ++k) {
          pd.normal[k] = p.normal[k];
          pd.point[k] = p.point[k];
        }
        cs.planes.push_back(pd);
      }
    } else if (auto* c = dynamic_cast<const NormalCylindricalShape*>(s)) {
      for (int k = 0; k < 3; ++k) {
        d.end0[k] = c->end0[k];
        d.end1[k] = c->end1[k];
      }
      d.radius = c->radius;
    } else if (auto* i = dynamic_cast<const NormalIntersectionShape*>(s)) {
      d.a = addShape(i->a);
      d.b = addShape(i->b);
    } else if (auto* u = dynamic_cast<const NormalUnionShape*>(s)) {
      d.a = addShape(u->a);
      d.b = addShape(u->b);
    } else if (auto* sp = dynamic_cast<const Sphere*>(s)) {
      for (int k = 0; k < 3; ++k) d.center[k] = sp->center[k];
      d.radius = sp->radius;
    } else {
      throw Error(MONSEL_E_UNSUPPORTED, "unsupported shape class");
    }
    int idx = (int)cs.shapes.size();
    cs.shapes.push_back(d);
    shape[s] = idx;
    return idx;
  }

  void addRegion(const Region* r, int parent) {
    int idx = (int)cs.regions.size();
    cs.regions.push_back(monsel_region_desc{parent, 0, 0, 0});
    cs.regionIndex[r] = idx;
    int s = addShape(r->getShape());
    int m = addMSM(r->getScatterModel());
    cs.regions[idx].shape = s;
    cs.regions[idx].msm = m;
    for (const Region* c : r->getSubRegions()) addRegion(c, idx);
  }
};
}  // namespace

/* ---------------------------------------------------------------- device -- */

static void check(int st, const char* what) {
  if (st != MONSEL_OK) throw Error(st, std::string(what) + ": " + monsel_last_error());
}

Device& Device::get(int ordinal) {
  static std::map<int, std::unique_ptr<Device>> devs;
  static std::mutex mu;
  std::lock_guard<std::mutex> g(mu);
  auto it = devs.find(ordinal);
  if (it == devs.end()) it = devs.emplace(ordinal, std::unique_ptr<Device>(new Device(ordinal))).first;
  return *it->second;
}

Device::Device(int ordinal) { check(monsel_create(ordinal, nullptr, &ctx_), "monsel_create"); }

Device::~Device() {
  if (ctx_) monsel_destroy(ctx_);
}

monsel_scene_id Device::upload(const CompiledScene& s) {
  auto it = cache_.find(s.hash);
  if (it != cache_.end()) return it->second;
  monsel_scene_id id;
  check(monsel_scene_upload(ctx_, &s.desc, &id), "monsel_scene_upload");
  cache_[s.hash] = id;
  return id;
}

/* --------------------------------------------------------- MonteCarloSS -- */

MonteCarloSS::MonteCarloSS(const ElectronGun* gun, Region* chamber, Electron* electron)
    : gun_(gun), chamber_(chamber), electron_(electron) {
  if (!gun_ || !chamber_) throw Error(MONSEL_E_INVALID, "MonteCarloSS: gun and chamber are required");
}

void MonteCarloSS::addActionListener(ActionListener& l) { listeners_.push_back(&l); }

void MonteCarloSS::removeActionListener(ActionListener& l) {
  auto it = std::find(listeners_.begin(), listeners_.end(), &l);
  if (it != listeners_.end()) listeners_.erase(it);
}

CompiledScene MonteCarloSS::compileScene() const {
  CompiledScene cs;
  Builder b(cs);
  b.addRegion(chamber_, -1);
  for (int z : b.mottZ) {
    MottData md = loadMottTable(z);
    cs.keep.push_back(md.spwem);
    cs.keep.push_back(md.x1);
    monsel_mott_table t;
    memset(&t, 0, sizeof(t));
    t.Z = z;
    t.spwem = md.spwem->data();
    t.x1 = md.x1->data();
    cs.mott.push_back(t);
  }
  double radius = ChamberRadius;
  if (auto* sp = dynamic_cast<const Sphere*>(chamber_->getShape())) radius = sp->radius;
  cs.finalize(compat_, radius);
  return cs;
}

PixelResults MonteCarloSS::runPixels(const std::vector<std::array<double, 3>>& centers, int nPerPixel,
                                     int nEnergyBins, double seCutoffEV) {
  if (nEnergyBins <= 0) {
    for (ActionListener* l : listeners_)
      if (l->energyBins() > 0) {
        nEnergyBins = l->energyBins();
        break;
      }
    if (nEnergyBins <= 0) nEnergyBins = std::max(1, (int)(FromSI::eV(gun_->getBeamEnergy()) / 10.));
  }
  for (ActionListener* l : listeners_)
    if (l->energyBins() > 0 && l->energyBins() != nEnergyBins)
      throw Error(MONSEL_E_INVALID, "all BackscatterStats of one MonteCarloSS must use the same bin count");
  CompiledScene cs = compileScene();
  Device& dev = Device::get(device_);
  monsel_scene_id sid = dev.upload(cs);
  PixelResults pr;
  pr.nPixels = (int)centers.size();
  pr.perPixel = nPerPixel;
  std::vector<double> flat(3 * centers.size());
  for (size_t i = 0; i < centers.size(); ++i)
    for (int k = 0; k < 3; ++k) flat[3 * i + k] = centers[i][k];
  monsel_job_desc job;
  memset(&job, 0, sizeof(job));
  job.beam_energy = gun_->getBeamEnergy();
  job.beam_width = gun_->getWidth();
  job.n_pixels = pr.nPixels;
  job.primaries_per_pixel = nPerPixel;
  job.pixel_center = flat.data();
  job.seed = seed_;
  job.n_energy_bins = nEnergyBins;
  job.se_cutoff_ev = seCutoffEV;
  job.max_steps = maxSteps_;
  size_t np = centers.size();
  pr.back.assign(np, 0);
  pr.fwd.assign(np, 0);
  pr.se.assign(np, 0);
  pr.seSq.assign(np, 0);
  pr.bseSq.assign(np, 0);
  pr.backEnergy.assign(np, 0);
  pr.backEnergyHist.assign(nEnergyBins + 2, 0);
  pr.fwdEnergyHist.assign(nEnergyBins + 2, 0);
  pr.elevationHist.assign(182, 0);
  pr.azimuthHist.assign(362, 0);
  monsel_tallies t;
  t.back = pr.back.data();
  t.fwd = pr.fwd.data();
  t.se = pr.se.data();
  t.se_sq = pr.seSq.data();
  t.bse_sq = pr.bseSq.data();
  t.back_energy = pr.backEnergy.data();
  t.back_energy_hist = pr.backEnergyHist.data();
  t.fwd_energy_hist = pr.fwdEnergyHist.data();
  t.elevation_hist = pr.elevationHist.data();
  t.azimuth_hist = pr.azimuthHist.data();
  int64_t logCap = 0;
  for (ActionListener* l : listeners_) logCap = std::max(logCap, l->detectedCapacity(nPerPixel * pr.nPixels));
  std::vector<monsel_detected> detected;
  int64_t nDetected = 0;
  if (logCap == 0) {
    check(monsel_run(dev.ctx(), sid, &job, &t, &pr.stats), "monsel_run");
  } else {
    job.log_detected = logCap;
    monsel_job_id jid = 0;
    check(monsel_job_submit(dev.ctx(), sid, &job, &jid), "monsel_job_submit");
    int st = monsel_job_wait(dev.ctx(), jid, -1);
    if (st == MONSEL_OK) st = monsel_job_read(dev.ctx(), jid, &t);
    if (st == MONSEL_OK) st = monsel_job_stats_get(dev.ctx(), jid, &pr.stats);
    if (st == MONSEL_OK) {
      /* the count first, then a buffer for the records kept */
      st = monsel_job_read_detected(dev.ctx(), jid, nullptr, 0, &nDetected);
      if (st == MONSEL_OK && nDetected > 0) {
        detected.resize((size_t)std::min<int64_t>(nDetected, logCap));
        st = monsel_job_read_detected(dev.ctx(), jid, detected.data(), (int64_t)detected.size(), &nDetected);
      }
    }
    std::string err = st == MONSEL_OK ? std::string() : monsel_last_error();
    monsel_job_release(dev.ctx(), jid);
    if (st != MONSEL_OK) throw Error(st, "monsel_job: " + err);
  }
  stats_ = pr.stats;
  /* advance the seed so successive runs draw fresh streams, like the
   * continuing rand() sequence of the reference */
  seed_ = seed_ * 6364136223846793005ull + 1442695040888963407ull;
  for (ActionListener* l : listeners_) {
    l->jobCompleted(*this, nPerPixel * pr.nPixels, t);
    if (logCap > 0) l->detectedLog(detected, nDetected);
  }
  return pr;
}

void MonteCarloSS::runMultipleTrajectories(int n) {
  auto c = gun_->getCenter();
  runPixels({c}, n);
}

}  // namespace monsel
