/*
 * monsel_host.hpp -- C++ host API of the MI355X transport engine.
 *
 * The classes keep the names, constructor arguments, units (SI) and meaning
 * of the reference's host objects (m1bcodes/CudaMONSEL, CudaMONSEL/CudaMONSEL/
 * gov/nist/...), so a driver written against the reference -- e.g.
 * nanoscalemetrology/JMONSELTests/LinesOnLayers.cu -- reads the same here.
 * They are descriptions: MonteCarloSS serialises the region tree into a
 * monsel_scene_desc and runs the trajectories on the GPU through the C-ABI in
 * include/monsel_gpu.h.  There is no CPU transport path in this library.
 */
#ifndef MONSEL_HOST_HPP
#define MONSEL_HOST_HPP

#include <stdint.h>

#include <array>
#include <map>
#include <memory>
#include <set>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/monsel_gpu.h"

namespace monsel {

/* Raised for every C-ABI failure and invalid host configuration. */
struct Error : std::runtime_error {
  int status;
  Error(int st, const std::string& m) : std::runtime_error(m), status(st) {}
};

/* Location of nistmott.bin / atomic_weights.txt (default: <lib dir>/data or
 * $MONSEL_DATA_DIR). */
void setDataDirectory(const std::string& dir);
std::string getDataDirectory();

/* ---- units: EPQLibrary/ToSI.cu, FromSI.cu, PhysicalConstants.cu */
namespace PhysicalConstants {
constexpr double ElectronCharge = 1.60217653e-19;
constexpr double AvagadroNumber = 6.0221415e23;
constexpr double BohrRadius = 5.291772083e-11;
constexpr double PI = 3.14159265358979323846;
}  // namespace PhysicalConstants
namespace ToSI {
inline double eV(double e) { return 1.60217653e-19 * e; }
inline double keV(double e) { return 1.60217653e-19 * 1.0e3 * e; }
}  // namespace ToSI
namespace FromSI {
inline double eV(double e) { return e * (1.0 / 1.60217653e-19); }
inline double keV(double e) { return e * (1.0 / (1.60217653e-19 * 1.0e3)); }
}  // namespace FromSI

/* ---- EPQLibrary/Element.cu: atomic number + FP32 atomic weight */
class Element {
 public:
  static const Element& byAtomicNumber(int Z);
  int getAtomicNumber() const { return Z_; }
  double getAtomicWeight() const; /* g/mol, read as float like Element.cu:606 */
  bool operator<(const Element& o) const { return Z_ < o.Z_; }
  bool operator==(const Element& o) const { return Z_ == o.Z_; }
  explicit Element(int Z) : Z_(Z) {}

 private:
  int Z_;
};
namespace Elements {
extern const Element& H;
extern const Element& C;
extern const Element& N;
extern const Element& O;
extern const Element& Si;
extern const Element& Cu;
}  // namespace Elements

/* ---- EPQLibrary/Composition.cu (weight fractions only) */
class Composition {
 public:
  Composition() = default;
  void defineByMoleFraction(const Element* const elms[], int n, const double moleFracs[], int nf);
  void defineByWeightFraction(const Element* const elms[], int n, const double wgtFracs[], int nf);
  double weightFraction(const Element& el, bool normalized) const;
  int getElementCount() const { return (int)elements_.size(); }
  const std::vector<int>& elementZ() const { return elements_; }

 protected:
  std::vector<int> elements_; /* insertion order */
  std::vector<double> wf_;    /* un-normalised weight fractions */
};

/* ---- EPQLibrary/Material.cu */
class Material : public Composition {
 public:
  Material() = default;
  Material(const Composition& comp, double density) : Composition(comp), density_(density) {}
  double getDensity() const { return density_; }
  void setDensity(double d) { density_ = d; }
  void setName(const std::string& n) { name_ = n; }
  const std::string& getName() const { return name_; }

 protected:
  double density_ = 0.;
  std::string name_;
};

/* ---- JMONSEL/SEmaterial.cu */
class SEmaterial : public Material {
 public:
  SEmaterial() = default; /* vacuum (SEmaterial.cu:11-13) */
  SEmaterial(const Composition& comp, double density) : Material(comp, density) {}
  SEmaterial(const Element* const elms[], int n, const double weightFracs[], int nw, double density,
             const char* name);
  void setWorkfunction(double w); /* clamps the CB bottom (SEmaterial.cu:392-399) */
  void setEnergyCBbottom(double e) { energyCBbottom_ = e; }
  void setBandgap(double g) { bandgap_ = g; }
  void setCoreEnergy(const double coreEnergy[], int len);
  void addCoreEnergy(double e) { core_.insert(e); }
  /* SEmaterial::setEstimatedCoreEnergy (SEmaterial.cu:347-366): the core
   * energies from the constituent elements' edge energies (NIST-Chantler
   * 2005, FFastEdgeDB) of their occupied shells, innermost first, down to
   * the first occupied shell whose edge is at or below the cutoff (J;
   * default 20 eV). */
  void setEstimatedCoreEnergy();
  void setEstimatedCoreEnergy(double cutoff);
  double getWorkfunction() const { return workfunction_; }
  double getEnergyCBbottom() const { return energyCBbottom_; }
  double getBandgap() const { return bandgap_; }
  double getEFermi() const { return -energyCBbottom_ - workfunction_; }
  const std::set<double>& getCoreEnergyArray() const { return core_; }

 private:
  double workfunction_ = 0., energyCBbottom_ = 0., bandgap_ = 0.;
  std::set<double> core_;
};

/* ---- slowing-down algorithms (JMONSEL/SlowingDownAlg.cuh) */
class SlowingDownAlg {
 public:
  virtual ~SlowingDownAlg() = default;
  virtual int kind() const = 0;
  virtual double breakE() const { return 0.; }
};
class ZeroCSD : public SlowingDownAlg {
 public:
  int kind() const override { return MONSEL_CSD_ZERO; }
};
/* JoyLuoNieminenCSD.cu:8-16 */
class JoyLuoNieminenCSD : public SlowingDownAlg {
 public:
  JoyLuoNieminenCSD(const SEmaterial& mat, double breakE);
  int kind() const override { return MONSEL_CSD_JOY_LUO_NIEMINEN; }
  double breakE() const override { return breakE_; }

 private:
  double breakE_;
};

/* ---- elastic cross-section factories (EPQLibrary/RandomizedScatterFactory.cuh) */
class RandomizedScatterFactory {
 public:
  RandomizedScatterFactory(int kind, int method) : kind_(kind), method_(method) {}
  int kind() const { return kind_; }
  int method() const { return method_; }

 private:
  int kind_, method_;
};
namespace NISTMottRS {
extern const RandomizedScatterFactory& Factory;        /* method 1 */
extern const RandomizedScatterFactory& Factory100;     /* method 2 */
extern const RandomizedScatterFactory& Factory100Lin;  /* method 3 */
}  // namespace NISTMottRS
namespace BrowningEmpiricalCrossSection {
extern const RandomizedScatterFactory& Factory;
}
namespace ScreenedRutherfordScatteringAngle {
extern const RandomizedScatterFactory& Factory;
}

/* ---- scatter mechanisms (JMONSEL/ScatterMechanism.cuh) */
class ScatterMechanism {
 public:
  virtual ~ScatterMechanism() = default;
  virtual monsel_mech_desc describe() const = 0;
  /* NU tables used by the mechanism, in table-slot order (-1 for unused). */
  virtual std::vector<std::string> tableFiles() const { return {}; }
};

/* SelectableElasticSM.cu:10-18 */
class SelectableElasticSM : public ScatterMechanism {
 public:
  SelectableElasticSM(const Material& mat, const RandomizedScatterFactory& rsf) : rsf_(rsf) { (void)mat; }
  explicit SelectableElasticSM(const Material& mat) : SelectableElasticSM(mat, NISTMottRS::Factory) {}
  monsel_mech_desc describe() const override;

 private:
  const RandomizedScatterFactory& rsf_;
};

/* FittedInelSM.cu:8-11 */
class FittedInelSM : public ScatterMechanism {
 public:
  FittedInelSM(const SEmaterial& mat, double energySEgen, const SlowingDownAlg& sdAlg);
  monsel_mech_desc describe() const override;

 private:
  double energySEgen_;
  double breakE_;
};

/* GanachaudMokraniPolaronTrapSM.cu:8-13 */
class GanachaudMokraniPolaronTrapSM : public ScatterMechanism {
 public:
  GanachaudMokraniPolaronTrapSM(double prefactor, double extinction);
  monsel_mech_desc describe() const override;

 private:
  double prefactor_, extinction_;
};

/* GanachaudMokraniPhononInelasticSM.cu:8-17 */
class GanachaudMokraniPhononInelasticSM : public ScatterMechanism {
 public:
  GanachaudMokraniPhononInelasticSM(double ratemultiplier, double phononE, double temperature, double eps0,
                                    double epsInfinity);
  monsel_mech_desc describe() const override;

 private:
  double p_[5];
};

/* TabulatedInelasticSM.cu:11-66; tables = {IIMFP, reduced deltaE, theta, SE E0}. */
class TabulatedInelasticSM : public ScatterMechanism {
 public:
  TabulatedInelasticSM(const SEmaterial& mat, int methodSE, const std::string tables[]);
  TabulatedInelasticSM(const SEmaterial& mat, int methodSE, const std::string tables[], double energyOffset);
  void setMinEgenSE(double e) { minEgenSE_ = e; }
  void setRateMult(double r) { rateMult_ = r; }
  void setE0fromDispersion(bool b) { e0FromDispersion_ = b; }
  monsel_mech_desc describe() const override;
  std::vector<std::string> tableFiles() const override { return tables_; }

 private:
  int methodSE_;
  std::vector<std::string> tables_;
  double energyOffset_ = 0., minEgenSE_ = 0., rateMult_ = 1.;
  bool e0FromDispersion_ = false;
};

/* ---- barriers */
class BarrierScatterMechanism {
 public:
  virtual ~BarrierScatterMechanism() = default;
  virtual int kind() const = 0;
  virtual double lambda() const { return 0.; }
};
/* ExpQMBarrierSM.cu:12-28 */
class ExpQMBarrierSM : public BarrierScatterMechanism {
 public:
  explicit ExpQMBarrierSM(const Material* mat) : classical_(true), lambda_(0.) { (void)mat; }
  ExpQMBarrierSM(const Material* mat, double lambda) : classical_(false), lambda_(lambda) { (void)mat; }
  int kind() const override { return classical_ ? MONSEL_BARRIER_EXPQM_CLASSICAL : MONSEL_BARRIER_EXPQM; }
  double lambda() const override { return lambda_; }

 private:
  bool classical_;
  double lambda_;
};

/* ---- material scatter models (NISTMonte/IMaterialScatterModel.cuh) */
class IMaterialScatterModel {
 public:
  virtual ~IMaterialScatterModel() = default;
  virtual const SEmaterial& getMaterial() const = 0;
  virtual double getMinEforTracking() const = 0;
  virtual void setMinEforTracking(double e) = 0;
  virtual monsel_msm_desc describe(int materialIndex) const = 0;
  virtual std::vector<const ScatterMechanism*> mechanisms() const { return {}; }
};

/* MONSEL_MaterialScatterModel.cu:13-37, 146-179 */
class MONSEL_MaterialScatterModel : public IMaterialScatterModel {
 public:
  MONSEL_MaterialScatterModel(const SEmaterial* mat, const BarrierScatterMechanism* bsm);
  bool addScatterMechanism(const ScatterMechanism* mech);
  bool removeScatterMechanism(const ScatterMechanism* mech);
  void setCSD(const SlowingDownAlg* csd) { csd_ = csd; }
  const SEmaterial& getMaterial() const override { return *mat_; }
  double getMinEforTracking() const override { return minE_; }
  void setMinEforTracking(double e) override { minE_ = e; }
  monsel_msm_desc describe(int materialIndex) const override;
  std::vector<const ScatterMechanism*> mechanisms() const override { return mech_; }

 private:
  const SEmaterial* mat_;
  const BarrierScatterMechanism* barrier_;
  const SlowingDownAlg* csd_ = nullptr;
  std::vector<const ScatterMechanism*> mech_;
  double minE_;
};

/* NullMaterialScatterModel.cu:10-50 */
class NullMaterialScatterModel : public IMaterialScatterModel {
 public:
  NullMaterialScatterModel();
  const SEmaterial& getMaterial() const override { return vacuum_; }
  double getMinEforTracking() const override { return minE_; }
  void setMinEforTracking(double e) override { minE_ = e; }
  monsel_msm_desc describe(int materialIndex) const override;

 private:
  SEmaterial vacuum_;
  double minE_;
};

/* ---- shapes (NISTMonte/Shape.cuh, JMONSEL/NormalShape.cuh) */
class Shape {
 public:
  virtual ~Shape() = default;
  virtual int kind() const = 0;
  virtual bool isNormalShape() const { return true; }
};
class NormalShape : public Shape {};

/* MultiPlaneShape.cu:9-16 */
class Plane {
 public:
  Plane(const double normal[], int normallen, const double point[], int pointlen);
  std::array<double, 3> normal, point;
};

/* NormalMultiPlaneShape.cu */
class NormalMultiPlaneShape : public NormalShape {
 public:
  NormalMultiPlaneShape() = default;
  void addPlane(const Plane& plane) { planes_.push_back(plane); }
  void addPlane(const double normal[], const double point[]);
  int getNumPlanes() const { return (int)planes_.size(); }
  const std::vector<Plane>& planes() const { return planes_; }
  void translate(const double d[]);
  int kind() const override { return MONSEL_SHAPE_PLANES; }

 private:
  std::vector<Plane> planes_;
};

/* NormalCylindricalShape.cu:8-17 */
class NormalCylindricalShape : public NormalShape {
 public:
  NormalCylindricalShape(const double end0[], const double end1[], double radius);
  void translate(const double d[]);
  int kind() const override { return MONSEL_SHAPE_CYLINDER; }
  std::array<double, 3> end0, end1;
  double radius;
};

/* NormalIntersectionShape.cu:5-7 */
class NormalIntersectionShape : public NormalShape {
 public:
  NormalIntersectionShape(NormalShape& a, NormalShape& b) : a(&a), b(&b) {}
  int kind() const override { return MONSEL_SHAPE_INTERSECTION; }
  NormalShape *a, *b;
};

/* NormalUnionShape.cu:8-10 */
class NormalUnionShape : public NormalShape {
 public:
  NormalUnionShape(NormalShape& a, NormalShape& b) : a(&a), b(&b) {}
  int kind() const override { return MONSEL_SHAPE_UNION; }
  NormalShape *a, *b;
};

/* Sphere.cu:7-9 (the chamber) */
class Sphere : public Shape {
 public:
  Sphere(const double center[], double radius);
  int kind() const override { return MONSEL_SHAPE_SPHERE; }
  bool isNormalShape() const override { return false; }
  std::array<double, 3> center;
  double radius;
};

/* JMONSEL/NShapes.cu:55-149.  With referenceCompat the left corner is built
 * from the right half-space as the reference does (NShapes.cu:140), so the
 * line has no left sidewall; otherwise the symmetric trapezoid is built.
 * The returned shape and its parts are owned by a process-wide pool. */
namespace NShapes {
NormalShape* createLine(double topz, double width, double length, double thetal, double thetar, double radl,
                        double radr, bool referenceCompat = true);
/* A step edge: the half-space z >= 0 joined with the slab x <= x0, -h <= z <= 0
 * (a plain NormalMultiPlaneShape/NormalUnionShape composition). */
NormalShape* createStepEdge(double x0, double h);
void translate(NormalShape* s, const double d[]);
}  // namespace NShapes

/* ---- regions (NISTMonte/RegionBase.cu:193-200) */
class Region {
 public:
  Region(Region* parent, IMaterialScatterModel* msm, Shape* shape);
  /* RegionBase.cu:21-28 */
  void updateMaterial(const IMaterialScatterModel& oldMat, IMaterialScatterModel& newMat);
  Region* getParent() const { return parent_; }
  IMaterialScatterModel* getScatterModel() const { return msm_; }
  Shape* getShape() const { return shape_; }
  const std::vector<Region*>& getSubRegions() const { return children_; }
  void removeSubRegion(Region& r);
  void clearSubRegions() { children_.clear(); }

 private:
  Region* parent_;
  IMaterialScatterModel* msm_;
  Shape* shape_;
  std::vector<Region*> children_;
};

/* ---- an electron on the host (NISTMonte/Electron.cu:15-60): a start
 * point, direction (theta from +z, phi) and kinetic energy.  Transport
 * creates and moves its electrons on the device; this class is the value
 * that ElectronGun::createElectron returns and MonteCarloSS's constructor
 * takes, as in the reference's drivers. */
class Electron {
 public:
  Electron(const double initialPos[], double kE) : Electron(initialPos, 0., 0., kE) {}
  Electron(const double initialPos[], double theta, double phi, double kE);
  std::array<double, 3> getPosition() const { return pos_; }
  void setPosition(const double p[]) { pos_ = {p[0], p[1], p[2]}; }
  double getEnergy() const { return energy_; }
  void setEnergy(double e) { energy_ = e; }
  double getTheta() const { return theta_; }
  double getPhi() const { return phi_; }
  void setDirection(double theta, double phi) {
    theta_ = theta;
    phi_ = phi;
  }
  long getStepCount() const { return 0; }
  long getIdent() const { return ident_; }

 private:
  std::array<double, 3> pos_;
  double theta_, phi_, energy_;
  long ident_;
};

/* ---- beam (NISTMonte/ElectronGun.cuh, GaussianBeam.cu) */
class ElectronGun {
 public:
  virtual ~ElectronGun() = default;
  /* a new electron at the gun (the caller owns it); this base places it at
   * the centre */
  virtual Electron* createElectron() const;
  virtual double getBeamEnergy() const = 0;
  virtual void setBeamEnergy(double e) = 0;
  virtual std::array<double, 3> getCenter() const = 0;
  virtual void setCenter(const double c[]) = 0;
  virtual double getWidth() const = 0;
};
class GaussianBeam : public ElectronGun {
 public:
  explicit GaussianBeam(double width);
  GaussianBeam(double width, double energy, const double center[]);
  void setWidth(double w) { width_ = w; }
  double getWidth() const override { return width_; }
  void setBeamEnergy(double e) override { energy_ = e; }
  double getBeamEnergy() const override { return energy_; }
  void setCenter(const double c[]) override { center_ = {c[0], c[1], c[2]}; }
  std::array<double, 3> getCenter() const override { return center_; }
  /* GaussianBeam.cu:52-65: the centre moved by a radius sqrt(-2 ln u1) *
   * width at the angle 2 pi u2, the beam energy, along +z (uniforms from a
   * host generator of the gun's own; device primaries use their Philox
   * streams) */
  Electron* createElectron() const override;

 private:
  double width_, energy_ = 0.;
  std::array<double, 3> center_;
  mutable uint64_t draws_ = 0;
};

/* ---- detectors (Utility/Histogram.cu, NISTMonte/BackscatterStats.cu) */
class Histogram {
 public:
  Histogram(double min, double max, int nBins);
  int binCount() const { return nbins_; }
  int counts(int bin) const { return (int)counts_[bin + 1]; }
  uint64_t counts64(int bin) const { return counts_[bin + 1]; }
  int overrange() const { return (int)counts_.back(); }
  int underrange() const { return (int)counts_.front(); }
  int totalCounts() const;
  double minValue(int bin) const;
  double maxValue(int bin) const;
  void clear();
  std::vector<uint64_t>& slots() { return counts_; }
  const std::vector<uint64_t>& slots() const { return counts_; }

 private:
  double min_, max_;
  int nbins_;
  std::vector<uint64_t> counts_;
};

class MonteCarloSS;

class ActionListener {
 public:
  virtual ~ActionListener() = default;
  /* Called once per run with the job's device tallies. */
  virtual void jobCompleted(const MonteCarloSS& mc, int nTrajectories, const monsel_tallies& t) = 0;
  virtual int energyBins() const { return 0; }
  /* Detected-electron records wanted from a run of nTrajectories (0: none),
   * and the records of the run (kept ones, and how many were detected). */
  virtual int64_t detectedCapacity(int nTrajectories) const { (void)nTrajectories; return 0; }
  virtual void detectedLog(const std::vector<monsel_detected>& log, int64_t nDetected) {
    (void)log;
    (void)nDetected;
  }
};

/* BackscatterStats::Datum (BackscatterStats.cu:40-48): one detected electron.
 * The electron's identity is its lineage within the cascade of `primary`. */
struct Datum {
  uint64_t electronID;
  uint32_t primary;
  long trajStep;
  double energyeV;
  std::array<double, 3> position; /* m */
  double theta, phi;
  long getElectronID() const { return (long)electronID; }
  long getTrajStep() const { return trajStep; }
  double getEnergyeV() const { return energyeV; }
  std::array<double, 3> getPosition() const { return position; }
  double getTheta() const { return theta; }
  double getPhi() const { return phi; }
};

/* BackscatterStats.cu:107-118, 166-168, 252-270 */
class BackscatterStats : public ActionListener {
 public:
  BackscatterStats(const MonteCarloSS& mcss, int nEnergyBins);
  /* BackscatterStats::setLogDetected: keep a Datum per detected electron
   * (at most 64 per trajectory and 2^26 per run are kept; logTruncated()
   * tells when a run detected more). */
  void setLogDetected(bool logDetected) { logDetected_ = logDetected; }
  bool getLogDetected() const { return logDetected_; }
  const std::vector<Datum>& getLog() const { return log_; }
  bool logTruncated() const { return logTruncated_; }
  int64_t detectedCapacity(int nTrajectories) const override {
    return logDetected_ ? std::min<int64_t>((int64_t)64 * nTrajectories, (int64_t)1 << 26) : 0;
  }
  void detectedLog(const std::vector<monsel_detected>& log, int64_t nDetected) override;
  double backscatterFraction() const;
  double forwardscatterFraction() const;
  double scatterFraction() const { return backscatterFraction() + forwardscatterFraction(); }
  const Histogram& backscatterEnergyHistogram() const { return back_; }
  const Histogram& forwardscatterEnergyHistogram() const { return fwd_; }
  const Histogram& elevationHistogram() const { return elev_; }
  const Histogram& azimuthalHistogram() const { return azim_; }
  int getEnergyBinCount() const { return nbins_; }
  void jobCompleted(const MonteCarloSS& mc, int nTrajectories, const monsel_tallies& t) override;
  int energyBins() const override { return nbins_; }

 private:
  int nbins_;
  double beamEnergyEV_;
  long eventCount_ = 0;
  Histogram elev_, azim_, fwd_, back_;
  bool logDetected_ = false;
  bool logTruncated_ = false;
  std::vector<Datum> log_;
};

/* Per-pixel results of a scan (the quantities LinesOnLayers.cu:366-376 derives). */
struct PixelResults {
  int nPixels = 0, perPixel = 0;
  std::vector<uint64_t> back, fwd, se;
  std::vector<double> seSq, bseSq, backEnergy;
  std::vector<uint64_t> backEnergyHist, fwdEnergyHist, elevationHist, azimuthHist;
  monsel_job_stats stats{};
  double seYield(int p) const { return (double)se[p] / perPixel; }
  double bseYield(int p) const { return (double)(back[p] - se[p]) / perPixel; }
};

/* A serialised scene: owns every array monsel_scene_desc points to. */
struct CompiledScene {
  monsel_scene_desc desc{};
  std::vector<monsel_material_desc> materials;
  std::vector<monsel_msm_desc> msms;
  std::vector<monsel_shape_desc> shapes;
  std::vector<monsel_plane_desc> planes;
  std::vector<monsel_region_desc> regions;
  std::vector<monsel_mott_table> mott;
  std::vector<monsel_nu_table> nu;
  std::vector<std::shared_ptr<const std::vector<double>>> keep; /* table storage */
  std::map<const Region*, int> regionIndex;
  uint64_t hash = 0;
  void finalize(uint32_t compat, double chamberRadius);
};

/* Process-wide device context (one per HIP device ordinal). */
class Device {
 public:
  static Device& get(int ordinal = 0);
  monsel_ctx ctx() const { return ctx_; }
  monsel_scene_id upload(const CompiledScene& s); /* cached by content hash */
  ~Device();

 private:
  explicit Device(int ordinal);
  monsel_ctx ctx_ = nullptr;
  std::map<uint64_t, monsel_scene_id> cache_;
};

/* NISTMonte/MonteCarloSS.cu:40-52, 181-192, 222-231 */
class MonteCarloSS {
 public:
  static constexpr double ChamberRadius = 0.1;  /* MonteCarloSS.cu:32 */
  /* the initial electron (MonteCarloSS.cu:40-52), owned from here on; the
   * trajectories themselves start from the gun on the device */
  MonteCarloSS(const ElectronGun* gun, Region* chamber, Electron* electron = nullptr);
  void addActionListener(ActionListener& l);
  void removeActionListener(ActionListener& l);
  /* runs n trajectories at the gun's current centre on the GPU */
  void runMultipleTrajectories(int n);
  /* batched extension: n trajectories at each of the given beam centres */
  PixelResults runPixels(const std::vector<std::array<double, 3>>& centers, int nPerPixel, int nEnergyBins = 0,
                         double seCutoffEV = 50.);
  double getBeamEnergy() const { return gun_->getBeamEnergy(); }
  Region* getChamber() const { return chamber_; }
  const ElectronGun* getElectronGun() const { return gun_; }
  void setElectronGun(const ElectronGun& g) { gun_ = &g; }
  /* Reference-compatibility flags (MONSEL_COMPAT_*); default REFERENCE. */
  void setRefCompat(uint32_t flags) { compat_ = flags; }
  uint32_t getRefCompat() const { return compat_; }
  void setSeed(uint64_t s) { seed_ = s; }
  uint64_t getSeed() const { return seed_; }
  void setDevice(int ordinal) { device_ = ordinal; }
  /* per-electron step limit (watchdog); 0 selects the engine default */
  void setMaxSteps(int n) { maxSteps_ = n; }
  int getMaxSteps() const { return maxSteps_; }
  /* Serialise the current object graph. */
  CompiledScene compileScene() const;
  const monsel_job_stats& lastStats() const { return stats_; }

 private:
  const ElectronGun* gun_;
  Region* chamber_;
  std::unique_ptr<Electron> electron_;
  std::vector<ActionListener*> listeners_;
  uint32_t compat_ = MONSEL_COMPAT_REFERENCE;
  uint64_t seed_ = 0x5EED5EEDull;
  int device_ = 0;
  int maxSteps_ = 0;
  monsel_job_stats stats_{};
};

/* Load every NIST SRD-64 table known to the data directory for Z. */
struct MottData {
  std::shared_ptr<const std::vector<double>> spwem, x1;
};
MottData loadMottTable(int Z);
/* NUTableInterpolation file reader (NUTableInterpolation.cu:117-228). */
struct NUTableData {
  int dim = 0;
  int n[4] = {0, 0, 0, 0};
  std::vector<double> x[4];
  std::vector<double> values;
};
NUTableData readNUTable(const std::string& path);

}  // namespace monsel

#endif
