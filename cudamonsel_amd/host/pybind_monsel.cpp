/*
 * pybind_monsel.cpp -- Python bindings of the host API (monsel_host.hpp), so
 * that Python drivers and tests build scenes with the reference's class names
 * and run them on the GPU through the C-ABI.
 */
#include <pybind11/numpy.h>
#include <pybind11/pybind11.h>
#include <pybind11/stl.h>

#include "monsel_host.hpp"

namespace py = pybind11;
using namespace monsel;

template <class T>
static py::array_t<T> to_np(const std::vector<T>& v) {
  py::array_t<T> a(v.size());
  std::copy(v.begin(), v.end(), a.mutable_data());
  return a;
}

static std::vector<const Element*> elements(const py::list& l) {
  std::vector<const Element*> r;
  for (auto h : l) r.push_back(h.cast<const Element*>());
  return r;
}

PYBIND11_MODULE(_monsel, m) {
  m.doc() = "MI355X electron-transport engine: host API mirroring CudaMONSEL";
  py::register_exception<Error>(m, "MonselError");

  m.def("setDataDirectory", &setDataDirectory);
  m.def("getDataDirectory", &getDataDirectory);
  m.attr("MONSEL_COMPAT_REFERENCE") = (uint32_t)MONSEL_COMPAT_REFERENCE;
  m.attr("MONSEL_COMPAT_FIXED") = (uint32_t)MONSEL_COMPAT_FIXED;
  m.attr("MONSEL_COMPAT_BARRIER_NORMAL") = (uint32_t)MONSEL_COMPAT_BARRIER_NORMAL;
  m.attr("MONSEL_COMPAT_HISTOGRAM_BIN") = (uint32_t)MONSEL_COMPAT_HISTOGRAM_BIN;
  m.attr("MONSEL_COMPAT_D3_ABSCISSAE") = (uint32_t)MONSEL_COMPAT_D3_ABSCISSAE;

  py::module tosi = m.def_submodule("ToSI");
  tosi.def("eV", &ToSI::eV);
  tosi.def("keV", &ToSI::keV);
  py::module fromsi = m.def_submodule("FromSI");
  fromsi.def("eV", &FromSI::eV);
  fromsi.def("keV", &FromSI::keV);

  py::class_<Element>(m, "Element")
      .def_static("byAtomicNumber", &Element::byAtomicNumber, py::return_value_policy::reference)
      .def("getAtomicNumber", &Element::getAtomicNumber)
      .def("getAtomicWeight", &Element::getAtomicWeight);
  py::module el = m.def_submodule("Elements");
  el.attr("H") = py::cast(&Elements::H, py::return_value_policy::reference);
  el.attr("C") = py::cast(&Elements::C, py::return_value_policy::reference);
  el.attr("N") = py::cast(&Elements::N, py::return_value_policy::reference);
  el.attr("O") = py::cast(&Elements::O, py::return_value_policy::reference);
  el.attr("Si") = py::cast(&Elements::Si, py::return_value_policy::reference);
  el.attr("Cu") = py::cast(&Elements::Cu, py::return_value_policy::reference);

  py::class_<Composition>(m, "Composition")
      .def(py::init<>())
      .def("defineByMoleFraction",
           [](Composition& c, const py::list& els, const std::vector<double>& f) {
             auto e = elements(els);
             c.defineByMoleFraction(e.data(), (int)e.size(), f.data(), (int)f.size());
           })
      .def("defineByWeightFraction",
           [](Composition& c, const py::list& els, const std::vector<double>& f) {
             auto e = elements(els);
             c.defineByWeightFraction(e.data(), (int)e.size(), f.data(), (int)f.size());
           })
      .def("weightFraction", &Composition::weightFraction)
      .def("getElementCount", &Composition::getElementCount);
  py::class_<Material, Composition>(m, "Material")
      .def(py::init<const Composition&, double>())
      .def("getDensity", &Material::getDensity)
      .def("setName", &Material::setName);
  py::class_<SEmaterial, Material>(m, "SEmaterial")
      .def(py::init<>())
      .def(py::init<const Composition&, double>())
      .def(py::init([](const py::list& els, const std::vector<double>& w, double density, const std::string& name) {
        auto e = elements(els);
        return new SEmaterial(e.data(), (int)e.size(), w.data(), (int)w.size(), density, name.c_str());
      }))
      .def("setWorkfunction", &SEmaterial::setWorkfunction)
      .def("setEnergyCBbottom", &SEmaterial::setEnergyCBbottom)
      .def("setBandgap", &SEmaterial::setBandgap)
      .def("setCoreEnergy",
           [](SEmaterial& s, const std::vector<double>& c) { s.setCoreEnergy(c.data(), (int)c.size()); })
      .def("setEstimatedCoreEnergy", py::overload_cast<>(&SEmaterial::setEstimatedCoreEnergy))
      .def("setEstimatedCoreEnergy", py::overload_cast<double>(&SEmaterial::setEstimatedCoreEnergy))
      .def("getCoreEnergyArray", [](const SEmaterial& m) {
        return std::vector<double>(m.getCoreEnergyArray().begin(), m.getCoreEnergyArray().end());
      })
      .def("getWorkfunction", &SEmaterial::getWorkfunction)
      .def("getEnergyCBbottom", &SEmaterial::getEnergyCBbottom)
      .def("getBandgap", &SEmaterial::getBandgap)
      .def("getEFermi", &SEmaterial::getEFermi);

  py::class_<SlowingDownAlg>(m, "SlowingDownAlg");
  py::class_<ZeroCSD, SlowingDownAlg>(m, "ZeroCSD").def(py::init<>());
  py::class_<JoyLuoNieminenCSD, SlowingDownAlg>(m, "JoyLuoNieminenCSD").def(py::init<const SEmaterial&, double>());

  py::class_<RandomizedScatterFactory>(m, "RandomizedScatterFactory");
  py::module nm = m.def_submodule("NISTMottRS");
  nm.attr("Factory") = py::cast(&NISTMottRS::Factory, py::return_value_policy::reference);
  nm.attr("Factory100") = py::cast(&NISTMottRS::Factory100, py::return_value_policy::reference);
  nm.attr("Factory100Lin") = py::cast(&NISTMottRS::Factory100Lin, py::return_value_policy::reference);
  py::module br = m.def_submodule("BrowningEmpiricalCrossSection");
  br.attr("Factory") = py::cast(&BrowningEmpiricalCrossSection::Factory, py::return_value_policy::reference);
  py::module sr = m.def_submodule("ScreenedRutherfordScatteringAngle");
  sr.attr("Factory") = py::cast(&ScreenedRutherfordScatteringAngle::Factory, py::return_value_policy::reference);

  py::class_<ScatterMechanism>(m, "ScatterMechanism");
  py::class_<SelectableElasticSM, ScatterMechanism>(m, "SelectableElasticSM")
      .def(py::init<const Material&, const RandomizedScatterFactory&>())
      .def(py::init<const Material&>());
  py::class_<FittedInelSM, ScatterMechanism>(m, "FittedInelSM")
      .def(py::init<const SEmaterial&, double, const SlowingDownAlg&>());
  py::class_<GanachaudMokraniPolaronTrapSM, ScatterMechanism>(m, "GanachaudMokraniPolaronTrapSM")
      .def(py::init<double, double>());
  py::class_<GanachaudMokraniPhononInelasticSM, ScatterMechanism>(m, "GanachaudMokraniPhononInelasticSM")
      .def(py::init<double, double, double, double, double>());
  py::class_<TabulatedInelasticSM, ScatterMechanism>(m, "TabulatedInelasticSM")
      .def(py::init([](const SEmaterial& mat, int method, const std::vector<std::string>& t, double offset) {
             return new TabulatedInelasticSM(mat, method, t.data(), offset);
           }),
           py::arg("mat"), py::arg("methodSE"), py::arg("tables"), py::arg("energyOffset") = 0.)
      .def("setMinEgenSE", &TabulatedInelasticSM::setMinEgenSE)
      .def("setRateMult", &TabulatedInelasticSM::setRateMult)
      .def("setE0fromDispersion", &TabulatedInelasticSM::setE0fromDispersion);

  py::class_<BarrierScatterMechanism>(m, "BarrierScatterMechanism");
  py::class_<ExpQMBarrierSM, BarrierScatterMechanism>(m, "ExpQMBarrierSM")
      .def(py::init<const Material*>())
      .def(py::init<const Material*, double>());

  py::class_<IMaterialScatterModel>(m, "IMaterialScatterModel")
      .def("getMinEforTracking", &IMaterialScatterModel::getMinEforTracking)
      .def("setMinEforTracking", &IMaterialScatterModel::setMinEforTracking);
  py::class_<MONSEL_MaterialScatterModel, IMaterialScatterModel>(m, "MONSEL_MaterialScatterModel")
      .def(py::init<const SEmaterial*, const BarrierScatterMechanism*>(), py::keep_alive<1, 2>(),
           py::keep_alive<1, 3>())
      .def("addScatterMechanism", &MONSEL_MaterialScatterModel::addScatterMechanism, py::keep_alive<1, 2>())
      .def("removeScatterMechanism", &MONSEL_MaterialScatterModel::removeScatterMechanism)
      .def("setCSD", &MONSEL_MaterialScatterModel::setCSD, py::keep_alive<1, 2>());
  py::class_<NullMaterialScatterModel, IMaterialScatterModel>(m, "NullMaterialScatterModel").def(py::init<>());

  py::class_<Shape>(m, "Shape");
  py::class_<NormalShape, Shape>(m, "NormalShape");
  py::class_<Plane>(m, "Plane").def(py::init([](const std::vector<double>& n, const std::vector<double>& p) {
    return new Plane(n.data(), (int)n.size(), p.data(), (int)p.size());
  }));
  py::class_<NormalMultiPlaneShape, NormalShape>(m, "NormalMultiPlaneShape")
      .def(py::init<>())
      .def("addPlane", [](NormalMultiPlaneShape& s, const Plane& p) { s.addPlane(p); })
      .def("addPlane", [](NormalMultiPlaneShape& s, const std::vector<double>& n,
                          const std::vector<double>& p) { s.addPlane(n.data(), p.data()); })
      .def("getNumPlanes", &NormalMultiPlaneShape::getNumPlanes)
      .def("translate", [](NormalMultiPlaneShape& s, const std::vector<double>& d) { s.translate(d.data()); });
  py::class_<NormalCylindricalShape, NormalShape>(m, "NormalCylindricalShape")
      .def(py::init([](const std::vector<double>& a, const std::vector<double>& b, double r) {
        return new NormalCylindricalShape(a.data(), b.data(), r);
      }))
      .def("translate", [](NormalCylindricalShape& s, const std::vector<double>& d) { s.translate(d.data()); });
  py::class_<NormalIntersectionShape, NormalShape>(m, "NormalIntersectionShape")
      .def(py::init<NormalShape&, NormalShape&>(), py::keep_alive<1, 2>(), py::keep_alive<1, 3>());
  py::class_<NormalUnionShape, NormalShape>(m, "NormalUnionShape")
      .def(py::init<NormalShape&, NormalShape&>(), py::keep_alive<1, 2>(), py::keep_alive<1, 3>());
  py::class_<Sphere, Shape>(m, "Sphere").def(
      py::init([](const std::vector<double>& c, double r) { return new Sphere(c.data(), r); }));
  py::module ns = m.def_submodule("NShapes");
  ns.def("createLine", &NShapes::createLine, py::arg("topz"), py::arg("width"), py::arg("length"), py::arg("thetal"),
         py::arg("thetar"), py::arg("radl"), py::arg("radr"), py::arg("referenceCompat") = true,
         py::return_value_policy::reference);
  ns.def("createStepEdge", &NShapes::createStepEdge, py::return_value_policy::reference);
  ns.def("translate", [](NormalShape* s, const std::vector<double>& d) { NShapes::translate(s, d.data()); });

  py::class_<Region>(m, "Region")
      .def(py::init<Region*, IMaterialScatterModel*, Shape*>(), py::keep_alive<1, 2>(), py::keep_alive<1, 3>(),
           py::keep_alive<1, 4>(), py::keep_alive<2, 1>())
      .def("updateMaterial", &Region::updateMaterial, py::keep_alive<1, 3>())
      .def("getParent", &Region::getParent, py::return_value_policy::reference)
      .def("getSubRegions", &Region::getSubRegions, py::return_value_policy::reference)
      .def("getScatterModel", &Region::getScatterModel, py::return_value_policy::reference);

  py::class_<Electron>(m, "Electron")
      .def(py::init([](const std::vector<double>& p, double kE) { return new Electron(p.data(), kE); }))
      .def(py::init([](const std::vector<double>& p, double th, double ph, double kE) {
        return new Electron(p.data(), th, ph, kE);
      }))
      .def("getPosition", &Electron::getPosition)
      .def("getEnergy", &Electron::getEnergy)
      .def("getTheta", &Electron::getTheta)
      .def("getPhi", &Electron::getPhi)
      .def("getStepCount", &Electron::getStepCount)
      .def("getIdent", &Electron::getIdent);

  py::class_<ElectronGun>(m, "ElectronGun")
      .def("createElectron", [](const ElectronGun& g) { return std::unique_ptr<Electron>(g.createElectron()); });
  py::class_<GaussianBeam, ElectronGun>(m, "GaussianBeam")
      .def(py::init<double>())
      .def(py::init([](double w, double e, const std::vector<double>& c) { return new GaussianBeam(w, e, c.data()); }))
      .def("setCenter", [](GaussianBeam& g, const std::vector<double>& c) { g.setCenter(c.data()); })
      .def("getCenter", &GaussianBeam::getCenter)
      .def("setBeamEnergy", &GaussianBeam::setBeamEnergy)
      .def("getBeamEnergy", &GaussianBeam::getBeamEnergy)
      .def("setWidth", &GaussianBeam::setWidth)
      .def("getWidth", &GaussianBeam::getWidth);

  py::class_<Histogram>(m, "Histogram")
      .def("binCount", &Histogram::binCount)
      .def("counts", &Histogram::counts)
      .def("overrange", &Histogram::overrange)
      .def("underrange", &Histogram::underrange)
      .def("totalCounts", &Histogram::totalCounts)
      .def("minValue", &Histogram::minValue)
      .def("maxValue", &Histogram::maxValue)
      .def("slots", [](const Histogram& h) { return to_np(h.slots()); });

  py::class_<ActionListener>(m, "ActionListener");
  py::class_<BackscatterStats, ActionListener>(m, "BackscatterStats")
      .def(py::init<const MonteCarloSS&, int>())
      .def("backscatterFraction", &BackscatterStats::backscatterFraction)
      .def("forwardscatterFraction", &BackscatterStats::forwardscatterFraction)
      .def("scatterFraction", &BackscatterStats::scatterFraction)
      .def("backscatterEnergyHistogram", &BackscatterStats::backscatterEnergyHistogram,
           py::return_value_policy::reference_internal)
      .def("forwardscatterEnergyHistogram", &BackscatterStats::forwardscatterEnergyHistogram,
           py::return_value_policy::reference_internal)
      .def("elevationHistogram", &BackscatterStats::elevationHistogram, py::return_value_policy::reference_internal)
      .def("azimuthalHistogram", &BackscatterStats::azimuthalHistogram, py::return_value_policy::reference_internal)
      .def("getEnergyBinCount", &BackscatterStats::getEnergyBinCount)
      .def("setLogDetected", &BackscatterStats::setLogDetected)
      .def("getLogDetected", &BackscatterStats::getLogDetected)
      .def("logTruncated", &BackscatterStats::logTruncated)
      .def("getLog", [](const BackscatterStats& b) {
        /* columns: electron id, primary, step, energy (eV), x, y, z (m), theta, phi */
        const auto& L = b.getLog();
        py::array_t<double> a({(py::ssize_t)L.size(), (py::ssize_t)9});
        auto r = a.mutable_unchecked<2>();
        for (size_t i = 0; i < L.size(); ++i) {
          const Datum& d = L[i];
          const double v[9] = {(double)d.electronID, (double)d.primary, (double)d.trajStep, d.energyeV,
                               d.position[0], d.position[1], d.position[2], d.theta, d.phi};
          for (int k = 0; k < 9; ++k) r(i, k) = v[k];
        }
        return a;
      });

  py::class_<monsel_job_stats>(m, "JobStats")
      .def_readonly("primaries", &monsel_job_stats::primaries)
      .def_readonly("electrons", &monsel_job_stats::electrons)
      .def_readonly("steps", &monsel_job_stats::steps)
      .def_readonly("escapes", &monsel_job_stats::escapes)
      .def_readonly("watchdog_drops", &monsel_job_stats::watchdog_drops)
      .def_readonly("queue_overflow", &monsel_job_stats::queue_overflow)
      .def_readonly("generations", &monsel_job_stats::generations)
      .def_readonly("device_ms", &monsel_job_stats::device_ms)
      .def_readonly("transport_ms", &monsel_job_stats::transport_ms)
      .def_readonly("transport_launches", &monsel_job_stats::transport_launches)
      .def_readonly("total_ms", &monsel_job_stats::total_ms);

  py::class_<PixelResults>(m, "PixelResults")
      .def_readonly("nPixels", &PixelResults::nPixels)
      .def_readonly("perPixel", &PixelResults::perPixel)
      .def_property_readonly("back", [](const PixelResults& r) { return to_np(r.back); })
      .def_property_readonly("fwd", [](const PixelResults& r) { return to_np(r.fwd); })
      .def_property_readonly("se", [](const PixelResults& r) { return to_np(r.se); })
      .def_property_readonly("seSq", [](const PixelResults& r) { return to_np(r.seSq); })
      .def_property_readonly("bseSq", [](const PixelResults& r) { return to_np(r.bseSq); })
      .def_property_readonly("backEnergy", [](const PixelResults& r) { return to_np(r.backEnergy); })
      .def_property_readonly("backEnergyHist", [](const PixelResults& r) { return to_np(r.backEnergyHist); })
      .def_property_readonly("fwdEnergyHist", [](const PixelResults& r) { return to_np(r.fwdEnergyHist); })
      .def_property_readonly("elevationHist", [](const PixelResults& r) { return to_np(r.elevationHist); })
      .def_property_readonly("azimuthHist", [](const PixelResults& r) { return to_np(r.azimuthHist); })
      .def_readonly("stats", &PixelResults::stats);

  py::class_<CompiledScene>(m, "CompiledScene")
      .def("desc_address", [](const CompiledScene& c) { return (uintptr_t)&c.desc; })
      .def_property_readonly("hash", [](const CompiledScene& c) { return c.hash; })
      .def("region_index", [](const CompiledScene& c, const Region* r) { return c.regionIndex.at(r); })
      .def_property_readonly("n_regions", [](const CompiledScene& c) { return (int)c.regions.size(); })
      .def_property_readonly("n_msm", [](const CompiledScene& c) { return (int)c.msms.size(); })
      .def("msm_index_of_region", [](const CompiledScene& c, int r) { return c.regions.at(r).msm; });

  py::class_<Device, std::unique_ptr<Device, py::nodelete>>(m, "Device")
      .def_static("get", &Device::get, py::arg("ordinal") = 0, py::return_value_policy::reference)
      .def("ctx_address", [](const Device& d) { return (uintptr_t)d.ctx(); })
      .def("upload", &Device::upload);

  py::class_<MonteCarloSS>(m, "MonteCarloSS")
      .def(py::init([](const ElectronGun* g, Region* c) { return new MonteCarloSS(g, c); }), py::keep_alive<1, 2>(),
           py::keep_alive<1, 3>())
      .def(py::init([](const ElectronGun* g, Region* c, const Electron& e) {
             return new MonteCarloSS(g, c, new Electron(e));
           }),
           py::keep_alive<1, 2>(), py::keep_alive<1, 3>())
      .def("addActionListener", &MonteCarloSS::addActionListener, py::keep_alive<1, 2>())
      .def("removeActionListener", &MonteCarloSS::removeActionListener)
      .def("runMultipleTrajectories", &MonteCarloSS::runMultipleTrajectories, py::call_guard<py::gil_scoped_release>())
      .def(
          "runPixels",
          [](MonteCarloSS& mc, py::array_t<double, py::array::c_style | py::array::forcecast> centers, int n,
             int nbins, double cutoff) {
            if (centers.ndim() != 2 || centers.shape(1) != 3) throw Error(MONSEL_E_INVALID, "centers must be (N, 3)");
            std::vector<std::array<double, 3>> c(centers.shape(0));
            auto a = centers.unchecked<2>();
            for (ssize_t i = 0; i < centers.shape(0); ++i) c[i] = {a(i, 0), a(i, 1), a(i, 2)};
            py::gil_scoped_release rel;
            return mc.runPixels(c, n, nbins, cutoff);
          },
          py::arg("centers"), py::arg("nPerPixel"), py::arg("nEnergyBins") = 0, py::arg("seCutoffEV") = 50.)
      .def("getBeamEnergy", &MonteCarloSS::getBeamEnergy)
      .def("setRefCompat", &MonteCarloSS::setRefCompat)
      .def("getRefCompat", &MonteCarloSS::getRefCompat)
      .def("setSeed", &MonteCarloSS::setSeed)
      .def("getSeed", &MonteCarloSS::getSeed)
      .def("setDevice", &MonteCarloSS::setDevice)
      .def("setMaxSteps", &MonteCarloSS::setMaxSteps)
      .def("getMaxSteps", &MonteCarloSS::getMaxSteps)
      .def("compileScene", &MonteCarloSS::compileScene)
      .def("lastStats", &MonteCarloSS::lastStats);
  m.attr("ChamberRadius") = MonteCarloSS::ChamberRadius;
}
