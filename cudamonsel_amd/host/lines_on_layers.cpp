/* A LinesOnLayers-style driver written against the host API alone
 * (monsel_host.hpp): the scene, the beam and the scan loop of
 * LinesOnLayers::run (JMONSELTests/LinesOnLayers.cu:37-399), one
 * runMultipleTrajectories call per landing position with a BackscatterStats
 * listener, and the reference's five-column output text.
 *
 * Scene (LinesOnLayers.cu:86-282): one PMMA line from NShapes::createLine
 * (height 120 nm, bottom width 80 nm, 3 deg sidewalls, 20 nm top corners,
 * 1000 nm long) standing on an 80 nm PMMA layer over Si.  The driver places
 * its third layer plane at the second layer's position (:260), so Si begins
 * right below the PMMA layer and its glassy-carbon layer never holds an
 * electron; this scene leaves that layer out.  Si uses NIST Mott elastic
 * scattering, the phonon mechanism of :198 and, because the reference's Si
 * tables are not shipped, FittedInelSM with Joy-Luo-Nieminen slowing down
 * (E_SEgen 63 eV, breakE 45 eV) in place of TabulatedInelasticSM.
 *
 * Usage: lines_on_layers [--traj N] [--xmax K] [--y Y1,Y2,...] [--seed S]
 *                        [--fixed] [--batched] [--out FILE]
 *   --traj   trajectories per landing position (100, as the reference)
 *   --xmax   the first K x positions of the reference's grid (all)
 *   --y      y positions in nm (0)
 *   --seed   MonteCarloSS seed (the class default)
 *   --fixed  the corrected geometry and physics (MONSEL_COMPAT_FIXED)
 *   --batched  the whole scan as one MonteCarloSS::runPixels job (the
 *            landing positions run concurrently on the GPU) instead of one
 *            runMultipleTrajectories call per position; the yields are the
 *            same estimates from other random streams
 *   --out    output text file (output.txt)
 */
#include <array>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <memory>
#include <sstream>
#include <string>
#include <vector>

#include "monsel_host.hpp"

using namespace monsel;

namespace {

struct Options {
  int traj = 100;
  int xmax = -1;
  std::vector<double> y{0.0};
  bool seeded = false;
  uint64_t seed = 0;
  bool fixed = false;
  bool batched = false;
  std::string out = "output.txt";
};

Options parse(int argc, char** argv) {
  Options o;
  for (int i = 1; i < argc; ++i) {
    const std::string a = argv[i];
    auto next = [&]() -> std::string {
      if (i + 1 >= argc) {
        fprintf(stderr, "%s needs a value\n", a.c_str());
        exit(2);
      }
      return argv[++i];
    };
    if (a == "--traj") {
      o.traj = std::atoi(next().c_str());
    } else if (a == "--xmax") {
      o.xmax = std::atoi(next().c_str());
    } else if (a == "--y") {
      o.y.clear();
      std::stringstream ss(next());
      std::string v;
      while (std::getline(ss, v, ',')) o.y.push_back(std::atof(v.c_str()));
    } else if (a == "--seed") {
      o.seeded = true;
      o.seed = std::strtoull(next().c_str(), nullptr, 0);
    } else if (a == "--fixed") {
      o.fixed = true;
    } else if (a == "--batched") {
      o.batched = true;
    } else if (a == "--out") {
      o.out = next();
    } else {
      fprintf(stderr, "unknown argument %s\n", a.c_str());
      exit(2);
    }
  }
  if (o.traj <= 0 || o.y.empty()) {
    fprintf(stderr, "--traj must be positive and --y non-empty\n");
    exit(2);
  }
  return o;
}

/* the x landing positions of LinesOnLayers.cu:296-329: 5 nm steps up to
 * 20.5 nm before the line's top edge, 1 nm steps across the right sidewall
 * to 20.5 nm past its foot, 5 nm steps beyond, and the last point */
std::vector<double> x_grid(double wnm, double hnm, double thetar) {
  const double xbottom = wnm / 2.;
  const double xtop = wnm / 2. - hnm * std::tan(thetar);
  const double xstart = xbottom - 300.5, xstop = xbottom + 100.5;
  const double xfinestart = xtop - 20.5;
  const double xfinestop = thetar < 0. ? xtop + 20.5 : wnm / 2. + 20.5;
  std::vector<double> xs;
  double x = xstart;
  for (; x < xfinestart; x += 5.) xs.push_back(x);
  for (x = xfinestart; x < xfinestop; x += 1.) xs.push_back(x);
  for (x = xfinestop; x < xstop; x += 5.) xs.push_back(x);
  xs.push_back(xstop);
  return xs;
}

std::string fmt(double v) { return std::to_string(v); }

}  // namespace

int main(int argc, char** argv) {
  const Options opt = parse(argc, argv);
  try {
    /* geometry (LinesOnLayers.cu:55-66) */
    const double hnm = 120., wnm = 80., linelengthnm = 1000., thetadeg = 3., radnm = 20.;
    const double layer1nm = 80., beamEeV = 500., beamsizenm = 0.5, binSizeEV = 10.;
    const double nm = 1.e-9;
    const double pi = 3.14159265358979323846;
    const double theta = thetadeg * pi / 180.;

    /* vacuum (:80-84) */
    SEmaterial vacuum;
    vacuum.setName("SE vacuum");
    ExpQMBarrierSM vacuumBarrier(&vacuum);
    MONSEL_MaterialScatterModel vacuumMSM(&vacuum, &vacuumBarrier);

    /* PMMA (:86-119): Mott elastic, fitted inelastic with JLN slowing down,
     * polaron trapping, classical exponential barrier */
    const Element* cohElems[] = {&Elements::C, &Elements::O, &Elements::H};
    const double cohMoles[] = {5., 2., 8.};
    Composition pmmaComp;
    pmmaComp.defineByMoleFraction(cohElems, 3, cohMoles, 3);
    SEmaterial pmma(pmmaComp, 1190.);
    pmma.setName("PMMA");
    pmma.setWorkfunction(ToSI::eV(5.5));
    pmma.setBandgap(ToSI::eV(5.));
    pmma.setEnergyCBbottom(ToSI::eV(-5.5 + 5.));
    SelectableElasticSM pmmaElastic(pmma, NISTMottRS::Factory);
    JoyLuoNieminenCSD pmmaCSD(pmma, ToSI::eV(45.));
    FittedInelSM pmmaInel(pmma, ToSI::eV(65.4), pmmaCSD);
    GanachaudMokraniPolaronTrapSM pmmaPolaron(2.e7, 1. / ToSI::eV(4.));
    ExpQMBarrierSM pmmaBarrier(&pmma);
    /* the line and the layer under it each have their own model, as the
     * reference copies PMMAMSM into ARCMSM (:121) */
    auto pmmaModel = [&]() {
      auto m = std::make_unique<MONSEL_MaterialScatterModel>(&pmma, &pmmaBarrier);
      m->addScatterMechanism(&pmmaElastic);
      m->addScatterMechanism(&pmmaInel);
      m->addScatterMechanism(&pmmaPolaron);
      m->setCSD(&pmmaCSD);
      return m;
    };
    auto arcMSM = pmmaModel();
    auto lineMSM = pmmaModel();

    /* Si (:172-212) with the fitted-inelastic stand-in for its tables */
    const Element* siElems[] = {&Elements::Si};
    const double siFrac[] = {1.};
    SEmaterial si(siElems, 1, siFrac, 1, 2330., "Silicon");
    si.setWorkfunction(ToSI::eV(4.85));
    si.setEnergyCBbottom(ToSI::eV(-4.85 + 1.1));
    si.setBandgap(ToSI::eV(1.1));
    const double siCore[] = {ToSI::eV(99.2), ToSI::eV(99.8), ToSI::eV(149.7), ToSI::eV(1839.)};
    si.setCoreEnergy(siCore, 4);
    SelectableElasticSM siElastic(si, NISTMottRS::Factory);
    JoyLuoNieminenCSD siCSD(si, ToSI::eV(45.));
    FittedInelSM siInel(si, ToSI::eV(63.), siCSD);
    GanachaudMokraniPhononInelasticSM siPhonon(3., ToSI::eV(0.063), 300., 11.7, 1.);
    ExpQMBarrierSM siBarrier(&si);
    MONSEL_MaterialScatterModel siMSM(&si, &siBarrier);
    siMSM.addScatterMechanism(&siElastic);
    siMSM.addScatterMechanism(&siInel);
    siMSM.addScatterMechanism(&siPhonon);
    siMSM.setCSD(&siCSD);

    /* chamber, layers and the line (:225-282) */
    const double origin[] = {0., 0., 0.};
    NullMaterialScatterModel nullMSM;
    Sphere sphere(origin, MonteCarloSS::ChamberRadius);
    GaussianBeam eg(beamsizenm * nm, ToSI::eV(beamEeV), origin);
    Region chamber(nullptr, &nullMSM, &sphere);
    chamber.updateMaterial(*chamber.getScatterModel(), vacuumMSM);
    const double down[] = {0., 0., -1.};
    NormalMultiPlaneShape layer1;
    layer1.addPlane(Plane(down, 3, origin, 3));
    Region layer1Region(&chamber, arcMSM.get(), &layer1);
    const double layer2Pos[] = {0., 0., layer1nm * nm};
    NormalMultiPlaneShape layer2;
    layer2.addPlane(Plane(down, 3, layer2Pos, 3));
    Region siRegion(&layer1Region, &siMSM, &layer2);
    const double h = hnm * nm;
    /* the shapes createLine builds belong to the library's shape pool */
    NormalShape* line =
        NShapes::createLine(-h, wnm * nm, linelengthnm * nm, theta, theta, radnm * nm, radnm * nm, !opt.fixed);
    Region lineRegion(&chamber, lineMSM.get(), line);

    MonteCarloSS monte(&eg, &chamber, eg.createElectron());
    monte.setRefCompat(opt.fixed ? MONSEL_COMPAT_FIXED : MONSEL_COMPAT_REFERENCE);
    if (opt.seeded) monte.setSeed(opt.seed);

    std::vector<double> xvals = x_grid(wnm, hnm, theta);
    if (opt.xmax >= 0 && opt.xmax < (int)xvals.size()) xvals.resize(opt.xmax);

    std::string output = "# LinesOnLayers scan on the MI355X transport engine";
    output += "\n# Trajectories at each landing position: " + std::to_string(opt.traj);
    output += "\n# lines: 1";
    output += "\nLine height (nm): " + fmt(hnm);
    output += "\nLine bottom width (nm): " + fmt(wnm);
    output += "\nLine length (nm): " + fmt(linelengthnm);
    output += "\nLeft and right sidewall angles (deg): " + fmt(thetadeg) + " " + fmt(thetadeg);
    output += "\nLeft and right top corner radii (nm): " + fmt(radnm) + " " + fmt(radnm);
    output += "\nThickness of the PMMA layer (nm): " + fmt(layer1nm);
    output += "\nBeam landing energies (eV): " + fmt(beamEeV);
    output += "\nBeam size (standard deviation, in nm): " + fmt(beamsizenm);
    output += "\n";
    output += "\nbeamE (eV)\tx(nm)\ty (nm)\tBSE yield\tSE yield";

    const int nbins = (int)(beamEeV / binSizeEV);
    const auto t0 = std::chrono::steady_clock::now();
    auto emit = [&](double xnm, double ynm, double bsf, double sef) {
      const std::string row = fmt(beamEeV) + " " + fmt(xnm) + " " + fmt(ynm) + " " + fmt(bsf) + " " + fmt(sef);
      printf("%s\n", row.c_str());
      output += "\n" + row;
    };
    if (opt.batched) {
      /* one job over every landing position; per-pixel SE counts use the
       * same histogram bins below 50 eV as the listener's histogram */
      std::vector<std::array<double, 3>> centers;
      for (double ynm : opt.y)
        for (double xnm : xvals) centers.push_back({xnm * nm, ynm * nm, -h - 20. * nm});
      PixelResults r = monte.runPixels(centers, opt.traj, nbins, 50.);
      int p = 0;
      for (double ynm : opt.y)
        for (double xnm : xvals) {
            emit(xnm, ynm, r.bseYield(p), r.seYield(p));
            ++p;
        }
    } else {
      for (double ynm : opt.y)
        for (double xnm : xvals) {
          const double c[] = {xnm * nm, ynm * nm, -h - 20. * nm};
          eg.setCenter(c);
          BackscatterStats back(monte, nbins);
          monte.addActionListener(back);
          monte.runMultipleTrajectories(opt.traj);
          /* SE: back-scattered electrons in the bins below 50 eV; BSE: the
           * rest of the back-scattered fraction (:360-371) */
          const Histogram& hist = back.backscatterEnergyHistogram();
          const double perBin = beamEeV / hist.binCount();
          const int seBins = (int)(50. / perBin);
          long totalSE = 0;
          for (int j = 0; j < seBins; ++j) totalSE += hist.counts(j);
          const double sef = (double)totalSE / opt.traj;
          const double bsf = back.backscatterFraction() - sef;
          emit(xnm, ynm, bsf, sef);
          monte.removeActionListener(back);
        }
    }
    const double elapsed = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    output += "\n" + std::to_string(elapsed);
    std::ofstream f(opt.out);
    f << output;
    if (!f) {
      fprintf(stderr, "cannot write %s\n", opt.out.c_str());
      return 1;
    }
    fprintf(stderr, "%zu landing positions in %.2f s\n", xvals.size() * opt.y.size(), elapsed);
  } catch (const std::exception& e) {
    fprintf(stderr, "lines_on_layers: %s\n", e.what());
    return 1;
  }
  return 0;
}
