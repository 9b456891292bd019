// C++ tests of the host layer above the C ABI: Mesh/Point, the CSV reader and writer, and
// BDFIntegratorHIP's init/integrate/clean contract. They cover the behaviours the
// reference's own tests check (tests/Mesh/testMesh.cpp, tests/Mesh/testPoint.cpp,
// tests/InputOutput/Reader/testCsvReader.cpp, tests/InputOutput/Writer/testCsvWriter.cpp,
// tests/Integrators/testCVodeIntegrator.cpp).
//
//   test_host cpu <golden-dir> <tmp-dir>    no device needed (mesh, CSV, init errors)
//   test_host gpu <golden-dir> <tmp-dir>    integrates on device 0
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <functional>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "BDFIntegratorHIP.hpp"
#include "csvIO.hpp"

static int g_fail = 0;
#define CHECK(c)                                                          \
  do {                                                                    \
    if (!(c)) {                                                           \
      std::fprintf(stderr, "FAILED %s:%d: %s\n", __FILE__, __LINE__, #c); \
      ++g_fail;                                                           \
    }                                                                     \
  } while (0)

template <class E>
static std::string message_of(const std::function<void()>& f) {
  try {
    f();
  } catch (const E& e) {
    return e.what();
  }
  return "<no exception>";
}

static Point full_point(double T) {
  return Point::create(3).addTemperature(T).addSpecies({0.2, 0.2, 0.6}).addEnthalpy(-1e4).addCoordinates({1, 2, 3});
}

static std::shared_ptr<Context> quiet_ctx() {
  return std::make_shared<Context>(std::make_shared<MemoryLogger>(), std::make_shared<Mesh>());
}

static void test_point_and_mesh() {
  Point p = full_point(100);
  CHECK(p.hasTemperature() && p.hasEnthalpy() && p.hasCoordinates() && p.hasSpecies() && p.isReady());
  Point bad = Point::create(3).addSpecies({0.2, 0.2, 0.6});
  CHECK(!bad.isReady());
  // the reference builder's checks (src/Mesh/PointBuilder.cpp:8-70)
  CHECK(message_of<std::runtime_error>([] { Point q = Point::create(3).addSpecies({1.0}); (void)q; }) ==
        "Vector size does not match the number of the species of the Point");
  CHECK(message_of<std::runtime_error>([] { Point q = Point::create(3).addSpecies({0.2, 0.2, 0.5}); (void)q; }) ==
        "Species vector breaks mass conservation law");
  CHECK(message_of<std::runtime_error>([] { Point q = Point::create(3).addSpecies({1.5, -0.5, 0.0}); (void)q; }) ==
        "Species vector breaks mass conservation law");
  CHECK(message_of<std::runtime_error>([] { Point q = Point::create(3).addSpecies({0.2, 0.2, 0.605}); (void)q; }) ==
        "<no exception>");
  CHECK(message_of<std::runtime_error>([] { Point q = Point::create(3).addTemperature(-1.0); (void)q; }) ==
        "Temperature cannot be negative");
  CHECK(message_of<std::logic_error>([] { Point q = Point::create(3).addTemperature(1.0).addTemperature(2.0); (void)q; }) ==
        "Temperature already set in the build process");
  CHECK(message_of<std::logic_error>([] { Point q = Point::create(3).addEnthalpy(1.0).addEnthalpy(2.0); (void)q; }) ==
        "Enthalpy already set in the build process");
  CHECK(message_of<std::logic_error>([] {
          Point q = Point::create(3).addSpecies({0.2, 0.2, 0.6}).addSpecies({0.2, 0.2, 0.6});
          (void)q;
        }) == "Species vector already set in the build process");
  CHECK(message_of<std::logic_error>([] { Point q = Point::create(3).addCoordinates({0, 0, 0}).addCoordinates({1, 1, 1}); (void)q; }) ==
        "Coordinates already set in the build process");

  Mesh m;
  for (int i = 0; i < 3; ++i) m.addPoint(full_point(100 + i));
  CHECK(m.totalSize() == 3 && m.numSpecies() == 3 && m.hasEnthalpy() && m.hasCoordinates());
  CHECK(m.getTemperatureVector()[2] == 102);
  CHECK(message_of<std::out_of_range>([&] { m.getSpeciesVector(4); }) != "<no exception>");
  // in-place mutation through the raw pointer is visible through the vector
  double* y0 = m.getSpeciesPointer(0);
  for (int i = 0; i < 3; ++i) y0[i] = 0.0;
  for (double v : m.getSpeciesVector(0)) CHECK(v == 0.0);
  CHECK(message_of<std::runtime_error>([&] { m.addPoint(bad); }) == "Trying to add an incompatible Point to Mesh");
  CHECK(message_of<std::runtime_error>([&] { m.addPoint(Point::create(3).addTemperature(1).addSpecies({1, 0, 0})); }) ==
        "Trying to add an incompatible Point to Mesh");
  m.clear();
  CHECK(m.totalSize() == 0 && !m.hasEnthalpy());
  CHECK(message_of<std::runtime_error>([&] { m.getEnthalpyVector(); }) == "Mesh does not have enthalpy");
  Mesh empty;
  CHECK(message_of<std::runtime_error>([&] { empty.addPoint(bad); }) == "Trying to add an incompatible Point to Mesh");
  CHECK(empty.totalSize() == 0);
}

static void test_csv(const std::string& gold, const std::string& tmp) {
  {
    auto ctx = quiet_ctx();
    csvReader r(ctx, gold + "/good_input_file.csv");
    r.read();
    CHECK(ctx->getMesh()->totalSize() > 0);
    CHECK(ctx->getMesh()->numSpecies() == 53);
    CHECK(ctx->getMesh()->hasEnthalpy() && ctx->getMesh()->hasCoordinates());
  }
  {
    auto ctx = quiet_ctx();
    csvReader r(ctx, gold + "/file_without_temperature.csv");
    CHECK(message_of<std::runtime_error>([&] { r.read(); }) == "Input file does not have temperature!!");
  }
  {
    auto ctx = quiet_ctx();
    csvReader r(ctx, "data/nothing.csv");
    CHECK(message_of<std::runtime_error>([&] { r.read(); }) == "Failing opening input file: data/nothing.csv");
    CHECK(message_of<std::invalid_argument>([&] { csvReader q(ctx, "data/dummmy.txt"); }) ==
          "Input file should end with '.csv', actual: data/dummmy.txt");
    CHECK(message_of<std::invalid_argument>([&] { csvWriter q(ctx, "out.txt"); }) ==
          "Output file should end with '.csv', actual: out.txt");
  }
  {
    // 10 cells in, header + 10 lines out; reading the written file gives the values back to
    // the writer's 6 digits
    auto ctx = quiet_ctx();
    csvReader(ctx, gold + "/res_gri_10.csv").read();
    CHECK(ctx->getMesh()->totalSize() == 10);
    ctx->outFolder = tmp;
    csvWriter w(ctx, "gri_10_results.csv");
    w.write();
    std::ifstream f(w.path());
    std::string line;
    int lines = 0;
    std::string head;
    while (std::getline(f, line)) {
      if (lines == 0) head = line;
      ++lines;
    }
    CHECK(lines == 11);
    CHECK(head.rfind("0CO0,CO1,CO2,", 0) == 0);
    CHECK(head.find(",CO52,ENTHA,TEMPE,Points:0,Points:1,Points:2") != std::string::npos);
    auto back = quiet_ctx();
    csvReader(back, w.path()).read();
    Mesh& a = *ctx->getMesh();
    Mesh& b = *back->getMesh();
    CHECK(b.totalSize() == 10 && b.numSpecies() == 53);
    for (int c = 0; c < 10 && c < b.totalSize(); ++c) {
      CHECK(std::fabs(a.getTemperatureVector()[c] - b.getTemperatureVector()[c]) <= 5e-6 * a.getTemperatureVector()[c]);
      for (int k = 0; k < 53; ++k) {
        double x = a.getSpeciesVector(c)[k], y = b.getSpeciesVector(c)[k];
        CHECK(std::fabs(x - y) <= 5e-7 * std::fabs(x));
      }
    }
  }
}

static void test_integrator_contract(const std::string& gold) {
  // species count of the mesh must equal the mechanism's (CVodeIntegrator.cpp:41-48)
  auto ctx = quiet_ctx();
  ctx->getMesh()->addPoint(Point::create(3).addTemperature(1000).addSpecies({0.2, 0.2, 0.6}));
  IntegratorConfig cfg;
  cfg.mechanism = "gri30.yaml";
  BDFIntegratorHIP integ;
  std::string msg = message_of<std::runtime_error>([&] { integ.init(ctx, cfg); });
  CHECK(msg == "mechanism size of: gri30.yaml (53) differs from the mesh size: 3");
  cfg.mechanism = "no_such_mechanism.yaml";
  CHECK(message_of<std::runtime_error>([&] { integ.init(ctx, cfg); }) ==
        "mechanism file not found: no_such_mechanism.yaml");
  CHECK(message_of<std::runtime_error>([&] { integ.integrate(0, 1e-3); }) ==
        "BDFIntegratorHIP::integrate called before init");
  (void)gold;
}

static void test_integrate_reference_cells(const std::string& gold) {
  // The reference's 100-cell GRI file integrated for 1e-3 s at 101325.15 Pa
  // (tests/Integrators/testCVodeIntegratorGPU.cpp:15-27): every cell is a cold 300 K mixture.
  auto ctx = quiet_ctx();
  csvReader(ctx, gold + "/res_gri_100.csv").read();
  Mesh& m = *ctx->getMesh();
  const int n = m.totalSize();
  std::vector<std::vector<double>> vin = m.getSpeciesMatrix();
  IntegratorConfig cfg;
  cfg.mechanism = "gri30.yaml";
  cfg.reltol = 1e-6;
  cfg.abstol = 1e-10;
  cfg.pressure = 101325.15;
  for (int lu = 0; lu < 2; ++lu) {
    for (int c = 0; c < n; ++c) m.getSpeciesVector(c) = vin[c];
    hpo_opts o;
    hpo_opts_default(&o);
    o.lu_precision = lu;
    BDFIntegratorHIP integ(&o);
    integ.init(ctx, cfg);
    integ.integrate(0, 1e-3);
    for (int32_t s : integ.cellStatus()) CHECK(s == HPO_CELL_SUCCESS);
    // the state changed (testCVodeIntegrator.cpp:41-49) ...
    bool changed = false;
    for (int k = 0; k < 53; ++k) changed |= std::fabs(m.getSpeciesVector(0)[k] - vin[0][k]) > 1e-70;
    CHECK(changed);
    // ... the last species closes the mass balance ...
    for (int c = 0; c < n; ++c) {
      double s = 0;
      for (int k = 0; k < 52; ++k) s += m.getSpeciesVector(c)[k];
      CHECK(m.getSpeciesVector(c)[52] == 1.0 - s);
    }
    integ.clean();
    CHECK(message_of<std::runtime_error>([&] { integ.integrate(0, 1e-3); }) ==
          "BDFIntegratorHIP::integrate called before init");
  }
}

int main(int argc, char** argv) {
  if (argc < 4) {
    std::fprintf(stderr, "usage: test_host cpu|gpu <golden-dir> <tmp-dir>\n");
    return 2;
  }
  std::string mode = argv[1], gold = argv[2], tmp = argv[3];
  try {
    if (mode == "cpu") {
      test_point_and_mesh();
      test_csv(gold, tmp);
      test_integrator_contract(gold);
    } else if (mode == "gpu") {
      test_integrate_reference_cells(gold);
    } else {
      return 2;
    }
  } catch (const std::exception& e) {
    std::fprintf(stderr, "FAILED with exception: %s\n", e.what());
    return 1;
  }
  std::printf("%s: %d failures\n", mode.c_str(), g_fail);
  return g_fail ? 1 : 0;
}
