// Reads a mesh CSV, advances every cell with BDFIntegratorHIP, writes the mesh CSV: the
// reader -> integrator -> writer sequence of the reference application (app/main.cpp:39-64),
// with the integrator replaced.
//
//   hpo_integrate_csv --input cells.csv --output result.csv [--mechanism gri30.yaml]
//       [--dt 1e-3] [--pressure 101325.15] [--rtol 1e-6] [--atol 1e-10] [--lu fp64|fp32]
//       [--sweeps 1] [--reactor CONP|CONV] [--raw state.bin] [--performance-yaml performance.yaml]
//       [--quiet]
//
// --reactor CONV integrates at constant volume: each cell keeps the density it has at
// --pressure (default CONP, constant pressure).
//
// --raw writes the integrated mesh in full precision: per cell, T then the nsp mass
// fractions, as little-endian doubles, followed by the int32 solver status of every cell.
// Prints one JSON line: cells, species, integrate seconds (wall, as the reference's
// integrateTimer), kernel milliseconds, failed cells. --performance-yaml writes the reference
// application's performance.yaml (readTime, integrationTime, writeTime, totalTime in seconds,
// app/main.cpp:65-85), which its Python runner reads (app/main.py:157-168).
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <stdexcept>
#include <string>

#include "BDFIntegratorHIP.hpp"
#include "csvIO.hpp"

int main(int argc, char** argv) {
  std::string in, out, raw, perf, mech = "gri30.yaml", lu = "fp64", reactor = "CONP";
  double dt = 1e-3, P = 101325.15, rtol = 1e-6, atol = 1e-10;
  int sweeps = 1;
  bool quiet = false;
  for (int i = 1; i < argc; ++i) {
    std::string a = argv[i];
    auto val = [&]() -> std::string {
      if (i + 1 >= argc) throw std::invalid_argument("missing value for " + a);
      return argv[++i];
    };
    if (a == "--input") in = val();
    else if (a == "--output") out = val();
    else if (a == "--mechanism") mech = val();
    else if (a == "--dt") dt = std::stod(val());
    else if (a == "--pressure") P = std::stod(val());
    else if (a == "--rtol") rtol = std::stod(val());
    else if (a == "--atol") atol = std::stod(val());
    else if (a == "--lu") lu = val();
    else if (a == "--sweeps") sweeps = std::stoi(val());
    else if (a == "--reactor") reactor = val();
    else if (a == "--raw") raw = val();
    else if (a == "--performance-yaml") perf = val();
    else if (a == "--quiet") quiet = true;
    else {
      std::fprintf(stderr, "unknown argument %s\n", a.c_str());
      return 2;
    }
  }
  if (reactor != "CONP" && reactor != "CONV") {
    std::fprintf(stderr, "--reactor must be CONP or CONV\n");
    return 2;
  }
  if (in.empty() || out.empty()) {
    std::fprintf(stderr, "usage: hpo_integrate_csv --input in.csv --output out.csv [options]\n");
    return 2;
  }
  try {
    auto logger = quiet ? std::shared_ptr<BaseLogger>(std::make_shared<ConsoleLogger>(LogLevel::ERROR))
                        : std::shared_ptr<BaseLogger>(std::make_shared<ConsoleLogger>(LogLevel::INFO));
    auto ctx = std::make_shared<Context>(logger, std::make_shared<Mesh>());
    auto r0 = std::chrono::steady_clock::now();
    csvReader(ctx, in).read();
    auto r1 = std::chrono::steady_clock::now();

    hpo_opts o;
    hpo_opts_default(&o);
    o.lu_precision = (lu == "fp32") ? 1 : 0;
    o.refine_sweeps = sweeps;
    o.reactor = reactor == "CONV" ? 1 : 0;
    IntegratorConfig cfg;
    cfg.mechanism = mech;
    cfg.reltol = rtol;
    cfg.abstol = atol;
    cfg.pressure = P;

    auto t0 = std::chrono::steady_clock::now();
    BDFIntegratorHIP integ(&o);
    integ.init(ctx, cfg);
    integ.integrate(0.0, dt);
    std::vector<int32_t> status = integ.cellStatus();
    const double kernel_ms = integ.lastTiming().kernel_ms;
    integ.clean();
    auto t1 = std::chrono::steady_clock::now();   // init + integrate + clean, as integrateTimer

    auto w0 = std::chrono::steady_clock::now();
    csvWriter(ctx, out).write();
    auto w1 = std::chrono::steady_clock::now();
    if (!perf.empty()) {
      const double tr = std::chrono::duration<double>(r1 - r0).count();
      const double ti = std::chrono::duration<double>(t1 - t0).count();
      const double tw = std::chrono::duration<double>(w1 - w0).count();
      std::FILE* f = std::fopen(perf.c_str(), "w");
      if (!f) throw std::runtime_error("cannot open " + perf);
      std::fprintf(f, "readTime: %.17g\nintegrationTime: %.17g\nwriteTime: %.17g\ntotalTime: %.17g\n", tr, ti, tw,
                   tr + ti + tw);
      std::fclose(f);
    }
    Mesh& m = *ctx->getMesh();
    if (!raw.empty()) {
      std::FILE* f = std::fopen(raw.c_str(), "wb");
      if (!f) throw std::runtime_error("cannot open " + raw);
      for (int c = 0; c < m.totalSize(); ++c) {
        std::fwrite(&m.getTemperatureVector()[c], sizeof(double), 1, f);
        std::fwrite(m.getSpeciesVector(c).data(), sizeof(double), m.numSpecies(), f);
      }
      if (!status.empty()) std::fwrite(status.data(), sizeof(int32_t), status.size(), f);
      std::fclose(f);
    }
    int failed = 0;
    for (int32_t s : status) failed += (s != HPO_CELL_SUCCESS);
    std::printf("{\"cells\": %d, \"species\": %d, \"integrate_s\": %.6f, \"kernel_ms\": %.3f, \"failed\": %d}\n",
                m.totalSize(), m.numSpecies(), std::chrono::duration<double>(t1 - t0).count(), kernel_ms, failed);
    return failed ? 3 : 0;
  } catch (const std::exception& e) {
    std::fprintf(stderr, "error: %s\n", e.what());
    return 1;
  }
}
