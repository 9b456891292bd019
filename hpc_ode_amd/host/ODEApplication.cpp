// The reference's ODEApplication (app/main.cpp:17-88) with the integrator replaced: reads the
// same config.yaml (app/config.yaml:1-23), creates the timestamped execution folder of its
// OutFileService (include/ODEIntegrator/Context/OutFileService.hpp:10-30), copies the config
// there, reads the mesh CSV, integrates every cell from 0 to integrator.dt with
// BDFIntegratorHIP, writes the results CSV and performance.yaml into the execution folder, and
// logs to out.log (FileLogger) or the console (ConsoleLogger). The reference's Python runner
// (app/main.py:151) can launch it unchanged: it passes the config path as the only argument and
// collects config.yaml, performance.yaml, out.log and the .csv from the execution folder.
//
//   ODEApplication <path to config.yaml>
//
// Integrator types: the reference's CVODE integrators (CVodeIntegrator, CVodeIntegratorOMP,
// CVodeIntegratorGPU) and BDFIntegratorHIP all select the GPU BDF; the Cantera integrators are
// not part of this build and are refused like an unknown type ("Integrator type is not valid").
// The omp block is read and logged; the GPU integrator does not use it.
// Optional keys of this build under integrator: luPrecision (fp64 | fp32), refineSweeps, and
// reactor (CONP, the default: constant pressure; CONV: constant volume at the density each cell
// has at the configured pressure).
#include <chrono>
#include <cstdio>
#include <ctime>
#include <filesystem>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <memory>
#include <sstream>
#include <stdexcept>
#include <string>

#include "BDFIntegratorHIP.hpp"
#include "csvIO.hpp"
#include "yaml_lite.hpp"

namespace fs = std::filesystem;

namespace {

// Execution folder: <outFolder>/<local date and time, %F_%T>, as the reference's service.
std::string make_execution_folder(const std::string& outFolder) {
  if (fs::exists(outFolder)) {
    if (!fs::is_directory(outFolder)) {
      fs::remove_all(outFolder);
      fs::create_directory(outFolder);
    }
  } else {
    fs::create_directories(outFolder);
  }
  const std::time_t now = std::chrono::system_clock::to_time_t(std::chrono::system_clock::now());
  std::stringstream date;
  date << std::put_time(std::localtime(&now), "%F_%T");
  fs::path exec = fs::path(outFolder) / date.str();
  fs::create_directory(exec);
  return exec.string();
}

class FileLogger : public BaseLogger {
 public:
  FileLogger(LogLevel level, const std::string& path) : BaseLogger(level), out_(path) {
    if (!out_.is_open()) throw std::runtime_error("cannot open log file " + path);
  }
  void log(std::string line) override {
    std::lock_guard<std::mutex> g(mu_);
    out_ << line << '\n';
    out_.flush();
  }

 private:
  std::ofstream out_;
  std::mutex mu_;
};

struct Timer {
  std::chrono::steady_clock::time_point t1, t2;
  void tic() { t1 = std::chrono::steady_clock::now(); }
  void toc() { t2 = std::chrono::steady_clock::now(); }
  double seconds() const { return std::chrono::duration<double>(t2 - t1).count(); }
};

std::string fmt(double v) {
  std::ostringstream s;
  s << v;
  return s.str();
}

void run(const std::string& configFile) {
  const hpo::YNode config = hpo::yaml_load_file(configFile);
  Timer readTimer, writeTimer, integrateTimer;

  // context: output service and logger (ODEIntegratorFactory.hpp:31-35, 74-99)
  const hpo::YNode* of = config["outFileService"].find("outFolder");
  const std::string outFolder = (of && !of->is_null()) ? of->as_string() : std::string("./out/");
  const std::string execFolder = make_execution_folder(outFolder);
  const std::string levelName = config["logger"]["logLevel"].as_string();
  const LogLevel level = levelName == "DEBUG" ? LogLevel::DEBUG : LogLevel::INFO;
  const std::string loggerType = config["logger"]["type"].as_string();
  std::shared_ptr<BaseLogger> logger;
  if (loggerType == "FileLogger")
    logger = std::make_shared<FileLogger>(level, (fs::path(execFolder) / "out.log").string());
  else if (loggerType == "ConsoleLogger")
    logger = std::make_shared<ConsoleLogger>(level);
  else
    throw std::runtime_error("Logger type is not valid");
  auto ctx = std::make_shared<Context>(logger, std::make_shared<Mesh>());
  ctx->outFolder = execFolder;
  logger->info("Starting ODEApplication...");

  logger->info("Copying config.yaml to results directory");
  fs::copy_file(configFile, fs::path(execFolder) / "config.yaml");

  // reader
  readTimer.tic();
  if (config["reader"]["type"].as_string() != "csvReader") throw std::runtime_error("Reader type is not valid");
  csvReader(ctx, config["reader"]["filename"].as_string()).read();
  readTimer.toc();

  // integrator
  integrateTimer.tic();
  const hpo::YNode& ic = config["integrator"];
  IntegratorConfig cfg;
  cfg.reltol = ic["reltol"].as_double();
  cfg.abstol = ic["abstol"].as_double();
  cfg.mechanism = ic["mechanism"].as_string();
  cfg.pressure = ic["pressure"].as_double();
  const double dt = ic["dt"].as_double();
  const std::string type = ic["type"].as_string();
  if (type != "BDFIntegratorHIP" && type != "CVodeIntegrator" && type != "CVodeIntegratorOMP" &&
      type != "CVodeIntegratorGPU")
    throw std::runtime_error("Integrator type is not valid");
  if (const hpo::YNode* omp = ic.find("omp"); omp && omp->has("cpus"))
    logger->debug("omp settings (not used by the GPU integrator): cpus " + (*omp)["cpus"].scalar);
  hpo_opts o;
  hpo_opts_default(&o);
  if (const hpo::YNode* lp = ic.find("luPrecision")) {
    if (lp->as_string() == "fp32") o.lu_precision = 1;
    else if (lp->as_string() != "fp64") throw std::runtime_error("luPrecision must be fp64 or fp32");
  }
  if (const hpo::YNode* rs = ic.find("refineSweeps")) o.refine_sweeps = static_cast<int32_t>(rs->as_long());
  if (const hpo::YNode* re = ic.find("reactor")) {
    if (re->as_string() == "CONV") o.reactor = 1;
    else if (re->as_string() != "CONP") throw std::runtime_error("reactor must be CONP or CONV");
  }
  logger->info("Integrator " + type + " runs as BDFIntegratorHIP");
  {
    BDFIntegratorHIP integrator(&o);
    integrator.init(ctx, cfg);
    integrator.integrate(0.0, dt);
    integrator.clean();
  }
  integrateTimer.toc();

  // writer
  writeTimer.tic();
  if (config["writer"]["type"].as_string() != "csvWriter") throw std::runtime_error("Writer type is not valid");
  csvWriter(ctx, config["writer"]["filename"].as_string()).write();
  writeTimer.toc();

  const double tr = readTimer.seconds(), ti = integrateTimer.seconds(), tw = writeTimer.seconds();
  logger->info("{ readTime: " + fmt(tr) + ", integrationTime: " + fmt(ti) + ", writeTime: " + fmt(tw) +
               ", totalTime: " + fmt(tr + ti + tw) + " }");
  logger->info("Writing performance.yaml to results directory");
  std::ofstream perf(fs::path(execFolder) / "performance.yaml");
  perf << std::setprecision(17) << "readTime: " << tr << "\nintegrationTime: " << ti << "\nwriteTime: " << tw
       << "\ntotalTime: " << tr + ti + tw << "\n";
  if (!perf) throw std::runtime_error("cannot write performance.yaml");
  logger->info("End of ODEApplication execution");
}

}  // namespace

int main(int argc, char* argv[]) {
  if (argc != 2) {
    std::cout << "Error: No argument provided" << std::endl;
    return 1;
  }
  const std::string arg1 = argv[1];
  if (arg1 == "-h" || arg1 == "--help") {
    std::cout << argv[0] << " <path to config.yaml>" << std::endl;
    return 0;
  }
  try {
    run(arg1);
  } catch (const std::exception& e) {
    std::cerr << "ODEApplication: " << e.what() << std::endl;
    return 1;
  }
  return 0;
}
