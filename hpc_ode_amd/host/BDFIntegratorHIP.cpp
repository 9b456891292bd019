#include "BDFIntegratorHIP.hpp"

#include <dlfcn.h>
#include <sys/stat.h>

#include <algorithm>
#include <cstdlib>
#include <map>
#include <sstream>
#include <stdexcept>
#include <string>

namespace {

bool file_exists(const std::string& p) {
  struct stat st;
  return !p.empty() && ::stat(p.c_str(), &st) == 0 && S_ISREG(st.st_mode);
}

std::string dir_of_this_library() {
  Dl_info info{};
  if (dladdr(reinterpret_cast<void*>(&BDFIntegratorHIP::resolveMechanism), &info) == 0 || !info.dli_fname)
    return std::string();
  std::string p(info.dli_fname);
  size_t s = p.rfind('/');
  return s == std::string::npos ? std::string(".") : p.substr(0, s);
}

void check(int rc, const char* what) {
  if (rc != HPO_OK) throw std::runtime_error(std::string(what) + ": " + hpo_last_error());
}

// Largest number of cells resident on the device at once; larger meshes run in chunks.
constexpr int64_t kMaxResidentCells = int64_t(1) << 22;

}  // namespace

std::string BDFIntegratorHIP::resolveMechanism(const std::string& name) {
  if (file_exists(name)) return name;
  if (const char* d = std::getenv("HPO_MECHANISM_DIR")) {
    std::string p = std::string(d) + "/" + name;
    if (file_exists(p)) return p;
  }
  std::string p = dir_of_this_library() + "/mechanisms/" + name;
  if (file_exists(p)) return p;
  throw std::runtime_error("mechanism file not found: " + name);
}

BDFIntegratorHIP::BDFIntegratorHIP(const hpo_opts* opts) {
  if (opts)
    opts_ = *opts;
  else
    hpo_opts_default(&opts_);
}

BDFIntegratorHIP::~BDFIntegratorHIP() { release(); }

void BDFIntegratorHIP::release() {
  if (batch_) hpo_batch_destroy(batch_);
  if (mech_) hpo_mech_destroy(mech_);
  batch_ = nullptr;
  mech_ = nullptr;
}

void BDFIntegratorHIP::init(std::shared_ptr<Context> c, IntegratorConfig config) {
  Integrator::init(std::move(c), config);
  release();
  opts_.rtol = reltol;
  opts_.atol = abstol;
  try {
    const std::string path = resolveMechanism(mechanism);
    check(hpo_mech_load(path.c_str(), &mech_), "loading the mechanism");
    int32_t nsp = 0, nstate = 0, nrxn = 0;
    check(hpo_mech_info(mech_, &nsp, &nstate, &nrxn), "mechanism info");
    if (nsp != systemSize) {
      std::ostringstream ss;
      ss << "mechanism size of: " << mechanism << " (" << nsp << ") differs from the mesh size: " << systemSize;
      throw std::runtime_error(ss.str());
    }
    W_.assign(nsp, 0.0);
    check(hpo_mech_molecular_weights(mech_, W_.data()), "molecular weights");
    const int64_t cap = std::max<int64_t>(1, std::min<int64_t>(totalSize, kMaxResidentCells));
    check(hpo_batch_create(mech_, cap, &batch_), "allocating the device batch");
  } catch (const std::exception& e) {
    release();
    logger->error(e.what());
    throw std::runtime_error(e.what());
  }
  P_.clear();
  dt_.clear();
  logger->info("BDFIntegratorHIP initialized");
}

void BDFIntegratorHIP::setCellPressure(std::vector<double> P) {
  if (!P.empty() && static_cast<int>(P.size()) != totalSize)
    throw std::invalid_argument("setCellPressure: one pressure per cell is required");
  P_ = std::move(P);
}

void BDFIntegratorHIP::setCellTimeStep(std::vector<double> dt) {
  if (!dt.empty() && static_cast<int>(dt.size()) != totalSize)
    throw std::invalid_argument("setCellTimeStep: one time step per cell is required");
  dt_ = std::move(dt);
}

void BDFIntegratorHIP::integrate(double t0, double t) {
  Integrator::integrate(t0, t);
  if (!batch_) throw std::runtime_error("BDFIntegratorHIP::integrate called before init");
  if (mesh->totalSize() != totalSize || mesh->numSpecies() != systemSize)
    throw std::runtime_error("mesh changed size since init");
  const int n = totalSize, ns = systemSize;
  status_.assign(n, HPO_CELL_SUCCESS);
  stats_.assign(static_cast<size_t>(n) * HPO_NSTATS, 0);
  timing_ = hpo_timing{};
  if (n == 0 || ns == 0) return;

  // Mesh -> rows y = [T, Y_0 .. Y_{nsp-2}] (CVodeIntegrator.cpp:64-77)
  y_.resize(static_cast<size_t>(n) * ns);
  std::vector<double>& T = mesh->getTemperatureVector();
  for (int c = 0; c < n; ++c) {
    double* row = &y_[static_cast<size_t>(c) * ns];
    const std::vector<double>& Y = mesh->getSpeciesVector(c);
    row[0] = T[c];
    std::copy(Y.begin(), Y.end() - 1, row + 1);
  }

  // constant volume: each cell keeps the density it has at its pressure, rho = P Wbar / (R T)
  if (opts_.reactor == 1) {
    rho_.resize(n);
    for (int c = 0; c < n; ++c) {
      const std::vector<double>& Y = mesh->getSpeciesVector(c);
      double S = 0.0;
      for (int k = 0; k < ns; ++k) S += Y[k] / W_[k];
      rho_[c] = (P_.empty() ? pressure : P_[c]) / (HPO_GAS_CONSTANT * T[c] * S);
    }
  }
  const double* par = opts_.reactor == 1 ? rho_.data() : (P_.empty() ? nullptr : P_.data());

  const int64_t cap = std::max<int64_t>(1, std::min<int64_t>(n, kMaxResidentCells));
  for (int64_t first = 0; first < n; first += cap) {
    const int64_t cnt = std::min<int64_t>(cap, n - first);
    double* rows = &y_[static_cast<size_t>(first) * ns];
    check(hpo_batch_put(batch_, 0, cnt, rows, ns, par ? par + first : nullptr, pressure,
                        dt_.empty() ? nullptr : &dt_[first], t - t0),
          "copying cells to the device");
    hpo_timing tm{};
    check(hpo_batch_integrate(batch_, cnt, &opts_, nullptr, &tm), "integrating");
    check(hpo_batch_get(batch_, 0, cnt, rows, ns, &status_[first], &stats_[static_cast<size_t>(first) * HPO_NSTATS]),
          "copying cells from the device");
    timing_.kernel_ms += tm.kernel_ms;
    timing_.total_ms += tm.total_ms;
    timing_.ncells += tm.ncells;
    timing_.grid = tm.grid;
    timing_.block = tm.block;
    timing_.lds_bytes = tm.lds_bytes;
  }

  // rows -> Mesh, last species = 1 - sum of the others (CVodeIntegrator.cpp:79-96)
  for (int c = 0; c < n; ++c) {
    const double* row = &y_[static_cast<size_t>(c) * ns];
    std::vector<double>& Y = mesh->getSpeciesVector(c);
    T[c] = row[0];
    std::copy(row + 1, row + ns, Y.begin());
    double s = 0.0;
    for (int i = 0; i < ns - 1; ++i) s += Y[i];
    Y[ns - 1] = 1.0 - s;
  }
  // ENTHA = h(T, Y) of the integrated state, J/kg (what CanteraIntegrator.cpp:93 writes)
  if (mesh->hasEnthalpy()) {
    std::vector<double>& h = mesh->getEnthalpyVector();
    for (int c = 0; c < n; ++c)
      check(hpo_mech_enthalpy_mass(mech_, 1, &T[c], mesh->getSpeciesVector(c).data(), ns, &h[c]), "enthalpy");
  }

  std::map<int32_t, int> failed;
  for (int32_t s : status_)
    if (s != HPO_CELL_SUCCESS) ++failed[s];
  if (!failed.empty()) {
    std::ostringstream ss;
    ss << "cells not integrated to the end of the interval:";
    for (auto& kv : failed) ss << " " << kv.second << " with status " << kv.first;
    logger->error(ss.str());
  }
}

void BDFIntegratorHIP::clean() {
  release();
  Integrator::clean();
}
