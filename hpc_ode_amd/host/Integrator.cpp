#include "Integrator.hpp"

#include <sstream>

void Integrator::init(std::shared_ptr<Context> c, IntegratorConfig config) {
  ctx = std::move(c);
  logger = ctx->getLogger();
  mesh = ctx->getMesh();
  totalSize = mesh->totalSize();
  systemSize = mesh->numSpecies();
  mechanism = config.mechanism;
  reltol = config.reltol;
  abstol = config.abstol;
  pressure = config.pressure;
  logger->info("Initializing integrator...");
  std::ostringstream ss;
  ss << "{ mechanism: " << mechanism << ", reltol: " << reltol << ", abstol: " << abstol
     << ", pressure: " << pressure << " }";
  logger->info(ss.str());
}

void Integrator::integrate(double t0, double t) {
  std::ostringstream ss;
  ss << "Integrating interval of: " << t - t0 << " seconds";
  logger->info(ss.str());
}

void Integrator::clean() { logger->info("Integrator clean completed"); }
