// The reference's integrator plugin interface (include/ODEIntegrator/Integrators/
// Integrator.hpp:13-40, src/Integrators/Integrator.cpp:8-37): init(ctx, config) /
// integrate(t0, t) / clean(), with the mesh mutated in place.
//
// IntegratorConfig carries the same four fields the reference reads
// (app/main.cpp:47-51). The reference's fifth field, a yaml-cpp node with OpenMP
// settings, configures its CPU thread pool and has no meaning for the GPU integrator.
#pragma once

#include <memory>
#include <string>

#include "Context.hpp"
#include "Mesh.hpp"

struct IntegratorConfig {
  std::string mechanism;   // mechanism file: a path, or a name looked up in the mechanism folder
  double reltol = 1e-6;
  double abstol = 1e-10;
  double pressure = 101325.0;   // Pa, one value for every cell
};

class Integrator {
 public:
  virtual ~Integrator() = default;
  // Stores the context, logger and mesh sizes and the configuration, and logs them.
  virtual void init(std::shared_ptr<Context> ctx, IntegratorConfig config);
  // Logs the interval; subclasses advance every cell of the mesh from t0 to t.
  virtual void integrate(double t0, double t);
  virtual void clean();

 protected:
  int totalSize = 0;    // cells in the mesh
  int systemSize = 0;   // species per cell (= state length: T plus nsp - 1 mass fractions)
  std::shared_ptr<Mesh> mesh;
  std::shared_ptr<BaseLogger> logger;
  std::shared_ptr<Context> ctx;

  std::string mechanism;
  double reltol = 0, abstol = 0, pressure = 0;
};
