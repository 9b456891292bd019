// GPU drop-in for the reference's CVodeIntegrator (src/Integrators/CVodeIntegrator.cpp:34-155):
// every mesh cell is integrated by its own variable-order BDF with CVODE's step, order and
// error control, on one MI355X, through the C ABI of include/hpc_ode.h.
//
// Differences a caller can see, all deliberate:
//  * The mechanism is read from its YAML or CTML file at init (the reference compiles pyJac code
//    for one mechanism). Kernels are instantiated for states of 9, 10 and 53 variables (one
//    wave per cell: h2_li_2004, h2o2, GRI-3.0 and other mechanisms of those sizes) and of 53
//    and 100 variables (one workgroup per cell: GRI-3.0, nDodecane_Reitz); integrate throws
//    for other sizes.
//  * A mesh with enthalpy gets ENTHA = h(T, Y) of the integrated state in J/kg, as the
//    reference's CanteraIntegrator writes it (CanteraIntegrator.cpp:93); its CVODE integrators
//    leave the input value. In the constant-pressure reactor h is conserved, so the two agree
//    to the integration tolerance when the input ENTHA was consistent with (T, Y).
//  * Device memory is owned from init to clean, so repeated integrate calls (substeps of a
//    CFD run) reuse it; the reference GPU path allocates per call (CVodeIntegratorGPU.cpp:57,121).
//  * The per-cell solver return code is kept (cellStatus()) and failed cells are logged; the
//    reference ignores it (CVodeIntegrator.cpp:143). A failed cell's mesh state is the state the
//    solver had reached, as in the reference.
//  * integrate(t0, t) advances each cell by t - t0; the reference integrates from 0 to t
//    whatever t0 is (CVodeIntegrator.cpp:59). The two agree for t0 = 0, the only call the
//    reference makes (app/main.cpp:56).
//  * Optional per-cell pressure and time step (setCellPressure / setCellTimeStep) extend the
//    single IntegratorConfig::pressure and the single interval.
//  * hpo_opts::reactor = 1 integrates each cell at constant volume (the reference's CONV
//    option, header.h:6-7, whose pyJac branch does not compile): the cell keeps the density it
//    has at the configured (or per-cell) pressure, rho = P Wbar / (R T), and its pressure
//    follows from the state.
#pragma once

#include <cstdint>
#include <vector>

#include "Integrator.hpp"
#include "hpc_ode.h"

class BDFIntegratorHIP : public Integrator {
 public:
  // opts: rtol/atol are taken from IntegratorConfig at init; the remaining fields (LU
  // precision, refinement sweeps, launch shape, max steps) from here. nullptr = defaults.
  explicit BDFIntegratorHIP(const hpo_opts* opts = nullptr);
  ~BDFIntegratorHIP() override;
  BDFIntegratorHIP(const BDFIntegratorHIP&) = delete;
  BDFIntegratorHIP& operator=(const BDFIntegratorHIP&) = delete;

  // Loads the mechanism, checks that its species count equals the mesh's (throws
  // std::runtime_error otherwise, as CVodeIntegrator.cpp:41-48) and allocates device memory.
  void init(std::shared_ptr<Context> ctx, IntegratorConfig config) override;
  void integrate(double t0, double t) override;
  void clean() override;

  void setCellPressure(std::vector<double> P);    // Pa, one per cell; empty = config pressure
  void setCellTimeStep(std::vector<double> dt);   // s, one per cell; empty = t - t0
  const std::vector<int32_t>& cellStatus() const { return status_; }   // HPO_CELL_* per cell
  const std::vector<int32_t>& cellStats() const { return stats_; }     // [cells][HPO_NSTATS]
  const hpo_timing& lastTiming() const { return timing_; }
  const hpo_opts& options() const { return opts_; }

  // Resolves a mechanism name: an existing path is used as is; otherwise the name is looked
  // up in $HPO_MECHANISM_DIR, then in the mechanisms folder installed beside the library.
  static std::string resolveMechanism(const std::string& name);

 private:
  void release();
  hpo_opts opts_;
  hpo_mech* mech_ = nullptr;
  hpo_batch* batch_ = nullptr;
  std::vector<double> y_;          // [cells][systemSize] staging rows y = [T, Y_0 .. Y_{nsp-2}]
  std::vector<double> P_, dt_;
  std::vector<double> W_, rho_;    // molecular weights; per-cell densities of the constant-volume reactor
  std::vector<int32_t> status_, stats_;
  hpo_timing timing_{};
};
