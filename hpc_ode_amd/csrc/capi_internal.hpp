// Internals shared by the C ABI implementation files (capi.cpp, server.cpp): error
// plumbing, device allocation, the mechanism handle and the launch of the BDF kernel.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <mutex>
#include <stdexcept>
#include <string>

#include "../../include/hpc_ode.h"
#include "kernels.hpp"
#include "mechanism.hpp"

namespace hpo {
namespace capi {

int fail(int code, const std::string& msg);
int map_exceptions();   // call inside a catch(...) block

struct HipError : std::runtime_error {
  hipError_t e;
  HipError(hipError_t err, const std::string& what) : std::runtime_error(what), e(err) {}
};

#define HPO_HIP(call)                                                                               \
  do {                                                                                              \
    hipError_t _e = (call);                                                                         \
    if (_e != hipSuccess)                                                                           \
      throw ::hpo::capi::HipError(_e, std::string(#call) + ": " + hipGetErrorString(_e));           \
  } while (0)

template <typename T>
T* dalloc(size_t n) {
  void* p = nullptr;
  HPO_HIP(hipMalloc(&p, std::max<size_t>(n, 1) * sizeof(T)));
  return static_cast<T*>(p);
}

// Work-queue counter and per-wave (wave layout) or per-workgroup (block layout) scratch of one
// launch context. Launches that may run at the same time need separate contexts; a mechanism
// handle owns one for its own entry points.
struct ExecRes {
  uint32_t* d_queue = nullptr;
  double* d_scratch = nullptr;
  int64_t scratch_doubles = 0;  // capacity
  int32_t share_pct = 100;      // largest share of n_CU x workgroups-per-CU the grid may take
  // copy of the input rows of an in-place launch, for its FP64 fallback (null: none)
  double* d_backup = nullptr;
  int64_t backup_doubles = 0;   // capacity
};

// Which kernel layout a launch of this mechanism with these options uses: true = one workgroup
// per cell (block_kernel.hip), false = one wave per cell.
bool use_block_layout(const hpo_mech* m, const hpo_opts& o);

void check_opts(const hpo_opts& o);

// Launches the persistent BDF kernel on arrays already resident on the device. With res ==
// nullptr the mechanism's own context is used (and grown as needed); an explicit context must
// hold the grid's wave slots already. With tm != nullptr the call waits and reports timing.
void launch_integrate(hpo_mech* m, int64_t ncells, const double* d_yin, double* d_yout, int64_t ld,
                      const double* d_P, const double* d_dt, int32_t* d_status, int32_t* d_stats,
                      const int32_t* d_order, int64_t* d_prof, const hpo_opts& o, hipStream_t st,
                      hpo_timing* tm, const ExecRes* res = nullptr);

// Scratch doubles the largest launch on the whole device can need, in either layout (sizing
// of explicit contexts).
int64_t max_scratch_doubles(hpo_mech* m);

}  // namespace capi
}  // namespace hpo

struct hpo_mech {
  hpo::Mechanism host;
  hpo::DevMechBlob blob;        // wave-per-cell tables (has_wave)
  hpo::DevBlkBlob bblob;        // block-per-cell tables (has_block)
  bool has_wave = false, has_block = false;
  int device = -1;
  int n_cu = 0;
  uint8_t* d_blob = nullptr;
  uint8_t* d_bblob = nullptr;
  double* d_scratch = nullptr;
  int64_t scratch_cap = 0;      // doubles
  int64_t scratch_stride = 0;   // doubles per wave slot of the wave layout
  uint32_t* d_queue = nullptr;
  double* d_backup = nullptr;   // input rows of in-place launches, for the FP64 fallback
  int64_t backup_cap = 0;       // doubles
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  // Launches on the handle's own queue counter and scratch are ordered on the device: each
  // waits for the previous one (ev_last, recorded on last_stream) when it is issued on another
  // stream, and host threads take launch_mu around the issue.
  std::mutex launch_mu;
  hipEvent_t ev_last = nullptr;
  hipStream_t last_stream = nullptr;
  bool launched = false;

  ~hpo_mech();
  void ensure_device();               // uploads the tables on first use (current device)
  void ensure_scratch(int64_t slots);  // wave slots of the wave layout
  void ensure_scratch_doubles(int64_t n);
  void ensure_backup_doubles(int64_t n);
  int state_size() const { return host.K(); }
};
