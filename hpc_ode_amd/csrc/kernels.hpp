// Kernel argument blocks and host launch entry points of the HIP path.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "mechanism.hpp"

// per-cell status codes, CVODE's return-code values (cvode.h)
#define HPO_SUCCESS 0
#define HPO_TOO_MUCH_WORK (-1)
#define HPO_TOO_MUCH_ACC (-2)
#define HPO_ERR_FAILURE (-3)
#define HPO_CONV_FAILURE (-4)
#define HPO_ILL_INPUT (-22)
#define HPO_TOO_CLOSE (-27)

#define HPO_NSTATS 8  // nst, nfe, nje, nsetups, nni, ncfn, netf, q_last
#define HPO_NPROF 5   // cycles in: RHS, Jacobian, LU factorization, LU solves, whole cell

namespace hpo {

struct BdfArgs {
  const uint8_t* mech;       // device table blob
  DevMechLayout lay;
  const double* y_in;        // [ncells][ld] state rows y = [T, Y_0..Y_{K-2}]
  double* y_out;             // [ncells][ld]
  const double* P;           // [ncells] Pa
  const double* dt;          // [ncells] s
  int32_t* status;           // [ncells]
  int32_t* stats;            // [ncells][HPO_NSTATS] or null
  const int32_t* order;      // work-queue order of cells, or null for 0..ncells-1
  uint32_t* queue;           // work-queue counter (zero before launch)
  double* scratch;           // per-wave scratch
  int64_t scratch_stride;    // doubles per wave slot
  int64_t ncells;
  int64_t ld;
  double rtol, atol;
  int32_t maxsteps;
  int32_t refine_sweeps;
  int64_t* prof;             // [ncells][HPO_NPROF] shader-clock cycles per phase, or null
  BlkLayout blay;            // block-per-cell kernel: layout of the block tables at `mech`
  int32_t reactor;           // 0 constant pressure (P = pressure, Pa); 1 constant volume (P = density, kg/m^3)
};

// Solver failures a cell may recover from in FP64: convergence and error-test failures, too
// much work or too much accuracy asked. Bad input (ILL_INPUT, TOO_CLOSE) is not retried.
__host__ __device__ inline bool hpo_fallback_status(int s) {
  return s == HPO_CONV_FAILURE || s == HPO_ERR_FAILURE || s == HPO_TOO_MUCH_WORK || s == HPO_TOO_MUCH_ACC;
}

// The work queue of a launch, four words: [0] the cell counter of a plain launch, [1] the cells
// an FP32 launch left failed (hpo_fallback_status), counted by count_failed after it, [2] the
// cell counter of the FP64 launch that retries them, [3] nonzero when that FP64 fallback is
// armed. Armed, an FP64 launch integrates only the failed cells, from the input rows it is
// given, one per pull from [2]; it does nothing when [1] is 0. The kernels read word 3 from
// memory where they need it rather than as an argument, so that no value of it stays live
// across a cell's integration.

struct EvalArgs {
  const uint8_t* mech;
  DevMechLayout lay;
  const double* y;           // [ncells][ld]
  const double* P;           // [ncells]
  double* out;               // mode 0: [ncells][ld] f; mode 1: [ncells][NS][NS] row-major J
  double* scratch;
  int64_t scratch_stride;
  int64_t ncells;
  int64_t ld;
  int32_t mode;
  BlkLayout blay;            // block-per-cell component kernels: layout of the block tables
  int32_t reactor;           // as BdfArgs::reactor
};

struct LuArgs {
  const double* M;           // [ncases][NS][NS] row-major
  const double* b;           // [ncases][NS]
  double* x;                 // [ncases][NS]
  int32_t* ok;               // [ncases]
};

bool kernel_supports(int ns);
int bdf_wave_lds_doubles(const DevMechLayout& L);
int eval_wave_lds_doubles(const DevMechLayout& L);
int64_t wave_scratch_doubles_host(const DevMechLayout& L);
// variant: 0 FP64 LU; 1 FP32 explicit inverse (Gauss-Jordan); 2 FP32 in-place LU
const void* bdf_kernel_fn(int ns, int variant, int prof = 0, int conv = 0);
hipError_t launch_bdf(int ns, const BdfArgs& A, int variant, int grid, int block, size_t shmem, hipStream_t st);
hipError_t launch_reset_counter(uint32_t* q, bool fallback_armed, hipStream_t st);
// q[1] += the cells of status[0..n) that ended in a solver failure an FP64 run may recover from
hipError_t launch_count_failed(const int32_t* status, int64_t n, uint32_t* q, hipStream_t st);
// Device-built work-queue order (queue_order.hip): order[0..n) = the cells sorted by descending
// initial temperature y[i*ld], by buckets of kOrderBuckets over [kOrderTmin, kOrderTmax) K.
// work: kOrderBuckets words of device memory.
constexpr int kOrderBuckets = 1024;
constexpr double kOrderTmin = 200.0, kOrderTmax = 4000.0;
hipError_t launch_t0_order(const double* y, int64_t n, int64_t ld, uint32_t* work, int32_t* order, hipStream_t st);
// The no-allocation entry point's fixed scratch: 192 MiB of static device memory per GPU (the
// counter in the first 256 bytes, then the wave slots), enough for a whole-device launch of a
// 53-variable FP32 kernel at two waves per SIMD. noalloc_scratch returns its address on the
// current device (hipGetSymbolAddress on the first call for that device, cached after).
constexpr size_t kNoallocScratchBytes = size_t(192) << 20;
constexpr int kNoallocMaxDevices = 64;
hipError_t noalloc_scratch(int device, void** ptr);
inline int bdf_variant(int lu_precision, int lu_layout) { return lu_precision == 0 ? 0 : (lu_layout == 0 ? 1 : 2); }
hipError_t launch_eval(int ns, const EvalArgs& A, int grid, int block, size_t shmem, hipStream_t st);
// Block-per-cell kernel (block_kernel.hip): one 256-thread workgroup per cell.
bool block_kernel_supports(int ns);
int blk_work_lds_bytes(const BlkLayout& L);           // LDS after the tables, per workgroup
int blk_work_lds_bytes_for(int K, int R);             // the same before the tables are built
int64_t blk_scratch_doubles_host(const BlkLayout& L); // global scratch per workgroup
const void* bdf_block_kernel_fn(int ns, int lu32, int prof);
hipError_t launch_bdf_block(int ns, const BdfArgs& A, int lu32, int grid, size_t shmem, hipStream_t st);
// the block kernel's RHS (mode 0) or Jacobian (mode 1) on one workgroup per cell; A.mech holds
// the block tables, A.scratch grid * blk_scratch_doubles_host doubles
hipError_t launch_eval_block(int ns, const EvalArgs& A, int grid, size_t shmem, hipStream_t st);
// the block kernel's Newton-matrix code: Gauss-Jordan inverse across one workgroup per matrix,
// applied to b (FP64, or FP32 without refinement)
hipError_t launch_gj_block(int ns, int fp32, const LuArgs& A, int64_t ncases, hipStream_t st);
hipError_t launch_lu(int ns, int fp32, const LuArgs& A, int64_t ncases, hipStream_t st);
// one 256-thread workgroup per matrix, FP64 (lu_block.hip)
hipError_t launch_lu_block(int ns, const LuArgs& A, int64_t ncases, hipStream_t st);

}  // namespace hpo
