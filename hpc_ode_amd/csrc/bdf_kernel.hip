// Persistent per-cell BDF kernel for gfx950.
//
// Each wave64 pulls cells from a device work queue and integrates one cell at a
// time from t = 0 to its dt with CVODE's variable-order BDF (the solver the reference
// configures at src/Integrators/CVodeIntegrator.cpp:98-155: BDF, scalar tolerances,
// dense direct linear solver, analytic Jacobian, maxsteps, one CV_NORMAL call).
// The whole integration is resident: Nordsieck history, error weights and the
// Newton matrix row of lane l stay in registers; the RHS is evaluated at one code
// site driven by a small state machine so the kernel is not replicated per caller.
#include <hip/hip_runtime.h>
#include <float.h>

#include <mutex>

#include "bdf_cell.hpp"

// Occupancy targets of the BDF kernel (waves per SIMD). At 1 a wave may use the whole
// 512-entry unified VGPR/AGPR file: the FP64 path keeps its 106-register Newton-matrix row,
// the Nordsieck history and the RHS working set there without scratch spills. The FP32
// path holds a 53-register row and fits 256 registers, so it runs two waves per SIMD (a few
// dozen bytes per lane spill, where the RHS and solve working sets peak together).
#ifndef HPO_WAVES_PER_EU_FP64
#define HPO_WAVES_PER_EU_FP64 1
#endif
// Most loads of the refinement's FP64 Jacobian row issued together. Each group costs one
// round trip to the (cache-missing) Jacobian, so fewer, larger groups wait less, at two
// registers per load in flight. GRI's 53 columns as two groups of 27 measured fastest on
// 262,144 cells (1,262 ms against 1,292 ms for seven groups of 8); one group of 53 spills.
#ifndef HPO_JX_GROUP
#define HPO_JX_GROUP 27
#endif
#ifndef HPO_WAVES_PER_EU_FP32
#define HPO_WAVES_PER_EU_FP32 2
#endif

namespace hpo {
namespace dev {

// The refinement's J x product reads its NS columns in jx_groups() equal groups of
// jx_group() loads, NS rounded up to jx_cols(). The padding columns lie inside the wave's
// 64 x 64 Jacobian block, after the NS x NS entries the Jacobian writes, and are zeroed at
// kernel entry.
template <int NS>
constexpr int jx_groups() {
  return (NS + HPO_JX_GROUP - 1) / HPO_JX_GROUP;
}
template <int NS>
constexpr int jx_group() {
  return (NS + jx_groups<NS>() - 1) / jx_groups<NS>();
}
template <int NS>
constexpr int jx_cols() {
  return jx_groups<NS>() * jx_group<NS>();
}

template <int NS, bool LU32>
struct LuStore;
template <int NS>
struct LuStore<NS, false> {
  LuReg<NS, double> L;
};
template <int NS>
struct LuStore<NS, true> {
  LuReg<NS, float> L;
};

// One wave per cell: lane l owns state component l, the matching Newton-matrix row (in
// registers) and the row of the Jacobian (in the wave's scratch).
template <int NS, bool LU32, bool INV>
struct WaveLayout {
  const Tabs& t;
  const WaveLds& w;
  const WaveScratch& g;
  const int tid;         // lane = owned component
  const int wave_lane;   // the same (one wave per cell)
  const double* exptab;
  LuStore<NS, LU32> lu;

  HPO_DI WaveLayout(const Tabs& t_, const WaveLds& w_, const WaveScratch& g_, int lane)
      : t(t_), w(w_), g(g_), tid(lane), wave_lane(lane), exptab(t_.exptab) {}
  HPO_DI bool any(bool c) const { return __ballot(c) != 0ull; }
  HPO_DI double sum(double v) const { return wsum(v); }
  HPO_DI double max(double v) const { return wmax(v); }
  HPO_DI double rhs(double y, double P) const { return dev::rhs(t, w, y, P, tid, NS); }

  // cvLsSetup / cvLsLinSys: a new Jacobian when jnew, then M = I - gamma J factored (FP64 LU)
  // or inverted (FP32 Gauss-Jordan) or factored in FP32. cj: cycles of the Jacobian (PROF).
  template <bool PROF>
  HPO_DI bool setup(double y, double P, double gamma, bool jnew, long long& cj) {
    const int ol = opaque(tid);
    // J's columns are packed at the state length: lanes >= NS read entries of other rows
    // (or never-written scratch, still inside the 64 x 64 J block). Their Newton-matrix rows
    // are never pivot candidates and are zeroed when the factors are permuted, and the
    // refinement's products on those lanes are discarded, so nothing they hold reaches the
    // live lanes.
    const double* Jrow = g.J + ol;
    if (jnew) {
      const long long c0 = PROF ? clock64() : 0;
      jacobian<NS>(t, w, g, y, P, tid);
      if (PROF) cj = clock64() - c0;
    }
    if constexpr (LU32) {
#pragma unroll
      for (int j = 0; j < NS; ++j) {
        double Jv = Jrow[j * NS];
        lu.L.a[j] = static_cast<float>(-gamma * Jv + ((j == ol) ? 1.0 : 0.0));
      }
    } else {
#pragma unroll
      for (int j = 0; j < NS; ++j) lu.L.a[j] = -gamma * Jrow[j * NS] + ((j == ol) ? 1.0 : 0.0);
    }
    if constexpr (INV)
      return gj_invert<NS>(lu.L, tid);
    else
      return lu_factor<NS>(lu.L, tid);
  }

  // x = M^-1 b: FP64 triangular solves, or the FP32 solve followed by `sweeps` FP64 residual
  // sweeps r = b - (I - gamma J) x against the FP64 Jacobian.
  HPO_DI double solve(double b, double gl, int sweeps) {
    if constexpr (LU32) {
      const double* Jrow = g.J + opaque(tid);
      float* vbuf = reinterpret_cast<float*>(w.cp);
      float x32 = solve32<NS, INV>(lu.L, static_cast<float>(b), tid, vbuf);
      double x = static_cast<double>(x32);
      for (int sw = 0; sw < sweeps; ++sw) {
        // r = b - (I - gamma J) x in FP64, with the FP64 Jacobian row kept in this
        // wave's scratch (cache-resident) rather than in registers. x is broadcast through
        // per-wave LDS (the enthalpy slots, which only the Jacobian uses, after writing them)
        wave_sync();
        w.hRT[tid] = x;
        wave_sync();
        // Whole groups of JG loads are issued before their products, one round trip per
        // group. Columns NS..JN-1 are zero padding (set at kernel entry) and x is zero on
        // lanes >= NS, so the padded products add nothing.
        constexpr int JG = jx_group<NS>();
        double Jx = 0.0;
#pragma unroll 1
        for (int j0 = 0; j0 < jx_cols<NS>(); j0 += JG) {
          double v[JG];
#pragma unroll
          for (int k = 0; k < JG; ++k) v[k] = Jrow[(j0 + k) * NS];
#pragma unroll
          for (int k = 0; k < JG; ++k) Jx = fma(v[k], w.hRT[j0 + k], Jx);
        }
        double r = b - x + gl * Jx;
        float d32 = solve32<NS, INV>(lu.L, static_cast<float>(r), tid, vbuf);
        x += static_cast<double>(d32);
      }
      return x;
    } else {
      return lu_solve<NS>(lu.L, b, tid);
    }
  }
};

// Small mechanisms (at most 16 state variables: a 16-register Newton row, one or two reaction
// passes) fit several waves per SIMD, which hides the latency one wave exposes. On 262,144
// H2/air cells (second substep, 17 steps per cell): 28.8 ms FP32 / 29.8 ms FP64 at two waves,
// 27.1 / 29.7 ms at three (168 registers), with identical results.
//
// Two is the default. Three-wave FP64 builds of the small kernels have twice returned a Newton
// convergence failure for every cell after a change to the cell loop that leaves the work of a
// cell unchanged: once with a never-taken per-workgroup exit test before each queue pull, and
// once (the constant-volume 9-variable kernel) with the FP64-fallback queue of this build. The
// first was thought to come from an inline-asm v_writelane (now a compare and select, wli); the
// second shows that was not the whole cause, and none has been shown. The two-wave builds of
// the same sources match the oracle.
#ifndef HPO_WAVES_PER_EU_SMALL
#define HPO_WAVES_PER_EU_SMALL 2
#endif
template <int NS, bool LU32>
constexpr int bdf_waves_per_eu() {
  return NS <= 16 ? HPO_WAVES_PER_EU_SMALL : (LU32 ? HPO_WAVES_PER_EU_FP32 : HPO_WAVES_PER_EU_FP64);
}

// CONV: the reactor as a compile-time constant (constant volume), so the constant-pressure
// instantiation carries none of the other reactor's selects.
template <int NS, bool LU32, bool PROF, bool INV, bool CONV>
__global__ __launch_bounds__(256) __attribute__((
    amdgpu_waves_per_eu(bdf_waves_per_eu<NS, LU32>(), bdf_waves_per_eu<NS, LU32>()))) void bdf_kernel(BdfArgs A) {
  extern __shared__ __align__(16) uint8_t smem[];
  const DevMechLayout& Lay = A.lay;
  for (int i = threadIdx.x * 16; i < Lay.lds_bytes; i += blockDim.x * 16)
    *reinterpret_cast<uint4*>(smem + i) = *reinterpret_cast<const uint4*>(A.mech + i);
  __syncthreads();
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int wpb = blockDim.x >> 6;
  Tabs t = make_tabs(smem, A.mech, Lay);
  t.conv = CONV ? 1 : 0;
  const int wl_doubles = wave_lds_doubles(Lay.R, Lay.n_partials) + kScalDoubles;
  double* wbase = reinterpret_cast<double*>(smem + Lay.lds_bytes) + wave * wl_doubles;
  WaveLds w = make_wave_lds(wbase, Lay.R, Lay.n_partials);
  Scal& S = *reinterpret_cast<Scal*>(wbase + wave_lds_doubles(Lay.R, Lay.n_partials));
  const int64_t slot = static_cast<int64_t>(blockIdx.x) * wpb + wave;
  WaveScratch g = make_wave_scratch(A.scratch + slot * A.scratch_stride, Lay);
  if (lane < 16) g.coefbuf[lane] = t.coef[lane];
  static_assert((jx_cols<NS>() - 1) * NS + 64 <= 64 * 64, "padded J x columns leave the Jacobian block");
  for (int i = NS * NS + lane; i < (jx_cols<NS>() - 1) * NS + 64; i += 64) g.J[i] = 0.0;
  wave_sync();
  // An FP64 launch with the fallback armed takes only the cells that the FP32 launch before it
  // left failed, one per pull from queue[2], from their input rows
  if (!LU32 && hpo_fallback_armed(A.queue) &&
      __builtin_amdgcn_readfirstlane(static_cast<int>(*reinterpret_cast<volatile uint32_t*>(A.queue + 1))) == 0)
    return;
  for (;;) {
    const bool fb = !LU32 && hpo_fallback_armed(A.queue);
    unsigned int c = 0;
    if (lane == 0) c = atomicAdd(A.queue + (fb ? 2 : 0), 1u);
    c = __builtin_amdgcn_readfirstlane(c);
    if (static_cast<int64_t>(c) >= A.ncells) break;
    const int64_t cell = (!fb && A.order) ? static_cast<int64_t>(A.order[c]) : static_cast<int64_t>(c);
    if (fb && !hpo_fallback_status(__builtin_amdgcn_readfirstlane(A.status[cell]))) continue;
    WaveLayout<NS, LU32, INV> Ly(t, w, g, lane);
    integrate_cell<NS, LU32, PROF>(A, Ly, S, cell);
  }
}

// ---- component kernels used by the parity tests ----
template <int NS, int MODE>
__global__ __launch_bounds__(256) void eval_kernel(EvalArgs A) {
  extern __shared__ __align__(16) uint8_t smem[];
  const DevMechLayout& Lay = A.lay;
  for (int i = threadIdx.x * 16; i < Lay.lds_bytes; i += blockDim.x * 16)
    *reinterpret_cast<uint4*>(smem + i) = *reinterpret_cast<const uint4*>(A.mech + i);
  __syncthreads();
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, wpb = blockDim.x >> 6;
  Tabs t = make_tabs(smem, A.mech, Lay);
  t.conv = A.reactor;
  const int wl_doubles = wave_lds_doubles(Lay.R, Lay.n_partials);
  WaveLds w = make_wave_lds(reinterpret_cast<double*>(smem + Lay.lds_bytes) + wave * wl_doubles, Lay.R, Lay.n_partials);
  const int64_t slot = static_cast<int64_t>(blockIdx.x) * wpb + wave;
  WaveScratch g = make_wave_scratch(A.scratch + slot * A.scratch_stride, Lay);
  if (lane < 16) g.coefbuf[lane] = t.coef[lane];
  for (int64_t c = slot; c < A.ncells; c += static_cast<int64_t>(gridDim.x) * wpb) {
    double y = lane < NS ? A.y[c * A.ld + lane] : 0.0;
    double P = A.P[c];
    if constexpr (MODE == 0) {
      double f = rhs(t, w, y, P, lane, NS);
      if (lane < NS) A.out[c * A.ld + lane] = f;
    } else {
      jacobian<NS>(t, w, g, y, P, lane);
      if (lane < NS)
        for (int j = 0; j < NS; ++j) A.out[(c * NS + lane) * NS + j] = g.J[j * NS + lane];
    }
  }
}

template <int NS, typename F>
__global__ __launch_bounds__(64) void lu_kernel(LuArgs A) {
  const int lane = threadIdx.x & 63;
  const int64_t c = blockIdx.x;
  LuReg<NS, F> L;
#pragma unroll
  for (int j = 0; j < NS; ++j) L.a[j] = lane < NS ? static_cast<F>(A.M[(c * NS + lane) * NS + j]) : F(0);
  bool ok;
  F x;
  F b = lane < NS ? static_cast<F>(A.b[c * NS + lane]) : F(0);
  if constexpr (sizeof(F) == 4) {
    __shared__ float buf[64];
    ok = gj_invert<NS>(L, lane);
    x = gj_apply<NS>(L, b, lane, buf);
  } else {
    ok = lu_factor<NS>(L, lane);
    x = lu_solve<NS>(L, b, lane);
  }
  if (lane < NS) A.x[c * NS + lane] = static_cast<double>(x);
  if (lane == 0) A.ok[c] = ok ? 1 : 0;
}

}  // namespace dev

// ---------------------------------------------------------------------------
// launch wrappers (host)
// ---------------------------------------------------------------------------
// Kernel variants: 0 FP64 LU; 1 FP32 explicit inverse (Gauss-Jordan); 2 FP32 in-place LU.
template <int NS, bool PROF, bool CONV>
static const void* bdf_fn_p(int variant) {
  if (variant == 1) return reinterpret_cast<const void*>(&dev::bdf_kernel<NS, true, PROF, true, CONV>);
  if (variant == 2) return reinterpret_cast<const void*>(&dev::bdf_kernel<NS, true, PROF, false, CONV>);
  return reinterpret_cast<const void*>(&dev::bdf_kernel<NS, false, PROF, false, CONV>);
}

template <int NS>
static const void* bdf_fn(int variant, int prof, int conv) {
  if (conv) return prof ? bdf_fn_p<NS, true, true>(variant) : bdf_fn_p<NS, false, true>(variant);
  return prof ? bdf_fn_p<NS, true, false>(variant) : bdf_fn_p<NS, false, false>(variant);
}

template <int NS, bool PROF, bool CONV>
static void launch_bdf_p(const BdfArgs& A, int variant, int grid, int block, size_t shmem, hipStream_t st) {
  if (variant == 1)
    hipLaunchKernelGGL((dev::bdf_kernel<NS, true, PROF, true, CONV>), dim3(grid), dim3(block), shmem, st, A);
  else if (variant == 2)
    hipLaunchKernelGGL((dev::bdf_kernel<NS, true, PROF, false, CONV>), dim3(grid), dim3(block), shmem, st, A);
  else
    hipLaunchKernelGGL((dev::bdf_kernel<NS, false, PROF, false, CONV>), dim3(grid), dim3(block), shmem, st, A);
}

template <int NS>
static hipError_t launch_bdf_ns(const BdfArgs& A, int variant, int grid, int block, size_t shmem, hipStream_t st) {
  const bool prof = A.prof != nullptr, conv = A.reactor == 1;
  if (prof && conv)
    launch_bdf_p<NS, true, true>(A, variant, grid, block, shmem, st);
  else if (prof)
    launch_bdf_p<NS, true, false>(A, variant, grid, block, shmem, st);
  else if (conv)
    launch_bdf_p<NS, false, true>(A, variant, grid, block, shmem, st);
  else
    launch_bdf_p<NS, false, false>(A, variant, grid, block, shmem, st);
  return hipGetLastError();
}

#define HPO_FOR_EACH_NS(X) X(9) X(10) X(53)

int bdf_wave_lds_doubles(const DevMechLayout& L) { return dev::wave_lds_doubles(L.R, L.n_partials) + dev::kScalDoubles; }
int eval_wave_lds_doubles(const DevMechLayout& L) { return dev::wave_lds_doubles(L.R, L.n_partials); }
int64_t wave_scratch_doubles_host(const DevMechLayout& L) { return dev::wave_scratch_doubles(L); }

bool kernel_supports(int ns) {
#define HPO_CASE(N) \
  if (ns == N) return true;
  HPO_FOR_EACH_NS(HPO_CASE)
#undef HPO_CASE
  return false;
}

const void* bdf_kernel_fn(int ns, int variant, int prof, int conv) {
#define HPO_CASE(N) \
  if (ns == N) return bdf_fn<N>(variant, prof, conv);
  HPO_FOR_EACH_NS(HPO_CASE)
#undef HPO_CASE
  return nullptr;
}

namespace dev {
// Sets a launch's four work-queue words on the launch's stream: the counters to zero and the
// FP64-fallback word. A kernel rather than a memset, so that a caller capturing the launch into
// a HIP graph records only kernel nodes.
__global__ __launch_bounds__(64) void reset_counter(uint32_t* q, uint32_t armed) {
  if (threadIdx.x < 4) q[threadIdx.x] = threadIdx.x == 3 ? armed : 0u;
}
}  // namespace dev

hipError_t launch_reset_counter(uint32_t* q, bool fallback_armed, hipStream_t st) {
  hipLaunchKernelGGL(dev::reset_counter, dim3(1), dim3(64), 0, st, q, fallback_armed ? 1u : 0u);
  return hipGetLastError();
}

namespace dev {
// Block-strided over the cells; every thread of a block runs the same iterations, so the
// ballot sees whole waves.
__global__ __launch_bounds__(256) void count_failed(const int32_t* status, int64_t n, uint32_t* q) {
  const int64_t stride = static_cast<int64_t>(blockDim.x) * gridDim.x;
  for (int64_t base = static_cast<int64_t>(blockIdx.x) * blockDim.x; base < n; base += stride) {
    const int64_t i = base + threadIdx.x;
    const unsigned long long b = __ballot(i < n && hpo_fallback_status(status[i]));
    if ((threadIdx.x & 63) == 0 && b) atomicAdd(q + 1, static_cast<uint32_t>(__popcll(b)));
  }
}
}  // namespace dev

hipError_t launch_count_failed(const int32_t* status, int64_t n, uint32_t* q, hipStream_t st) {
  if (n <= 0) return hipSuccess;
  const int64_t blocks = std::min<int64_t>((n + 255) / 256, 1024);
  hipLaunchKernelGGL(dev::count_failed, dim3(static_cast<unsigned>(blocks)), dim3(256), 0, st, status, n, q);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Fixed device scratch of the no-allocation entry point (hpo_integrate_device_noalloc): the
// work-queue counter in its first 256 bytes, then the per-wave slots. One copy per GPU, held
// until the process ends. In host code `scratch` names a host-side stand-in, not device
// memory: only the address hipGetSymbolAddress returns for the current device is used.
// ---------------------------------------------------------------------------
static __device__ __align__(256) double scratch[kNoallocScratchBytes / sizeof(double)];

namespace {
std::mutex g_noalloc_mu;
void* g_noalloc_ptr[kNoallocMaxDevices] = {};
}  // namespace

hipError_t noalloc_scratch(int device, void** ptr) {
  if (device < 0 || device >= kNoallocMaxDevices) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> g(g_noalloc_mu);
  if (!g_noalloc_ptr[device]) {
    void* p = nullptr;
    hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    g_noalloc_ptr[device] = p;
  }
  *ptr = g_noalloc_ptr[device];
  return hipSuccess;
}

hipError_t launch_bdf(int ns, const BdfArgs& A, int variant, int grid, int block, size_t shmem, hipStream_t st) {
#define HPO_CASE(N) \
  if (ns == N) return launch_bdf_ns<N>(A, variant, grid, block, shmem, st);
  HPO_FOR_EACH_NS(HPO_CASE)
#undef HPO_CASE
  return hipErrorInvalidValue;
}

hipError_t launch_eval(int ns, const EvalArgs& A, int grid, int block, size_t shmem, hipStream_t st) {
#define HPO_CASE(N)                                                                         \
  if (ns == N) {                                                                            \
    if (A.mode == 0)                                                                        \
      hipLaunchKernelGGL((dev::eval_kernel<N, 0>), dim3(grid), dim3(block), shmem, st, A);  \
    else                                                                                    \
      hipLaunchKernelGGL((dev::eval_kernel<N, 1>), dim3(grid), dim3(block), shmem, st, A);  \
    return hipGetLastError();                                                               \
  }
  HPO_FOR_EACH_NS(HPO_CASE)
#undef HPO_CASE
  return hipErrorInvalidValue;
}

hipError_t launch_lu(int ns, int fp32, const LuArgs& A, int64_t ncases, hipStream_t st) {
#define HPO_CASE(N)                                                                                        \
  if (ns == N) {                                                                                           \
    if (fp32)                                                                                              \
      hipLaunchKernelGGL((dev::lu_kernel<N, float>), dim3(static_cast<unsigned>(ncases)), dim3(64), 0, st, A);  \
    else                                                                                                   \
      hipLaunchKernelGGL((dev::lu_kernel<N, double>), dim3(static_cast<unsigned>(ncases)), dim3(64), 0, st, A); \
    return hipGetLastError();                                                                              \
  }
  HPO_FOR_EACH_NS(HPO_CASE)
#undef HPO_CASE
  return hipErrorInvalidValue;
}

}  // namespace hpo
