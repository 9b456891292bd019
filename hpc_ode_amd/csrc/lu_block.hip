// Batched dense solve M x = b with one 256-thread workgroup per matrix: the block-per-matrix
// layout of the batched LU that the reference's GPU path takes from MAGMA
// (magma_dgetrf_batched / magma_dgetrs_batched through SUNLinSol_MagmaDense,
// src/Integrators/CVodeIntegratorGPU.cpp:86-91).
//
// The augmented matrix [M | b] lives in LDS, one row of NS + 1 doubles per matrix row. Step k
// of the elimination:
//   1. wave 0 finds the pivot row p >= k of column k (largest |a|, the first such row on ties,
//      as LAPACK's idamax; the BDF kernel's register LU uses the same 64-bit key);
//   2. rows k and p swap their columns k..NS (b included);
//   3. every row i > k subtracts l_i = a[i][k] / a[k][k] times row k from its columns k+1..NS,
//      the 256 threads spread as 4 row groups x 64 columns.
// Column NS then holds the forward-substituted right-hand side, and wave 0 back-substitutes
// with row i of U in the registers of lane i. The arithmetic is SUNDIALS' denseGETRF/denseGETRS
// in the same order (mult = 1 / pivot, l_i = a[i][k] * mult, col_j -= a[k][j] * l_i;
// b[k] /= U[k][k] then b[i] -= U[i][k] b[k]).
//
// FP64 only. A zero or NaN pivot marks the matrix singular (ok = 0); the elimination then
// continues with a zero multiplier so every thread reaches every barrier.
#include <hip/hip_runtime.h>

#include "bdf_device.hpp"
#include "kernels.hpp"

namespace hpo {
namespace dev {

template <int NS>
__global__ __launch_bounds__(256) void lu_block_kernel(LuArgs A) {
  static_assert(NS >= 1 && NS <= 64, "one wave holds a pivot column and a solution vector");
  constexpr int LD = NS + 1;  // row stride: the NS columns of M and b
  __shared__ double a[NS * LD];
  __shared__ int s_piv;
  __shared__ int s_bad;
  const int64_t c = blockIdx.x;
  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int grp = tid >> 6;  // row group 0..3
  const double* M = A.M + c * NS * NS;
  for (int e = tid; e < NS * NS; e += 256) a[(e / NS) * LD + (e % NS)] = M[e];
  for (int i = tid; i < NS; i += 256) a[i * LD + NS] = A.b[c * NS + i];
  if (tid == 0) s_bad = 0;
  __syncthreads();

  for (int k = 0; k < NS; ++k) {
    if (tid < 64) {
      // rows k..NS-1 are the candidates; lanes past the last row hold key 0
      const int r = k + lane;
      const double key = (r < NS) ? pivot_key(a[r * LD + k], lane) : 0.0;
      const double kmax = wmax(key);
      if (lane == 0) {
        const int p = k + key_pos(kmax);
        s_piv = (p < NS) ? p : k;
        const double pv = a[s_piv * LD + k];
        if (!(pv != 0.0) || pv != pv) s_bad = 1;
      }
    }
    __syncthreads();
    const int p = s_piv;
    if (p != k) {
      for (int j = k + tid; j <= NS; j += 256) {
        const double t = a[k * LD + j];
        a[k * LD + j] = a[p * LD + j];
        a[p * LD + j] = t;
      }
    }
    __syncthreads();
    const double piv = a[k * LD + k];
    const double mult = (piv != 0.0 && piv == piv) ? 1.0 / piv : 0.0;
    const int j = k + 1 + lane;  // this thread's column of the trailing block (NS = b)
    if (j <= NS) {
      const double akj = a[k * LD + j];
      for (int i = k + 1 + grp; i < NS; i += 4) {
        const double l = a[i * LD + k] * mult;
        a[i * LD + j] = fma(-akj, l, a[i * LD + j]);
      }
    }
    __syncthreads();
  }

  if (tid < 64) {
    // back substitution: lane i holds row i of U and y_i
    double u[NS];
#pragma unroll
    for (int q = 0; q < NS; ++q) u[q] = (lane < NS) ? a[lane * LD + q] : 0.0;
    double y = (lane < NS) ? a[lane * LD + NS] : 0.0;
#pragma unroll
    for (int q = NS - 1; q >= 0; --q) {
      const double xq = rl(y, q) / rl(u[q], q);
      y = (lane == q) ? xq : ((lane < q) ? fma(-u[q], xq, y) : y);
    }
    if (lane < NS) A.x[c * NS + lane] = y;
    if (lane == 0) A.ok[c] = s_bad ? 0 : 1;
  }
}

}  // namespace dev

hipError_t launch_lu_block(int ns, const LuArgs& A, int64_t ncases, hipStream_t st) {
#define HPO_CASE(N)                                                                                  \
  if (ns == N) {                                                                                     \
    hipLaunchKernelGGL((dev::lu_block_kernel<N>), dim3(static_cast<unsigned>(ncases)), dim3(256), 0, st, A); \
    return hipGetLastError();                                                                        \
  }
  HPO_CASE(9)
  HPO_CASE(10)
  HPO_CASE(53)
#undef HPO_CASE
  return hipErrorInvalidValue;
}

}  // namespace hpo
