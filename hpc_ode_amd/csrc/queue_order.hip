// Work-queue order by descending initial temperature, built on the device: a counting sort of
// the cells into temperature buckets. The persistent kernel pulls cells in this order, so the
// hot cells (the most BDF steps) start first and the queue's tail holds short ones. Used by the
// serving loop, whose launches pack cells already resident on the device (the batch entry
// point sorts on the host). Within a bucket the order is whatever the atomics give: it changes
// which wave integrates a cell, never the cell's result.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <algorithm>

#include "kernels.hpp"

namespace hpo {
namespace dev {

__device__ inline int t0_bucket(double T) {
  // descending: the hottest bucket first; NaN and out-of-range temperatures go to the ends
  const double x = (kOrderTmax - T) * (kOrderBuckets / (kOrderTmax - kOrderTmin));
  if (!(x == x)) return kOrderBuckets - 1;
  return static_cast<int>(fmin(fmax(x, 0.0), static_cast<double>(kOrderBuckets - 1)));
}

__global__ __launch_bounds__(1024) void order_clear(uint32_t* count) {
  if (threadIdx.x < kOrderBuckets) count[threadIdx.x] = 0u;
}

// Bucket sizes: a histogram in LDS per workgroup, added to the global one.
__global__ __launch_bounds__(256) void order_hist(const double* y, int64_t n, int64_t ld, uint32_t* count) {
  __shared__ uint32_t h[kOrderBuckets];
  for (int b = threadIdx.x; b < kOrderBuckets; b += blockDim.x) h[b] = 0u;
  __syncthreads();
  const int64_t stride = static_cast<int64_t>(blockDim.x) * gridDim.x;
  for (int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride)
    atomicAdd(&h[t0_bucket(y[i * ld])], 1u);
  __syncthreads();
  for (int b = threadIdx.x; b < kOrderBuckets; b += blockDim.x)
    if (h[b]) atomicAdd(count + b, h[b]);
}

// Exclusive prefix sum of the bucket sizes in place (one workgroup of kOrderBuckets threads).
__global__ __launch_bounds__(1024) void order_scan(uint32_t* count) {
  __shared__ uint32_t s[kOrderBuckets];
  const int t = threadIdx.x;
  const uint32_t own = count[t];
  s[t] = own;
  __syncthreads();
  for (int d = 1; d < kOrderBuckets; d <<= 1) {
    const uint32_t v = t >= d ? s[t - d] : 0u;
    __syncthreads();
    s[t] += v;
    __syncthreads();
  }
  count[t] = s[t] - own;
}

__global__ __launch_bounds__(256) void order_scatter(const double* y, int64_t n, int64_t ld, uint32_t* next,
                                                     int32_t* order) {
  const int64_t stride = static_cast<int64_t>(blockDim.x) * gridDim.x;
  for (int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride)
    order[atomicAdd(next + t0_bucket(y[i * ld]), 1u)] = static_cast<int32_t>(i);
}

}  // namespace dev

hipError_t launch_t0_order(const double* y, int64_t n, int64_t ld, uint32_t* work, int32_t* order, hipStream_t st) {
  if (n <= 0) return hipSuccess;
  const unsigned blocks = static_cast<unsigned>(std::min<int64_t>((n + 255) / 256, 256));
  hipLaunchKernelGGL(dev::order_clear, dim3(1), dim3(kOrderBuckets), 0, st, work);
  hipLaunchKernelGGL(dev::order_hist, dim3(blocks), dim3(256), 0, st, y, n, ld, work);
  hipLaunchKernelGGL(dev::order_scan, dim3(1), dim3(kOrderBuckets), 0, st, work);
  hipLaunchKernelGGL(dev::order_scatter, dim3(blocks), dim3(256), 0, st, y, n, ld, work, order);
  return hipGetLastError();
}

}  // namespace hpo
