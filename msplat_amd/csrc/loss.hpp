// Squared-error loss of a rendered image against a target, forward and
// adjoint, each one launch: the training step's loss next to the rasterizer
// (the reference's training loops compute it with separate torch kernels,
// tutorials/gs_2d.py:73-90).
#pragma once

#include "common.hpp"

namespace msplat {

constexpr int kLossThreads = 256;
constexpr int kLossBlocks = 1024;  // partial sums: one per workgroup, summed in order by the last one

// partial[b] = sum over the grid-stride slice of workgroup b of (a - t)^2
// (fp32 per thread, a fixed-order workgroup tree); the last workgroup to
// finish (ticket) adds the partials in fp64 in a fixed order and writes the
// total, so the result does not depend on scheduling.  ticket must be 0 on
// entry; it is reset for the next launch.
__global__ __launch_bounds__(kLossThreads) void sq_diff_sum_kernel(const float *__restrict__ a,
                                                                   const float *__restrict__ t, int64_t n,
                                                                   float *__restrict__ partial,
                                                                   uint32_t *__restrict__ ticket,
                                                                   float *__restrict__ out) {
  __shared__ float ws[kLossThreads / kWave];
  __shared__ bool last;
  float acc = 0.f;
  const int64_t stride = int64_t(gridDim.x) * kLossThreads * 4;
  for (int64_t i = (int64_t(blockIdx.x) * kLossThreads + threadIdx.x) * 4; i < n; i += stride) {
    if (i + 4 <= n && (reinterpret_cast<uintptr_t>(a + i) & 15) == 0 && (reinterpret_cast<uintptr_t>(t + i) & 15) == 0) {
      const float4 x = *reinterpret_cast<const float4 *>(a + i);
      const float4 y = *reinterpret_cast<const float4 *>(t + i);
      const float d0 = x.x - y.x, d1 = x.y - y.y, d2 = x.z - y.z, d3 = x.w - y.w;
      acc += (d0 * d0 + d1 * d1) + (d2 * d2 + d3 * d3);
    } else {
      for (int64_t j = i; j < i + 4 && j < n; ++j) {
        const float d = a[j] - t[j];
        acc += d * d;
      }
    }
  }
  acc = wave_sum(acc);
  const int lane = threadIdx.x % kWave, wave = threadIdx.x / kWave;
  if (lane == 0) ws[wave] = acc;
  __syncthreads();
  if (threadIdx.x == 0) {
    float s = 0.f;
#pragma unroll
    for (int w = 0; w < kLossThreads / kWave; ++w) s += ws[w];
    partial[blockIdx.x] = s;
    __threadfence();
    last = atomicAdd(ticket, 1u) == gridDim.x - 1;
  }
  __syncthreads();
  if (!last) return;
  __threadfence();
  // every thread adds its 4 consecutive partials, then a fixed shuffle tree
  // and the waves in order: the same order on every launch
  double v = 0.0;
#pragma unroll
  for (int q = 0; q < kLossBlocks / kLossThreads; ++q) {
    const unsigned b = threadIdx.x * (kLossBlocks / kLossThreads) + q;
    if (b < gridDim.x) v += double(__hip_atomic_load(partial + b, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
  }
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, kWave);
  __shared__ double wd[kLossThreads / kWave];
  if (lane == 0) wd[wave] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    double total = 0.0;
#pragma unroll
    for (int w = 0; w < kLossThreads / kWave; ++w) total += wd[w];
    out[0] = float(total);
    __hip_atomic_store(ticket, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
}

// grad = 2 (a - t) * dL (dL: the upstream gradient of the loss, one float in device memory).
__global__ __launch_bounds__(kLossThreads) void sq_diff_grad_kernel(const float *__restrict__ a,
                                                                    const float *__restrict__ t, int64_t n,
                                                                    const float *__restrict__ dL,
                                                                    float *__restrict__ grad) {
  const float s = 2.f * dL[0];
  const int64_t stride = int64_t(gridDim.x) * kLossThreads;
  for (int64_t i = int64_t(blockIdx.x) * kLossThreads + threadIdx.x; i < n; i += stride) grad[i] = (a[i] - t[i]) * s;
}

}  // namespace msplat
