// Squared-error loss of a rendered image against a target, forward and
// adjoint, each one launch: the training step's loss next to the rasterizer
// (the reference's training loops compute it with separate torch kernels,
// tutorials/gs_2d.py:73-90).
#pragma once

#include "common.hpp"

namespace msplat {

constexpr int kLossThreads = 256;   // gradient kernel
constexpr int kSumThreads = 1024;   // sum kernel: one 16-wave workgroup per CU
constexpr int kSumBlocks = 256;     // partial sums: one per workgroup, summed in order by the last one
constexpr int kSumUnroll = 4;       // float4 pairs loaded together per lane
static_assert(kSumBlocks <= kSumThreads, "the last workgroup reads one partial per thread");

MSPLAT_DEV float sq_diff4(float4 x, float4 y) {
  const float d0 = x.x - y.x, d1 = x.y - y.y, d2 = x.z - y.z, d3 = x.w - y.w;
  return (d0 * d0 + d1 * d1) + (d2 * d2 + d3 * d3);
}

// partial[b] = sum over the grid-stride slice of workgroup b of (a - t)^2
// (fp32 per thread, a fixed-order workgroup tree); the last workgroup to
// finish (ticket) adds the partials in fp64 in a fixed order and writes the
// total, so the result does not depend on scheduling.  ticket must be 0 on
// entry; it is reset for the next launch.
//
// Few, large workgroups: every workgroup's ticket is an atomic on one word,
// and those serialise (2048 workgroups spent more time on their tickets
// than on reading the arrays); each lane keeps kSumUnroll pairs of 16-byte
// loads in flight instead.
__global__ __launch_bounds__(kSumThreads) void sq_diff_sum_kernel(const float *__restrict__ a,
                                                                  const float *__restrict__ t, int64_t n,
                                                                  float *__restrict__ partial,
                                                                  uint32_t *__restrict__ ticket,
                                                                  float *__restrict__ out) {
  __shared__ float ws[kSumThreads / kWave];
  __shared__ double wd[kSumThreads / kWave];
  __shared__ bool last;
  const int64_t tid = int64_t(blockIdx.x) * kSumThreads + threadIdx.x;
  const int64_t nth = int64_t(gridDim.x) * kSumThreads;
  // 16-byte loads when both arrays allow them; the remainder element-wise
  const bool vec = ((reinterpret_cast<uintptr_t>(a) | reinterpret_cast<uintptr_t>(t)) & 15) == 0;
  const int64_t n4 = vec ? n / 4 : 0;
  const float4 *a4 = reinterpret_cast<const float4 *>(a), *t4 = reinterpret_cast<const float4 *>(t);
  float acc0 = 0.f, acc1 = 0.f;
  int64_t i = tid;
  for (; i + (kSumUnroll - 1) * nth < n4; i += kSumUnroll * nth) {
    float4 x[kSumUnroll], y[kSumUnroll];
#pragma unroll
    for (int u = 0; u < kSumUnroll; ++u) {
      x[u] = a4[i + u * nth];
      y[u] = t4[i + u * nth];
    }
#pragma unroll
    for (int u = 0; u < kSumUnroll; ++u) {
      if (u & 1) acc1 += sq_diff4(x[u], y[u]);
      else acc0 += sq_diff4(x[u], y[u]);
    }
  }
  for (; i < n4; i += nth) acc0 += sq_diff4(a4[i], t4[i]);
  for (int64_t j = n4 * 4 + tid; j < n; j += nth) {
    const float d = a[j] - t[j];
    acc1 += d * d;
  }
  float acc = acc0 + acc1;
  acc = wave_sum(acc);
  const int lane = threadIdx.x % kWave, wave = threadIdx.x / kWave;
  if (lane == 0) ws[wave] = acc;
  __syncthreads();
  if (threadIdx.x == 0) {
    float s = 0.f;
#pragma unroll
    for (int w = 0; w < kSumThreads / kWave; ++w) s += ws[w];
    partial[blockIdx.x] = s;
    __threadfence();
    last = atomicAdd(ticket, 1u) == gridDim.x - 1;
  }
  __syncthreads();
  if (!last) return;
  __threadfence();
  // thread b takes partial b, then a fixed shuffle tree and the waves in
  // order: the same order on every launch
  double v = 0.0;
  if (threadIdx.x < gridDim.x)
    v = double(__hip_atomic_load(partial + threadIdx.x, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, kWave);
  if (lane == 0) wd[wave] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    double total = 0.0;
#pragma unroll
    for (int w = 0; w < kSumThreads / kWave; ++w) total += wd[w];
    out[0] = float(total);
    __hip_atomic_store(ticket, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
}

// out = the sum of n floats, in fp64 and a fixed order (one workgroup):
// thread t adds p[t + 1024 (8 k + u)] into running sum u, then the sums,
// a fixed shuffle tree and the waves in order.  The fused rasterizer loss's
// per-wave partials.
__global__ __launch_bounds__(kSumThreads) void sum_floats_kernel(const float *__restrict__ p, int64_t n,
                                                                 float *__restrict__ out) {
  __shared__ double wd[kSumThreads / kWave];
  // 8 loads in flight per thread, 8 running sums, added in a fixed order
  constexpr int kU = 8;
  double acc[kU];
#pragma unroll
  for (int u = 0; u < kU; ++u) acc[u] = 0.0;
  for (int64_t base = threadIdx.x; base < n; base += int64_t(kU) * kSumThreads) {
    float x[kU];
#pragma unroll
    for (int u = 0; u < kU; ++u) {
      const int64_t i = base + int64_t(u) * kSumThreads;
      x[u] = i < n ? p[i] : 0.f;
    }
#pragma unroll
    for (int u = 0; u < kU; ++u) acc[u] += double(x[u]);
  }
  double v = ((acc[0] + acc[1]) + (acc[2] + acc[3])) + ((acc[4] + acc[5]) + (acc[6] + acc[7]));
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, kWave);
  const int lane = threadIdx.x % kWave, wave = threadIdx.x / kWave;
  if (lane == 0) wd[wave] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    double total = 0.0;
#pragma unroll
    for (int w = 0; w < kSumThreads / kWave; ++w) total += wd[w];
    out[0] = float(total);
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
