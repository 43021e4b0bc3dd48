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

// out = the sum of n floats in a fixed order: workgroup b sums its slice
// p[b * 1024 * kSumRounds ...] (each thread kSumRounds strided values in
// fp64, a fixed shuffle tree, the waves in order) into part[b]; the last
// workgroup to finish (ticket) adds part[0..) in order and writes out.
// Many workgroups, one load per thread per round: the partials of a whole
// frame in about one memory latency.  ticket must be 0 on entry; it is
// reset for the next launch.  The fused rasterizer loss's per-wave partials.
constexpr int kSumRounds = 4;
constexpr int kSumParts = 256;  // workgroups at most; more values take more rounds each

inline unsigned sum_floats_blocks(int64_t n) {
  const int64_t per = int64_t(kSumThreads) * kSumRounds;
  const int64_t b = (n + per - 1) / per;
  return unsigned(b < 1 ? 1 : (b > kSumParts ? kSumParts : b));
}

__global__ __launch_bounds__(kSumThreads) void sum_floats_kernel(const float *__restrict__ p, int64_t n,
                                                                 double *__restrict__ part,
                                                                 uint32_t *__restrict__ ticket,
                                                                 float *__restrict__ out) {
  __shared__ double wd[kSumThreads / kWave];
  __shared__ bool last;
  const int lane = threadIdx.x % kWave, wave = threadIdx.x / kWave;
  // this workgroup's slice of the values, in kSumRounds strided rounds
  const int64_t per = (n + gridDim.x - 1) / gridDim.x;
  const int64_t lo = int64_t(blockIdx.x) * per, hi = lo + per < n ? lo + per : n;
  double acc[kSumRounds];
#pragma unroll
  for (int u = 0; u < kSumRounds; ++u) acc[u] = 0.0;
  for (int64_t base = lo + threadIdx.x; base < hi; base += int64_t(kSumRounds) * kSumThreads) {
    float x[kSumRounds];
#pragma unroll
    for (int u = 0; u < kSumRounds; ++u) {
      const int64_t i = base + int64_t(u) * kSumThreads;
      x[u] = i < hi ? p[i] : 0.f;
    }
#pragma unroll
    for (int u = 0; u < kSumRounds; ++u) acc[u] += double(x[u]);
  }
  double v = (acc[0] + acc[1]) + (acc[2] + acc[3]);
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, kWave);
  if (lane == 0) wd[wave] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    double s = 0.0;
#pragma unroll
    for (int w = 0; w < kSumThreads / kWave; ++w) s += wd[w];
    part[blockIdx.x] = s;
    __threadfence();
    last = atomicAdd(ticket, 1u) == gridDim.x - 1;
  }
  __syncthreads();
  if (!last) return;
  __threadfence();
  if (wave == 0) {  // the parts in order: lane l adds parts l, l + 64, ..., then a fixed tree
    double t = 0.0;
    for (unsigned b = lane; b < gridDim.x; b += kWave) {
      const unsigned long long bits =
          __hip_atomic_load(reinterpret_cast<unsigned long long *>(part) + b, __ATOMIC_RELAXED,
                            __HIP_MEMORY_SCOPE_AGENT);
      t += __longlong_as_double(static_cast<long long>(bits));
    }
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) t += __shfl_xor(t, off, kWave);
    if (lane == 0) {
      out[0] = float(t);
      __hip_atomic_store(ticket, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
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
