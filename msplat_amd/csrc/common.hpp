// Shared device helpers for the gfx950 rasterizer kernels.
//
// The library is one translation unit (msplat_amd.hip includes every
// component header), so kernels defined in headers are compiled exactly once.
//
// The whole library is compiled with -ffp-contract=off and correctly rounded
// fp32 division/sqrt, so every per-Gaussian formula below rounds exactly like
// the same sequence of IEEE operations on the host.  The fp32 numpy oracle
// reproduces projection / EWA / keys bit for bit from that.
#pragma once

#include <hip/hip_runtime.h>

#include <climits>
#include <cstdint>

#define MSPLAT_DEV __device__ __forceinline__
#define MSPLAT_HD __host__ __device__ __forceinline__

namespace msplat {

// 16x16 pixel tiles; the tile id is part of the sort key, so this is fixed
// (msplat/include/config.h:7-10, msplat/src/sort_gaussian.cu:37).
constexpr int kTile = 16;
constexpr int kWave = 64;

MSPLAT_HD int tiles_x(int W) { return (W + kTile - 1) / kTile; }
MSPLAT_HD int tiles_y(int H) { return (H + kTile - 1) / kTile; }

// float -> int32 the way the GPU converts: saturating, NaN -> 0.  Written out
// so that out-of-range inputs are not undefined behaviour for the compiler.
MSPLAT_HD int f2i_sat(float v) {
  if (!(v == v)) return 0;
  if (v >= 2147483648.0f) return INT_MAX;
  if (v <= -2147483648.0f) return INT_MIN;
  return static_cast<int>(v);
}

MSPLAT_HD int imin(int a, int b) { return a < b ? a : b; }
MSPLAT_HD int imax(int a, int b) { return a > b ? a : b; }

struct TileRect {
  int x0, y0, x1, y1;
  MSPLAT_HD int area() const { return (x1 - x0) * (y1 - y0); }
};

// Tiles overlapped by a splat of integer radius r centred at (px, py):
// msplat/include/utils.h:17-37.  Every step is an fp32 operation in the
// reference's order: (p - r)/16 and ((p + r) + 16 - 1)/16, truncated, clamped
// to [0, grid].
MSPLAT_HD TileRect tile_rect(float px, float py, int r, int gx, int gy) {
  const float rf = static_cast<float>(r);
  TileRect t;
  t.x0 = imin(gx, imax(0, f2i_sat((px - rf) / 16.0f)));
  t.y0 = imin(gy, imax(0, f2i_sat((py - rf) / 16.0f)));
  t.x1 = imin(gx, imax(0, f2i_sat((((px + rf) + 16.0f) - 1.0f) / 16.0f)));
  t.y1 = imin(gy, imax(0, f2i_sat((((py + rf) + 16.0f) - 1.0f) / 16.0f)));
  return t;
}

// ---- wave-level helpers (wave64) -------------------------------------------

MSPLAT_DEV int lane_id() { return __lane_id(); }

MSPLAT_DEV float wave_sum(float v) {
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, kWave);
  return v;
}

MSPLAT_DEV int wave_max(int v) {
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) v = imax(v, __shfl_xor(v, off, kWave));
  return v;
}

MSPLAT_DEV bool wave_any(bool p) { return __ballot(p) != 0ull; }

// Sum a block of `kVals` per-thread values over a 256-thread workgroup and add
// the result to global memory with one atomic per value per workgroup.
template <int kVals, int kThreads>
MSPLAT_DEV void block_reduce_atomic(float (&v)[kVals], float *smem, float *dst) {
  constexpr int kWaves = kThreads / kWave;
  const int lane = threadIdx.x % kWave;
  const int wave = threadIdx.x / kWave;
#pragma unroll
  for (int i = 0; i < kVals; ++i) {
    float s = wave_sum(v[i]);
    if (lane == 0) smem[wave * kVals + i] = s;
  }
  __syncthreads();
  if (threadIdx.x < kVals) {
    float s = 0.f;
#pragma unroll
    for (int w = 0; w < kWaves; ++w) s += smem[w * kVals + threadIdx.x];
    unsafeAtomicAdd(dst + threadIdx.x, s);
  }
}

// Grid-stride zero fill used before atomic accumulation (also safe inside a
// captured graph, unlike a memset node).
template <typename T>
__global__ void fill_zero(T *p, int64_t n) {
  const int64_t stride = int64_t(gridDim.x) * blockDim.x;
  for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride)
    p[i] = T(0);
}

template <typename T>
inline void launch_fill_zero(T *p, int64_t n, hipStream_t s) {
  if (n <= 0) return;
  int64_t blocks = (n + 255) / 256;
  if (blocks > 8192) blocks = 8192;
  fill_zero<T><<<dim3(unsigned(blocks)), dim3(256), 0, s>>>(p, n);
}

inline unsigned grid_1d(int64_t n, int threads) {
  return unsigned((n + threads - 1) / threads);
}

}  // namespace msplat
