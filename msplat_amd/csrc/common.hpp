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
#include <initializer_list>
#include <type_traits>
#include <utility>

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

// Tile rectangle packed into 8 bytes (x0 | y0 << 16, x1 | y1 << 16); the
// tile grid is at most 4096 x 4096.
MSPLAT_HD uint2 pack_rect(const TileRect &t) {
  return make_uint2(uint32_t(t.x0) | (uint32_t(t.y0) << 16), uint32_t(t.x1) | (uint32_t(t.y1) << 16));
}

MSPLAT_DEV bool rect_covers(uint2 r, int tx, int ty) {
  const int x0 = int(r.x & 0xffffu), y0 = int(r.x >> 16), x1 = int(r.y & 0xffffu), y1 = int(r.y >> 16);
  return x0 <= tx && tx < x1 && y0 <= ty && ty < y1;
}

// ---- views of one launch ---------------------------------------------------
// Several cameras (of different image sizes, possibly of different scenes)
// render in one launch over one global tile space: view b owns global tiles
// [tile0[b], tile0[b+1]), cells [cell0[b], cell0[b+1]), pixels [pix0[b],
// pix0[b] + W H) of every concatenated per-pixel plane, and per-view
// Gaussian rows [vbase[b], vbase[b+1]); its row vbase[b] + g is Gaussian
// gbase[b] + g of the shared (per-Gaussian) arrays.  Passed by value as a
// kernel argument (about 3 KB).
constexpr int kMaxViews = 64;

struct Views {
  int count;
  int W[kMaxViews], H[kMaxViews];
  uint16_t gx[kMaxViews], gy[kMaxViews];  // tile grid (<= 4096 x 4096)
  uint16_t ncx[kMaxViews];                // cells per row (cell lists)
  int tile0[kMaxViews + 1];
  int group0[kMaxViews + 1];  // 2 x 2 tile groups (the grouped blending adjoint)
  int cell0[kMaxViews + 1];
  int64_t pix0[kMaxViews + 1];
  int64_t vbase[kMaxViews + 1];
  int64_t gbase[kMaxViews];
};

// View owning global tile t (binary search over tile0).
MSPLAT_DEV int view_of_tile(const Views &v, int t) {
  int lo = 0, hi = v.count - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (v.tile0[mid] <= t) lo = mid; else hi = mid - 1;
  }
  return lo;
}

// View owning global 2 x 2 tile group q (binary search over group0).
MSPLAT_DEV int view_of_group(const Views &v, int q) {
  int lo = 0, hi = v.count - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (v.group0[mid] <= q) lo = mid; else hi = mid - 1;
  }
  return lo;
}

// View owning per-view row i (binary search over vbase).
MSPLAT_DEV int view_of_row(const Views &v, int64_t i) {
  int lo = 0, hi = v.count - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (v.vbase[mid] <= i) lo = mid; else hi = mid - 1;
  }
  return lo;
}

// ---- wave-level helpers (wave64) -------------------------------------------

MSPLAT_DEV int lane_id() { return __lane_id(); }

// A 64-bit value of the first active lane (uniform_u64) or of lane l
// (lane_u64): readfirstlane and readlane return 32-bit ints, so each half is
// taken as unsigned before the halves are joined (an int low half would be
// sign-extended over the high one).
MSPLAT_DEV uint64_t uniform_u64(uint64_t v) {
  const uint32_t lo = uint32_t(__builtin_amdgcn_readfirstlane(uint32_t(v)));
  const uint32_t hi = uint32_t(__builtin_amdgcn_readfirstlane(uint32_t(v >> 32)));
  return (uint64_t(hi) << 32) | uint64_t(lo);
}
MSPLAT_DEV uint64_t lane_u64(uint64_t v, int l) {
  const uint32_t lo = uint32_t(__builtin_amdgcn_readlane(uint32_t(v), l));
  const uint32_t hi = uint32_t(__builtin_amdgcn_readlane(uint32_t(v >> 32), l));
  return (uint64_t(hi) << 32) | uint64_t(lo);
}

// Workgroup barrier for LDS traffic only: each wave waits for its own LDS
// operations, then for the other waves; its global loads stay in flight
// across it (__syncthreads' fence would wait for them too).
MSPLAT_DEV void lds_barrier() { asm volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory"); }

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

// ---- transposed multi-value wave reduction (gfx950) ---------------------------
// Sums NV values (NV a multiple of 4) over the 64 lanes at once.  Lanes form
// 4 rows of 16.  v_permlane32_swap pairs value 2i with 2i+1 across the two
// wave halves, v_permlane16_swap pairs the results across rows, and four DPP
// row rotations finish each row.  On return out[j] holds, in every lane of
// row r, the total of value 4 j + kRowValue[r].  About 1.75 + 1 instructions
// per value instead of 12 for a separate butterfly per value.
constexpr int kRowValue[4] = {0, 2, 1, 3};

// kRowValue[r] as arithmetic (a table indexed by a lane-dependent row would
// be a memory load).
MSPLAT_HD int row_value(int r) { return ((r & 1) << 1) | (r >> 1); }

MSPLAT_DEV float dpp_row_ror_add(float v, int ctrl) {
  // ctrl: 0x121..0x12F = row_ror:1..15 (rotate within each row of 16 lanes;
  // every lane has a source, so bound_ctrl changes nothing but lets the move
  // fold into the add)
  int r;
  switch (ctrl) {
    case 8: r = __builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x128, 0xf, 0xf, true); break;
    case 4: r = __builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x124, 0xf, 0xf, true); break;
    case 2: r = __builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x122, 0xf, 0xf, true); break;
    default: r = __builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x121, 0xf, 0xf, true); break;
  }
  return v + __int_as_float(r);
}

// NV need not be a multiple of 4: values past NV count as zero, and pairs
// made only of them cost nothing.
template <int NV>
MSPLAT_DEV void wave_sum_transposed(const float (&in)[NV], float (&out)[(NV + 3) / 4]) {
  constexpr int NP = (NV + 3) / 4 * 4;
  float a[NP / 2];
#pragma unroll
  for (int i = 0; i < NP / 2; ++i) {
    if (2 * i >= NV) {
      a[i] = 0.f;
      continue;
    }
    const float hi = 2 * i + 1 < NV ? in[2 * i + 1 < NV ? 2 * i + 1 : 0] : 0.f;
    const auto s = __builtin_amdgcn_permlane32_swap(__float_as_uint(in[2 * i]), __float_as_uint(hi), false, false);
    a[i] = __uint_as_float(s[0]) + __uint_as_float(s[1]);
  }
#pragma unroll
  for (int j = 0; j < NP / 4; ++j) {
    const auto s = __builtin_amdgcn_permlane16_swap(__float_as_uint(a[2 * j]), __float_as_uint(a[2 * j + 1]),
                                                    false, false);
    float b = __uint_as_float(s[0]) + __uint_as_float(s[1]);
    b = dpp_row_ror_add(b, 8);
    b = dpp_row_ror_add(b, 4);
    b = dpp_row_ror_add(b, 2);
    b = dpp_row_ror_add(b, 1);
    out[j] = b;
  }
}

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

// Up to kZeroRanges ranges of 32-bit words zeroed by one launch (blockIdx.y
// picks the range): one launch instead of one per buffer.
constexpr int kZeroRanges = 6;
struct ZeroRanges {
  uint32_t *p[kZeroRanges];
  int64_t n[kZeroRanges];
};

__global__ void fill_zero_ranges(ZeroRanges r) {
  uint32_t *p = r.p[blockIdx.y];
  const int64_t n = r.n[blockIdx.y];
  const int64_t stride = int64_t(gridDim.x) * blockDim.x;
  for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride) p[i] = 0u;
}

// ranges: (pointer, 32-bit words); null pointers and empty ranges are
// skipped; more than kZeroRanges ranges take further launches.
inline void launch_fill_zero_words(std::initializer_list<std::pair<void *, int64_t>> ranges, hipStream_t s) {
  ZeroRanges r{};
  unsigned k = 0;
  int64_t most = 0;
  auto flush = [&]() {
    if (k == 0) return;
    int64_t blocks = (most + 255) / 256;
    if (blocks > 4096) blocks = 4096;
    fill_zero_ranges<<<dim3(unsigned(blocks), k), dim3(256), 0, s>>>(r);
    r = ZeroRanges{};
    k = 0;
    most = 0;
  };
  for (const auto &pr : ranges) {
    if (!pr.first || pr.second <= 0) continue;
    r.p[k] = static_cast<uint32_t *>(pr.first);
    r.n[k] = pr.second;
    most = pr.second > most ? pr.second : most;
    if (++k == kZeroRanges) flush();
  }
  flush();
}

// ---- expired waits ----------------------------------------------------------
// Every look-back wait is bounded.  When one expires, its pass still finishes
// and stays inside its outputs (a partial prefix only lowers the offsets), but
// the order it wrote is undefined, so the expiry has to reach the caller:
//  * the pass sets the error word of its workspace;
//  * the last launch of the binning ORs the error words into wait_expired,
//    one word per device that stays set until the host clears it;
//  * while wait_expired is set, the cell-list blending forward writes NaN
//    images and transmittances, so undefined lists never yield a
//    plausible-looking image;
//  * the next binning's counting pass copies wait_expired into the second
//    word of its count, which the host reads with the count, and raises.
// force_expiry (set by msplat_debug_force_expiry, for tests only) makes every
// look-back give up at once without reading its predecessors, so the whole
// path can be exercised while every kernel still runs to its end.
__device__ uint32_t wait_expired;
__device__ uint32_t force_expiry;

MSPLAT_DEV bool expiry_forced() {
  return __hip_atomic_load(&force_expiry, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0u;
}
MSPLAT_DEV uint32_t expired_state() {
  return __hip_atomic_load(&wait_expired, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}


inline unsigned grid_1d(int64_t n, int threads) {
  return unsigned((n + threads - 1) / threads);
}

}  // namespace msplat
