// The training step next to the rasterizer: the Adam update of the Gaussian
// parameters and the densification that consumes dL/dndc.
//
// The reference trains with torch.optim.Adam over the Gaussian parameter
// tensors, one learning rate per tensor (tutorials/gs_3d.py:165-172,
// tutorials/gs_2d.py:32-36), and every few hundred iterations prunes
// low-opacity Gaussians and refills their slots with copies of the largest
// ones, keeping the count fixed (tutorials/gs_3d.py:193-225).  dL/dndc, the
// screen-space position gradient the rasterizer hands to the caller through
// `ndc` (msplat/alpha_blending.py:107-110), is what densification schemes
// accumulate per Gaussian.
//
// Both are HBM-bound element-wise passes; nothing here is a contraction.
#pragma once

#include "common.hpp"
#include "sort.hpp"

namespace msplat {
namespace train {

// ---- Adam -------------------------------------------------------------------
// One launch updates up to kAdamGroups parameter arrays (xyz, scale, rotate,
// opacity, features / SH coefficients ...), each with its own learning rate
// and step count.  Per element, in torch.optim.Adam's order of fp32
// operations (its multi-tensor path):
//   m <- m + (1 - b1) (g - m)                 (lerp towards g)
//   v <- v b2;  v <- v + (1 - b2) g g
//   p <- p + (-lr / (1 - b1^t)) (m / (sqrt(v) / sqrt(1 - b2^t) + eps))
// with the bias corrections formed on the host in double and rounded to
// fp32, as torch passes them.  28 bytes of HBM traffic per element.

constexpr int kAdamGroups = 8;
constexpr int kAdamThreads = 256;
constexpr int kAdamItems = 4;  // elements per lane (one 16-byte access per array when aligned)
constexpr int64_t kAdamBlockElems = int64_t(kAdamThreads) * kAdamItems;

struct AdamGroup {
  float *p;
  const float *g;
  float *m, *v;
  int64_t n;
  int64_t blk0;       // first workgroup of this array
  float neg_step;     // -lr / (1 - b1^t)
  float bc2_sqrt;     // sqrt(1 - b2^t)
  int vec;            // all four arrays 16-byte aligned
};

struct AdamArgs {
  AdamGroup grp[kAdamGroups];
  int count;
  float w1;   // 1 - b1
  float b2;
  float w2;   // 1 - b2
  float eps;
};

MSPLAT_DEV void adam_elem(float &p, float g, float &m, float &v, float w1, float b2, float w2, float eps,
                          float neg_step, float bc2_sqrt) {
  m = m + w1 * (g - m);
  v = v * b2;
  v = v + (w2 * g) * g;
  const float denom = sqrtf(v) / bc2_sqrt + eps;
  p = p + neg_step * (m / denom);
}

__global__ __launch_bounds__(kAdamThreads) void adam_kernel(AdamArgs a) {
  int gi = 0;  // uniform: the array this workgroup updates
  while (gi + 1 < a.count && int64_t(blockIdx.x) >= a.grp[gi + 1].blk0) ++gi;
  const AdamGroup &G = a.grp[gi];
  const int64_t i0 = (int64_t(blockIdx.x) - G.blk0) * kAdamBlockElems + int64_t(threadIdx.x) * kAdamItems;
  if (i0 >= G.n) return;
  if (G.vec && i0 + kAdamItems <= G.n) {
    float4 p = *reinterpret_cast<const float4 *>(G.p + i0);
    const float4 g = *reinterpret_cast<const float4 *>(G.g + i0);
    float4 m = *reinterpret_cast<const float4 *>(G.m + i0);
    float4 v = *reinterpret_cast<const float4 *>(G.v + i0);
    adam_elem(p.x, g.x, m.x, v.x, a.w1, a.b2, a.w2, a.eps, G.neg_step, G.bc2_sqrt);
    adam_elem(p.y, g.y, m.y, v.y, a.w1, a.b2, a.w2, a.eps, G.neg_step, G.bc2_sqrt);
    adam_elem(p.z, g.z, m.z, v.z, a.w1, a.b2, a.w2, a.eps, G.neg_step, G.bc2_sqrt);
    adam_elem(p.w, g.w, m.w, v.w, a.w1, a.b2, a.w2, a.eps, G.neg_step, G.bc2_sqrt);
    *reinterpret_cast<float4 *>(G.p + i0) = p;
    *reinterpret_cast<float4 *>(G.m + i0) = m;
    *reinterpret_cast<float4 *>(G.v + i0) = v;
    return;
  }
  for (int k = 0; k < kAdamItems; ++k) {
    const int64_t i = i0 + k;
    if (i >= G.n) break;
    float p = G.p[i], m = G.m[i], v = G.v[i];
    adam_elem(p, G.g[i], m, v, a.w1, a.b2, a.w2, a.eps, G.neg_step, G.bc2_sqrt);
    G.p[i] = p;
    G.m[i] = m;
    G.v[i] = v;
  }
}

// ---- densification statistics (the consumer of dL/dndc) ---------------------
// For every Gaussian seen by the view: accum += |dL/dndc| (the Euclidean
// norm of its two components), denom += 1.  Seen means radius > 0 when the
// radii are given, else a non-zero gradient.

__global__ __launch_bounds__(256) void densify_accumulate_kernel(const float2 *__restrict__ dndc,
                                                                 const int32_t *__restrict__ radius, int64_t n,
                                                                 float *__restrict__ accum,
                                                                 float *__restrict__ denom) {
  const int64_t i = int64_t(blockIdx.x) * 256 + threadIdx.x;
  if (i >= n) return;
  const float2 d = dndc[i];
  const bool seen = radius ? radius[i] > 0 : (d.x != 0.f || d.y != 0.f);
  if (!seen) return;
  accum[i] = accum[i] + sqrtf(d.x * d.x + d.y * d.y);
  denom[i] = denom[i] + 1.f;
}

// ---- prune and refill at a fixed count (tutorials/gs_3d.py:193-225) ----------
// 1. prune = { g : opacity[g] < threshold } in ascending index order (M rows);
// 2. the M donors are the Gaussians of largest score, score = the largest of
//    the row's three scale parameters (or a caller-given score, such as the
//    mean dL/dndc norm), in descending score order, ties by ascending index
//    (a stable descending argsort);
// 3. pruned row prune[i] receives donor i's row of every attribute array.
// A pruned row that is itself a donor keeps its own values: that is what the
// reference's assignment gives when executed in order (its index list holds
// the row twice, last as a donor of itself); the donors' values are read
// before any row is written, and no donor row is written, so the copies
// do not race.

constexpr int kDensifyArrays = 8;

struct DensifyRows {
  float *a[kDensifyArrays];
  int w[kDensifyArrays];  // floats per row
  int count;
};

// Sort key of a score for a descending order: larger scores first, NaN
// before everything (torch's sort places NaN above +inf), -0 equal to +0.
MSPLAT_DEV uint32_t descending_key(float s) {
  if (s != s) return 0u;
  if (s == 0.f) s = 0.f;
  uint32_t u = __float_as_uint(s);
  u = (u & 0x80000000u) ? ~u : (u | 0x80000000u);  // ascending order of the floats
  return ~u;
}

// Largest of a row of three, NaN if any is NaN (torch.amax).
MSPLAT_DEV float amax3(float a, float b, float c) {
  if (a != a || b != b || c != c) return __builtin_nanf("");
  return fmaxf(fmaxf(a, b), c);
}

__global__ __launch_bounds__(256) void densify_prepare_kernel(const float *__restrict__ opacity, float threshold,
                                                              const float *__restrict__ score,
                                                              const float *__restrict__ scale, int64_t n,
                                                              uint32_t *__restrict__ flag,
                                                              uint32_t *__restrict__ key,
                                                              uint32_t *__restrict__ val) {
  const int64_t i = int64_t(blockIdx.x) * 256 + threadIdx.x;
  if (i >= n) return;
  flag[i] = opacity[i] < threshold ? 1u : 0u;
  const float s = score ? score[i] : amax3(scale[3 * i], scale[3 * i + 1], scale[3 * i + 2]);
  key[i] = descending_key(s);
  val[i] = uint32_t(i);
}

// rank[order[j]] = j; prune list in ascending index order.
__global__ __launch_bounds__(256) void densify_lists_kernel(const uint32_t *__restrict__ flag,
                                                            const uint32_t *__restrict__ pos,
                                                            const uint32_t *__restrict__ order, int64_t n,
                                                            uint32_t *__restrict__ rank,
                                                            uint32_t *__restrict__ prune) {
  const int64_t i = int64_t(blockIdx.x) * 256 + threadIdx.x;
  if (i >= n) return;
  const uint32_t o = order[i];
  if (int64_t(o) < n) rank[o] = uint32_t(i);
  if (flag[i]) prune[pos[i]] = uint32_t(i);
}

// The pruned-row count, or -2 when the sort's bounded wait expired (its order
// is then undefined: nothing is copied, and the device's wait word is set).
__global__ void densify_count_kernel(const int64_t *__restrict__ total, const uint32_t *__restrict__ error,
                                     int64_t *__restrict__ count) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    const bool bad = *error != 0u;
    if (bad) atomicOr(&wait_expired, 1u);
    *count = bad ? -2 : *total;
  }
}

// Waves stride over the pruned rows (the count lives on the device, so the
// grid is fixed); each wave's lanes copy the donor's row, array by array.
constexpr int kCopyBlocks = 2048;

__global__ __launch_bounds__(256) void densify_copy_kernel(const uint32_t *__restrict__ prune,
                                                           const uint32_t *__restrict__ order,
                                                           const uint32_t *__restrict__ rank, int64_t n,
                                                           const int64_t *__restrict__ count, DensifyRows rows) {
  const int lane = threadIdx.x % kWave;
  const int64_t M = *count;  // negative (an expired wait): nothing is copied
  const int64_t waves = int64_t(gridDim.x) * (256 / kWave);
  for (int64_t i = (int64_t(blockIdx.x) * 256 + threadIdx.x) / kWave; i < M; i += waves) {
    const int64_t dst = prune[i], src = order[i];
    if (dst >= n || src >= n || int64_t(rank[dst]) < M) continue;  // a donor keeps its own row
    for (int k = 0; k < rows.count; ++k) {
      float *a = rows.a[k];
      const int w = rows.w[k];
      for (int c = lane; c < w; c += kWave) a[dst * w + c] = a[src * w + c];
    }
  }
}

}  // namespace train
}  // namespace msplat
