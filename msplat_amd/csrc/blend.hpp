// Front-to-back alpha blending over the sorted per-tile Gaussian lists,
// forward and adjoint (msplat/src/alpha_blending.cu:16-246).
//
// Work decomposition (gfx950): a 16x16 tile is split into four 8x8
// quadrants, each handled by a one-wavefront workgroup (64 lanes = 64
// pixels).  A wave stages 64 Gaussians of the tile list at a time in LDS
// (position, conic, opacity and all C feature channels) and every lane walks
// them; the wave stops as soon as all 64 of its pixels are saturated (a
// ballot), independent of the other quadrants, and needs no workgroup-wide
// barriers.  The four quadrants of one tile read the same list, so their
// workgroup ids are chosen congruent modulo 8: workgroups are dispatched
// round-robin over the 8 XCDs, which puts the four on one XCD and its L2.
//
// All C channels are blended in one walk (the reference re-walks the list
// for every 32-channel chunk).
#pragma once

#include "common.hpp"

namespace msplat {

constexpr int kBlendLanes = 64;

struct QuadPix {
  int tile, px, py;
  bool active;  // workgroup maps to a real tile
};

// blockIdx.x -> (tile, quadrant): xcd = b % 8 selects the tile within a group
// of 8 tiles, (b / 8) % 4 the quadrant, b / 32 the group.
MSPLAT_DEV QuadPix quad_pixel(int gx, int ntiles) {
  const int b = blockIdx.x;
  const int tile = (b >> 5) * 8 + (b & 7);
  const int quad = (b >> 3) & 3;
  QuadPix q;
  q.tile = tile;
  q.active = tile < ntiles;
  const int tx = q.active ? tile % gx : 0, ty = q.active ? tile / gx : 0;
  const int lane = threadIdx.x;
  q.px = tx * kTile + (quad & 1) * 8 + (lane & 7);
  q.py = ty * kTile + (quad >> 1) * 8 + (lane >> 3);
  return q;
}

inline unsigned blend_grid(int ntiles) { return unsigned(((ntiles + 7) / 8) * 32); }

// exp(x) as 2^(x log2 e) on the hardware exp2 (v_exp_f32).
MSPLAT_DEV float fast_exp(float x) { return __builtin_amdgcn_exp2f(x * 1.44269504088896341f); }

// One staged Gaussian is a 16-byte aligned record (u, v, a, b | c, o, f0, f1 |
// f2 ...), so the inner loops fetch it with 128-bit LDS broadcast reads.
template <int CMAX>
struct BlendStage {
  static constexpr int kQuads = (6 + CMAX + 3) / 4;  // float4s per record
  float4 rec[kBlendLanes * kQuads];
};

// Stage Gaussian g into LDS slot `slot`: channels [c0, c0 + cc) of its row.
template <int CMAX>
MSPLAT_DEV void stage_gaussian(BlendStage<CMAX> &s, int slot, int g, const float *__restrict__ uv,
                               const float *__restrict__ conic, const float *__restrict__ opacity,
                               const float *__restrict__ feature, int C, int c0, int cc) {
  constexpr int NQ = BlendStage<CMAX>::kQuads;
  float r[NQ * 4];
  r[0] = uv[2 * int64_t(g)];
  r[1] = uv[2 * int64_t(g) + 1];
  r[2] = conic[3 * int64_t(g)];
  r[3] = conic[3 * int64_t(g) + 1];
  r[4] = conic[3 * int64_t(g) + 2];
  r[5] = opacity[g];
  const float *fr = feature + int64_t(g) * C + c0;
#pragma unroll
  for (int i = 6; i < NQ * 4; ++i) r[i] = (i - 6 < cc) ? fr[i - 6] : 0.f;
#pragma unroll
  for (int q = 0; q < NQ; ++q) s.rec[slot * NQ + q] = make_float4(r[4 * q], r[4 * q + 1], r[4 * q + 2], r[4 * q + 3]);
}

// Geometry of record j: (u, v, a, b) and (c, o, f0, f1).
template <int CMAX>
MSPLAT_DEV void stage_geom(const BlendStage<CMAX> &s, int j, float4 &g0, float4 &g1) {
  constexpr int NQ = BlendStage<CMAX>::kQuads;
  g0 = s.rec[j * NQ];
  g1 = s.rec[j * NQ + 1];
}

// All CMAX channels of record j (channels 0, 1 come from g1).
template <int CMAX>
MSPLAT_DEV void stage_features(const BlendStage<CMAX> &s, int j, const float4 &g1, float (&f)[CMAX]) {
  constexpr int NQ = BlendStage<CMAX>::kQuads;
  f[0] = g1.z;
  if constexpr (CMAX > 1) f[1] = g1.w;
#pragma unroll
  for (int q = 2; q < NQ; ++q) {
    const float4 v = s.rec[j * NQ + q];
    const int b = 4 * q - 6;
    if (b < CMAX) f[b] = v.x;
    if (b + 1 < CMAX) f[b + 1] = v.y;
    if (b + 2 < CMAX) f[b + 2] = v.z;
    if (b + 3 < CMAX) f[b + 3] = v.w;
  }
}

// power = -0.5 (a dx^2 + c dy^2) - b dx dy with d = mean - pixel
// (alpha_blending.cu:74-82).
MSPLAT_DEV float gauss_power(float a, float b, float c, float dx, float dy) {
  return -0.5f * ((a * dx) * dx + (c * dy) * dy) - (b * dx) * dy;
}

template <int CMAX>
__global__ __launch_bounds__(kBlendLanes) void blend_fwd_kernel(
    const float *__restrict__ uv, const float *__restrict__ conic, const float *__restrict__ opacity,
    const float *__restrict__ feature, int C, int c0, const int32_t *__restrict__ idx_sorted,
    const int32_t *__restrict__ tile_range, float bg, int W, int H, int gx, int ntiles,
    float *__restrict__ out, float *__restrict__ final_T, int32_t *__restrict__ ncontrib) {
  __shared__ BlendStage<CMAX> s;
  const QuadPix q = quad_pixel(gx, ntiles);
  if (!q.active) return;
  const int lane = threadIdx.x;
  const bool inside = q.px < W && q.py < H;
  const float pxf = float(q.px), pyf = float(q.py);
  const int start = tile_range[2 * q.tile], end = tile_range[2 * q.tile + 1];
  const int cc = C - c0 < CMAX ? C - c0 : CMAX;  // channels of this pass

  bool done = !inside;
  float T = 1.f;
  int contributor = 0, last = 0;
  float F[CMAX];
#pragma unroll
  for (int ch = 0; ch < CMAX; ++ch) F[ch] = 0.f;

  for (int base = start; base < end; base += kBlendLanes) {
    if (__ballot(!done) == 0ull) break;
    const int cnt = end - base < kBlendLanes ? end - base : kBlendLanes;
    __syncthreads();
    if (lane < cnt) stage_gaussian<CMAX>(s, lane, idx_sorted[base + lane], uv, conic, opacity, feature, C, c0, cc);
    __syncthreads();
    for (int j = 0; j < cnt; ++j) {
      if (!done) {
        ++contributor;
        float4 g0, g1;
        stage_geom(s, j, g0, g1);
        const float dx = g0.x - pxf, dy = g0.y - pyf;
        const float power = gauss_power(g0.z, g0.w, g1.x, dx, dy);
        if (!(power > 0.f)) {  // a NaN power is not skipped, as in the reference
          const float alpha = fminf(0.99f, g1.y * fast_exp(power));
          if (!(alpha < 1.f / 255.f)) {
            const float nT = T * (1.f - alpha);
            if (nT < 0.0001f) {
              done = true;
            } else {
              float f[CMAX];
              stage_features(s, j, g1, f);
#pragma unroll
              for (int ch = 0; ch < CMAX; ++ch)
                if (ch < cc) F[ch] += f[ch] * alpha * T;
              T = nT;
              last = contributor;
            }
          }
        }
      }
    }
  }
  if (inside) {
    const int64_t pix = int64_t(q.py) * W + q.px;
    const int64_t hw = int64_t(H) * W;
    if (c0 == 0) {
      final_T[pix] = T;
      ncontrib[pix] = last;
    }
#pragma unroll
    for (int ch = 0; ch < CMAX; ++ch)
      if (ch < cc) out[(c0 + ch) * hw + pix] = F[ch] + T * bg;
  }
}

// Adjoint: walk each pixel's contributors back to front, recovering T by
// division (alpha_blending.cu:155-245); per Gaussian, the wave sums its 64
// lanes' contributions and issues one atomic per gradient value.
template <int CMAX>
__global__ __launch_bounds__(kBlendLanes) void blend_bwd_kernel(
    const float *__restrict__ uv, const float *__restrict__ conic, const float *__restrict__ opacity,
    const float *__restrict__ feature, int C, int c0, const int32_t *__restrict__ idx_sorted,
    const int32_t *__restrict__ tile_range, float bg, int W, int H, int gx, int ntiles,
    const float *__restrict__ final_T, const int32_t *__restrict__ ncontrib, const float *__restrict__ dout,
    float *__restrict__ duv, float *__restrict__ dconic, float *__restrict__ dopac, float *__restrict__ dfeat) {
  __shared__ BlendStage<CMAX> s;
  const QuadPix q = quad_pixel(gx, ntiles);
  if (!q.active) return;
  const int lane = threadIdx.x;
  const bool inside = q.px < W && q.py < H;
  const float pxf = float(q.px), pyf = float(q.py);
  const int start = tile_range[2 * q.tile];
  const int64_t pix = int64_t(q.py) * W + q.px;
  const int64_t hw = int64_t(H) * W;

  const int cc = C - c0 < CMAX ? C - c0 : CMAX;  // channels of this pass
  const float T_final = inside ? final_T[pix] : 0.f;
  const int last = inside ? ncontrib[pix] : 0;
  float g[CMAX];
  float bg_dot = 0.f;
#pragma unroll
  for (int ch = 0; ch < CMAX; ++ch) {
    g[ch] = (inside && ch < cc) ? dout[(c0 + ch) * hw + pix] : 0.f;
    if (ch < cc) bg_dot += bg * g[ch];
  }
  const int wave_last = wave_max(last);

  float T = T_final;
  float accum[CMAX], last_f[CMAX];
#pragma unroll
  for (int ch = 0; ch < CMAX; ++ch) accum[ch] = last_f[ch] = 0.f;
  float last_alpha = 0.f;

  for (int hi = start + wave_last; hi > start; hi -= kBlendLanes) {
    const int lo = hi - kBlendLanes > start ? hi - kBlendLanes : start;
    const int cnt = hi - lo;
    __syncthreads();
    if (lane < cnt) stage_gaussian<CMAX>(s, lane, idx_sorted[lo + lane], uv, conic, opacity, feature, C, c0, cc);
    __syncthreads();
    for (int j = cnt - 1; j >= 0; --j) {
      const int rel = lo + j - start;  // 0-based position in the tile list
      bool contrib = rel < last;
      float4 g0, g1;
      stage_geom(s, j, g0, g1);
      const float dx = g0.x - pxf, dy = g0.y - pyf;
      const float a = g0.z, b = g0.w, c = g1.x, op = g1.y;
      float G = 0.f, alpha = 0.f;
      if (contrib) {
        const float power = gauss_power(a, b, c, dx, dy);
        if (power > 0.f) {
          contrib = false;
        } else {
          G = fast_exp(power);
          alpha = fminf(0.99f, op * G);
          if (alpha < 1.f / 255.f) contrib = false;
        }
      }
      const uint64_t any = __ballot(contrib);
      if (any == 0ull) continue;
      float vuv0 = 0.f, vuv1 = 0.f, vc0 = 0.f, vc1 = 0.f, vc2 = 0.f, vo = 0.f;
      float vf[CMAX];
#pragma unroll
      for (int ch = 0; ch < CMAX; ++ch) vf[ch] = 0.f;
      if (contrib) {
        // 1 / (1 - alpha) on the hardware reciprocal (the reference divides
        // with fast-math); 1 - alpha >= 0.01, so it is well conditioned.
        const float inv = __builtin_amdgcn_rcpf(1.f - alpha);
        T = T * inv;
        const float dcol = alpha * T;
        float dL_dalpha = 0.f;
        float fe[CMAX];
        stage_features(s, j, g1, fe);
#pragma unroll
        for (int ch = 0; ch < CMAX; ++ch) {
          if (ch < cc) {
            const float f = fe[ch];
            accum[ch] = last_alpha * last_f[ch] + (1.f - last_alpha) * accum[ch];
            last_f[ch] = f;
            dL_dalpha += (f - accum[ch]) * g[ch];
            vf[ch] = dcol * g[ch];
          }
        }
        dL_dalpha *= T;
        last_alpha = alpha;
        dL_dalpha += (-T_final * inv) * bg_dot;
        const float dL_dG = op * dL_dalpha;
        const float gdx = ((-G) * dx) * a - (G * dy) * b;
        const float gdy = ((-G) * dy) * c - (G * dx) * b;
        vuv0 = dL_dG * gdx;
        vuv1 = dL_dG * gdy;
        vc0 = (((-0.5f * G) * dx) * dx) * dL_dG;
        vc1 = (((-G) * dx) * dy) * dL_dG;
        vc2 = (((-0.5f * G) * dy) * dy) * dL_dG;
        vo = G * dL_dalpha;
      }
      // Sum the 6 + cc per-lane contributions over the wave in one transposed
      // reduction; lane 16 r + k then owns value 4 k' + kRowValue[r] of
      // vector k' = k + 16 q in atomic round q and issues its atomic.
      const int gidx = idx_sorted[lo + j];
      constexpr int NV = ((6 + CMAX) + 3) / 4 * 4;
      float vals[NV];
      vals[0] = vuv0;
      vals[1] = vuv1;
      vals[2] = vc0;
      vals[3] = vc1;
      vals[4] = vc2;
      vals[5] = vo;
#pragma unroll
      for (int ch = 0; ch < NV - 6; ++ch) vals[6 + ch] = ch < CMAX ? vf[ch < CMAX ? ch : 0] : 0.f;
      float tot[NV / 4];
      wave_sum_transposed<NV>(vals, tot);
      const int row = lane >> 4, k = lane & 15;
#pragma unroll
      for (int q = 0; q * 16 < NV / 4; ++q) {
        const int vec = k + 16 * q;
        float mine = 0.f;
#pragma unroll
        for (int jj = 0; jj < NV / 4; ++jj)
          if (jj == vec) mine = tot[jj];
        const int v = 4 * vec + kRowValue[row];
        float *dst = nullptr;
        if (vec < NV / 4 && v < 6 + cc) {
          if (v < 2) dst = duv + 2 * int64_t(gidx) + v;
          else if (v < 5) dst = dconic + 3 * int64_t(gidx) + (v - 2);
          else if (v == 5) dst = dopac + gidx;
          else dst = dfeat + int64_t(gidx) * C + c0 + (v - 6);
        }
        if (dst) unsafeAtomicAdd(dst, mine);
      }
    }
  }
}

}  // namespace msplat
