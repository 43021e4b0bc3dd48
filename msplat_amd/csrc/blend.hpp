// Front-to-back alpha blending, forward and adjoint
// (msplat/src/alpha_blending.cu:16-246), over two kinds of Gaussian lists.
//
//  * Tile lists (kWalk = false): idx_sorted / tile_range as sort_gaussian
//    returns them, the reference's format (alpha_blending API).
//  * Cell lists (kWalk = true): every visible Gaussian binned into cells of
//    S x S tiles, each cell's list in (depth bits, id) order.  A tile walks its
//    cell's list and keeps the Gaussians whose tile rectangle covers it: that
//    is exactly its reference tile list, in the same order, so images,
//    final_T and ncontrib are bit-identical -- but the K tile x Gaussian
//    pairs (most of which lie behind saturated pixels and are never read) are
//    neither written nor sorted.
//
// Work decomposition (gfx950): a one-wavefront workgroup covers 64 * PPL
// pixels of a 16x16 tile: PPL = 1 an 8x8 quadrant, 2 a 16x8 half, 4 the whole
// tile with 4 pixels per lane (rows r, r + 4, r + 8, r + 12).  The wave stages
// up to 64 Gaussians at a time in LDS as 16-byte records and every lane walks
// them for its PPL pixels; it stops when all its pixels are saturated (a
// ballot) and needs no workgroup barrier.  More pixels per lane amortise the
// staging reads and, in the adjoint, divide the per-Gaussian wave reductions
// and global atomics by PPL.  Workgroups are dispatched round-robin over the
// 8 XCDs; wave w = (b % 8) * per_xcd + b / 8 gives each XCD a contiguous band
// of tiles, so neighbouring tiles (which share most Gaussians) share an L2.
//
// All C channels of a pass are blended in one walk (the reference re-walks
// the list for every 32-channel chunk).
#pragma once

#include "common.hpp"

namespace msplat {

constexpr int kBlendLanes = 64;

// Largest pixels-per-lane a channel bucket is instantiated for (the adjoint
// keeps 3 * CMAX + 5 registers per pixel).
MSPLAT_HD constexpr int blend_max_ppl(int cmax) { return cmax <= 8 ? 4 : (cmax <= 16 ? 2 : 1); }

// Everything a blending kernel reads.  Per-view arrays (uv, conic, rect, the
// lists' entries and gradients of those) are indexed by the per-view row vg;
// opacity by the Gaussian gbase[b] + (vg - vbase[b]); feature rows by vg when
// feat_per_view, else by that Gaussian.
struct BlendIn {
  const float *uv, *conic, *opacity, *feature;
  int C, c0;               // channels, first channel of this pass
  const int32_t *list;     // idx_sorted, or cell lists
  const int32_t *range;    // tile_range [T, 2], or cell_range [cells, 2]
  const uint32_t *rect;    // cell lists: tile rectangle per Gaussian (x0 | y0 << 16, x1 | y1 << 16)
  int cell;                // cell lists: cell edge in tiles
  int cull;                // cell lists: skip Gaussians that reach no pixel of the tile (may_touch)
  float bg;
  int feat_per_view;
  int ntiles, per_xcd;     // global tiles of all views; workgroups per XCD band
  int ngroups;             // global 2 x 2 tile groups (grouped adjoint)
  Views views;
  // Fused squared-error loss against a target image of the same layout
  // (optional).  Forward: each wave stores sum (rendered - target)^2 over
  // its pixels and the pass's channels in l2_part[pass * l2_slots + wave].
  // Adjoint: dL/drendered = 2 (rendered - target) dL[0], plus dout when given.
  const float *l2_target;
  float *l2_part;
  int64_t l2_slots;
  const float *l2_rendered, *l2_dL;
};

template <int PPL>
struct PixBlock {
  int tile, view, tx, ty;  // global tile, its view, tile coordinates in the view
  int sub;                 // which of the tile's 4 / PPL pixel blocks
  int W, H;
  int64_t pix0;            // first pixel of the view in the concatenated planes
  int64_t vbase;           // first per-view row of the view
  int64_t gshift;          // Gaussian of per-view row vg = vg + gshift
  bool active;
  int px[PPL], py[PPL];
};

template <int PPL>
MSPLAT_DEV PixBlock<PPL> pix_block(const BlendIn &in) {
  constexpr int R = 4 / PPL;  // waves per tile
  const int b = blockIdx.x;
  const int w = (b & 7) * in.per_xcd + (b >> 3);
  PixBlock<PPL> p;
  p.tile = w / R;
  const int sub = w % R;
  p.sub = sub;
  p.active = p.tile < in.ntiles;
  p.view = p.active ? view_of_tile(in.views, p.tile) : 0;
  const int local = p.active ? p.tile - in.views.tile0[p.view] : 0;
  const int gx = in.views.gx[p.view];
  p.tx = local % gx;
  p.ty = local / gx;
  p.W = in.views.W[p.view];
  p.H = in.views.H[p.view];
  p.pix0 = in.views.pix0[p.view];
  p.vbase = in.views.vbase[p.view];
  p.gshift = in.views.gbase[p.view] - p.vbase;
  const int lane = threadIdx.x;
#pragma unroll
  for (int i = 0; i < PPL; ++i) {
    if constexpr (PPL == 1) {
      p.px[i] = p.tx * kTile + (sub & 1) * 8 + (lane & 7);
      p.py[i] = p.ty * kTile + (sub >> 1) * 8 + (lane >> 3);
    } else {
      p.px[i] = p.tx * kTile + (lane & 15);
      p.py[i] = p.ty * kTile + sub * (kTile / R) + (lane >> 4) + 4 * i;
    }
  }
  return p;
}

// Pixel block of wave `wave` (0..3) in a workgroup covering a 2 x 2 tile
// group: the whole tile (PPL 4) at (2 gx + wave % 2, 2 gy + wave / 2).
MSPLAT_DEV PixBlock<4> pix_block_group(const BlendIn &in, int wave) {
  const int b = blockIdx.x;
  const int q = (b & 7) * in.per_xcd + (b >> 3);  // XCD bands, as for single tiles
  PixBlock<4> p;
  const bool live = q < in.ngroups;
  p.view = live ? view_of_group(in.views, q) : 0;
  const int gx = in.views.gx[p.view], gy = in.views.gy[p.view];
  const int local = live ? q - in.views.group0[p.view] : 0;
  const int ngx = (gx + 1) / 2;
  p.tx = 2 * (local % ngx) + (wave & 1);
  p.ty = 2 * (local / ngx) + (wave >> 1);
  p.active = live && p.tx < gx && p.ty < gy;
  p.tile = in.views.tile0[p.view] + p.ty * gx + p.tx;
  p.sub = 0;
  p.W = in.views.W[p.view];
  p.H = in.views.H[p.view];
  p.pix0 = in.views.pix0[p.view];
  p.vbase = in.views.vbase[p.view];
  p.gshift = in.views.gbase[p.view] - p.vbase;
  const int lane = threadIdx.x % kWave;
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    p.px[i] = p.tx * kTile + (lane & 15);
    p.py[i] = p.ty * kTile + (lane >> 4) + 4 * i;
  }
  return p;
}

// Synchronisation of one wave's staging: a workgroup barrier for one-wave
// workgroups (free), a wave-level one when several waves work independently.
template <int TG>
MSPLAT_DEV void stage_sync() {
  if constexpr (TG == 1) {
    __syncthreads();
  } else {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  }
}

// Workgroups per launch for a given pixels-per-lane (8 XCD bands of per_xcd).
inline int blend_per_xcd(int ntiles, int ppl) { return (ntiles * (4 / ppl) + 7) / 8; }

// 0.0f in a fresh VGPR that the compiler cannot treat as a constant.
MSPLAT_DEV float opaque_zero() {
  float z;
  asm volatile("v_mov_b32 %0, 0" : "=v"(z));
  return z;
}

// exp(x) as 2^(x log2 e) on the hardware exp2 (v_exp_f32).
MSPLAT_DEV float fast_exp(float x) { return __builtin_amdgcn_exp2f(x * 1.44269504088896341f); }

// One staged Gaussian is a 16-byte aligned record (u, v, a, b | c, o, f0, f1 |
// f2 ...), so the inner loops fetch it with 128-bit LDS broadcast reads.
template <int CMAX>
struct BlendStage {
  static constexpr int kQuads = (6 + CMAX + 3) / 4;  // float4s per record
  float4 rec[kBlendLanes * kQuads];
  int32_t id[kBlendLanes];   // Gaussian of each slot
  int32_t pos[kBlendLanes];  // cell lists: list index of each slot
  uint32_t strips[kBlendLanes];  // strips of the tile each slot may touch
};

// Feature row of per-view row vg (gshift: Gaussian = vg + gshift).
MSPLAT_DEV int64_t feature_row(const BlendIn &in, int64_t vg, int64_t gshift) {
  return in.feat_per_view ? vg : vg + gshift;
}

// Stage list entry vg into LDS slot `slot`: channels [c0, c0 + cc) of its row.
// Returns whether all those channels are finite.
template <int CMAX>
MSPLAT_DEV bool stage_gaussian(BlendStage<CMAX> &s, int slot, int vg, const BlendIn &in, int cc, int64_t gshift) {
  constexpr int NQ = BlendStage<CMAX>::kQuads;
  float r[NQ * 4];
  r[0] = in.uv[2 * int64_t(vg)];
  r[1] = in.uv[2 * int64_t(vg) + 1];
  r[2] = in.conic[3 * int64_t(vg)];
  r[3] = in.conic[3 * int64_t(vg) + 1];
  r[4] = in.conic[3 * int64_t(vg) + 2];
  r[5] = in.opacity[vg + gshift];
  const float *fr = in.feature + feature_row(in, vg, gshift) * in.C + in.c0;
#pragma unroll
  for (int i = 6; i < NQ * 4; ++i) r[i] = (i - 6 < cc) ? fr[i - 6] : 0.f;
#pragma unroll
  for (int q = 0; q < NQ; ++q) s.rec[slot * NQ + q] = make_float4(r[4 * q], r[4 * q + 1], r[4 * q + 2], r[4 * q + 3]);
  s.id[slot] = vg;
  bool finite = true;
#pragma unroll
  for (int i = 6; i < NQ * 4; ++i) finite = finite && __builtin_isfinite(r[i]);
  return finite;
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

// Lanes below this one in a ballot mask.
MSPLAT_DEV int mask_rank(uint64_t m) {
  return __builtin_amdgcn_mbcnt_hi(uint32_t(m >> 32), __builtin_amdgcn_mbcnt_lo(uint32_t(m), 0u));
}

// Exact culling for cell lists.  A Gaussian is skipped for a pixel box (a
// tile, or one of its four 16x4 strips) only if no pixel centre in the box can
// reach alpha >= 1/255: its largest power over the box's pixel-centre
// rectangle -- q = a dx^2 + 2 b dx dy + c dy^2 is convex, so off-centre its
// minimum lies on an edge of the box -- must stay below -ln(255 o) by a margin
// (0.05 in power, alpha x 0.95) far above the fp32 rounding of the blending's
// own power and exp2; degenerate or non-finite parameters are never culled.
// Culled Gaussians change neither colour nor transmittance, so images,
// final_T and gradients are unchanged; only the list positions counted in
// ncontrib differ.
struct CullGeom {
  float u, v, a, b, c, lim;
  float ba, bc;  // b / a, b / c: where q is least along a vertical / horizontal edge
  bool regular;  // false: never cull
};

MSPLAT_DEV CullGeom cull_geom(const BlendIn &in, int vg, int64_t gshift) {
  CullGeom g;
  g.u = in.uv[2 * int64_t(vg)];
  g.v = in.uv[2 * int64_t(vg) + 1];
  g.a = in.conic[3 * int64_t(vg)];
  g.b = in.conic[3 * int64_t(vg) + 1];
  g.c = in.conic[3 * int64_t(vg) + 2];
  const float o = in.opacity[vg + gshift];
  g.regular = g.a > 0.f && g.c > 0.f && g.a * g.c - g.b * g.b > 0.f && o > 0.f && o < 1e30f &&
              fabsf(g.u) < 1e30f && fabsf(g.v) < 1e30f;
  // alpha < 1/255  <=  power = -q / 2 < -ln(255 o) - 0.05  <=>  q > lim
  g.lim = g.regular ? 2.f * logf(255.f * o) + 0.1f : 0.f;
  g.ba = g.regular ? g.b / g.a : 0.f;
  g.bc = g.regular ? g.b / g.c : 0.f;
  return g;
}

// May a pixel centre p in [px0, px1] x [py0, py1] get alpha >= 1/255?
MSPLAT_DEV bool box_may_touch(const CullGeom &g, float px0, float px1, float py0, float py1) {
  if (!g.regular) return true;
  const float x0 = g.u - px1, x1 = g.u - px0;  // d = mean - pixel
  const float y0 = g.v - py1, y1 = g.v - py0;
  if (x0 <= 0.f && x1 >= 0.f && y0 <= 0.f && y1 >= 0.f) return true;
  const float a = g.a, b = g.b, c = g.c;
  auto q = [&](float dx, float dy) { return a * dx * dx + 2.f * b * dx * dy + c * dy * dy; };
  // the least q on an edge: at the clamped stationary point (the divisions are
  // hoisted into cull_geom; a rounding difference in where q is evaluated
  // changes it only to second order, far inside the margin)
  auto on_x = [&](float dx) { return q(dx, fminf(fmaxf(-g.bc * dx, y0), y1)); };
  auto on_y = [&](float dy) { return q(fminf(fmaxf(-g.ba * dy, x0), x1), dy); };
  const float qmin = fminf(fminf(on_x(x0), on_x(x1)), fminf(on_y(y0), on_y(y1)));
  return !(qmin > g.lim);
}

// Cell-list filter: 0 if the Gaussian is skipped for tile (tx, ty), else the
// mask of the tile's 16x4 strips (bit s: rows 4s..4s+3) it may touch -- all
// four without culling.
MSPLAT_DEV uint32_t walk_mask(const BlendIn &in, int vg, int64_t gshift, int tx, int ty) {
  if (!rect_covers(reinterpret_cast<const uint2 *>(in.rect)[vg], tx, ty)) return 0u;
  if (!in.cull) return 0xfu;
  const CullGeom g = cull_geom(in, vg, gshift);
  const float px0 = float(kTile * tx), px1 = float(kTile * tx + kTile - 1), py0 = float(kTile * ty);
  if (!box_may_touch(g, px0, px1, py0, py0 + float(kTile - 1))) return 0u;
  uint32_t m = 0u;
#pragma unroll
  for (int s = 0; s < 4; ++s)
    if (box_may_touch(g, px0, px1, py0 + float(4 * s), py0 + float(4 * s + 3))) m |= 1u << s;
  return m;
}

// Strips (of the mask above) holding pixel slot i of a wave's lanes.
template <int PPL>
MSPLAT_DEV uint32_t slot_strips(int sub, int i) {
  if constexpr (PPL == 1) return 3u << (2 * (sub >> 1));  // an 8x8 quadrant: rows of two strips
  else return 1u << ((PPL == 2 ? 2 * sub : 0) + i);  // 16x4 strips: rows 4 i (+ 8 sub)
}

// The list a tile walks: [start, end) of its tile list or of its cell's list.
template <bool kWalk, int PPL>
MSPLAT_DEV int2 tile_list(const BlendIn &in, const PixBlock<PPL> &pb) {
  const int c = kWalk ? in.views.cell0[pb.view] + (pb.ty / in.cell) * in.views.ncx[pb.view] + pb.tx / in.cell
                      : pb.tile;
  return make_int2(in.range[2 * c], in.range[2 * c + 1]);
}

// power = -0.5 (a dx^2 + c dy^2) - b dx dy with d = mean - pixel
// (alpha_blending.cu:74-82), with two explicit fused multiply-adds (the
// blending kernels are bound by VALU issue).  Written out rather than left to
// contraction, so that every kernel instantiation -- forward and adjoint, any
// pixels-per-lane, tile or cell lists -- rounds it identically and they all
// agree on every alpha and every skip.
MSPLAT_DEV float gauss_power(float a, float b, float c, float dx, float dy) {
  const float q = __builtin_fmaf(a * dx, dx, (c * dy) * dy);
  return __builtin_fmaf(-0.5f, q, -((b * dx) * dy));
}

// gauss_power for the pixels of one column (a lane's pixels share dx): the
// dx products computed once, the same roundings as gauss_power.
struct PowerCol {
  float dx, adx, bdx;
};
MSPLAT_DEV PowerCol power_col(float a, float b, float dx) { return PowerCol{dx, a * dx, b * dx}; }
MSPLAT_DEV float gauss_power_col(const PowerCol &p, float c, float dy) {
  const float q = __builtin_fmaf(p.adx, p.dx, (c * dy) * dy);
  return __builtin_fmaf(-0.5f, q, -(p.bdx * dy));
}

// Forward.  Outputs rendered [C, H, W] (+ bg * T), final_T and ncontrib
// [H, W] (the 1-based list position of each pixel's last contributor).  Cell
// lists also record, per tile, the end of the walk that the adjoint starts
// from: walk[2 t] = list index after the last contributor of any pixel of the
// tile, walk[2 t + 1] = that contributor's position in the tile's list
// (atomic maxima; the caller zero-fills walk).
// Up to 4 channels both blending kernels are held to 80 VGPRs (6 waves per
// SIMD instead of 5 at the compiler's own 86-89): no spills, and the extra
// waves hide more of the staged loads' latency (the 1080p training step
// measured 13 us shorter).  More channels would spill under that bound.
constexpr int blend_min_waves(int cmax) { return cmax <= 4 ? 6 : 1; }

// The forward's output when a sort or binning on this device had an expired
// look-back wait (its lists are undefined and are not read): NaN image and
// transmittance, no contributor, NaN loss part.
template <int PPL>
MSPLAT_DEV void blend_fwd_undefined(const BlendIn &in, const PixBlock<PPL> &pb, int cc, float *out, float *final_T,
                                    int32_t *ncontrib) {
  const float nan = __builtin_nanf("");
  const int64_t hw = int64_t(pb.H) * pb.W;
#pragma unroll
  for (int i = 0; i < PPL; ++i) {
    if (!(pb.px[i] < pb.W && pb.py[i] < pb.H)) continue;
    const int64_t pix = int64_t(pb.py[i]) * pb.W + pb.px[i];
    if (in.c0 == 0) {
      final_T[pb.pix0 + pix] = nan;
      ncontrib[pb.pix0 + pix] = 0;
    }
    for (int ch = 0; ch < cc; ++ch) out[int64_t(in.C) * pb.pix0 + (in.c0 + ch) * hw + pix] = nan;
  }
  if (in.l2_part && threadIdx.x == 0)
    in.l2_part[int64_t(in.c0 / 64) * in.l2_slots + (blockIdx.x & 7) * in.per_xcd + (blockIdx.x >> 3)] = nan;
}

template <int CMAX, int PPL, bool kWalk>
__global__ __launch_bounds__(kBlendLanes) __attribute__((amdgpu_waves_per_eu(blend_min_waves(CMAX)))) void
blend_fwd_kernel(BlendIn in, float *__restrict__ out, float *__restrict__ final_T, int32_t *__restrict__ ncontrib,
                 int32_t *__restrict__ walk) {
  __shared__ BlendStage<CMAX> s;
  const PixBlock<PPL> pb = pix_block<PPL>(in);
  if (!pb.active) return;
  const int lane = threadIdx.x;
  const int cc = in.C - in.c0 < CMAX ? in.C - in.c0 : CMAX;  // channels of this pass
  if (expired_state() != 0u) {  // uniform: lists left undefined by an expired wait
    blend_fwd_undefined<PPL>(in, pb, cc, out, final_T, ncontrib);
    return;
  }
  const int2 lst = tile_list<kWalk>(in, pb);
  int pos = lst.x;
  const int end = lst.y;

  // T[i] > 0 is the pixel's transmittance; a saturated pixel (and one
  // outside the image) holds -T, its final transmittance negated, so one
  // compare tells whether the pixel still blends.
  float T[PPL], fx[PPL], fy[PPL];
  int last[PPL];
  int we = 0;  // cell lists: list index after the wave's last contribution
  float F[PPL][CMAX];
#pragma unroll
  for (int i = 0; i < PPL; ++i) {
    fx[i] = float(pb.px[i]);
    fy[i] = float(pb.py[i]);
    T[i] = (pb.px[i] < pb.W && pb.py[i] < pb.H) ? 1.f : -1.f;
    last[i] = 0;
#pragma unroll
    for (int ch = 0; ch < CMAX; ++ch) F[i][ch] = 0.f;
  }

  int fpos = 0;  // position in the tile list of the first staged Gaussian
  while (pos < end) {
    bool alive = false;
#pragma unroll
    for (int i = 0; i < PPL; ++i) alive |= T[i] > 0.f;
    if (__ballot(alive) == 0ull) break;
    int cnt;
    bool nonfinite = false;  // a staged Gaussian has a non-finite feature
    __syncthreads();
    if constexpr (kWalk) {
      const int e = pos + lane;
      int g = 0;
      bool pass = false;
      uint32_t strips = 0u;
      if (e < end) {
        g = in.list[e];
        strips = walk_mask(in, g, pb.gshift, pb.tx, pb.ty);
        pass = strips != 0u;
      }
      const uint64_t m = __ballot(pass);
      cnt = __popcll(m);
      if (pass) {
        const int slot = mask_rank(m);
        nonfinite = !stage_gaussian<CMAX>(s, slot, g, in, cc, pb.gshift);
        s.pos[slot] = e;
        s.strips[slot] = strips;
      }
      pos += kBlendLanes;
    } else {
      cnt = end - pos < kBlendLanes ? end - pos : kBlendLanes;
      if (lane < cnt) {
        nonfinite = !stage_gaussian<CMAX>(s, lane, in.list[pos + lane], in, cc, pb.gshift);
        s.strips[lane] = 0xfu;
      }
      pos += cnt;
    }
    __syncthreads();
    // Branch-free per pixel: every quantity is computed and the reference's
    // early-outs become selects, so each update happens exactly where the
    // reference's would (with the same operations) and nowhere else.  With
    // finite features, f w' + F with w' = 0 leaves F unchanged (F starts at
    // +0 and is never -0), so one select of the weight serves all channels; a
    // batch holding a non-finite feature (f 0 = NaN) selects per channel.
    auto walk_batch = [&](auto finite) {
    for (int j = 0; j < cnt; ++j) {
      if ((j & 7) == 7) {  // every 8 Gaussians: stop once every pixel of the wave is saturated
        bool live = false;
#pragma unroll
        for (int i = 0; i < PPL; ++i) live |= T[i] > 0.f;
        if (__ballot(live) == 0ull) break;
      }
      float4 g0, g1;
      stage_geom(s, j, g0, g1);
      float f[CMAX];
      stage_features(s, j, g1, f);
      const int contributor = fpos + j + 1;
      const uint32_t strips = __builtin_amdgcn_readfirstlane(s.strips[j]);
      const PowerCol col = power_col(g0.z, g0.w, g0.x - fx[0]);  // a lane's pixels share one column
#pragma unroll
        for (int i = 0; i < PPL; ++i) {
          if (!(strips & slot_strips<PPL>(pb.sub, i))) continue;  // uniform: culled for these rows
          const float power = gauss_power_col(col, g1.x, g0.y - fy[i]);
          const float alpha = fminf(0.99f, g1.y * fast_exp(power));
          // a NaN power is not skipped, as in the reference
          const bool hit = T[i] > 0.f && !(power > 0.f) && !(alpha < 1.f / 255.f);
          const float nT = T[i] * (1.f - alpha);
          const bool low = nT < 0.0001f;  // with hit: the pixel saturates here
          const bool add = hit && !low;
          const float w = alpha * T[i];
          if constexpr (decltype(finite)::value) {
            const float wa = add ? w : 0.f;
#pragma unroll
            for (int ch = 0; ch < CMAX; ++ch)  // channels >= cc are staged as 0 and never stored
              F[i][ch] = __builtin_fmaf(f[ch], wa, F[i][ch]);
          } else {
#pragma unroll
            for (int ch = 0; ch < CMAX; ++ch)
              F[i][ch] = add ? __builtin_fmaf(f[ch], w, F[i][ch]) : F[i][ch];
          }
          T[i] = hit ? (low ? -T[i] : nT) : T[i];
          last[i] = add ? contributor : last[i];
        }
    }
    };
    if (__ballot(nonfinite) == 0ull)  // uniform
      walk_batch(std::true_type{});
    else
      walk_batch(std::false_type{});
    if constexpr (kWalk) {  // the wave's last contributor so far: its list index, if staged in this batch
      int lm = 0;
#pragma unroll
      for (int i = 0; i < PPL; ++i) lm = imax(lm, last[i]);
      lm = wave_max(lm);
      if (lm > fpos) we = s.pos[lm - 1 - fpos] + 1;
    }
    fpos += cnt;
  }
  const int64_t hw = int64_t(pb.H) * pb.W;
  float *vout = out + int64_t(in.C) * pb.pix0;  // this view's [C, H, W]
  const float *vtgt = in.l2_target ? in.l2_target + int64_t(in.C) * pb.pix0 : nullptr;
  float lsum = 0.f;
#pragma unroll
  for (int i = 0; i < PPL; ++i) {
    if (!(pb.px[i] < pb.W && pb.py[i] < pb.H)) continue;
    const int64_t pix = int64_t(pb.py[i]) * pb.W + pb.px[i];
    const float Tf = fabsf(T[i]);
    if (in.c0 == 0) {
      final_T[pb.pix0 + pix] = Tf;
      ncontrib[pb.pix0 + pix] = last[i];
    }
#pragma unroll
    for (int ch = 0; ch < CMAX; ++ch) {
      if (ch >= cc) continue;
      const int64_t off = (in.c0 + ch) * hw + pix;
      const float val = F[i][ch] + Tf * in.bg;
      vout[off] = val;
      if (vtgt) {
        const float d = val - vtgt[off];
        lsum += d * d;
      }
    }
  }
  if (in.l2_part) {  // uniform
    lsum = wave_sum(lsum);
    if (lane == 0) in.l2_part[int64_t(in.c0 / 64) * in.l2_slots + (blockIdx.x & 7) * in.per_xcd + (blockIdx.x >> 3)] = lsum;
  }
  if constexpr (kWalk) {
    int wl = 0;
#pragma unroll
    for (int i = 0; i < PPL; ++i) wl = imax(wl, last[i]);
    wl = wave_max(wl);
    if (lane == 0 && in.c0 == 0 && wl > 0) {
      atomicMax(walk + 2 * pb.tile, we);
      atomicMax(walk + 2 * pb.tile + 1, wl);
    }
  }
}

// Adjoint: walk each pixel's contributors back to front, recovering T by
// division (alpha_blending.cu:155-245).  Per Gaussian, every lane first sums
// its PPL pixels' contributions, then the wave sums its 64 lanes.
//
// TG = 1: one wave per workgroup; each wave adds its sums to the global
// gradients, one atomic per value per Gaussian.  Those atomics -- the same
// Gaussians from every tile it covers -- cost about a third of the kernel.
// TG = 4 (cell lists, PPL 4, even cells): a workgroup's 4 waves cover a 2 x 2
// group of tiles of one cell, so they walk the same cell list; each wave
// adds its sums into an LDS accumulator indexed by cell-list position (the
// first kGroupAcc positions of the cell) and the last wave to finish adds
// the group's totals to the global gradients: up to 4 times fewer global
// atomics.  The waves stage and walk independently; only the accumulator is
// shared.
constexpr int kGroupAcc = 256;

template <int CMAX, int PPL, bool kWalk, int TG = 1>
__global__ __launch_bounds__(kBlendLanes * TG) __attribute__((amdgpu_waves_per_eu(blend_min_waves(CMAX)))) void
blend_bwd_kernel(
    BlendIn in, const float *__restrict__ final_T, const int32_t *__restrict__ ncontrib,
    const int32_t *__restrict__ walk, const float *__restrict__ dout, float *__restrict__ duv,
    float *__restrict__ dconic, float *__restrict__ dopac, float *__restrict__ dfeat) {
  // The gradient arithmetic may contract into fused multiply-adds (gradients
  // are compared to a tolerance); every value the forward also computes
  // (power, alpha, T) goes through explicitly rounded code.
#pragma clang fp contract(fast)
  static_assert(TG == 1 || (TG == 4 && PPL == 4 && kWalk), "tile groups: cell lists, 4 pixels per lane");
  // lists left undefined by an expired wait are not read: the gradients keep
  // their zero fill (the forward wrote NaN)
  if (expired_state() != 0u) return;  // uniform over the workgroup
  constexpr int NVA = 6 + CMAX;                     // accumulated values per Gaussian
  constexpr int kAcc = TG > 1 ? kGroupAcc : 1;
  __shared__ BlendStage<CMAX> stage_all[TG];
  __shared__ float acc[kAcc * NVA];
  __shared__ int32_t acc_id[kAcc];
  __shared__ uint32_t waves_done;
  const int wave = TG > 1 ? int(threadIdx.x / kBlendLanes) : 0;
  const int lane = TG > 1 ? int(threadIdx.x % kBlendLanes) : int(threadIdx.x);
  BlendStage<CMAX> &s = stage_all[wave];
  const int cc = in.C - in.c0 < CMAX ? in.C - in.c0 : CMAX;  // channels of this pass
  PixBlock<PPL> pb;
  if constexpr (TG > 1) {
    pb = pix_block_group(in, wave);
    for (int k = threadIdx.x; k < kAcc * NVA; k += kBlendLanes * TG) acc[k] = 0.f;
    for (int k = threadIdx.x; k < kAcc; k += kBlendLanes * TG) acc_id[k] = -1;
    if (threadIdx.x == 0) waves_done = 0u;
    __syncthreads();  // the accumulator is clear before any wave adds to it
  } else {
    pb = pix_block<PPL>(in);
  }
  // The last wave of a group to finish adds the accumulated totals to the
  // global gradients (every wave's LDS additions precede its increment).
  auto finish_group = [&]() {
    if constexpr (TG > 1) {
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
      uint32_t before = 0u;
      if (lane == 0) before = atomicAdd(&waves_done, 1u);
      before = __shfl(before, 0, kWave);
      if (before != uint32_t(TG - 1)) return;
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
      for (int k = lane; k < kAcc * NVA; k += kWave) {
        const int slot = k / NVA, v = k - slot * NVA;
        const int gidx = acc_id[slot];
        const float val = acc[k];
        if (gidx < 0 || v >= 6 + cc || val == 0.f) continue;
        float *dst;
        if (v < 2) dst = duv + 2 * int64_t(gidx) + v;
        else if (v < 5) dst = dconic + 3 * int64_t(gidx) + (v - 2);
        else if (v == 5) dst = dopac + (gidx + pb.gshift);
        else dst = dfeat + feature_row(in, gidx, pb.gshift) * in.C + in.c0 + (v - 6);
        unsafeAtomicAdd(dst, val);
      }
    }
  };
  if (!pb.active) {
    finish_group();
    return;
  }
  const int64_t hw = int64_t(pb.H) * pb.W;
  const float *vdout = dout ? dout + int64_t(in.C) * pb.pix0 : nullptr;  // this view's [C, H, W]
  const float *vrend = in.l2_dL ? in.l2_rendered + int64_t(in.C) * pb.pix0 : nullptr;
  const float *vtgt = in.l2_dL ? in.l2_target + int64_t(in.C) * pb.pix0 : nullptr;
  const float l2s = in.l2_dL ? 2.f * in.l2_dL[0] : 0.f;

  // Per pixel: T (recovered back to front), tb = -T_final (bg . dL/dpixel)
  // for the background term, the channels' dL/dpixel g and
  // A = accum . g, the colour accumulated behind the current Gaussian
  // (alpha_blending.cu:213-214) dotted with g.  The adjoint needs accum only
  // through dL/dalpha = sum_c (f_c - accum_c) g_c = f . g - A, and the
  // recurrence accum' = accum + alpha (f - accum) carries over to A' = A +
  // alpha (f . g - A): one scalar per pixel instead of C colour sums.
  float T[PPL], tb[PPL], fx[PPL], fy[PPL], A[PPL];
  int last[PPL];
  float g[PPL][CMAX];
  int wl = 0;
#pragma unroll
  for (int i = 0; i < PPL; ++i) {
    const bool inside = pb.px[i] < pb.W && pb.py[i] < pb.H;
    const int64_t pix = int64_t(pb.py[i]) * pb.W + pb.px[i];
    const float T_final = inside ? final_T[pb.pix0 + pix] : 0.f;
    fx[i] = float(pb.px[i]);
    fy[i] = float(pb.py[i]);
    last[i] = inside ? ncontrib[pb.pix0 + pix] : 0;
    wl = imax(wl, last[i]);
    float bg_dot = 0.f;
#pragma unroll
    for (int ch = 0; ch < CMAX; ++ch) {
      float gv = 0.f;
      if (inside && ch < cc) {
        const int64_t off = (in.c0 + ch) * hw + pix;
        if (vrend) {  // the fused loss's gradient (+ the image's own, as autograd would add them)
          gv = (vrend[off] - vtgt[off]) * l2s;
          if (vdout) gv += vdout[off];
        } else {
          gv = vdout[off];
        }
      }
      g[i][ch] = gv;
      if (ch < cc) bg_dot += in.bg * g[i][ch];
    }
    A[i] = 0.f;
    T[i] = T_final;
    tb[i] = -T_final * bg_dot;
  }
  // Wave-uniform bounds are read into scalar registers: with several waves
  // per workgroup the compiler cannot tell they are uniform, and the
  // per-Gaussian and per-slot skips below would become per-lane masks.
  wl = __builtin_amdgcn_readfirstlane(wave_max(wl));
  int wl_slot[PPL];  // per pixel slot: the wave's largest list end
#pragma unroll
  for (int i = 0; i < PPL; ++i) wl_slot[i] = __builtin_amdgcn_readfirstlane(wave_max(last[i]));

  const int2 lst = tile_list<kWalk>(in, pb);
  const int lo_bound = __builtin_amdgcn_readfirstlane(lst.x);
  int hi, fcount;  // walk backwards over [lo_bound, hi); fcount Gaussians of the tile list lie below hi
  if constexpr (kWalk) {
    hi = __builtin_amdgcn_readfirstlane(imin(walk[2 * pb.tile], lst.y));
    fcount = __builtin_amdgcn_readfirstlane(walk[2 * pb.tile + 1]);
  } else {
    hi = __builtin_amdgcn_readfirstlane(imin(lst.x + wl, lst.y));
    fcount = hi - lo_bound;
  }

  // After the transposed reduction, lane 16 r + k holds gradient value
  // v = 4 (k + 16 q) + kRowValue[r] in round q.  Its target array, offset and
  // row stride depend on the lane only: the address of a Gaussian's value is
  // then base + gidx * stride.
  constexpr int NV = ((6 + CMAX) + 3) / 4 * 4;
  constexpr int NQ = (NV / 4 + 15) / 16;  // atomic rounds
  float *tbase[NQ];
  int tstride[NQ], tval[NQ];
  bool tvalid[NQ];
#pragma unroll
  for (int q = 0; q < NQ; ++q) {
    const int vec = (lane & 15) + 16 * q;
    const int v = 4 * vec + row_value(lane >> 4);
    tval[q] = v;
    tvalid[q] = vec < NV / 4 && v < 6 + cc;
    const int64_t gsh = pb.gshift;
    if (v < 2) {
      tbase[q] = duv + v;
      tstride[q] = 2;
    } else if (v < 5) {
      tbase[q] = dconic + (v - 2);
      tstride[q] = 3;
    } else if (v == 5) {
      tbase[q] = dopac + gsh;
      tstride[q] = 1;
    } else {
      tbase[q] = dfeat + (in.feat_per_view ? 0 : gsh * in.C) + in.c0 + (v - 6);
      tstride[q] = in.C;
    }
  }

  while (fcount > 0 && hi > lo_bound) {
    const int lo = hi - kBlendLanes > lo_bound ? hi - kBlendLanes : lo_bound;
    int cnt;
    stage_sync<TG>();
    if constexpr (kWalk) {
      const int e = lo + lane;
      int gi = 0;
      bool pass = false;
      uint32_t strips = 0u;
      if (e < hi) {
        gi = in.list[e];
        strips = walk_mask(in, gi, pb.gshift, pb.tx, pb.ty);
        pass = strips != 0u;
      }
      const uint64_t m = __ballot(pass);
      cnt = __popcll(m);
      if (pass) {
        const int slot = mask_rank(m);
        stage_gaussian<CMAX>(s, slot, gi, in, cc, pb.gshift);
        s.strips[slot] = strips;
        if constexpr (TG > 1) s.pos[slot] = e;
      }
    } else {
      cnt = hi - lo;
      if (lane < cnt) {
        stage_gaussian<CMAX>(s, lane, in.list[lo + lane], in, cc, pb.gshift);
        s.strips[lane] = 0xfu;
      }
    }
    const int fbase = fcount - cnt;  // tile-list position of slot 0
    stage_sync<TG>();
    for (int j = cnt - 1; j >= 0; --j) {
      const int rel = fbase + j;  // 0-based position in the tile list
      if (rel >= wl) continue;    // uniform: no pixel of this wave reaches it
      float4 g0, g1;
      stage_geom(s, j, g0, g1);
      const float a = g0.z, b = g0.w, c = g1.x, op = g1.y;
      float fe[CMAX];
      stage_features(s, j, g1, fe);
      // The lane's pixels share one column (PPL 2 and 4 stack rows), so dx is
      // common to them and the conic and mean gradients reduce to three
      // per-lane moments of w = G dL/dG over its pixels: sum w, sum w dy and
      // sum w dy^2.  Accumulators start from an opaque zero: with a literal 0
      // the compiler rematerialises the constant on every skip path.
      const float dx = g0.x - fx[0];
      // Per lane: the moments of t = G dL/dalpha (= dL/dG / opacity) over its
      // pixels, sum t, sum t dy and sum t dy^2, and the colour sums.  The
      // opacity factor of the conic and mean gradients is applied once, to
      // the wave totals.
      float m0 = opaque_zero(), m1 = opaque_zero(), m2 = opaque_zero();
      float vf[CMAX];
#pragma unroll
      for (int ch = 0; ch < CMAX; ++ch) vf[ch] = opaque_zero();
      bool any = false;  // the lane took a contribution
      const uint32_t strips = __builtin_amdgcn_readfirstlane(s.strips[j]);
#pragma unroll
      for (int i = 0; i < PPL; ++i) {
        if (!(strips & slot_strips<PPL>(pb.sub, i))) continue;  // uniform: culled for these rows
        if (rel >= wl_slot[i]) continue;                         // uniform: every pixel of the slot ended
        // The reference's three early-outs (position, power > 0, alpha <
        // 1/255) folded into one branch: alpha is cheap next to the gradient
        // terms, and one level of divergence keeps the accumulators out of
        // per-level zero copies.
        const float dy = g0.y - fy[i];
        const float power = gauss_power(a, b, c, dx, dy);
        const float G = fast_exp(power);
        const float alpha = fminf(0.99f, op * G);
        if (!(rel < last[i]) || power > 0.f || alpha < 1.f / 255.f) continue;
        any = true;
        // 1 / (1 - alpha) on the hardware reciprocal (the reference divides
        // with fast-math); 1 - alpha >= 0.01, so it is well conditioned.
        const float inv = __builtin_amdgcn_rcpf(1.f - alpha);
        T[i] = T[i] * inv;
        const float dcol = alpha * T[i];
        float fg = 0.f;  // f . g
#pragma unroll
        for (int ch = 0; ch < CMAX; ++ch) {  // channels >= cc carry zeros
          fg += fe[ch] * g[i][ch];
          vf[ch] += dcol * g[i][ch];
        }
        const float d = fg - A[i];
        A[i] += alpha * d;
        float dL_dalpha = d * T[i];
        dL_dalpha += tb[i] * inv;
        const float t = G * dL_dalpha;
        const float ty = t * dy;
        m0 += t;
        m1 += ty;
        m2 += ty * dy;
      }
      if (__ballot(any) == 0ull) continue;
      // With w = o t (= G dL/dG): dL/duv = -(sum w (a dx + b dy), sum w (b dx
      // + c dy)), dL/dconic = -(sum w dx^2 / 2, sum w dx dy, sum w dy^2 / 2),
      // dL/do = sum t.  The 6 + CMAX per-lane values are summed over the wave
      // in one transposed reduction, then each lane scales its total (-o for
      // the first five) and adds it to its target.
      const int gidx = s.id[j];
      constexpr int NR = 6 + CMAX;
      float vals[NR];
      vals[0] = a * dx * m0 + b * m1;
      vals[1] = b * dx * m0 + c * m1;
      vals[2] = 0.5f * dx * dx * m0;
      vals[3] = dx * m1;
      vals[4] = 0.5f * m2;
      vals[5] = m0;
#pragma unroll
      for (int ch = 0; ch < CMAX; ++ch) vals[6 + ch] = vf[ch];
      float tot[(NR + 3) / 4];
      wave_sum_transposed<NR>(vals, tot);
#pragma unroll
      for (int q = 0; q < NQ; ++q) {
        const int vec = (lane & 15) + 16 * q;
        float mine = 0.f;
#pragma unroll
        for (int jj = 0; jj < (NR + 3) / 4; ++jj)
          if (jj == vec) mine = tot[jj];
        if (!tvalid[q]) continue;
        mine = tval[q] < 5 ? -op * mine : mine;
        if constexpr (TG > 1) {
          const int key = s.pos[j] - lo_bound;  // position in the group's cell list
          if (key < kAcc) {
            if (tval[q] == 0) acc_id[key] = gidx;
            atomicAdd(&acc[key * NVA + tval[q]], mine);
            continue;
          }
        }
        unsafeAtomicAdd(tbase[q] + int64_t(gidx) * tstride[q], mine);
      }
    }
    fcount = fbase;
    hi = lo;
  }
  finish_group();
}

}  // namespace msplat
