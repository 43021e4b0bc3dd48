// Real spherical harmonics of degree 0..10, forward and adjoint.
//
// The reference writes each of the 121 basis functions out as a polynomial
// in the direction (x, y, z) (msplat/src/compute_sh.cu:116-503) and
// differentiates those polynomials literally (:505-1598).  Every one of them
// has the structure
//
//     Y_lm = K_lm * A_m(x, y) * Q_l|m|(z, rho2)
//     A_m  = Re (x + i y)^m      (m >= 0)      Im (x + i y)^|m|   (m < 0)
//     K_lm = sqrt((2l+1)/(4 pi) * (l-|m|)!/(l+|m|)!)  (times sqrt 2 if m != 0)
//     Q    = (-1)^m d^m P_l / dz^m  with every r^2 replaced by rho2,
//
// where rho2 = x^2 + y^2 + z^2 in bands 0..3 (homogeneous forms) and rho2 = 1
// in bands 4..10.  We evaluate that structure with the associated-Legendre
// recurrence in z and the complex-power recurrence in (x, y) instead of 121
// expanded polynomials, and carry the partial derivatives through the same
// recurrences, so values and dL/ddir equal the reference's polynomials (and
// their partials) without requiring |dir| = 1.
#pragma once

#include "common.hpp"

namespace msplat {
namespace shc {

constexpr int kMaxDeg = 10;

constexpr double csqrt(double v) {
  if (v <= 0.0) return 0.0;
  double x = v > 1.0 ? v : 1.0;
  for (int i = 0; i < 200; ++i) x = 0.5 * (x + v / x);
  return x;
}
constexpr double cfact(int n) {
  double f = 1.0;
  for (int i = 2; i <= n; ++i) f *= i;
  return f;
}
constexpr double kPi = 3.14159265358979323846;
// K_lm including the sqrt(2) of the real combinations.
constexpr float knorm(int l, int m) {
  const double k = csqrt((2.0 * l + 1.0) / (4.0 * kPi) * cfact(l - m) / cfact(l + m));
  return static_cast<float>(m == 0 ? k : k * csqrt(2.0));
}
// Q_mm = (-1)^m (2m-1)!!
constexpr float qmm(int m) {
  double v = 1.0;
  for (int i = 1; i <= m; ++i) v *= -(2.0 * i - 1.0);
  return static_cast<float>(v);
}

// The constants, evaluated once at compile time (called with loop variables,
// knorm/qmm would otherwise run as double-precision code in every thread).
struct Table {
  float k[kMaxDeg + 1][kMaxDeg + 1];
  float q[kMaxDeg + 1];
};
constexpr Table make_table() {
  Table t{};
  for (int l = 0; l <= kMaxDeg; ++l)
    for (int m = 0; m <= l; ++m) t.k[l][m] = knorm(l, m);
  for (int m = 0; m <= kMaxDeg; ++m) t.q[m] = qmm(m);
  return t;
}
constexpr Table kTable = make_table();

}  // namespace shc

// Calls emit(k, Y, dY/dx, dY/dy, dY/dz) for every basis index k = l^2 + l + m
// of degree <= L.  With all loops unrolled, k is a compile-time constant at
// every call, so callers can index register arrays with it.
template <int L, typename Emit>
MSPLAT_DEV void sh_basis(float x, float y, float z, Emit &&emit) {
  float cr[L + 1], ci[L + 1];  // Re / Im of (x + i y)^m
  cr[0] = 1.f;
  ci[0] = 0.f;
#pragma unroll
  for (int m = 1; m <= L; ++m) {
    cr[m] = x * cr[m - 1] - y * ci[m - 1];
    ci[m] = x * ci[m - 1] + y * cr[m - 1];
  }
  const float rho2 = (x * x + y * y) + z * z;
#pragma unroll
  for (int m = 0; m <= L; ++m) {
    float q1 = 0.f, q2 = 0.f, d1 = 0.f, d2 = 0.f;  // Q_{l-1}, Q_{l-2} and d/dz
#pragma unroll
    for (int l = m; l <= L; ++l) {
      float q, dq;
      if (l == m) {
        q = shc::kTable.q[m];
        dq = 0.f;
      } else if (l == m + 1) {
        q = float(2 * m + 1) * z * q1;
        dq = float(2 * m + 1) * q1;
      } else {
        const float a = float(2 * l - 1) / float(l - m);
        const float b = float(l + m - 1) / float(l - m);
        q = a * z * q1 - b * q2;
        dq = a * (q1 + z * d1) - b * d2;
      }
      // Bands 2 and 3 use r^2 where the recurrence has 1 (only the three
      // functions with l - m >= 2 differ): (2,0), (3,1), (3,0).
      float qv = q, qx = 0.f, qy = 0.f, qz = dq;
      if (l == 2 && m == 0) {
        qv = 1.5f * z * z - 0.5f * rho2;
        qx = -x;
        qy = -y;
        qz = 2.f * z;
      } else if (l == 3 && m == 1) {
        qv = -7.5f * z * z + 1.5f * rho2;
        qx = 3.f * x;
        qy = 3.f * y;
        qz = -12.f * z;
      } else if (l == 3 && m == 0) {
        qv = 2.5f * z * z * z - 1.5f * z * rho2;
        qx = -3.f * x * z;
        qy = -3.f * y * z;
        qz = 4.5f * z * z - 1.5f * rho2;
      }
      const float K = shc::kTable.k[l][m];
      if (m == 0) {
        emit(l * l + l, K * qv, K * qx, K * qy, K * qz);
      } else {
        const float fm = float(m);
        // cos part: A = Re z^m, dA/dx = m Re z^{m-1}, dA/dy = -m Im z^{m-1}
        emit(l * l + l + m, K * (cr[m] * qv), K * (fm * cr[m - 1] * qv + cr[m] * qx),
             K * (-fm * ci[m - 1] * qv + cr[m] * qy), K * (cr[m] * qz));
        // sin part: A = Im z^m, dA/dx = m Im z^{m-1}, dA/dy = m Re z^{m-1}
        emit(l * l + l - m, K * (ci[m] * qv), K * (fm * ci[m - 1] * qv + ci[m] * qx),
             K * (fm * cr[m - 1] * qv + ci[m] * qy), K * (ci[m] * qz));
      }
      q2 = q1;
      q1 = q;
      d2 = d1;
      d1 = dq;
    }
  }
}

constexpr int kShThreads = 256;

// dL/ddir = sum_c sum_k (g_c row[c, k]) grad Y_k: one basis walk per channel,
// the inner sum in basis order and channels added first to last (the
// rounding order every SH adjoint kernel shares).
template <int L>
MSPLAT_DEV void sh_dir_grad(float x, float y, float z, const float *row, const float *g, int C, float &gx,
                            float &gy, float &gz) {
  constexpr int D = (L + 1) * (L + 1);
  for (int c = 0; c < C; ++c) {
    const float gc = g[c];
    const float *rc = row + c * D;
    float px = 0.f, py = 0.f, pz = 0.f;
    sh_basis<L>(x, y, z, [&](int k, float, float dx, float dy, float dz) {
      const float w = gc * rc[k];
      px += w * dx;
      py += w * dy;
      pz += w * dz;
    });
    gx += px;
    gy += py;
    gz += pz;
  }
}

// ---- direction sources --------------------------------------------------------
// The kernels read Gaussian i's direction through `load` and hand back the
// direction gradient through `store_grad` (`store_zero` for an invisible
// Gaussian).

// Directions given as an [n,3] array (msplat/compute_sh.py:8-33); the
// gradient is dL/ddir [n,3].
struct ShDirArray {
  const float *dirs;
  float *grad;
  MSPLAT_DEV void load(int64_t i, float &x, float &y, float &z) const {
    x = dirs[3 * i];
    y = dirs[3 * i + 1];
    z = dirs[3 * i + 2];
  }
  MSPLAT_DEV void store_grad(int64_t i, float gx, float gy, float gz) const {
    grad[3 * i] = gx;
    grad[3 * i + 1] = gy;
    grad[3 * i + 2] = gz;
  }
  MSPLAT_DEV void store_zero(int64_t i) const { store_grad(i, 0.f, 0.f, 0.f); }
  // a Gaussian whose upstream gradient is zero in every channel
  MSPLAT_DEV void store_unreached(int64_t i) const { store_zero(i); }
};

// View directions from the Gaussian centres and one camera centre,
// d = (p - c) / |p - c|, |v| = sqrt((vx vx + vy vy) + vz vz); the gradient is
// dL/dp [n,3] through the normalisation, (g - d (d . g)) / |p - c|.
struct ShDirFromPoints {
  const float *xyz;
  const float *campos;  // [3] in device memory
  float *grad;
  bool add = false;  // add the xyz gradient into grad (one fp32 addition per element) instead of storing it
  MSPLAT_DEV float offset(int64_t i, float &vx, float &vy, float &vz) const {
    vx = xyz[3 * i] - campos[0];
    vy = xyz[3 * i + 1] - campos[1];
    vz = xyz[3 * i + 2] - campos[2];
    return sqrtf(vx * vx + vy * vy + vz * vz);
  }
  MSPLAT_DEV void load(int64_t i, float &x, float &y, float &z) const {
    float vx, vy, vz;
    const float r = offset(i, vx, vy, vz);
    x = vx / r;
    y = vy / r;
    z = vz / r;
  }
  MSPLAT_DEV void store_grad(int64_t i, float gx, float gy, float gz) const {
    float vx, vy, vz;
    const float r = offset(i, vx, vy, vz);
    const float x = vx / r, y = vy / r, z = vz / r;
    const float dot = x * gx + y * gy + z * gz;
    put(i, (gx - x * dot) / r, (gy - y * dot) / r, (gz - z * dot) / r);
  }
  MSPLAT_DEV void store_zero(int64_t i) const { put(i, 0.f, 0.f, 0.f); }
  // A Gaussian whose upstream gradient is zero in every channel: its
  // direction gradient is zero (for finite directions), which added to a
  // given gradient leaves it unchanged (a -0 stays -0), so it is not read
  // or written.
  MSPLAT_DEV void store_unreached(int64_t i) const {
    if (!add) put(i, 0.f, 0.f, 0.f);
  }
  MSPLAT_DEV void put(int64_t i, float gx, float gy, float gz) const {
    if (add) {  // uniform
      gx = grad[3 * i] + gx;
      gy = grad[3 * i + 1] + gy;
      gz = grad[3 * i + 2] + gz;
    }
    grad[3 * i] = gx;
    grad[3 * i + 1] = gy;
    grad[3 * i + 2] = gz;
  }
};

// ---- LDS-staged kernels ------------------------------------------------------
// A workgroup of G threads owns G consecutive Gaussians, whose coefficient
// rows (R = C*D floats each) form one contiguous block of HBM.  The block is
// moved between HBM and LDS with fully coalesced loads/stores (consecutive
// lanes, consecutive addresses) and each thread then works on its own row
// in LDS.  Rows are stored with an odd stride S so that the 64 lanes reading
// element k of 64 different rows hit 64 different banks.

struct ShGeom {
  int R;      // floats per Gaussian row (C*D)
  int S;      // LDS row stride (R rounded up to odd)
  int G;      // Gaussians (threads) per workgroup
  int gdivR;  // G / R and G % R: step of (row, col) when the flat index advances by G
  int gmodR;
};

// Copy rows [g0, g0 + rows) of a row-major [*, R] array between HBM and LDS.
template <bool kToLds>
MSPLAT_DEV void sh_block_copy(float *__restrict__ lds, float *__restrict__ glob, int64_t g0, int rows,
                              const ShGeom &sg) {
  const int total = rows * sg.R;
  int e = threadIdx.x;
  int row = e / sg.R, col = e % sg.R;
  float *base = glob + g0 * sg.R;
  for (; e < total; e += sg.G) {
    if constexpr (kToLds)
      lds[row * sg.S + col] = base[e];
    else
      base[e] = lds[row * sg.S + col];
    row += sg.gdivR;
    col += sg.gmodR;
    if (col >= sg.R) {
      col -= sg.R;
      ++row;
    }
  }
}

template <int L, class Dir>
__global__ __launch_bounds__(256) void sh_fwd_lds_kernel(int64_t n, int C, const float *__restrict__ shs,
                                                         const Dir dir,
                                                         const uint8_t *__restrict__ vis,
                                                         float *__restrict__ value, int ld, ShGeom sg) {
  constexpr int D = (L + 1) * (L + 1);
  extern __shared__ float lds[];
  const int64_t g0 = int64_t(blockIdx.x) * sg.G;
  const int rows = n - g0 < sg.G ? int(n - g0) : sg.G;
  sh_block_copy<true>(lds, const_cast<float *>(shs), g0, rows, sg);
  __syncthreads();
  const int t = threadIdx.x;
  if (t >= rows) return;
  const int64_t i = g0 + t;
  float *out = value + i * ld;
  if (vis && !vis[i]) {
    for (int c = 0; c < C; ++c) out[c] = 0.f;
    return;
  }
  float dx, dy, dz;
  dir.load(i, dx, dy, dz);
  float Y[D];
  sh_basis<L>(dx, dy, dz, [&](int k, float v, float, float, float) { Y[k] = v; });
  const float *row = lds + t * sg.S;
  for (int c = 0; c < C; ++c) {
    float acc = 0.f;
#pragma unroll
    for (int k = 0; k < D; ++k) acc += Y[k] * row[c * D + k];
    out[c] = acc;
  }
}

template <int L, class Dir>
__global__ __launch_bounds__(256) void sh_bwd_lds_kernel(int64_t n, int C, const float *__restrict__ shs,
                                                         const Dir dir,
                                                         const uint8_t *__restrict__ vis,
                                                         const float *__restrict__ dval, int ld,
                                                         float *__restrict__ dshs, ShGeom sg) {
  constexpr int D = (L + 1) * (L + 1);
  extern __shared__ float lds[];
  const int64_t g0 = int64_t(blockIdx.x) * sg.G;
  const int rows = n - g0 < sg.G ? int(n - g0) : sg.G;
  sh_block_copy<true>(lds, const_cast<float *>(shs), g0, rows, sg);
  __syncthreads();
  const int t = threadIdx.x;
  if (t < rows) {
    const int64_t i = g0 + t;
    float *row = lds + t * sg.S;
    const bool live = !vis || vis[i];
    float gx = 0.f, gy = 0.f, gz = 0.f;
    if (live) {
      float x, y, z;
      dir.load(i, x, y, z);
      const float *g = dval + i * ld;
      sh_dir_grad<L>(x, y, z, row, g, C, gx, gy, gz);
      // the row is consumed: overwrite it with dL/dshs = Y_k g_c
      float Y[D];
      sh_basis<L>(x, y, z, [&](int k, float v, float, float, float) { Y[k] = v; });
      for (int c = 0; c < C; ++c) {
        const float gc = g[c];
#pragma unroll
        for (int k = 0; k < D; ++k) row[c * D + k] = Y[k] * gc;
      }
    } else {
      for (int r = 0; r < sg.R; ++r) row[r] = 0.f;
    }
    if (live)
      dir.store_grad(i, gx, gy, gz);
    else
      dir.store_zero(i);
  }
  __syncthreads();
  sh_block_copy<false>(lds, dshs, g0, rows, sg);
}

// ---- one thread per (Gaussian, channel) ----------------------------------------
// Thread (i, c) evaluates the basis at Gaussian i's direction and works on row
// c of its coefficients: D consecutive floats.  Consecutive threads own
// consecutive rows, so a wave reads and writes one contiguous block (16-byte
// accesses when D is a multiple of 4 and the arrays are 16-byte aligned) with
// no LDS and full occupancy.  The basis is evaluated C times per Gaussian,
// which costs less than the memory traffic it avoids staging.  The adjoint
// keeps the C threads of a Gaussian in one wave (G = 64 / C Gaussians per
// wave) and sums their dL/ddir partials, channel 0 first, with lane shuffles.
// Used for degrees <= 6 and C <= 64 (the basis and a row stay in registers).

constexpr int kShGcMaxDeg = 6;

template <int D>
MSPLAT_DEV void sh_load_row(const float *__restrict__ row, float (&r)[D]) {
  if constexpr (D % 4 == 0) {
#pragma unroll
    for (int q = 0; q < D / 4; ++q) {
      const float4 v = reinterpret_cast<const float4 *>(row)[q];
      r[4 * q] = v.x;
      r[4 * q + 1] = v.y;
      r[4 * q + 2] = v.z;
      r[4 * q + 3] = v.w;
    }
  } else {
#pragma unroll
    for (int k = 0; k < D; ++k) r[k] = row[k];
  }
}

template <int D>
MSPLAT_DEV void sh_store_row(float *__restrict__ row, const float (&r)[D]) {
  if constexpr (D % 4 == 0) {
#pragma unroll
    for (int q = 0; q < D / 4; ++q)
      reinterpret_cast<float4 *>(row)[q] = make_float4(r[4 * q], r[4 * q + 1], r[4 * q + 2], r[4 * q + 3]);
  } else {
#pragma unroll
    for (int k = 0; k < D; ++k) row[k] = r[k];
  }
}

// value[i, c] = sum_k Y_k(dir_i) shs[i, c, k], k ascending.
// Rows per thread of sh_fwd_gc_kernel: 2 up to degree 3; higher degrees keep
// one (two rows of 25+ coefficients would cost occupancy).
constexpr int sh_fwd_rows(int L) { return L <= 3 ? 2 : 1; }
template <int L, class Dir>
__global__ __launch_bounds__(256) void sh_fwd_gc_kernel(int64_t n, int C, const float *__restrict__ shs,
                                                        const Dir dir,
                                                        const uint8_t *__restrict__ vis,
                                                        float *__restrict__ value, int ld) {
  constexpr int D = (L + 1) * (L + 1);
  // R rows per thread, 256 apart: the visibility loads of all of them
  // first, then all their coefficient loads, so each thread has R dependent
  // load chains in flight (the kernel is bound by the latency of
  // visibility -> coefficients more than by bandwidth).
  constexpr int R = sh_fwd_rows(L);
  const int64_t rows = n * C;
  int64_t t[R], i[R];
  bool drawn[R];
#pragma unroll
  for (int u = 0; u < R; ++u) {
    t[u] = (int64_t(blockIdx.x) * R + u) * 256 + threadIdx.x;  // row (i, c) = i * C + c
    i[u] = t[u] < rows ? t[u] / C : 0;
    drawn[u] = t[u] < rows && (!vis || vis[i[u]]);
  }
  float r[R][D], dx[R], dy[R], dz[R];
#pragma unroll
  for (int u = 0; u < R; ++u) {
    if (drawn[u]) {
      dir.load(i[u], dx[u], dy[u], dz[u]);
      sh_load_row<D>(shs + t[u] * D, r[u]);
    }
  }
#pragma unroll
  for (int u = 0; u < R; ++u) {
    if (t[u] >= rows) continue;
    float *out = value + i[u] * ld + (t[u] - i[u] * C);
    if (!drawn[u]) {
      *out = 0.f;
      continue;
    }
    float Y[D];
    sh_basis<L>(dx[u], dy[u], dz[u], [&](int k, float v, float, float, float) { Y[k] = v; });
    float acc = 0.f;
#pragma unroll
    for (int k = 0; k < D; ++k) acc += Y[k] * r[u][k];
    *out = acc;
  }
}

// dL/dshs[i,c,k] = Y_k g_c;  dL/ddir = sum_c sum_k (g_c shs[i,c,k]) grad Y_k,
// the inner sum in basis order.
template <int L, class Dir>
__global__ __launch_bounds__(256) void sh_bwd_gc_kernel(int64_t n, int C, const float *__restrict__ shs,
                                                        const Dir dir,
                                                        const uint8_t *__restrict__ vis,
                                                        const float *__restrict__ dval, int ld,
                                                        float *__restrict__ dshs) {
  constexpr int D = (L + 1) * (L + 1);
  // The wave's rows are contiguous in dL/dshs: for D a multiple of 4 (up to
  // 16) they are staged in LDS and stored a 16-byte piece per lane, so each
  // store instruction writes 1 KB in one run instead of 64 pieces 4 D bytes
  // apart (the kernel is bound by these writes).
  constexpr bool kStage = D % 4 == 0 && D <= 16;
  constexpr int kQ = kStage ? D / 4 : 1;
  __shared__ float4 stage[kStage ? 256 / kWave : 1][kStage ? kWave * kQ : 1];
  const int G = kWave / C;  // Gaussians per wave
  const int lane = threadIdx.x % kWave;
  const int wv = threadIdx.x / kWave;
  const int64_t wave = (int64_t(blockIdx.x) * 256 + threadIdx.x) / kWave;
  const int j = lane / C, c = lane - j * C;
  const int64_t i = wave * G + j;
  const bool active = j < G && i < n;
  float gx = 0.f, gy = 0.f, gz = 0.f;
  auto put_row = [&](int64_t t, const float (&r)[D]) {
    if constexpr (kStage) {
#pragma unroll
      for (int q = 0; q < kQ; ++q)
        stage[wv][lane * kQ + q] = make_float4(r[4 * q], r[4 * q + 1], r[4 * q + 2], r[4 * q + 3]);
    } else {
      sh_store_row<D>(dshs + t * D, r);
    }
  };
  // A zero upstream gradient (most rows: Gaussians nobody sees) gives a zero
  // dL/dshs row (Y_k 0) and adds nothing to dL/ddir: neither the direction
  // nor the coefficients are read and the basis is not evaluated (for a
  // finite direction and coefficients the results are the same).
  const bool vis_i = active && !(vis && !vis[i]);
  const float gc = vis_i ? dval[i * ld + c] : 0.f;
  if (active) {
    const int64_t t = i * C + c;
    if (gc == 0.f) {
      float z[D];
#pragma unroll
      for (int k = 0; k < D; ++k) z[k] = 0.f;
      put_row(t, z);
    } else {
      float x, y, zz;
      dir.load(i, x, y, zz);
      float r[D];
      sh_load_row<D>(shs + t * D, r);
      float Y[D];
      sh_basis<L>(x, y, zz, [&](int k, float v, float dx, float dy, float dz) {
        Y[k] = v * gc;
        const float w = gc * r[k];
        gx += w * dx;
        gy += w * dy;
        gz += w * dz;
      });
      put_row(t, Y);
    }
  }
  if constexpr (kStage) {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    const int64_t row0 = wave * G * C;  // the wave's first row
    const int64_t rows = n * C - row0 < int64_t(G * C) ? n * C - row0 : int64_t(G * C);
    float4 *dst = reinterpret_cast<float4 *>(dshs) + row0 * kQ;
#pragma unroll
    for (int k = 0; k < kQ; ++k) {
      const int idx = lane + kWave * k;
      if (idx < rows * kQ) dst[idx] = stage[wv][idx];
    }
  }
  // channel partials of one Gaussian sit in lanes j*C .. j*C + C - 1
  float sx = gx, sy = gy, sz = gz;
  bool reached = gc != 0.f;
  for (int cc = 1; cc < C; ++cc) {
    const int src = imin(lane + cc, kWave - 1);
    const float ox = __shfl(gx, src, kWave), oy = __shfl(gy, src, kWave), oz = __shfl(gz, src, kWave);
    sx += ox;
    sy += oy;
    sz += oz;
    reached = reached || (__shfl(gc, src, kWave) != 0.f && lane + cc < kWave);
  }
  if (active && c == 0) {
    if (reached)
      dir.store_grad(i, sx, sy, sz);
    else  // invisible, or no channel has an upstream gradient
      dir.store_unreached(i);
  }
}

// value[i, c] = sum_k Y_k(dir_i) * shs[i, c, k]  (compute_sh.cu:1600-1637)
template <int L, class Dir>
__global__ __launch_bounds__(kShThreads) void sh_fwd_kernel(int64_t n, int C, const float *__restrict__ shs,
                                                            const Dir dir,
                                                            const uint8_t *__restrict__ vis,
                                                            float *__restrict__ value, int ld) {
  constexpr int D = (L + 1) * (L + 1);
  const int64_t i = int64_t(blockIdx.x) * kShThreads + threadIdx.x;
  if (i >= n) return;
  float *out = value + i * ld;
  if (vis && !vis[i]) {
    for (int c = 0; c < C; ++c) out[c] = 0.f;
    return;
  }
  float dx, dy, dz;
  dir.load(i, dx, dy, dz);
  float Y[D];
  sh_basis<L>(dx, dy, dz, [&](int k, float v, float, float, float) { Y[k] = v; });
  const float *row = shs + i * int64_t(C) * D;
  for (int c = 0; c < C; ++c) {
    float acc = 0.f;
#pragma unroll
    for (int k = 0; k < D; ++k) acc += Y[k] * row[c * D + k];
    out[c] = acc;
  }
}

// dL/dshs[i,c,k] = Y_k g_c;  dL/ddir = sum_k (sum_c g_c shs[i,c,k]) grad Y_k
// (compute_sh.cu:1639-1694).
template <int L, class Dir>
__global__ __launch_bounds__(kShThreads) void sh_bwd_kernel(int64_t n, int C, const float *__restrict__ shs,
                                                            const Dir dir,
                                                            const uint8_t *__restrict__ vis,
                                                            const float *__restrict__ dval, int ld,
                                                            float *__restrict__ dshs) {
  constexpr int D = (L + 1) * (L + 1);
  const int64_t i = int64_t(blockIdx.x) * kShThreads + threadIdx.x;
  if (i >= n) return;
  float *drow = dshs + i * int64_t(C) * D;
  if (vis && !vis[i]) {
    for (int c = 0; c < C; ++c)
#pragma unroll
      for (int k = 0; k < D; ++k) drow[c * D + k] = 0.f;
    dir.store_zero(i);
    return;
  }
  float x, y, z;
  dir.load(i, x, y, z);
  const float *row = shs + i * int64_t(C) * D;
  const float *g = dval + i * ld;
  {
    float Y[D];
    sh_basis<L>(x, y, z, [&](int k, float v, float, float, float) { Y[k] = v; });
    for (int c = 0; c < C; ++c) {
      const float gc = g[c];
#pragma unroll
      for (int k = 0; k < D; ++k) drow[c * D + k] = Y[k] * gc;
    }
  }
  float gx = 0.f, gy = 0.f, gz = 0.f;
  sh_dir_grad<L>(x, y, z, row, g, C, gx, gy, gz);
  dir.store_grad(i, gx, gy, gz);
}

}  // namespace msplat
