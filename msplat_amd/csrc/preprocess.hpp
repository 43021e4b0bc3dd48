// Per-Gaussian geometry: pinhole projection, 3D covariance and EWA splatting,
// forward and adjoint, as separate kernels (the msplat step API) and as one
// fused kernel per direction (what rasterization() runs).  Both use the same
// device functions below, so the fused pass reproduces the chain bit for bit.
//
// Matrices follow glm's storage (m[col][row]) where the reference uses glm,
// and every sum is evaluated left to right in the reference's term order.
#pragma once

#include "common.hpp"

namespace msplat {

constexpr int kGeomThreads = 256;

struct Camera {
  float fx, fy, cx, cy;
  float e[12];  // row-major [R|t]
};

// intr/extr are tiny and read by every thread from the same address; the
// compiler turns these into scalar (SMEM) loads.
MSPLAT_DEV Camera load_camera(const float *__restrict__ intr, const float *__restrict__ extr) {
  Camera c;
  c.fx = intr[0];
  c.fy = intr[1];
  c.cx = intr[2];
  c.cy = intr[3];
#pragma unroll
  for (int i = 0; i < 12; ++i) c.e[i] = extr[i];
  return c;
}

// Camera-space point t = R p + t0 (project_point.cu:27-30, ewa_project.cu:33-37).
MSPLAT_DEV float3 to_camera(const Camera &c, float px, float py, float pz) {
  const float *e = c.e;
  return make_float3(((e[0] * px + e[1] * py) + e[2] * pz) + e[3],
                     ((e[4] * px + e[5] * py) + e[6] * pz) + e[7],
                     ((e[8] * px + e[9] * py) + e[10] * pz) + e[11]);
}

// ---- projection forward (project_point.cu:13-57) ----------------------------
// u = fx*tx*(1/(tz+1e-7)) + cx - 0.5, the reciprocal taken in double as the
// reference does; culled points keep uv = depth = 0.
MSPLAT_DEV bool project_core(const Camera &c, float3 t, int W, int H, float nearest,
                             float extent, float &u, float &v) {
  const float norm1 = static_cast<float>(1.0 / (static_cast<double>(t.z) + 1e-7));
  u = ((c.fx * t.x) * norm1 + c.cx) - 0.5f;
  v = ((c.fy * t.y) * norm1 + c.cy) - 0.5f;
  bool cull = false;
  if (nearest > 0.f) cull = t.z <= nearest;
  if (extent > 0.f) {
    const float xmin = ((1.f - extent) * static_cast<float>(W)) * 0.5f;
    const float xmax = ((1.f + extent) * static_cast<float>(W)) * 0.5f;
    const float ymin = ((1.f - extent) * static_cast<float>(H)) * 0.5f;
    const float ymax = ((1.f + extent) * static_cast<float>(H)) * 0.5f;
    cull = cull || u < xmin || u > xmax || v < ymin || v > ymax;
  }
  return !cull;
}

// ---- projection adjoint (project_point.cu:59-145) ---------------------------
// Adds dL/dxyz and, when kCam, the 16 camera partials (fx,fy,cx,cy,e0..e11).
template <bool kCam>
MSPLAT_DEV void project_adjoint(const Camera &c, float px, float py, float pz, float gu,
                                float gv, float gd, float (&dxyz)[3], float (&cam)[16]) {
  const float3 t = to_camera(c, px, py, pz);
  const float *e = c.e;
  const float norm1 = static_cast<float>(1.0 / static_cast<double>(t.z));
  const float norm2 = static_cast<float>(1.0 / static_cast<double>(t.z * t.z));
  const float fx = c.fx, fy = c.fy;
  dxyz[0] = (((fx * (e[0] * t.z - t.x * e[8])) * norm2) * gu +
             ((fy * (e[4] * t.z - t.y * e[8])) * norm2) * gv) + e[8] * gd;
  dxyz[1] = (((fx * (e[1] * t.z - t.x * e[9])) * norm2) * gu +
             ((fy * (e[5] * t.z - t.y * e[9])) * norm2) * gv) + e[9] * gd;
  dxyz[2] = (((fx * (e[2] * t.z - t.x * e[10])) * norm2) * gu +
             ((fy * (e[6] * t.z - t.y * e[10])) * norm2) * gv) + e[10] * gd;
  if constexpr (kCam) {
    cam[0] += (t.x * norm1) * gu;
    cam[1] += (t.y * norm1) * gv;
    cam[2] += gu;
    cam[3] += gv;
    cam[4 + 0] += ((fx * px) * norm1) * gu;
    cam[4 + 1] += ((fx * py) * norm1) * gu;
    cam[4 + 2] += ((fx * pz) * norm1) * gu;
    cam[4 + 3] += (fx * norm1) * gu;
    cam[4 + 4] += ((fy * px) * norm1) * gv;
    cam[4 + 5] += ((fy * py) * norm1) * gv;
    cam[4 + 6] += ((fy * pz) * norm1) * gv;
    cam[4 + 7] += (fy * norm1) * gv;
    cam[4 + 8] += (((((-fx) * px) * t.x) * norm2) * gu + ((((-fy) * px) * t.y) * norm2) * gv) + px * gd;
    cam[4 + 9] += (((((-fx) * py) * t.x) * norm2) * gu + ((((-fy) * py) * t.y) * norm2) * gv) + py * gd;
    cam[4 + 10] += (((((-fx) * pz) * t.x) * norm2) * gu + ((((-fy) * pz) * t.y) * norm2) * gv) + pz * gd;
    cam[4 + 11] += ((((-fx) * t.x) * norm2) * gu + (((-fy) * t.y) * norm2) * gv) + gd;
  }
}

// ---- 3D covariance (compute_cov3d.cu:14-58) ---------------------------------
// Rotation from the quaternion (w,x,y,z) without normalisation, stored as glm
// would store mat3(...) of compute_cov3d.cu:31-39: rg[col][row].
MSPLAT_DEV void quat_rot(float4 q, float (&rg)[3][3]) {
  const float r = q.x, x = q.y, y = q.z, z = q.w;
  rg[0][0] = 1.f - 2.f * (y * y + z * z);
  rg[0][1] = 2.f * (x * y - r * z);
  rg[0][2] = 2.f * (x * z + r * y);
  rg[1][0] = 2.f * (x * y + r * z);
  rg[1][1] = 1.f - 2.f * (x * x + z * z);
  rg[1][2] = 2.f * (y * z - r * x);
  rg[2][0] = 2.f * (x * z - r * y);
  rg[2][1] = 2.f * (y * z + r * x);
  rg[2][2] = 1.f - 2.f * (x * x + y * y);
}

// Sigma = (S R)^T (S R); out = [00,01,02,11,12,22].  M[j][i] = s_i * R[j][i].
MSPLAT_DEV void cov3d_core(float3 s, float4 q, float (&cov)[6]) {
  float rg[3][3];
  quat_rot(q, rg);
  const float sv[3] = {s.x, s.y, s.z};
  float m[3][3];
#pragma unroll
  for (int j = 0; j < 3; ++j)
#pragma unroll
    for (int i = 0; i < 3; ++i) m[j][i] = sv[i] * rg[j][i];
  auto dot = [&](int a, int b) { return (m[a][0] * m[b][0] + m[a][1] * m[b][1]) + m[a][2] * m[b][2]; };
  cov[0] = dot(0, 0);
  cov[1] = dot(1, 0);
  cov[2] = dot(2, 0);
  cov[3] = dot(1, 1);
  cov[4] = dot(2, 1);
  cov[5] = dot(2, 2);
}

// Adjoint of cov3d_core (compute_cov3d.cu:60-117): with G the symmetric
// gradient (off-diagonals halved), dL/dM = 2 M G, dL/ds_i = sum_j R[j][i]
// dM[j][i], and dL/dq from dL/dR[j][i] = s_i * dM[j][i].
MSPLAT_DEV void cov3d_adjoint(float3 s, float4 q, const float (&dc)[6], float3 &ds, float4 &dq) {
  float rg[3][3];
  quat_rot(q, rg);
  const float sv[3] = {s.x, s.y, s.z};
  float m2[3][3];  // 2*M
#pragma unroll
  for (int j = 0; j < 3; ++j)
#pragma unroll
    for (int i = 0; i < 3; ++i) m2[j][i] = 2.0f * (sv[i] * rg[j][i]);
  float g[3][3];
  g[0][0] = dc[0];
  g[0][1] = 0.5f * dc[1];
  g[0][2] = 0.5f * dc[2];
  g[1][0] = 0.5f * dc[1];
  g[1][1] = dc[3];
  g[1][2] = 0.5f * dc[4];
  g[2][0] = 0.5f * dc[2];
  g[2][1] = 0.5f * dc[4];
  g[2][2] = dc[5];
  // dM = (2M) * G, glm product: dM[j][i] = sum_k m2[k][i] * g[j][k]
  float dm[3][3];
#pragma unroll
  for (int j = 0; j < 3; ++j)
#pragma unroll
    for (int i = 0; i < 3; ++i) dm[j][i] = (m2[0][i] * g[j][0] + m2[1][i] * g[j][1]) + m2[2][i] * g[j][2];
  // dL/ds_c = dot(column c of R^T, column c of dM^T) = sum_i R[i][c] * dM[i][c]
  ds.x = (rg[0][0] * dm[0][0] + rg[1][0] * dm[1][0]) + rg[2][0] * dm[2][0];
  ds.y = (rg[0][1] * dm[0][1] + rg[1][1] * dm[1][1]) + rg[2][1] * dm[2][1];
  ds.z = (rg[0][2] * dm[0][2] + rg[1][2] * dm[1][2]) + rg[2][2] * dm[2][2];
  // P[j][i] = dL/d rg[j][i] = s_i * dM[j][i]; then contract with the
  // derivatives of quat_rot's entries with respect to (r, x, y, z).
  float P[3][3];
#pragma unroll
  for (int j = 0; j < 3; ++j)
#pragma unroll
    for (int i = 0; i < 3; ++i) P[j][i] = dm[j][i] * sv[i];
  const float r = q.x, x = q.y, y = q.z, z = q.w;
  dq.x = (2.f * z * (P[1][0] - P[0][1]) + 2.f * y * (P[0][2] - P[2][0])) + 2.f * x * (P[2][1] - P[1][2]);
  dq.y = ((2.f * y * (P[0][1] + P[1][0]) + 2.f * z * (P[0][2] + P[2][0])) + 2.f * r * (P[2][1] - P[1][2])) -
         4.f * x * (P[2][2] + P[1][1]);
  dq.z = ((2.f * x * (P[0][1] + P[1][0]) + 2.f * r * (P[0][2] - P[2][0])) + 2.f * z * (P[2][1] + P[1][2])) -
         4.f * y * (P[2][2] + P[0][0]);
  dq.w = ((2.f * r * (P[1][0] - P[0][1]) + 2.f * x * (P[0][2] + P[2][0])) + 2.f * y * (P[2][1] + P[1][2])) -
         4.f * z * (P[1][1] + P[0][0]);
}

// ---- EWA forward (ewa_project.cu:16-83) -------------------------------------
struct EwaT {
  float j00, j11, j20, j21;  // nonzero entries of the glm Jacobian J[col][row]
  float t[3][2];             // T = J W, rows 0..1 of every column (row 2 is 0)
  float cx, cy, cz;          // 2D covariance + 0.3 on the diagonal
  float det;
};

MSPLAT_DEV EwaT ewa_cov2d(const Camera &c, float3 t, const float (&v)[6]) {
  EwaT r;
  const float tz2 = t.z * t.z;
  r.j00 = c.fx / t.z;
  r.j11 = c.fy / t.z;
  r.j20 = -(c.fx * t.x) / tz2;
  r.j21 = -(c.fy * t.y) / tz2;
  const float *e = c.e;
  // T[j][i] = sum_k J[k][i] W[j][k] with W[j][k] = e[4k + j]; zero terms dropped.
#pragma unroll
  for (int j = 0; j < 3; ++j) {
    r.t[j][0] = r.j00 * e[j] + r.j20 * e[8 + j];
    r.t[j][1] = r.j11 * e[4 + j] + r.j21 * e[8 + j];
  }
  const float V[3][3] = {{v[0], v[1], v[2]}, {v[1], v[3], v[4]}, {v[2], v[4], v[5]}};
  // TV[k][i] = sum_m T[m][i] V[k][m];  cov[j][i] = sum_k TV[k][i] T[k][j]
  float tv[3][2];
#pragma unroll
  for (int k = 0; k < 3; ++k)
#pragma unroll
    for (int i = 0; i < 2; ++i) tv[k][i] = (r.t[0][i] * V[k][0] + r.t[1][i] * V[k][1]) + r.t[2][i] * V[k][2];
  const float c00 = (tv[0][0] * r.t[0][0] + tv[1][0] * r.t[1][0]) + tv[2][0] * r.t[2][0];
  const float c01 = (tv[0][1] * r.t[0][0] + tv[1][1] * r.t[1][0]) + tv[2][1] * r.t[2][0];
  const float c11 = (tv[0][1] * r.t[0][1] + tv[1][1] * r.t[1][1]) + tv[2][1] * r.t[2][1];
  r.cx = c00 + 0.3f;
  r.cy = c01;
  r.cz = c11 + 0.3f;
  r.det = r.cx * r.cz - r.cy * r.cy;
  return r;
}

// Returns false (outputs stay zero) when det == 0 or the tile rectangle is
// empty, as the reference's early returns do.
MSPLAT_DEV bool ewa_core(const Camera &c, float3 t, const float (&v)[6], float uvx, float uvy, int gx,
                         int gy, float3 &conic, int &radius, int &tiles) {
  const EwaT w = ewa_cov2d(c, t, v);
  if (w.det == 0.0f) return false;
  const float mid = 0.5f * (w.cx + w.cz);
  const float disc = fmaxf(0.1f, mid * mid - w.det);
  const float l1 = mid + sqrtf(disc);
  const float l2 = mid - sqrtf(disc);
  const float rf = ceilf(3.f * sqrtf(fmaxf(l1, l2)));
  const int r = f2i_sat(rf);
  const TileRect rect = tile_rect(uvx, uvy, r, gx, gy);
  const int area = rect.area();
  if (area == 0) return false;
  const float det_inv = 1.f / w.det;
  conic = make_float3(w.cz * det_inv, -w.cy * det_inv, w.cx * det_inv);
  radius = r;
  tiles = area;
  return true;
}

// ---- EWA adjoint (ewa_project.cu:85-252) ------------------------------------
// gc = dL/dconic.  Writes dL/dcov3d, dL/dt (camera-space point) and, when
// kCam, adds the fx, fy and extrinsic partials into cam[0..1], cam[4..15].
template <bool kCam>
MSPLAT_DEV bool ewa_adjoint(const Camera &c, float px, float py, float pz, const float (&v)[6], float3 gc,
                            float (&dcov)[6], float (&dxyz)[3], float (&cam)[16]) {
  const float3 t = to_camera(c, px, py, pz);
  const EwaT w = ewa_cov2d(c, t, v);
  if (w.det == 0.0f) return false;
  const float cx = w.cx, cy = w.cy, cz = w.cz, det = w.det;
  const float nom = 1.0f / (det * det);
  const float dcx = nom * (((-cz) * cz * gc.x + (cy * cz) * gc.y) + (det - cx * cz) * gc.z);
  const float dcy = nom * (((2.f * cy) * cz * gc.x - (det + (2.f * cy) * cy) * gc.y) + ((2.f * cx) * cy) * gc.z);
  const float dcz = nom * (((det - cx * cz) * gc.x + (cx * cy) * gc.y) - (cx * cx) * gc.z);
  const float(&T)[3][2] = w.t;
  // dL/dV for the symmetric 6-vector (off-diagonals count both entries).
  dcov[0] = ((T[0][0] * T[0][0]) * dcx + (T[0][0] * T[0][1]) * dcy) + (T[0][1] * T[0][1]) * dcz;
  dcov[1] = (((2.f * T[0][0]) * T[1][0]) * dcx + (T[0][0] * T[1][1] + T[0][1] * T[1][0]) * dcy) +
            ((2.f * T[0][1]) * T[1][1]) * dcz;
  dcov[2] = (((2.f * T[0][0]) * T[2][0]) * dcx + (T[0][0] * T[2][1] + T[0][1] * T[2][0]) * dcy) +
            ((2.f * T[0][1]) * T[2][1]) * dcz;
  dcov[3] = ((T[1][0] * T[1][0]) * dcx + (T[1][0] * T[1][1]) * dcy) + (T[1][1] * T[1][1]) * dcz;
  dcov[4] = (((2.f * T[1][0]) * T[2][0]) * dcx + (T[1][0] * T[2][1] + T[1][1] * T[2][0]) * dcy) +
            ((2.f * T[1][1]) * T[2][1]) * dcz;
  dcov[5] = ((T[2][0] * T[2][0]) * dcx + (T[2][0] * T[2][1]) * dcy) + (T[2][1] * T[2][1]) * dcz;
  // dL/dT[j][i] (glm indices), i = row 0 or 1.
  float dT[3][2];
  const int vi[3][3] = {{0, 1, 2}, {1, 3, 4}, {2, 4, 5}};
#pragma unroll
  for (int j = 0; j < 3; ++j) {
    const float a0 = (T[0][0] * v[vi[j][0]] + T[1][0] * v[vi[j][1]]) + T[2][0] * v[vi[j][2]];
    const float a1 = (T[0][1] * v[vi[j][0]] + T[1][1] * v[vi[j][1]]) + T[2][1] * v[vi[j][2]];
    dT[j][0] = (2.f * a0) * dcx + a1 * dcy;
    dT[j][1] = a0 * dcy + (2.f * a1) * dcz;
  }
  const float *e = c.e;
  const float dj00 = (e[0] * dT[0][0] + e[1] * dT[1][0]) + e[2] * dT[2][0];
  const float dj20 = (e[8] * dT[0][0] + e[9] * dT[1][0]) + e[10] * dT[2][0];
  const float dj11 = (e[4] * dT[0][1] + e[5] * dT[1][1]) + e[6] * dT[2][1];
  const float dj21 = (e[8] * dT[0][1] + e[9] * dT[1][1]) + e[10] * dT[2][1];
  const float tz = 1.f / t.z;
  const float tz2 = tz * tz;
  const float tz3 = tz2 * tz;
  const float fx = c.fx, fy = c.fy;
  const float dtx = ((-fx) * tz2) * dj20;
  const float dty = ((-fy) * tz2) * dj21;
  const float dtz = ((((-fx) * tz2) * dj00 - (fy * tz2) * dj11) + (((2.f * fx) * t.x) * tz3) * dj20) +
                    (((2.f * fy) * t.y) * tz3) * dj21;
  if constexpr (kCam) {
    cam[0] += tz * dj00 + ((-t.x) * tz2) * dj20;
    cam[1] += tz * dj11 + ((-t.y) * tz2) * dj21;
    cam[4 + 0] += w.j00 * dT[0][0] + px * dtx;
    cam[4 + 1] += w.j00 * dT[1][0] + py * dtx;
    cam[4 + 2] += w.j00 * dT[2][0] + pz * dtx;
    cam[4 + 3] += dtx;
    cam[4 + 4] += w.j11 * dT[0][1] + px * dty;
    cam[4 + 5] += w.j11 * dT[1][1] + py * dty;
    cam[4 + 6] += w.j11 * dT[2][1] + pz * dty;
    cam[4 + 7] += dty;
    cam[4 + 8] += (w.j20 * dT[0][0] + w.j21 * dT[0][1]) + px * dtz;
    cam[4 + 9] += (w.j20 * dT[1][0] + w.j21 * dT[1][1]) + py * dtz;
    cam[4 + 10] += (w.j20 * dT[2][0] + w.j21 * dT[2][1]) + pz * dtz;
    cam[4 + 11] += dtz;
  }
  dxyz[0] = (e[0] * dtx + e[4] * dty) + e[8] * dtz;
  dxyz[1] = (e[1] * dtx + e[5] * dty) + e[9] * dtz;
  dxyz[2] = (e[2] * dtx + e[6] * dty) + e[10] * dtz;
  return true;
}

// ---- camera-gradient reduction -----------------------------------------------
// 16 partials per thread -> one atomic per value per workgroup.
// Workgroups of the camera-gradient adjoints: the grid is capped here and the
// rows strided over it.
constexpr int64_t kCamGrid = 2048;

MSPLAT_DEV void reduce_camera(float (&cam)[16], float *dintr, float *dextr) {
  __shared__ float red[(kGeomThreads / kWave) * 16];
  const int lane = threadIdx.x % kWave;
  const int wave = threadIdx.x / kWave;
  // one transposed wave reduction of the 16 partials: lane 16 r + j (j < 4)
  // then holds the wave total of value 4 j + kRowValue[r]
  float tot[4];
  wave_sum_transposed<16>(cam, tot);
  const int row = lane >> 4, j = lane & 15;
#pragma unroll
  for (int jj = 0; jj < 4; ++jj)
    if (j == jj) red[wave * 16 + 4 * jj + row_value(row)] = tot[jj];
  lds_barrier();  // orders LDS only: the rows' stores need not complete
  if (threadIdx.x < 16) {
    float s = 0.f;
#pragma unroll
    for (int w = 0; w < kGeomThreads / kWave; ++w) s += red[w * 16 + threadIdx.x];
    if (threadIdx.x < 4) {
      if (dintr) unsafeAtomicAdd(dintr + threadIdx.x, s);
    } else if (dextr) {
      unsafeAtomicAdd(dextr + (threadIdx.x - 4), s);
    }
  }
  lds_barrier();  // red may be reused by a later call
}

MSPLAT_DEV float3 ld3(const float *p, int64_t i) { return make_float3(p[3 * i], p[3 * i + 1], p[3 * i + 2]); }
// Scalar loads/stores: caller buffers are only guaranteed 4-byte aligned.
MSPLAT_DEV float4 ld4(const float *p, int64_t i) {
  return make_float4(p[4 * i], p[4 * i + 1], p[4 * i + 2], p[4 * i + 3]);
}
MSPLAT_DEV void st4(float *p, int64_t i, float4 v) {
  p[4 * i] = v.x;
  p[4 * i + 1] = v.y;
  p[4 * i + 2] = v.z;
  p[4 * i + 3] = v.w;
}
MSPLAT_DEV void st3(float *p, int64_t i, float a, float b, float c) {
  p[3 * i] = a;
  p[3 * i + 1] = b;
  p[3 * i + 2] = c;
}

// ---- kernels -----------------------------------------------------------------

__global__ __launch_bounds__(kGeomThreads) void project_fwd_kernel(
    const float *__restrict__ xyz, int64_t n, const float *__restrict__ intr, const float *__restrict__ extr,
    int W, int H, float nearest, float extent, float *__restrict__ uv, float *__restrict__ depth) {
  const int64_t i = int64_t(blockIdx.x) * kGeomThreads + threadIdx.x;
  if (i >= n) return;
  const Camera c = load_camera(intr, extr);
  const float3 p = ld3(xyz, i);
  const float3 t = to_camera(c, p.x, p.y, p.z);
  float u, v;
  const bool keep = project_core(c, t, W, H, nearest, extent, u, v);
  uv[2 * i] = keep ? u : 0.f;
  uv[2 * i + 1] = keep ? v : 0.f;
  depth[i] = keep ? t.z : 0.f;
}

template <bool kCam>
__global__ __launch_bounds__(kGeomThreads) void project_bwd_kernel(
    const float *__restrict__ xyz, int64_t n, const float *__restrict__ intr, const float *__restrict__ extr,
    const float *__restrict__ depth, const float *__restrict__ duv, const float *__restrict__ ddepth,
    float *__restrict__ dxyz, float *__restrict__ dintr, float *__restrict__ dextr) {
  const int64_t i = int64_t(blockIdx.x) * kGeomThreads + threadIdx.x;
  const Camera c = load_camera(intr, extr);
  float cam[16] = {0.f};
  if (i < n) {
    float d[3] = {0.f, 0.f, 0.f};
    if (depth[i] != 0.f) {
      const float3 p = ld3(xyz, i);
      project_adjoint<kCam>(c, p.x, p.y, p.z, duv[2 * i], duv[2 * i + 1], ddepth ? ddepth[i] : 0.f, d, cam);
    }
    st3(dxyz, i, d[0], d[1], d[2]);
  }
  if constexpr (kCam) reduce_camera(cam, dintr, dextr);
}

__global__ __launch_bounds__(kGeomThreads) void cov3d_fwd_kernel(const float *__restrict__ scale,
                                                                  const float *__restrict__ rot,
                                                                  const uint8_t *__restrict__ vis, int64_t n,
                                                                  float *__restrict__ cov) {
  const int64_t i = int64_t(blockIdx.x) * kGeomThreads + threadIdx.x;
  if (i >= n) return;
  float v[6] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
  if (!vis || vis[i]) cov3d_core(ld3(scale, i), ld4(rot, i), v);
#pragma unroll
  for (int k = 0; k < 6; ++k) cov[6 * i + k] = v[k];
}

__global__ __launch_bounds__(kGeomThreads) void cov3d_bwd_kernel(
    const float *__restrict__ scale, const float *__restrict__ rot, const uint8_t *__restrict__ vis,
    const float *__restrict__ dcov, int64_t n, float *__restrict__ dscale, float *__restrict__ drot) {
  const int64_t i = int64_t(blockIdx.x) * kGeomThreads + threadIdx.x;
  if (i >= n) return;
  float3 ds = make_float3(0.f, 0.f, 0.f);
  float4 dq = make_float4(0.f, 0.f, 0.f, 0.f);
  if (!vis || vis[i]) {
    float g[6];
#pragma unroll
    for (int k = 0; k < 6; ++k) g[k] = dcov[6 * i + k];
    cov3d_adjoint(ld3(scale, i), ld4(rot, i), g, ds, dq);
  }
  st3(dscale, i, ds.x, ds.y, ds.z);
  st4(drot, i, dq);
}

__global__ __launch_bounds__(kGeomThreads) void ewa_fwd_kernel(
    const float *__restrict__ xyz, const float *__restrict__ cov3d, const float *__restrict__ intr,
    const float *__restrict__ extr, const float *__restrict__ uv, const uint8_t *__restrict__ vis, int64_t n,
    int gx, int gy, float *__restrict__ conic, int32_t *__restrict__ radius, int32_t *__restrict__ tiles) {
  const int64_t i = int64_t(blockIdx.x) * kGeomThreads + threadIdx.x;
  if (i >= n) return;
  float3 co = make_float3(0.f, 0.f, 0.f);
  int r = 0, nt = 0;
  if (!vis || vis[i]) {
    const Camera c = load_camera(intr, extr);
    const float3 p = ld3(xyz, i);
    float v[6];
#pragma unroll
    for (int k = 0; k < 6; ++k) v[k] = cov3d[6 * i + k];
    if (!ewa_core(c, to_camera(c, p.x, p.y, p.z), v, uv[2 * i], uv[2 * i + 1], gx, gy, co, r, nt)) {
      co = make_float3(0.f, 0.f, 0.f);
      r = 0;
      nt = 0;
    }
  }
  st3(conic, i, co.x, co.y, co.z);
  radius[i] = r;
  tiles[i] = nt;
}

template <bool kCam>
__global__ __launch_bounds__(kGeomThreads) void ewa_bwd_kernel(
    const float *__restrict__ xyz, const float *__restrict__ cov3d, const float *__restrict__ intr,
    const float *__restrict__ extr, const int32_t *__restrict__ radius, const float *__restrict__ dconic,
    int64_t n, float *__restrict__ dxyz, float *__restrict__ dcov3d, float *__restrict__ dintr,
    float *__restrict__ dextr) {
  const int64_t i = int64_t(blockIdx.x) * kGeomThreads + threadIdx.x;
  const Camera c = load_camera(intr, extr);
  float cam[16] = {0.f};
  if (i < n) {
    float dc[6] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
    float d[3] = {0.f, 0.f, 0.f};
    if (radius[i] > 0) {
      const float3 p = ld3(xyz, i);
      float v[6];
#pragma unroll
      for (int k = 0; k < 6; ++k) v[k] = cov3d[6 * i + k];
      if (!ewa_adjoint<kCam>(c, p.x, p.y, p.z, v, ld3(dconic, i), dc, d, cam)) {
#pragma unroll
        for (int k = 0; k < 6; ++k) dc[k] = 0.f;
        d[0] = d[1] = d[2] = 0.f;
      }
    }
#pragma unroll
    for (int k = 0; k < 6; ++k) dcov3d[6 * i + k] = dc[k];
    st3(dxyz, i, d[0], d[1], d[2]);
  }
  if constexpr (kCam) reduce_camera(cam, dintr, dextr);
}

// Fused forward: projection, covariance (for depth != 0) and EWA in one pass.
__global__ __launch_bounds__(kGeomThreads) void preprocess_fwd_kernel(
    const float *__restrict__ xyz, const float *__restrict__ scale, const float *__restrict__ rot, int64_t n,
    const float *__restrict__ intr, const float *__restrict__ extr, int W, int H, int gx, int gy,
    float nearest, float extent, float *__restrict__ uv, float *__restrict__ depth, float *__restrict__ conic,
    int32_t *__restrict__ radius, int32_t *__restrict__ tiles, float *__restrict__ depth_col, int ld_col,
    uint8_t *__restrict__ visible) {
  const int64_t i = int64_t(blockIdx.x) * kGeomThreads + threadIdx.x;
  if (i >= n) return;
  const Camera c = load_camera(intr, extr);
  const float3 p = ld3(xyz, i);
  const float3 t = to_camera(c, p.x, p.y, p.z);
  float u, v;
  const bool keep = project_core(c, t, W, H, nearest, extent, u, v);
  u = keep ? u : 0.f;
  v = keep ? v : 0.f;
  const float d = keep ? t.z : 0.f;
  uv[2 * i] = u;
  uv[2 * i + 1] = v;
  depth[i] = d;
  if (depth_col) depth_col[i * ld_col] = d;  // a feature column (the depth channel)
  float3 co = make_float3(0.f, 0.f, 0.f);
  int r = 0, nt = 0;
  if (d != 0.f) {
    float cv[6];
    cov3d_core(ld3(scale, i), ld4(rot, i), cv);
    if (!ewa_core(c, t, cv, u, v, gx, gy, co, r, nt)) {
      co = make_float3(0.f, 0.f, 0.f);
      r = 0;
      nt = 0;
    }
  }
  st3(conic, i, co.x, co.y, co.z);
  radius[i] = r;
  if (visible) visible[i] = r > 0 ? 1 : 0;  // projected: in some tile
  tiles[i] = nt;
}

// Fused adjoint: EWA -> covariance -> scale/rotation, plus projection; the
// xyz gradient is the projection part + the EWA part, the same single fp32
// addition autograd performs when it sums the two branches.
template <bool kCam>
__global__ __launch_bounds__(kGeomThreads) void preprocess_bwd_kernel(
    const float *__restrict__ xyz, const float *__restrict__ scale, const float *__restrict__ rot, int64_t n,
    const float *__restrict__ intr, const float *__restrict__ extr, const float *__restrict__ depth,
    const int32_t *__restrict__ radius, const float *__restrict__ duv, const float *__restrict__ ddepth,
    const float *__restrict__ dconic, float *__restrict__ dxyz, float *__restrict__ dscale,
    float *__restrict__ drot, float *__restrict__ dintr, float *__restrict__ dextr, int ld_ddepth) {
  const Camera c = load_camera(intr, extr);
  float cam[16] = {0.f};
  auto row = [&](int64_t i) {
    float dp[3] = {0.f, 0.f, 0.f};
    float de[3] = {0.f, 0.f, 0.f};
    float3 ds = make_float3(0.f, 0.f, 0.f);
    float4 dq = make_float4(0.f, 0.f, 0.f, 0.f);
    if (depth[i] != 0.f) {
      const float3 p = ld3(xyz, i);
      const float dd = ddepth ? ddepth[i * ld_ddepth] : 0.f;
      project_adjoint<kCam>(c, p.x, p.y, p.z, duv[2 * i], duv[2 * i + 1], dd, dp, cam);
      const float3 s = ld3(scale, i);
      const float4 q = ld4(rot, i);
      float dc[6] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
      if (radius[i] > 0) {
        float cv[6];
        cov3d_core(s, q, cv);
        if (!ewa_adjoint<kCam>(c, p.x, p.y, p.z, cv, ld3(dconic, i), dc, de, cam)) {
#pragma unroll
          for (int k = 0; k < 6; ++k) dc[k] = 0.f;
          de[0] = de[1] = de[2] = 0.f;
        }
      }
      cov3d_adjoint(s, q, dc, ds, dq);
    }
    st3(dxyz, i, dp[0] + de[0], dp[1] + de[1], dp[2] + de[2]);
    st3(dscale, i, ds.x, ds.y, ds.z);
    st4(drot, i, dq);
  };
  const int64_t first = int64_t(blockIdx.x) * kGeomThreads + threadIdx.x;
  if constexpr (kCam) {
    // a capped grid strides over the rows: each workgroup's one reduction
    // covers several rows, and fewer workgroups add into the 16 camera
    // values (atomics on one address are serialised)
    for (int64_t i = first; i < n; i += int64_t(gridDim.x) * kGeomThreads) row(i);
    reduce_camera(cam, dintr, dextr);
  } else if (first < n) {
    row(first);
  }
}

}  // namespace msplat

namespace msplat {

// ---- many views in one launch -------------------------------------------------
// Forward: grid.y = view; thread (b, g) projects Gaussian gbase[b] + g into
// view b and writes per-view row vbase[b] + g.  Per view the values are those
// of preprocess_fwd_kernel (same device code, same order).
__global__ __launch_bounds__(kGeomThreads) void preprocess_fwd_views_kernel(
    const float *__restrict__ xyz, const float *__restrict__ scale, const float *__restrict__ rot,
    const float *__restrict__ intr, const float *__restrict__ extr, Views vw, float nearest, float extent,
    float *__restrict__ uv, float *__restrict__ depth, float *__restrict__ conic, int32_t *__restrict__ radius,
    int32_t *__restrict__ tiles) {
  const int b = blockIdx.y;
  const int64_t g = int64_t(blockIdx.x) * kGeomThreads + threadIdx.x;
  if (g >= vw.vbase[b + 1] - vw.vbase[b]) return;
  const int64_t i = vw.gbase[b] + g;
  const int64_t o = vw.vbase[b] + g;
  const Camera c = load_camera(intr + 4 * b, extr + 12 * b);
  const float3 p = ld3(xyz, i);
  const float3 t = to_camera(c, p.x, p.y, p.z);
  float u, v;
  const bool keep = project_core(c, t, vw.W[b], vw.H[b], nearest, extent, u, v);
  u = keep ? u : 0.f;
  v = keep ? v : 0.f;
  const float d = keep ? t.z : 0.f;
  uv[2 * o] = u;
  uv[2 * o + 1] = v;
  depth[o] = d;
  float3 co = make_float3(0.f, 0.f, 0.f);
  int r = 0, nt = 0;
  if (d != 0.f) {
    float cv[6];
    cov3d_core(ld3(scale, i), ld4(rot, i), cv);
    if (!ewa_core(c, t, cv, u, v, vw.gx[b], vw.gy[b], co, r, nt)) {
      co = make_float3(0.f, 0.f, 0.f);
      r = 0;
      nt = 0;
    }
  }
  st3(conic, o, co.x, co.y, co.z);
  radius[o] = r;
  tiles[o] = nt;
}

// Adjoint over views that share one scene (per-view rows b * n + g): thread g
// sums Gaussian g's gradients over the views (view 0 first) in registers and
// stores them once; the covariance gradient of every view is summed before
// the one covariance adjoint.  The view loop is a latency chain (a row's depth
// decides whether its gradients are read), so it is software-pipelined: while
// view b is computed, the gradients of view b + 1 and the depth of view b + 2
// are in flight.  Camera partials (dintr + 4 b, dextr + 12 b): each wave adds
// its per-view totals into its own LDS slice (no workgroup barrier inside the
// loop, which would also wait for the loads in flight), and the workgroup adds
// the waves' slices, wave 0 first, and issues one atomic per view and value
// at the end -- the same per-workgroup totals as one reduction per view.
struct ViewGrad {
  float u0, u1, dd, c0, c1, c2;
  int32_t r;
  bool drawn;  // depth != 0: the row was projected
};

template <bool kCam>
__global__ __launch_bounds__(kGeomThreads) void preprocess_bwd_views_kernel(
    const float *__restrict__ xyz, const float *__restrict__ scale, const float *__restrict__ rot, int64_t n,
    const float *__restrict__ intr, const float *__restrict__ extr, int nviews, const float *__restrict__ depth,
    const int32_t *__restrict__ radius, const float *__restrict__ duv, const float *__restrict__ ddepth,
    const float *__restrict__ dconic, float *__restrict__ dxyz, float *__restrict__ dscale,
    float *__restrict__ drot, float *__restrict__ dintr, float *__restrict__ dextr) {
  constexpr int kWaves = kGeomThreads / kWave;
  __shared__ float camacc[kCam ? kWaves * kMaxViews * 16 : 1];
  const int lane = threadIdx.x % kWave, wave = threadIdx.x / kWave;
  if constexpr (kCam) {  // this wave's slice, touched by this wave only until the end
    for (int k = lane; k < nviews * 16; k += kWave) camacc[wave * kMaxViews * 16 + k] = 0.f;
  }
  // Gaussians strided over a capped grid (each workgroup's camera totals then
  // cover several rows; every workgroup adds them with one atomic per value)
  for (int64_t i = int64_t(blockIdx.x) * kGeomThreads + threadIdx.x;
       i - threadIdx.x < n; i += int64_t(gridDim.x) * kGeomThreads) {
    // uniform over the workgroup: rows past n take part in the wave reductions
    const bool live = i < n;
    float3 p = make_float3(0.f, 0.f, 0.f), s = p;
    float4 q = make_float4(0.f, 0.f, 0.f, 0.f);
    if (live) {
      p = ld3(xyz, i);
      s = ld3(scale, i);
      q = ld4(rot, i);
    }
    // the gradients of view b's row, read only where it was projected
    auto load_view = [&](int b, float d) {
      ViewGrad g{0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0, false};
      if (live && d != 0.f) {
        const int64_t o = int64_t(b) * n + i;
        g.drawn = true;
        g.u0 = duv[2 * o];
        g.u1 = duv[2 * o + 1];
        g.dd = ddepth ? ddepth[o] : 0.f;
        g.r = radius[o];
        g.c0 = dconic[3 * o];
        g.c1 = dconic[3 * o + 1];
        g.c2 = dconic[3 * o + 2];
      }
      return g;
    };
    auto load_depth = [&](int b) { return live && b < nviews ? depth[int64_t(b) * n + i] : 0.f; };
    ViewGrad cur = load_view(0, load_depth(0));
    float dnext = load_depth(1);
    float cv[6];
    bool have_cov = false;
    float gp[3] = {0.f, 0.f, 0.f};
    float gc[6] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
    bool seen = false;  // visible in some view
    for (int b = 0; b < nviews; ++b) {
      // in flight during view b: view b + 1's gradients, view b + 2's depth
      const ViewGrad nxt = load_view(b + 1, b + 1 < nviews ? dnext : 0.f);
      dnext = load_depth(b + 2);
      const Camera c = load_camera(intr + 4 * b, extr + 12 * b);
      float cam[16] = {0.f};
      if (cur.drawn) {
        seen = true;
        float dp[3] = {0.f, 0.f, 0.f};
        float de[3] = {0.f, 0.f, 0.f};
        project_adjoint<kCam>(c, p.x, p.y, p.z, cur.u0, cur.u1, cur.dd, dp, cam);
        if (cur.r > 0) {
          if (!have_cov) {
            cov3d_core(s, q, cv);
            have_cov = true;
          }
          float dc[6] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
          if (ewa_adjoint<kCam>(c, p.x, p.y, p.z, cv, make_float3(cur.c0, cur.c1, cur.c2), dc, de, cam)) {
#pragma unroll
            for (int k = 0; k < 6; ++k) gc[k] += dc[k];
          } else {
            de[0] = de[1] = de[2] = 0.f;
          }
        }
#pragma unroll
        for (int k = 0; k < 3; ++k) gp[k] += dp[k] + de[k];
      }
      if constexpr (kCam) {
        // lane 16 r + j (j < 4) holds the wave total of value 4 j + kRowValue[r]
        float tot[4];
        wave_sum_transposed<16>(cam, tot);
        const int row = lane >> 4, j = lane & 15;
        float *acc = camacc + wave * kMaxViews * 16 + b * 16;
#pragma unroll
        for (int jj = 0; jj < 4; ++jj)
          if (j == jj) acc[4 * jj + row_value(row)] += tot[jj];
      }
      cur = nxt;
    }
    if (live) {
      float3 ds = make_float3(0.f, 0.f, 0.f);
      float4 dq = make_float4(0.f, 0.f, 0.f, 0.f);
      if (seen) cov3d_adjoint(s, q, gc, ds, dq);
      st3(dxyz, i, gp[0], gp[1], gp[2]);
      st3(dscale, i, ds.x, ds.y, ds.z);
      st4(drot, i, dq);
    }
  }
  if constexpr (kCam) {
    lds_barrier();  // every wave's slice is complete (orders LDS only)
    for (int t = threadIdx.x; t < nviews * 16; t += kGeomThreads) {
      float sum = 0.f;
#pragma unroll
      for (int w = 0; w < kWaves; ++w) sum += camacc[w * kMaxViews * 16 + t];
      const int b = t / 16, v = t % 16;
      if (v < 4) {
        if (dintr) unsafeAtomicAdd(dintr + 4 * b + v, sum);
      } else if (dextr) {
        unsafeAtomicAdd(dextr + 12 * b + (v - 4), sum);
      }
    }
  }
}

}  // namespace msplat
