/*
 * msplat_amd.h -- C ABI of the MI355X (gfx950) Gaussian-splatting rasterizer.
 *
 * Every entry point takes plain device pointers, element counts and a
 * hipStream_t (passed as void*), launches its kernels on that stream and
 * returns 0 on success or a negative msplat_status code.  The library never
 * allocates device memory: outputs and workspaces belong to the caller.
 * Outputs are written in full (culled / skipped elements get zeros, matching
 * the zero-filled tensors the reference returns), so callers need no memset.
 *
 * Layouts (row-major, fp32 unless noted):
 *   xyz [n,3]  scale [n,3]  rotate [n,4] quaternion (w,x,y,z), used unnormalised
 *   opacity [n]  feature [n,C]  intr [4] = fx,fy,cx,cy
 *   extr: the first 12 floats are read as the row-major 3x4 [R|t] (a 4x4 works)
 *   uv [n,2]  depth [n]  cov3d [n,6] = upper triangle 00,01,02,11,12,22
 *   conic [n,3] = inverse 2D covariance a,b,c  radius/tiles [n] int32
 *   idx_sorted [K] int32  tile_range [T,2] int32, T = ceil(W/16)*ceil(H/16)
 *   rendered [C,H,W]  final_T [H,W]  ncontrib [H,W] int32
 *   visible: one byte per Gaussian (torch.bool), may be NULL = all visible
 *
 * Reference files are cited relative to the msplat source tree.
 */
#ifndef MSPLAT_AMD_H
#define MSPLAT_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum msplat_status {
  MSPLAT_OK = 0,
  MSPLAT_ERR_ARGUMENT = -1,  /* bad size / null pointer / unsupported value */
  MSPLAT_ERR_LAUNCH = -2,    /* a HIP launch or runtime call failed */
  MSPLAT_ERR_WORKSPACE = -3, /* workspace smaller than the size query asked for */
  MSPLAT_ERR_CAPACITY = -4   /* intersection count does not fit int32 */
};

/* Library version (major*10000 + minor*100 + patch). No reference counterpart. */
int msplat_version(void);

/* Message for the last non-zero status returned on the calling thread.
 * Replaces the TORCH_CHECK text of msplat/include/utils.h:9-10. */
const char *msplat_last_error(void);

/* ---- expired look-back waits (no reference counterpart: the reference's
 * sort is torch.sort, msplat/sort_gaussian.py:49, which has no such wait).
 * Every look-back wait of the sorts and the cell binning is bounded.  When
 * one expires (a predecessor workgroup made no progress for 2^16 polls), the
 * lists it wrote are undefined; the sort's or binning's last launch then sets
 * a word per device that stays set until cleared.  While it is set, the
 * cell-list blending forward writes NaN images and transmittances, and every
 * binning's counting pass copies it into d_count[1], which the host reads with
 * the count.  msplat_wait_status writes the word to d_out[0] (device int64;
 * may be NULL) in stream order and, when `clear` is non-zero, clears it.
 * msplat_debug_force_expiry (tests only) makes every look-back wait on the
 * device give up at once, without reading its predecessors, until it is
 * called again with on = 0; every kernel still runs to its end and writes
 * only inside its outputs. */
int msplat_wait_status(int64_t *d_out, int clear, void *stream);
int msplat_debug_force_expiry(int on, void *stream);

/* ---- projection: msplat/src/project_point.cu:147-179 (_C.project_point_forward,
 *      bound at msplat/src/ext.cpp:14, called from msplat/project_point.py:59) */
int msplat_project_point_forward(const float *xyz, int64_t n, const float *intr,
                                 const float *extr, int W, int H, float nearest,
                                 float extent, float *uv, float *depth,
                                 void *stream);

/* project_point.cu:181-228 (_C.project_point_backward, ext.cpp:15,
 * project_point.py:77).  dL_ddepth may be NULL (= zeros).  dL_dintr [4] and
 * dL_dextr [12] may be NULL (= not requested, project_point.cu:202-208);
 * when given they are overwritten with the reduced camera gradient. */
int msplat_project_point_backward(const float *xyz, int64_t n,
                                  const float *intr, const float *extr,
                                  const float *depth, const float *dL_duv,
                                  const float *dL_ddepth, float *dL_dxyz,
                                  float *dL_dintr, float *dL_dextr,
                                  void *stream);

/* ---- 3D covariance: msplat/src/compute_cov3d.cu:149-170 (ext.cpp:16,
 *      compute_cov3d.py:39) */
int msplat_compute_cov3d_forward(const float *scale, const float *rotate,
                                 const uint8_t *visible, int64_t n,
                                 float *cov3d, void *stream);

/* compute_cov3d.cu:172-199 (ext.cpp:17, compute_cov3d.py:51) */
int msplat_compute_cov3d_backward(const float *scale, const float *rotate,
                                  const uint8_t *visible,
                                  const float *dL_dcov3d, int64_t n,
                                  float *dL_dscale, float *dL_drotate,
                                  void *stream);

/* ---- EWA splatting: msplat/src/ewa_project.cu:254-297 (ext.cpp:18,
 *      ewa_project.py:59) */
int msplat_ewa_project_forward(const float *xyz, const float *cov3d,
                               const float *intr, const float *extr,
                               const float *uv, const uint8_t *visible,
                               int64_t n, int W, int H, float *conic,
                               int32_t *radius, int32_t *tiles, void *stream);

/* ewa_project.cu:299-345 (ext.cpp:19, ewa_project.py:71).  Camera gradients
 * as in msplat_project_point_backward. */
int msplat_ewa_project_backward(const float *xyz, const float *cov3d,
                                const float *intr, const float *extr,
                                const int32_t *radius, const float *dL_dconic,
                                int64_t n, float *dL_dxyz, float *dL_dcov3d,
                                float *dL_dintr, float *dL_dextr, void *stream);

/* ---- fused projection + cov3d + EWA.  Same outputs as the chain
 * project_point -> compute_cov3d(visible = depth != 0) -> ewa_project used by
 * msplat/__init__.py:71-81, in one pass over the Gaussians. */
int msplat_preprocess_forward(const float *xyz, const float *scale,
                              const float *rotate, int64_t n,
                              const float *intr, const float *extr, int W,
                              int H, float nearest, float extent, float *uv,
                              float *depth, float *conic, int32_t *radius,
                              int32_t *tiles, void *stream);

/* Adjoint of msplat_preprocess_forward: the sum of the three reference
 * backward kernels (project_point.cu:59-145, compute_cov3d.cu:131-147,
 * ewa_project.cu:85-252) as autograd composes them.  dL_ddepth may be NULL. */
int msplat_preprocess_backward(const float *xyz, const float *scale,
                               const float *rotate, int64_t n,
                               const float *intr, const float *extr,
                               const float *depth, const int32_t *radius,
                               const float *dL_duv, const float *dL_ddepth,
                               const float *dL_dconic, float *dL_dxyz,
                               float *dL_dscale, float *dL_drotate,
                               float *dL_dintr, float *dL_dextr, void *stream);

/* The same two, for a depth that is also a feature channel (no reference
 * counterpart; rasterization_sh's depth channel): the forward also writes
 * each depth to depth_col[i * ld_col] and visible[i] = radius[i] > 0 (a
 * uint8 mask, e.g. for msplat_compute_sh_points_forward; each NULL:
 * nothing), the adjoint reads dL_ddepth[i * ld_ddepth], so neither needs a
 * strided copy. */
int msplat_preprocess_forward_col(const float *xyz, const float *scale,
                                  const float *rotate, int64_t n,
                                  const float *intr, const float *extr, int W,
                                  int H, float nearest, float extent,
                                  float *uv, float *depth, float *conic,
                                  int32_t *radius, int32_t *tiles,
                                  float *depth_col, int ld_col,
                                  uint8_t *visible, void *stream);
int msplat_preprocess_backward_ld(const float *xyz, const float *scale,
                                  const float *rotate, int64_t n,
                                  const float *intr, const float *extr,
                                  const float *depth, const int32_t *radius,
                                  const float *dL_duv, const float *dL_ddepth,
                                  int ld_ddepth, const float *dL_dconic,
                                  float *dL_dxyz, float *dL_dscale,
                                  float *dL_drotate, float *dL_dintr,
                                  float *dL_dextr, void *stream);

/* ---- spherical harmonics, degree 0..10: msplat/src/compute_sh.cu:1696-1722
 *      (ext.cpp:22, compute_sh.py:39).  shs [n,C,D], D = (deg+1)^2. */
int msplat_compute_sh_forward(const float *shs, const float *dirs,
                              const uint8_t *visible, int64_t n, int C, int D,
                              float *value, void *stream);

/* compute_sh.cu:1724-1753 (ext.cpp:23, compute_sh.py:51) */
int msplat_compute_sh_backward(const float *shs, const float *dirs,
                               const uint8_t *visible, const float *dL_dvalue,
                               int64_t n, int C, int D, float *dL_dshs,
                               float *dL_ddirs, void *stream);

/* SH of the view direction from one camera centre, fused: value =
 * compute_sh(shs, (xyz - campos) / |xyz - campos|), with |v| =
 * sqrtf((vx vx + vy vy) + vz vz).  No counterpart in the reference, whose
 * compute_sh takes the directions (compute_sh.py:8-33) and leaves their
 * computation to user code; this removes that torch code and its adjoint.
 * campos is [3] in device memory.  value row i starts at value + i * ld_value
 * (ld_value >= C: the SH channels can be written straight into a wider
 * feature buffer), and likewise dL_dvalue.  The adjoint writes dL/dshs and
 * dL/dxyz (through the normalisation); dL/dcampos = -sum_i dL/dxyz_i. */
int msplat_compute_sh_points_forward(const float *shs, const float *xyz,
                                     const float *campos,
                                     const uint8_t *visible, int64_t n,
                                     int C, int D, float *value, int ld_value,
                                     void *stream);
int msplat_compute_sh_points_backward(const float *shs, const float *xyz,
                                      const float *campos,
                                      const uint8_t *visible,
                                      const float *dL_dvalue, int ld_dvalue,
                                      int64_t n, int C, int D, float *dL_dshs,
                                      float *dL_dxyz, void *stream);
/* As msplat_compute_sh_points_backward, but the xyz gradient is added into
 * dL_dxyz (one fp32 addition per element, zeros included: the same values as
 * storing it and adding the two arrays; with C = 0 nothing is added).  No
 * reference counterpart: autograd's sum of the xyz gradients that reach a
 * training step's xyz through compute_sh (msplat/compute_sh.py:47) and
 * through project_point (msplat/project_point.py:71), done inside the SH
 * adjoint. */
int msplat_compute_sh_points_backward_add(const float *shs, const float *xyz,
                                          const float *campos,
                                          const uint8_t *visible,
                                          const float *dL_dvalue, int ld_dvalue,
                                          int64_t n, int C, int D,
                                          float *dL_dshs, float *dL_dxyz,
                                          void *stream);

/* ---- kernel timing (no reference counterpart; for benchmarks).  While
 * enabled, the main kernel of each entry point is bracketed by HIP events on
 * the caller's stream.  Kernel names: preprocess_fwd, preprocess_bwd, sh_fwd,
 * sh_bwd, depth_sort, emit, class_sort, blend_fwd, blend_bwd,
 * rasterize_fwd, rasterize_bwd.  msplat_timing_enable returns the previous
 * state; msplat_timing_read waits for the recorded events of one kernel,
 * returns their total milliseconds and launch count, and forgets them.
 * msplat_timing_only restricts the timing to the comma-separated kernel
 * names given (NULL or "" times every kernel): each timed launch adds two
 * event records to the stream, which cost the stream some time of their own. */
int msplat_timing_enable(int on);
int msplat_timing_read(const char *kernel, double *total_ms, int64_t *launches);
int msplat_timing_only(const char *kernels);

/* ---- tile sort.  Replaces msplat/sort_gaussian.py:42-52 (cumsum, keygen
 * sort_gaussian.cu:74-113, torch.sort, gather, tile range
 * sort_gaussian.cu:115-142) with two calls around one 8-byte device read.
 *
 * Step 1 computes, for every Gaussian, its tile rectangle (utils.h:17-37),
 * orders Gaussians stably by depth bits, scans their tile counts and counts
 * keys per tile.  It writes the intersection count K to *d_count (device
 * int64).  If `tiles` is non-NULL it is checked against the rectangles and
 * *d_count is set to -1 on any disagreement (the reference would write out
 * of bounds); -2 reports an expired look-back wait of the depth sort.  Size
 * queries return 0 for an invalid image size. */
size_t msplat_sort_prepare_workspace_bytes(int64_t n, int W, int H);
int msplat_sort_prepare(const float *uv, const float *depth,
                        const int32_t *radius, const int32_t *tiles,
                        int64_t n, int W, int H, void *workspace,
                        size_t workspace_bytes, int64_t *d_count,
                        void *stream);

/* Step 2 emits the K (tile, Gaussian) pairs, sorts them stably by tile and
 * writes idx_sorted [K] and tile_range [T,2], bit-exact with the reference's
 * stable signed-int64 key order.  key_sorted [K] may be NULL; when given it
 * receives the reference's 64-bit keys (tile<<32 | depth bits) in sorted
 * order.  `prepared` is the step-1 workspace, unchanged. */
size_t msplat_sort_emit_workspace_bytes(int64_t n, int64_t k, int W, int H);
int msplat_sort_emit(const float *uv, const float *depth,
                     const int32_t *radius, int64_t n, int W, int H,
                     const void *prepared, int64_t k, void *workspace,
                     size_t workspace_bytes, int32_t *idx_sorted,
                     int32_t *tile_range, int64_t *key_sorted, void *stream);

/* The reference keygen in its own order: sort_gaussian.cu:17-43,74-113
 * (_C.compute_gaussian_key, ext.cpp:20).  `offsets` is the inclusive cumsum
 * of tiles; keys [k] int64 and idx [k] int32 are fully written (slots the
 * reference leaves untouched get 0, as its torch::zeros would). */
int msplat_gaussian_keys(const float *uv, const float *depth,
                         const int32_t *radius, const int32_t *offsets,
                         int64_t n, int W, int H, int64_t k, int64_t *keys,
                         int32_t *idx, void *stream);

/* Tile ranges from sorted keys: sort_gaussian.cu:45-71,115-142
 * (_C.compute_tile_gaussian_range, ext.cpp:21), without the out-of-bounds
 * write for negative-depth keys (tile id -1). */
int msplat_tile_range_from_keys(const int64_t *key_sorted, int64_t k, int W,
                                int H, int32_t *tile_range, void *stream);

/* ---- alpha blending: msplat/src/alpha_blending.cu:248-394 (ext.cpp:24,
 *      alpha_blending.py:59).  All C channels in one walk per pixel. */
int msplat_alpha_blending_forward(const float *uv, const float *conic,
                                  const float *opacity, const float *feature,
                                  int64_t n, int C, const int32_t *idx_sorted,
                                  const int32_t *tile_range, float bg, int W,
                                  int H, float *rendered, float *final_T,
                                  int32_t *ncontrib, void *stream);

/* alpha_blending.cu:396-573 (ext.cpp:25, alpha_blending.py:92).  Gradient
 * outputs are overwritten (zeroed, then accumulated). */
int msplat_alpha_blending_backward(
    const float *uv, const float *conic, const float *opacity,
    const float *feature, int64_t n, int C, const int32_t *idx_sorted,
    const int32_t *tile_range, float bg, int W, int H, const float *final_T,
    const int32_t *ncontrib, const float *dL_drendered, float *dL_duv,
    float *dL_dconic, float *dL_dopacity, float *dL_dfeature, void *stream);

/* ---- rasterizer over cell lists: replaces the sort_gaussian + alpha_blending
 * pair inside rasterization (msplat/__init__.py:83-91) with images, final_T,
 * ncontrib and gradients identical to that pair's.
 *
 * Binning groups the tile grid into cells of `cell` x `cell` tiles and lists,
 * per cell, every Gaussian whose tile rectangle (utils.h:17-37) touches it,
 * in the reference's (depth bits, id) order; negative depths are left out, as
 * the reference's keys put them in no tile.  A tile's reference list is its
 * cell's list filtered by rectangle, so the K (tile, Gaussian) pairs are never
 * materialised.  Two calls around one 8-byte device read, like the sort.
 *
 * Step 1 writes each Gaussian's packed tile rectangle to rect [n, 2] uint32
 * (x0 | y0 << 16, x1 | y1 << 16; 8-byte aligned), the number of cell
 * entries to d_count[0] and the device's expired-wait word (see
 * msplat_wait_status) to d_count[1]: d_count holds two int64. */
size_t msplat_bin_prepare_workspace_bytes(int64_t n, int W, int H, int cell);
int msplat_bin_prepare(const float *uv, const float *depth,
                       const int32_t *radius, int64_t n, int W, int H,
                       int cell, uint32_t *rect, void *workspace,
                       size_t workspace_bytes, int64_t *d_count, void *stream);

/* Step 2 writes cell_ids [count] and cell_range [cells, 2] (cells =
 * ceil(ceil(W/16)/cell) * ceil(ceil(H/16)/cell), row-major).  With d_count
 * NULL, `count` is the step-1 count read back by the host.  With d_count =
 * step 1's device counter, `count` is a capacity: grids are sized for it and
 * the kernels read the true count on the device (no host synchronisation, so
 * the two steps can be captured in a graph); if the true count exceeds the
 * capacity, entries past it are dropped and *d_count still holds the true
 * count, for the caller to detect. */
size_t msplat_bin_emit_workspace_bytes(int64_t n, int64_t count, int W, int H,
                                       int cell);
int msplat_bin_emit(const float *uv, const float *depth, const int32_t *radius,
                    int64_t n, int W, int H, int cell, const void *prepared,
                    int64_t count, const int64_t *d_count, void *workspace,
                    size_t workspace_bytes, int32_t *cell_ids,
                    int32_t *cell_range, void *stream);

/* Blending forward over cell lists (alpha_blending.cu:248-394 semantics).
 * `ppl` = pixels per lane (1, 2 or 4; 0 picks the default).  `cull` != 0
 * also skips, per tile, Gaussians that provably give every pixel of the tile
 * alpha < 1/255 (rendered and final_T are unchanged; ncontrib then counts
 * positions in the culled list).  walk [T, 2] int32 receives, per tile,
 * where the adjoint's walk starts.  The adjoint must use the same cell and
 * cull. */
int msplat_rasterize_forward(const float *uv, const float *conic,
                             const float *opacity, const float *feature,
                             int64_t n, int C, const uint32_t *rect,
                             const int32_t *cell_ids,
                             const int32_t *cell_range, int cell, float bg,
                             int W, int H, int ppl, int cull, float *rendered,
                             float *final_T, int32_t *ncontrib, int32_t *walk,
                             void *stream);

/* Adjoint over cell lists (alpha_blending.cu:396-573 semantics); gradient
 * outputs are overwritten. */
int msplat_rasterize_backward(
    const float *uv, const float *conic, const float *opacity,
    const float *feature, int64_t n, int C, const uint32_t *rect,
    const int32_t *cell_ids, const int32_t *cell_range, int cell, float bg,
    int W, int H, int ppl, int cull, const float *final_T, const int32_t *ncontrib,
    const int32_t *walk, const float *dL_drendered, float *dL_duv,
    float *dL_dconic, float *dL_dopacity, float *dL_dfeature, void *stream);

/* The same pair with a squared-error loss against target [C, H, W] fused in
 * (no reference counterpart; the training loss of the reference's tutorials,
 * tutorials/gs_2d.py:73-90, is separate torch kernels).  The forward also
 * writes *loss = sum (rendered - target)^2 (per-wave partial sums in the
 * workspace, added in a fixed order in fp64); the adjoint takes dL/drendered
 * = 2 (rendered - target) dL_dloss[0], plus dL_drendered when that is not
 * NULL (the image's gradient from other terms), without the separate
 * gradient image of msplat_sq_diff_grad. */
size_t msplat_rasterize_l2_workspace_bytes(int W, int H, int C);
int msplat_rasterize_forward_l2(const float *uv, const float *conic,
                                const float *opacity, const float *feature,
                                int64_t n, int C, const uint32_t *rect,
                                const int32_t *cell_ids,
                                const int32_t *cell_range, int cell, float bg,
                                int W, int H, int ppl, int cull,
                                const float *target, void *workspace,
                                size_t workspace_bytes, float *rendered,
                                float *final_T, int32_t *ncontrib,
                                int32_t *walk, float *loss, void *stream);
int msplat_rasterize_backward_l2(
    const float *uv, const float *conic, const float *opacity,
    const float *feature, int64_t n, int C, const uint32_t *rect,
    const int32_t *cell_ids, const int32_t *cell_range, int cell, float bg,
    int W, int H, int ppl, int cull, const float *final_T, const int32_t *ncontrib,
    const int32_t *walk, const float *dL_drendered, const float *rendered,
    const float *target, const float *dL_dloss, float *dL_duv,
    float *dL_dconic, float *dL_dopacity, float *dL_dfeature, void *stream);

/* ---- many views in one launch (no reference counterpart: the reference
 * renders one camera per rasterization call, msplat/__init__.py:22-93).
 *
 * B <= 64 cameras, view b of W[b] x H[b] pixels (W, H: host arrays), intr
 * [B,4] and extr [B,12] (row-major [R|t]) on the device.  View b draws the
 * Gaussians [first[b], first[b] + count[b]) of the shared per-Gaussian arrays
 * (xyz, scale, rotate, opacity, shared features; host arrays first / count,
 * so views of several resident scenes can share a launch), or all n of them
 * when first = count = NULL.  Per-view arrays (uv, depth, conic, radius,
 * tiles, rect, their gradients) hold one row per (view, Gaussian): view b's
 * rows start at R_b = sum of the earlier views' counts (b * n with NULL).
 * Feature rows are shared ([n, C], row first[b] + g) or per view ([R, C]) when
 * feat_per_view.  Images are concatenated: view b's rendered [C, H_b, W_b]
 * starts at C * P_b and its final_T / ncontrib [H_b, W_b] at P_b, with P_b =
 * sum over earlier views of W*H.  Cells and tiles of all views form one
 * global space; walk has one row per global tile.  Each view's results equal
 * the single-view calls' on that view.  The adjoints take views of one scene
 * (first = count = NULL); gradients of the shared arrays are summed over the
 * views. */
int msplat_preprocess_views_forward(const float *xyz, const float *scale,
                                    const float *rotate, int64_t n, int B,
                                    const float *intr, const float *extr,
                                    const int *W, const int *H,
                                    const int64_t *first, const int64_t *count,
                                    float nearest, float extent, float *uv,
                                    float *depth, float *conic, int32_t *radius,
                                    int32_t *tiles, void *stream);
/* dL_dintr [B,4] and dL_dextr [B,12] may be NULL (= not requested). */
int msplat_preprocess_views_backward(
    const float *xyz, const float *scale, const float *rotate, int64_t n, int B,
    const float *intr, const float *extr, const float *depth,
    const int32_t *radius, const float *dL_duv, const float *dL_ddepth,
    const float *dL_dconic, float *dL_dxyz, float *dL_dscale,
    float *dL_drotate, float *dL_dintr, float *dL_dextr, void *stream);
size_t msplat_bin_views_prepare_workspace_bytes(int64_t rows);
int msplat_bin_views_prepare(const float *uv, const float *depth,
                             const int32_t *radius, int64_t n, int B,
                             const int *W, const int *H, const int64_t *first,
                             const int64_t *count, int cell, uint32_t *rect,
                             void *workspace, size_t workspace_bytes,
                             int64_t *d_count, void *stream);
/* msplat_bin_views_prepare in two calls with its arguments, on one stream:
 * _count writes rect and *d_count (the host may copy the count as soon as
 * _count's work is done), _order then puts the listed Gaussians in depth
 * order (the bulk of step 1), so that the count's round trip to the host
 * overlaps the sort.  d_count[0] = -1 reports an expired look-back wait of
 * this counting pass, d_count[1] != 0 one of an earlier sort or binning on
 * the device (see msplat_wait_status); d_count[0] = 2^31 means the entries
 * reach 2^31 - 1 (more than int32 lists index).  d_count holds two int64. */
int msplat_bin_views_count(const float *uv, const float *depth,
                           const int32_t *radius, int64_t n, int B,
                           const int *W, const int *H, const int64_t *first,
                           const int64_t *count, int cell, uint32_t *rect,
                           void *workspace, size_t workspace_bytes,
                           int64_t *d_count, void *stream);
int msplat_bin_views_order(const float *uv, const float *depth,
                           const int32_t *radius, int64_t n, int B,
                           const int *W, const int *H, const int64_t *first,
                           const int64_t *count, int cell, uint32_t *rect,
                           void *workspace, size_t workspace_bytes,
                           int64_t *d_count, void *stream);
size_t msplat_bin_views_emit_workspace_bytes(int64_t count);
int msplat_bin_views_emit(const float *uv, const float *depth,
                          const int32_t *radius, int64_t n, int B,
                          const int *W, const int *H, const int64_t *first,
                          const int64_t *vcount, int cell, const void *prepared,
                          int64_t count, const int64_t *d_count,
                          void *workspace, size_t workspace_bytes,
                          int32_t *cell_ids, int32_t *cell_range, void *stream);
int msplat_rasterize_views_forward(
    const float *uv, const float *conic, const float *opacity,
    const float *feature, int64_t n, int C, int feat_per_view, int B,
    const int *W, const int *H, const int64_t *first, const int64_t *count,
    const uint32_t *rect, const int32_t *cell_ids,
    const int32_t *cell_range, int cell, float bg, int ppl, int cull,
    float *rendered, float *final_T, int32_t *ncontrib, int32_t *walk,
    void *stream);
int msplat_rasterize_views_backward(
    const float *uv, const float *conic, const float *opacity,
    const float *feature, int64_t n, int C, int feat_per_view, int B,
    const int *W, const int *H, const uint32_t *rect, const int32_t *cell_ids,
    const int32_t *cell_range, int cell, float bg, int ppl, int cull,
    const float *final_T, const int32_t *ncontrib, const int32_t *walk,
    const float *dL_drendered, float *dL_duv, float *dL_dconic,
    float *dL_dopacity, float *dL_dfeature, void *stream);

/* ---- squared-error loss next to the rasterizer (no reference counterpart:
 * the reference's training loops, tutorials/gs_2d.py:73-90, compute it with
 * separate torch kernels).  msplat_sq_diff_sum writes sum_i (a_i - t_i)^2
 * to out[0] in one launch (plus a one-word clear), independent of
 * scheduling: fixed per-workgroup trees, the partials added in order in
 * fp64.  Calls that share a workspace must not overlap.
 * msplat_sq_diff_grad writes grad_i = 2 (a_i - t_i) dL[0]. */
size_t msplat_sq_diff_sum_workspace_bytes(void);
int msplat_sq_diff_sum(const float *a, const float *t, int64_t n,
                       void *workspace, size_t workspace_bytes, float *out,
                       void *stream);
int msplat_sq_diff_grad(const float *a, const float *t, int64_t n,
                        const float *dL, float *grad, void *stream);

/* ---- training step next to the rasterizer --------------------------------
 *
 * msplat_adam_step: one Adam step (torch.optim.Adam, no weight decay, no
 * amsgrad) over `count` fp32 arrays in one launch per 8 arrays, replacing the
 * optimizer of tutorials/gs_2d.py:32-36,73-90 and tutorials/gs_3d.py:165-172
 * (torch.optim.Adam with one learning rate per Gaussian attribute).  Array k
 * has numel[k] elements, learning rate lr[k] and is at step step[k] (1 on
 * the first update); params, exp_avg and exp_avg_sq are updated in place. */
int msplat_adam_step(int count, float *const *params,
                     const float *const *grads, float *const *exp_avg,
                     float *const *exp_avg_sq, const int64_t *numel,
                     const float *lr, const int64_t *step, float beta1,
                     float beta2, float eps, void *stream);

/* msplat_densify_accumulate: the consumer of dL/dndc, the gradient the
 * rasterizer hands back through `ndc` (msplat/alpha_blending.py:107-110,
 * msplat/__init__.py:33,90).  For every Gaussian seen by the view (radius >
 * 0, or a non-zero gradient when radius is NULL): accum += |dL/dndc|,
 * denom += 1.  dndc is [n,2]. */
int msplat_densify_accumulate(const float *dndc, const int32_t *radius,
                              int64_t n, float *accum, float *denom,
                              void *stream);

/* msplat_prune_clone: prune and refill at a fixed count, as
 * tutorials/gs_3d.py:193-225 does with torch ops: the M Gaussians with
 * opacity[g] < threshold (ascending index order) receive the rows of the M
 * Gaussians of largest score (a stable descending order; score = the
 * largest of the row's three `scale` entries when `score` is NULL), in every
 * one of the `narrays` row-major fp32 arrays (widths[k] floats per row).  A
 * pruned Gaussian that is also a donor keeps its row.  *d_count receives M
 * (-2 when a bounded wait of the sort expired: nothing is copied).  No host
 * synchronisation. */
size_t msplat_prune_clone_workspace_bytes(int64_t n);
int msplat_prune_clone(const float *opacity, float threshold,
                       const float *score, const float *scale, int64_t n,
                       int narrays, float *const *arrays, const int *widths,
                       void *workspace, size_t workspace_bytes,
                       int64_t *d_count, void *stream);

/* ---- the fused-loss forward without a workspace (no reference counterpart:
 * for callers that cannot allocate, such as code under HIP graph capture or
 * serving loops that forbid hipMalloc).  Same as msplat_rasterize_forward_l2
 * (same kernels, same results), with the loss partials in a scratch array of
 * msplat_noalloc_scratch_bytes() (4 MiB) that the library holds on every
 * GPU it runs on until the process ends; an image needing more
 * (msplat_rasterize_l2_workspace_bytes) fails with MSPLAT_ERR_CAPACITY
 * before anything is launched.  Calls on one GPU share the array: calls that
 * overlap on one GPU race.  Allocates nothing. */
size_t msplat_noalloc_scratch_bytes(void);
int msplat_rasterize_forward_l2_noalloc(
    const float *uv, const float *conic, const float *opacity,
    const float *feature, int64_t n, int C, const uint32_t *rect,
    const int32_t *cell_ids, const int32_t *cell_range, int cell, float bg,
    int W, int H, int ppl, int cull, const float *target, float *rendered,
    float *final_T, int32_t *ncontrib, int32_t *walk, float *loss,
    void *stream);

#ifdef __cplusplus
}
#endif

#endif /* MSPLAT_AMD_H */
