// C ABI of the gfx950 rasterizer (declared in include/msplat_amd.h).
//
// This is the library's only translation unit: it includes every kernel
// header once.  Each entry point validates its arguments, launches on the
// caller's stream and reports launch errors; nothing here allocates.
#include "../../include/msplat_amd.h"

#include <atomic>
#include <cmath>
#include <cstdlib>
#include <cstdarg>
#include <cstdio>
#include <map>
#include <set>
#include <mutex>
#include <string>
#include <utility>
#include <vector>

#include "blend.hpp"
#include "loss.hpp"
#include "common.hpp"
#include "preprocess.hpp"
#include "sh.hpp"
#include "sort.hpp"
#include "train.hpp"

using namespace msplat;

namespace {

thread_local std::string g_error;

int fail(int code, const char *fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  g_error = buf;
  return code;
}

int launched(const char *what) {
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) return fail(MSPLAT_ERR_LAUNCH, "%s: %s", what, hipGetErrorString(e));
  return MSPLAT_OK;
}

hipStream_t as_stream(void *s) { return static_cast<hipStream_t>(s); }

// ---- optional kernel timing: HIP events on the launch stream ---------------
// While enabled, the main kernel of each entry point is bracketed by a pair of
// events recorded on the caller's stream; msplat_timing_read synchronises on
// them afterwards.  Disabled (the default), a launch pays one atomic load.

struct TimingState {
  std::atomic<bool> on{false};
  std::mutex mu;
  std::set<std::string> only;  // kernels to time; empty: all
  std::map<std::string, std::vector<std::pair<hipEvent_t, hipEvent_t>>> recorded;
  std::vector<hipEvent_t> pool;
};

TimingState &timing() {
  static TimingState t;
  return t;
}

hipEvent_t pooled_event(TimingState &t) {  // caller holds t.mu
  if (!t.pool.empty()) {
    hipEvent_t e = t.pool.back();
    t.pool.pop_back();
    return e;
  }
  hipEvent_t e = nullptr;
  return hipEventCreate(&e) == hipSuccess ? e : nullptr;
}

class Timed {
 public:
  Timed(const char *name, hipStream_t s) : name_(name), s_(s) {
    TimingState &t = timing();
    if (!t.on.load(std::memory_order_relaxed)) return;
    std::lock_guard<std::mutex> g(t.mu);
    if (!t.only.empty() && !t.only.count(name)) return;
    a_ = pooled_event(t);
    b_ = pooled_event(t);
    if (a_ && b_ && hipEventRecord(a_, s_) == hipSuccess) return;
    if (a_) t.pool.push_back(a_);
    if (b_) t.pool.push_back(b_);
    a_ = b_ = nullptr;
  }
  ~Timed() {
    if (!a_) return;
    TimingState &t = timing();
    std::lock_guard<std::mutex> g(t.mu);
    if (hipEventRecord(b_, s_) == hipSuccess) {
      t.recorded[name_].emplace_back(a_, b_);
    } else {
      t.pool.push_back(a_);
      t.pool.push_back(b_);
    }
  }
  Timed(const Timed &) = delete;
  Timed &operator=(const Timed &) = delete;

 private:
  const char *name_;
  hipStream_t s_;
  hipEvent_t a_ = nullptr, b_ = nullptr;
};

#define REQUIRE(cond, ...)                                         \
  do {                                                             \
    if (!(cond)) return fail(MSPLAT_ERR_ARGUMENT, __VA_ARGS__);    \
  } while (0)

bool valid_image(int W, int H) { return W > 0 && H > 0 && W <= (1 << 16) && H <= (1 << 16); }

size_t align256(size_t b) { return (b + 255) & ~size_t(255); }

// Workspace carving: sequential 256-byte aligned slices.
struct Carver {
  size_t off = 0;
  size_t take(size_t bytes) {
    const size_t o = off;
    off += align256(bytes);
    return o;
  }
};

struct PrepLayout {
  size_t area, ka, va, kb, vb, offs, hist, sums, ghist, status, flag, nlisted, area_s, rect_s, bstatus, ctotal,
      dtotal, total;
};

// Workgroups of a camera-gradient adjoint over n rows: at most kCamGrid.
unsigned cam_grid(int64_t n) {
  const unsigned g = grid_1d(n, kGeomThreads);
  return g < unsigned(kCamGrid) ? g : unsigned(kCamGrid);
}

// Workgroups (of sortk::kEmitThreads threads) of the cell-list emission over
// k entries, and so its look-back words; at least one word is always carved.
// The status buffer is sized for that kernel's workgroup only.
int64_t emit_blocks(int64_t k) { return k > 0 ? (k + sortk::kEmitThreads - 1) / sortk::kEmitThreads : 1; }

// (W, H are part of the size query so that the layout may depend on the
// tile grid; it currently does not.)
PrepLayout prep_layout(int64_t n, int W, int H) {
  Carver c;
  PrepLayout L;
  const size_t un = size_t(n);
  L.area = c.take(un * 4);
  L.ka = c.take(un * 4);
  L.va = c.take(un * 4);
  L.kb = c.take(un * 4);
  L.vb = c.take(un * 4);
  L.offs = c.take(un * 8);
  const int64_t nt = sortk::radix_tiles(n);
  L.hist = c.take(size_t(256 * nt) * 4);
  int64_t ns = sortk::scan_tiles(256 * nt);
  if (sortk::scan_tiles(n) > ns) ns = sortk::scan_tiles(n);
  L.sums = c.take(size_t(ns > 0 ? ns : 1) * 8);
  L.ghist = c.take(size_t(sortk::ghist_words()) * 4);
  L.status = c.take(size_t(sortk::status_words(n)) * 4);
  L.flag = c.take(4);
  L.nlisted = c.take(8);
  L.area_s = c.take(un * 4);  // cell counts and tile rectangles in depth order (cell binning)
  L.rect_s = c.take(un * 8);
  L.bstatus = c.take(size_t(sortk::prep_blocks(n) + 1) * 8);  // look-back words of the counting pass, its ticket
  L.ctotal = c.take(size_t(sortk::kCtSlices) * sortk::kFusedCells * 4);  // entries per cell (fused emission)
  L.dtotal = c.take(size_t(sortk::kCtSlices) * sortk::kMaxPasses * 256 * 4);  // depth digit totals, sliced
  (void)W;
  (void)H;
  L.total = c.off;
  return L;
}

struct EmitLayout {
  size_t ca, ga, cb, gb, owner, hist, sums, ghist, status, estatus, cstatus, total;
};

EmitLayout emit_layout(int64_t k, int W, int H) {
  Carver c;
  EmitLayout L;
  const size_t uk = size_t(k);
  (void)W;
  (void)H;
  L.ca = c.take(uk * 4);
  L.ga = c.take(uk * 4);
  L.cb = c.take(uk * 4);
  L.gb = c.take(uk * 4);
  L.owner = c.take(size_t(sortk::emit_chunks(k)) * 8);
  const int64_t nt = sortk::radix_tiles(k);
  L.hist = c.take(size_t(256 * nt) * 4);
  const int64_t ns = sortk::scan_tiles(256 * nt);
  L.sums = c.take(size_t(ns > 0 ? ns : 1) * 8);
  L.ghist = c.take(size_t(sortk::ghist_words()) * 4);
  L.status = c.take(size_t(sortk::status_words(k)) * 4);
  L.estatus = c.take(size_t(emit_blocks(k) + 1) * 8);  // look-back words of the cell-list emission, its ticket
  // the fused emission: per-cell look-back words of each workgroup, then its ticket
  L.cstatus = c.take((size_t(emit_blocks(k)) * sortk::kFusedCells + 1) * 4);
  L.total = c.off;
  return L;
}

template <typename T>
T *at(void *base, size_t off) {
  return reinterpret_cast<T *>(static_cast<char *>(base) + off);
}
template <typename T>
const T *at(const void *base, size_t off) {
  return reinterpret_cast<const T *>(static_cast<const char *>(base) + off);
}

int bit_length(uint32_t v) { return v ? 32 - __builtin_clz(v) : 0; }

// Channel bucket for one blending pass (rem <= 64).
int channel_bucket(int rem) {
  if (rem <= 1) return 1;
  if (rem <= 2) return 2;
  if (rem <= 3) return 3;
  if (rem <= 4) return 4;
  if (rem <= 8) return 8;
  if (rem <= 16) return 16;
  if (rem <= 32) return 32;
  return 64;
}

int sh_degree(int D) {
  for (int l = 0; l <= shc::kMaxDeg; ++l)
    if ((l + 1) * (l + 1) == D) return l;
  return -1;
}

bool aligned16(const void *p) { return (reinterpret_cast<uintptr_t>(p) & 15) == 0; }

int64_t imax64(int64_t a, int64_t b) { return a > b ? a : b; }

// Workgroup shape of the LDS-staged SH kernels: G rows of odd stride S in at
// most 64 KiB of LDS; G = 0 when a row is too long (direct kernels instead).
ShGeom sh_geom(int C, int D) {
  ShGeom g;
  g.R = C * D;
  g.S = g.R | 1;
  g.G = 0;
  for (int G = 256; G >= 32; G >>= 1)
    if (size_t(G) * g.S * 4 <= 65536) {
      g.G = G;
      break;
    }
  g.gdivR = g.G ? g.G / g.R : 0;
  g.gmodR = g.G ? g.G % g.R : 0;
  return g;
}

template <int L, class Dir>
void sh_fwd_launch(int64_t n, int C, const float *shs, const Dir &dir, const uint8_t *vis, float *val, int ld,
                   hipStream_t s) {
  const ShGeom g = sh_geom(C, (L + 1) * (L + 1));
  Timed timed_("sh_fwd", s);
  if constexpr (L <= kShGcMaxDeg) {
    if (C <= kWave && aligned16(shs)) {
      sh_fwd_gc_kernel<L, Dir><<<grid_1d(n * C, 256 * sh_fwd_rows(L)), 256, 0, s>>>(n, C, shs, dir, vis, val, ld);
      return;
    }
  }
  if (g.G)
    sh_fwd_lds_kernel<L, Dir><<<grid_1d(n, g.G), g.G, size_t(g.G) * g.S * 4, s>>>(n, C, shs, dir, vis, val, ld,
                                                                                   g);
  else
    sh_fwd_kernel<L, Dir><<<grid_1d(n, kShThreads), kShThreads, 0, s>>>(n, C, shs, dir, vis, val, ld);
}
template <int L, class Dir>
void sh_bwd_launch(int64_t n, int C, const float *shs, const Dir &dir, const uint8_t *vis, const float *dval,
                   int ld, float *dshs, hipStream_t s) {
  const ShGeom g = sh_geom(C, (L + 1) * (L + 1));
  Timed timed_("sh_bwd", s);
  if constexpr (L <= kShGcMaxDeg) {
    if (C <= kWave && aligned16(shs) && aligned16(dshs)) {
      const int64_t waves = (n + kWave / C - 1) / (kWave / C);
      sh_bwd_gc_kernel<L, Dir><<<grid_1d(waves * kWave, 256), 256, 0, s>>>(n, C, shs, dir, vis, dval, ld, dshs);
      return;
    }
  }
  if (g.G)
    sh_bwd_lds_kernel<L, Dir><<<grid_1d(n, g.G), g.G, size_t(g.G) * g.S * 4, s>>>(n, C, shs, dir, vis, dval,
                                                                                   ld, dshs, g);
  else
    sh_bwd_kernel<L, Dir><<<grid_1d(n, kShThreads), kShThreads, 0, s>>>(n, C, shs, dir, vis, dval, ld, dshs);
}

#define MSPLAT_SH_DISPATCH(L, CALL) \
  switch (L) {                      \
    case 0: CALL(0); break;         \
    case 1: CALL(1); break;         \
    case 2: CALL(2); break;         \
    case 3: CALL(3); break;         \
    case 4: CALL(4); break;         \
    case 5: CALL(5); break;         \
    case 6: CALL(6); break;         \
    case 7: CALL(7); break;         \
    case 8: CALL(8); break;         \
    case 9: CALL(9); break;         \
    default: CALL(10); break;       \
  }

sortk::RadixBuffers prep_radix(void *ws, const PrepLayout &L) {
  return sortk::RadixBuffers{at<uint32_t>(ws, L.hist), at<int64_t>(ws, L.sums), at<uint32_t>(ws, L.ghist),
                             at<uint32_t>(ws, L.status)};
}
sortk::RadixBuffers emit_radix(void *ws, const EmitLayout &L) {
  return sortk::RadixBuffers{at<uint32_t>(ws, L.hist), at<int64_t>(ws, L.sums), at<uint32_t>(ws, L.ghist),
                             at<uint32_t>(ws, L.status)};
}

// Stable order of the prepared (depth bits, id) pairs by depth -- four 8-bit
// LSD passes, ids end in L.va; only the first *dlen pairs when dlen is given
// -- and the exclusive scan of the per-Gaussian box counts in that order
// (total -> *d_count).
void depth_order(void *ws, const PrepLayout &L, int64_t n, int64_t *d_count, hipStream_t s,
                 const int64_t *dlen = nullptr) {
  uint32_t *ka = at<uint32_t>(ws, L.ka), *va = at<uint32_t>(ws, L.va);
  uint32_t *kb = at<uint32_t>(ws, L.kb), *vb = at<uint32_t>(ws, L.vb);
  const sortk::RadixBuffers rb = prep_radix(ws, L);
  int32_t *area = at<int32_t>(ws, L.area);
  {
    Timed timed_("depth_sort", s);
    sortk::radix_sort(ka, va, kb, vb, n, 32, 4, rb, s, dlen);  // a -> b -> a -> b -> a
  }
  sortk::exclusive_scan<int32_t, int64_t>(area, va, n, at<int64_t>(ws, L.sums), at<int64_t>(ws, L.offs), d_count,
                                          s, dlen);
}

// Emits the (class, id) pairs of every entry in depth order -- classes are
// 1 + global cell (cells of S x S tiles, per-view grids and rows from `vw`)
// -- and sorts them stably by class.  The ids end
// in `ids`; returns the sorted classes (in the workspace).
const uint32_t *emit_by_class(const float *uv, const float *depth, const int32_t *radius, int64_t n,
                              const Views &vw, int S, const void *prepared, int64_t k, const int64_t *dlen,
                              void *ws, const EmitLayout &L, uint32_t *ids, hipStream_t s, bool direct = false) {
  const int ncls = vw.cell0[vw.count];
  const PrepLayout P = prep_layout(n, 1, 1);
  const uint32_t *order = at<uint32_t>(prepared, P.va);
  // the counting pass's and depth sort's error word: the emission reads
  // nothing of an order it marks undefined
  const uint32_t *upstream = sortk::radix_error_word(prep_radix(const_cast<void *>(prepared), P));
  const int64_t *offs = at<int64_t>(prepared, P.offs);
  uint32_t *ca = at<uint32_t>(ws, L.ca), *ga = at<uint32_t>(ws, L.ga);
  uint32_t *cb = at<uint32_t>(ws, L.cb), *gb = at<uint32_t>(ws, L.gb);
  int64_t *owner = at<int64_t>(ws, L.owner);
  const sortk::RadixBuffers rb = emit_radix(ws, L);
  if (direct) {  // cell lists: one thread per listed row, offsets by a single-pass scan
    Timed timed_("emit", s);
    const int64_t rows = n < k ? n : k;  // every listed row has a cell: at most k rows hold slots below k
    sortk::emit_direct_kernel<<<grid_1d(rows, sortk::kEmitThreads), sortk::kEmitThreads, 0, s>>>(
        order, at<int32_t>(prepared, P.area_s), at<uint2>(prepared, P.rect_s), n, at<int64_t>(prepared, P.nlisted),
        k, dlen, vw, S, ca, ga, at<uint64_t>(ws, L.estatus), emit_blocks(k), sortk::radix_error_word(rb), upstream);
  } else {  // equal output chunks per workgroup (tile rectangles of thousands of tiles)
    sortk::emit_owner_kernel<<<grid_1d(n, sortk::kThreads), sortk::kThreads, 0, s>>>(
        at<int32_t>(prepared, P.area), order, offs, n, sortk::emit_chunks(k), owner, upstream);
    Timed timed_("emit", s);
    sortk::emit_chunk_kernel<<<dim3(unsigned(sortk::emit_chunks(k))), sortk::kThreads, 0, s>>>(
        uv, depth, radius, order, offs, owner, n, k, dlen, vw, S, ca, ga, upstream);
  }
  const int bits = bit_length(uint32_t(ncls));  // classes 0..ncls
  const int npass = (bits + 7) / 8;
  Timed timed_("class_sort", s);
  // cell lists: the emission's launch zeroed ghist
  return sortk::radix_sort(ca, ga, cb, gb, k, bits, npass, rb, s, dlen, ids, {}, direct) ? cb : ca;
}

// The emission writes straight into the cell lists when there are at most
// kFusedCells cells (the default; MSPLAT_FUSED_EMIT=0 keeps the emission +
// class sort + range launches).
bool fused_emit_enabled() {
  static const bool on = [] {
    const char *e = std::getenv("MSPLAT_FUSED_EMIT");
    return !(e && e[0] == '0');
  }();
  return on;
}
bool fused_emit(const Views &vw) { return fused_emit_enabled() && vw.cell0[vw.count] <= sortk::kFusedCells; }

// Cells of S x S tiles on the tile grid of a W x H image.
int cells_x(int W, int S) { return (tiles_x(W) + S - 1) / S; }
int cells_y(int H, int S) { return (tiles_y(H) + S - 1) / S; }
bool valid_cell(int S) { return S >= 1 && S <= 4096; }

// The Views table of B images (cells of S tiles; S = 1 makes cells = tiles).
// View b draws Gaussians [first[b], first[b] + count[b]) of the shared arrays
// into per-view rows vbase[b]..; first = count = NULL means every view draws
// all n Gaussians (rows b * n + g).  Returns an empty string or what is wrong.
std::string make_views(int B, const int *W, const int *H, int S, const int64_t *first, const int64_t *count,
                       int64_t n, Views &v) {
  if (B < 1 || B > kMaxViews) return "view count " + std::to_string(B) + " outside 1.." + std::to_string(kMaxViews);
  if (!W || !H) return "null image sizes";
  if ((first == nullptr) != (count == nullptr)) return "first and count must both be given or both be NULL";
  v = Views{};
  v.count = B;
  int64_t tiles = 0, cells = 0, pix = 0, rows = 0, groups = 0;
  for (int b = 0; b < B; ++b) {
    const int64_t f = first ? first[b] : 0, c = count ? count[b] : n;
    if (f < 0 || c < 0) return "negative Gaussian range of view " + std::to_string(b);
    if (f > n || c > n - f)
      return "Gaussian range [" + std::to_string(f) + ", " + std::to_string(f) + " + " + std::to_string(c) +
             ") of view " + std::to_string(b) + " exceeds the " + std::to_string(n) + " Gaussians";
    v.vbase[b] = rows;
    v.gbase[b] = f;
    rows += c;
    if (!valid_image(W[b], H[b]))
      return "bad image size " + std::to_string(W[b]) + "x" + std::to_string(H[b]) + " of view " + std::to_string(b);
    v.W[b] = W[b];
    v.H[b] = H[b];
    v.gx[b] = uint16_t(tiles_x(W[b]));
    v.gy[b] = uint16_t(tiles_y(H[b]));
    v.ncx[b] = uint16_t(cells_x(W[b], S));
    v.tile0[b] = int(tiles);
    v.group0[b] = int(groups);
    v.cell0[b] = int(cells);
    v.pix0[b] = pix;
    tiles += int64_t(v.gx[b]) * v.gy[b];
    groups += int64_t((v.gx[b] + 1) / 2) * ((v.gy[b] + 1) / 2);
    cells += int64_t(v.ncx[b]) * cells_y(H[b], S);
    pix += int64_t(W[b]) * H[b];
    if (tiles >= INT32_MAX || cells >= INT32_MAX) return "too many tiles";
  }
  if (rows > int64_t(INT32_MAX)) return "more than 2^31 per-view rows";
  v.tile0[B] = int(tiles);
  v.group0[B] = int(groups);
  v.cell0[B] = int(cells);
  v.pix0[B] = pix;
  v.vbase[B] = rows;
  return "";
}

Views one_view(int W, int H, int S, int64_t n) {
  Views v;
  make_views(1, &W, &H, S, nullptr, nullptr, n, v);
  return v;
}

// ---- blending launchers (tile lists and cell lists) -----------------------

// The grouped adjoint is the default; MSPLAT_GROUPED_ADJOINT=0 selects one
// wave per tile (for comparison).
bool grouped_adjoint() {
  static const bool on = [] {
    const char *e = std::getenv("MSPLAT_GROUPED_ADJOINT");
    return !(e && e[0] == '0');
  }();
  return on;
}

int pick_ppl(int requested, int cm) {
  int p = requested <= 0 ? 4 : (requested >= 4 ? 4 : (requested >= 2 ? 2 : 1));
  const int mx = blend_max_ppl(cm);
  return p < mx ? p : mx;
}

template <int CM, bool kWalk>
void fwd_launch(int p, BlendIn in, float *out, float *final_T, int32_t *ncontrib, int32_t *walk, hipStream_t s) {
  in.per_xcd = blend_per_xcd(in.ntiles, p);
  Timed timed_(kWalk ? "rasterize_fwd" : "blend_fwd", s);
  const dim3 grid(unsigned(8 * in.per_xcd));
  if constexpr (blend_max_ppl(CM) >= 4) {
    if (p == 4) {
      blend_fwd_kernel<CM, 4, kWalk><<<grid, kBlendLanes, 0, s>>>(in, out, final_T, ncontrib, walk);
      return;
    }
  }
  if constexpr (blend_max_ppl(CM) >= 2) {
    if (p == 2) {
      blend_fwd_kernel<CM, 2, kWalk><<<grid, kBlendLanes, 0, s>>>(in, out, final_T, ncontrib, walk);
      return;
    }
  }
  blend_fwd_kernel<CM, 1, kWalk><<<grid, kBlendLanes, 0, s>>>(in, out, final_T, ncontrib, walk);
}

template <int CM, bool kWalk>
void bwd_launch(int p, BlendIn in, const float *final_T, const int32_t *ncontrib, const int32_t *walk,
                const float *dout, float *duv, float *dconic, float *dopac, float *dfeat, hipStream_t s) {
  in.per_xcd = blend_per_xcd(in.ntiles, p);
  Timed timed_(kWalk ? "rasterize_bwd" : "blend_bwd", s);
  const dim3 grid(unsigned(8 * in.per_xcd));
  if constexpr (kWalk && blend_max_ppl(CM) >= 4 && CM <= 8) {
    // 2 x 2 tile groups sharing one LDS gradient accumulator (cells of an
    // even number of tiles, so that a group lies in one cell)
    if (p == 4 && in.cell % 2 == 0 && grouped_adjoint()) {
      in.per_xcd = (in.ngroups + 7) / 8;
      blend_bwd_kernel<CM, 4, true, 4><<<dim3(unsigned(8 * in.per_xcd)), 4 * kBlendLanes, 0, s>>>(
          in, final_T, ncontrib, walk, dout, duv, dconic, dopac, dfeat);
      return;
    }
  }
  if constexpr (blend_max_ppl(CM) >= 4) {
    if (p == 4) {
      blend_bwd_kernel<CM, 4, kWalk><<<grid, kBlendLanes, 0, s>>>(in, final_T, ncontrib, walk, dout, duv, dconic,
                                                                  dopac, dfeat);
      return;
    }
  }
  if constexpr (blend_max_ppl(CM) >= 2) {
    if (p == 2) {
      blend_bwd_kernel<CM, 2, kWalk><<<grid, kBlendLanes, 0, s>>>(in, final_T, ncontrib, walk, dout, duv, dconic,
                                                                  dopac, dfeat);
      return;
    }
  }
  blend_bwd_kernel<CM, 1, kWalk><<<grid, kBlendLanes, 0, s>>>(in, final_T, ncontrib, walk, dout, duv, dconic,
                                                              dopac, dfeat);
}

#define MSPLAT_BLEND_DISPATCH(CM, CALL) \
  switch (CM) {                         \
    case 1: CALL(1); break;             \
    case 2: CALL(2); break;             \
    case 3: CALL(3); break;             \
    case 4: CALL(4); break;             \
    case 8: CALL(8); break;             \
    case 16: CALL(16); break;           \
    case 32: CALL(32); break;           \
    default: CALL(64); break;           \
  }

// Forward over all channels, 64 per pass (each pass walks the lists again).
// Per-wave slots of the fused loss per 64-channel pass (any pixels-per-lane);
// after them the workspace holds the sum's per-workgroup parts and ticket.
int64_t l2_slots(int64_t ntiles) { return 4 * ntiles + 8; }
int64_t l2_passes(int C) { return C > 0 ? (C + 63) / 64 : 1; }
int64_t l2_part_floats(int64_t ntiles, int C) { return (l2_slots(ntiles) * l2_passes(C) + 63) / 64 * 64; }
size_t l2_workspace(int64_t ntiles, int C) {
  return size_t(l2_part_floats(ntiles, C)) * sizeof(float) + size_t(kSumParts) * sizeof(double) + 256;
}

template <bool kWalk>
void blend_forward(BlendIn in, int ppl, float *out, float *final_T, int32_t *ncontrib, int32_t *walk,
                   hipStream_t s, float *loss = nullptr) {
  // (the sum's ticket, after its parts, is zeroed with the partials)
  const int64_t l2_words = in.l2_part ? l2_part_floats(in.ntiles, in.C) + 2 * kSumParts + 1 : 0;
  launch_fill_zero_words({{kWalk ? walk : nullptr, 2 * int64_t(in.ntiles)}, {in.l2_part, l2_words}}, s);
  int c0 = 0;
  do {
    in.c0 = c0;
    const int rem = in.C - c0 < 64 ? in.C - c0 : 64;
    const int cm = channel_bucket(rem);
    const int p = pick_ppl(ppl, cm);
#define CALL(CMV) fwd_launch<CMV, kWalk>(p, in, out, final_T, ncontrib, walk, s)
    MSPLAT_BLEND_DISPATCH(cm, CALL)
#undef CALL
    c0 += 64;
  } while (c0 < in.C);
  if (in.l2_part && loss) {
    const int64_t nv = in.l2_slots * l2_passes(in.C);
    double *parts = reinterpret_cast<double *>(in.l2_part + l2_part_floats(in.ntiles, in.C));
    uint32_t *ticket = reinterpret_cast<uint32_t *>(parts + kSumParts);
    sum_floats_kernel<<<sum_floats_blocks(nv), kSumThreads, 0, s>>>(in.l2_part, nv, parts, ticket, loss);
  }
}

// Adjoint: zero-fills the gradients, then accumulates them pass by pass.
// m = per-view entries (uv / conic rows), n = Gaussians (opacity rows),
// nf = feature rows.
template <bool kWalk>
void blend_backward(BlendIn in, int ppl, int64_t m, int64_t n, int64_t nf, const float *final_T,
                    const int32_t *ncontrib, const int32_t *walk, const float *dout, float *duv, float *dconic,
                    float *dopac, float *dfeat, hipStream_t s) {
  launch_fill_zero_words({{duv, 2 * m}, {dconic, 3 * m}, {dopac, n}, {in.C ? dfeat : nullptr, nf * in.C}}, s);
  if (in.C == 0) return;  // no channel, no gradient
  for (int c0 = 0; c0 < in.C; c0 += 64) {
    in.c0 = c0;
    const int rem = in.C - c0 < 64 ? in.C - c0 : 64;
    const int cm = channel_bucket(rem);
    const int p = pick_ppl(ppl, cm);
#define CALL(CMV) bwd_launch<CMV, kWalk>(p, in, final_T, ncontrib, walk, dout, duv, dconic, dopac, dfeat, s)
    MSPLAT_BLEND_DISPATCH(cm, CALL)
#undef CALL
  }
}

BlendIn blend_inputs(const float *uv, const float *conic, const float *opacity, const float *feature, int C,
                     const int32_t *list, const int32_t *range, const uint32_t *rect, int cell, int cull, float bg,
                     const Views &vw, int feat_per_view) {
  BlendIn in{};
  in.uv = uv;
  in.conic = conic;
  in.opacity = opacity;
  in.feature = feature;
  in.C = C;
  in.c0 = 0;
  in.list = list;
  in.range = range;
  in.rect = rect;
  in.cell = cell;
  in.cull = cell > 0 && cull != 0;
  in.bg = bg;
  in.feat_per_view = feat_per_view;
  in.views = vw;
  in.ntiles = vw.tile0[vw.count];
  in.ngroups = vw.group0[vw.count];
  in.per_xcd = 0;
  return in;
}
}  // namespace

extern "C" {

// ---------------------------------------------------------------- loss

size_t msplat_sq_diff_sum_workspace_bytes(void) { return align256(sizeof(float) * kSumBlocks) + 256; }

int msplat_sq_diff_sum(const float *a, const float *t, int64_t n, void *workspace, size_t workspace_bytes,
                       float *out, void *stream) {
  REQUIRE(n >= 0, "sq_diff_sum: n = %lld", (long long)n);
  REQUIRE(out, "sq_diff_sum: null output");
  const size_t need = msplat_sq_diff_sum_workspace_bytes();
  if (workspace_bytes < need)
    return fail(MSPLAT_ERR_WORKSPACE, "sq_diff_sum: workspace %zu < %zu bytes", workspace_bytes, need);
  REQUIRE(workspace && (n == 0 || (a && t)), "sq_diff_sum: null pointer");
  hipStream_t s = as_stream(stream);
  float *partial = static_cast<float *>(workspace);
  uint32_t *ticket = reinterpret_cast<uint32_t *>(static_cast<char *>(workspace) + align256(sizeof(float) * kSumBlocks));
  launch_fill_zero_words({{ticket, 1}, {out, 1}}, s);
  if (n == 0) return launched("sq_diff_sum");
  const int64_t want = (n + 4 * kSumThreads - 1) / (4 * kSumThreads);
  const unsigned blocks = unsigned(want < kSumBlocks ? want : kSumBlocks);
  sq_diff_sum_kernel<<<blocks, kSumThreads, 0, s>>>(a, t, n, partial, ticket, out);
  return launched("sq_diff_sum");
}

int msplat_sq_diff_grad(const float *a, const float *t, int64_t n, const float *dL, float *grad, void *stream) {
  REQUIRE(n >= 0, "sq_diff_grad: n = %lld", (long long)n);
  if (n == 0) return MSPLAT_OK;
  REQUIRE(a && t && dL && grad, "sq_diff_grad: null pointer");
  const int64_t want = (n + kLossThreads - 1) / kLossThreads;
  const unsigned blocks = unsigned(want < 8192 ? want : 8192);
  sq_diff_grad_kernel<<<blocks, kLossThreads, 0, as_stream(stream)>>>(a, t, n, dL, grad);
  return launched("sq_diff_grad");
}

// ---- training step: Adam, densification statistics, prune and refill --------

int msplat_adam_step(int count, float *const *params, const float *const *grads, float *const *exp_avg,
                     float *const *exp_avg_sq, const int64_t *numel, const float *lr, const int64_t *step,
                     float beta1, float beta2, float eps, void *stream) {
  REQUIRE(count >= 0, "adam_step: count = %d", count);
  if (count == 0) return MSPLAT_OK;
  REQUIRE(params && grads && exp_avg && exp_avg_sq && numel && lr && step, "adam_step: null array");
  REQUIRE(beta1 >= 0.f && beta1 < 1.f && beta2 >= 0.f && beta2 < 1.f, "adam_step: betas (%g, %g) outside [0, 1)",
          double(beta1), double(beta2));
  hipStream_t s = as_stream(stream);
  for (int first = 0; first < count; first += train::kAdamGroups) {
    train::AdamArgs a{};
    // torch passes 1 - beta as a Python double rounded to fp32
    a.w1 = float(1.0 - double(beta1));
    a.b2 = beta2;
    a.w2 = float(1.0 - double(beta2));
    a.eps = eps;
    int64_t blocks = 0;
    for (int k = first; k < count && k < first + train::kAdamGroups; ++k) {
      REQUIRE(numel[k] >= 0, "adam_step: array %d has %lld elements", k, (long long)numel[k]);
      REQUIRE(step[k] >= 1, "adam_step: array %d at step %lld (the first step is 1)", k, (long long)step[k]);
      if (numel[k] == 0) continue;
      REQUIRE(params[k] && grads[k] && exp_avg[k] && exp_avg_sq[k], "adam_step: null pointer in array %d", k);
      train::AdamGroup &G = a.grp[a.count++];
      G.p = params[k];
      G.g = grads[k];
      G.m = exp_avg[k];
      G.v = exp_avg_sq[k];
      G.n = numel[k];
      G.blk0 = blocks;
      const double t = double(step[k]);
      const double bc1 = 1.0 - std::pow(double(beta1), t), bc2 = 1.0 - std::pow(double(beta2), t);
      G.neg_step = float(-(double(lr[k]) / bc1));
      G.bc2_sqrt = float(std::sqrt(bc2));
      G.vec = ((reinterpret_cast<uintptr_t>(G.p) | reinterpret_cast<uintptr_t>(G.g) |
                reinterpret_cast<uintptr_t>(G.m) | reinterpret_cast<uintptr_t>(G.v)) & 15) == 0;
      blocks += (G.n + train::kAdamBlockElems - 1) / train::kAdamBlockElems;
    }
    REQUIRE(blocks < (int64_t(1) << 31), "adam_step: %lld workgroups", (long long)blocks);
    if (blocks > 0) train::adam_kernel<<<unsigned(blocks), train::kAdamThreads, 0, s>>>(a);
  }
  return launched("adam_step");
}

int msplat_densify_accumulate(const float *dndc, const int32_t *radius, int64_t n, float *accum, float *denom,
                              void *stream) {
  REQUIRE(n >= 0, "densify_accumulate: n = %lld", (long long)n);
  if (n == 0) return MSPLAT_OK;
  REQUIRE(dndc && accum && denom, "densify_accumulate: null pointer");
  REQUIRE((reinterpret_cast<uintptr_t>(dndc) & 7) == 0, "densify_accumulate: dndc must be 8-byte aligned");
  train::densify_accumulate_kernel<<<grid_1d(n, 256), 256, 0, as_stream(stream)>>>(
      reinterpret_cast<const float2 *>(dndc), radius, n, accum, denom);
  return launched("densify_accumulate");
}

namespace {
struct PruneLayout {
  size_t flag, pos, ka, va, kb, vb, rank, prune, total, sums, hist, rsums, ghist, status, end;
};
PruneLayout prune_layout(int64_t n) {
  Carver c;
  PruneLayout L;
  const size_t un = size_t(n);
  L.flag = c.take(un * 4);
  L.pos = c.take(un * 4);
  L.ka = c.take(un * 4);
  L.va = c.take(un * 4);
  L.kb = c.take(un * 4);
  L.vb = c.take(un * 4);
  L.rank = c.take(un * 4);
  L.prune = c.take(un * 4);
  L.total = c.take(8);
  L.sums = c.take(size_t(sortk::scan_tiles(n) > 0 ? sortk::scan_tiles(n) : 1) * 8);
  const int64_t nt = sortk::radix_tiles(n);
  L.hist = c.take(size_t(256 * nt) * 4);
  const int64_t ns = sortk::scan_tiles(256 * nt);
  L.rsums = c.take(size_t(ns > 0 ? ns : 1) * 8);
  L.ghist = c.take(size_t(sortk::ghist_words()) * 4);
  L.status = c.take(size_t(sortk::status_words(n)) * 4);
  L.end = c.off;
  return L;
}
}  // namespace

size_t msplat_prune_clone_workspace_bytes(int64_t n) { return n < 0 ? 0 : prune_layout(n).end; }

int msplat_prune_clone(const float *opacity, float threshold, const float *score, const float *scale, int64_t n,
                       int narrays, float *const *arrays, const int *widths, void *workspace,
                       size_t workspace_bytes, int64_t *d_count, void *stream) {
  REQUIRE(n >= 0 && n <= int64_t(INT32_MAX), "prune_clone: n = %lld", (long long)n);
  REQUIRE(narrays >= 0 && narrays <= train::kDensifyArrays, "prune_clone: %d arrays (at most %d)", narrays,
          train::kDensifyArrays);
  REQUIRE(d_count, "prune_clone: null count pointer");
  const PruneLayout L = prune_layout(n);
  if (workspace_bytes < L.end)
    return fail(MSPLAT_ERR_WORKSPACE, "prune_clone: workspace %zu < %zu bytes", workspace_bytes, L.end);
  hipStream_t s = as_stream(stream);
  if (n == 0) {
    launch_fill_zero_words({{d_count, 2}}, s);
    return launched("prune_clone");
  }
  REQUIRE(workspace && opacity && (score || scale), "prune_clone: null pointer");
  REQUIRE(narrays == 0 || (arrays && widths), "prune_clone: null array table");
  train::DensifyRows rows{};
  for (int k = 0; k < narrays; ++k) {
    REQUIRE(arrays[k] && widths[k] >= 1, "prune_clone: array %d (width %d)", k, widths[k]);
    rows.a[k] = arrays[k];
    rows.w[k] = widths[k];
  }
  rows.count = narrays;
  uint32_t *flag = at<uint32_t>(workspace, L.flag), *pos = at<uint32_t>(workspace, L.pos);
  uint32_t *ka = at<uint32_t>(workspace, L.ka), *va = at<uint32_t>(workspace, L.va);
  uint32_t *kb = at<uint32_t>(workspace, L.kb), *vb = at<uint32_t>(workspace, L.vb);
  uint32_t *rank = at<uint32_t>(workspace, L.rank), *prune = at<uint32_t>(workspace, L.prune);
  int64_t *total = at<int64_t>(workspace, L.total);
  const sortk::RadixBuffers rb{at<uint32_t>(workspace, L.hist), at<int64_t>(workspace, L.rsums),
                               at<uint32_t>(workspace, L.ghist), at<uint32_t>(workspace, L.status)};
  launch_fill_zero_words({{rb.ghist, sortk::ghist_words()}}, s);
  const unsigned g = grid_1d(n, 256);
  train::densify_prepare_kernel<<<g, 256, 0, s>>>(opacity, threshold, score, scale, n, flag, ka, va);
  sortk::exclusive_scan<uint32_t, uint32_t>(flag, nullptr, n, at<int64_t>(workspace, L.sums), pos, total, s);
  const int r = sortk::radix_sort(ka, va, kb, vb, n, 32, 4, rb, s, nullptr, nullptr, {}, true);
  const uint32_t *order = r ? vb : va;
  train::densify_lists_kernel<<<g, 256, 0, s>>>(flag, pos, order, n, rank, prune);
  train::densify_count_kernel<<<1, 1, 0, s>>>(total, sortk::radix_error_word(rb), d_count);
  const int64_t cw = (n + 3) / 4 < train::kCopyBlocks ? (n + 3) / 4 : train::kCopyBlocks;  // 4 waves each
  if (narrays > 0) train::densify_copy_kernel<<<unsigned(cw), 256, 0, s>>>(prune, order, rank, n, d_count, rows);
  return launched("prune_clone");
}

int msplat_version(void) { return 200; }

int msplat_wait_status(int64_t *d_out, int clear, void *stream) {
  sortk::wait_status_kernel<<<1, 1, 0, as_stream(stream)>>>(d_out, clear);
  return launched("wait_status");
}

int msplat_debug_force_expiry(int on, void *stream) {
  sortk::set_force_expiry_kernel<<<1, 1, 0, as_stream(stream)>>>(on ? 1u : 0u);
  return launched("debug_force_expiry");
}

const char *msplat_last_error(void) { return g_error.c_str(); }

int msplat_timing_enable(int on) {
  return timing().on.exchange(on != 0) ? 1 : 0;
}

int msplat_timing_only(const char *kernels) {
  TimingState &t = timing();
  std::lock_guard<std::mutex> g(t.mu);
  t.only.clear();
  if (!kernels) return MSPLAT_OK;
  std::string cur;
  for (const char *c = kernels;; ++c) {
    if (*c == ',' || *c == '\0') {
      if (!cur.empty()) t.only.insert(cur);
      cur.clear();
      if (*c == '\0') break;
    } else if (*c != ' ') {
      cur += *c;
    }
  }
  return MSPLAT_OK;
}

int msplat_timing_read(const char *kernel, double *total_ms, int64_t *launches) {
  REQUIRE(kernel && total_ms && launches, "timing_read: null pointer");
  TimingState &t = timing();
  std::lock_guard<std::mutex> g(t.mu);
  double ms = 0.0;
  int64_t cnt = 0;
  auto it = t.recorded.find(kernel);
  if (it != t.recorded.end()) {
    for (auto &ev : it->second) {
      float e = 0.f;
      if (hipEventSynchronize(ev.second) == hipSuccess && hipEventElapsedTime(&e, ev.first, ev.second) == hipSuccess) {
        ms += e;
        ++cnt;
      }
      t.pool.push_back(ev.first);
      t.pool.push_back(ev.second);
    }
    t.recorded.erase(it);
  }
  *total_ms = ms;
  *launches = cnt;
  return MSPLAT_OK;
}

// ---------------------------------------------------------------- projection

int msplat_project_point_forward(const float *xyz, int64_t n, const float *intr, const float *extr, int W,
                                 int H, float nearest, float extent, float *uv, float *depth, void *stream) {
  REQUIRE(n >= 0, "project_point_forward: n = %lld", (long long)n);
  if (n == 0) return MSPLAT_OK;
  REQUIRE(xyz && intr && extr && uv && depth, "project_point_forward: null pointer");
  project_fwd_kernel<<<grid_1d(n, kGeomThreads), kGeomThreads, 0, as_stream(stream)>>>(xyz, n, intr, extr, W, H,
                                                                                       nearest, extent, uv, depth);
  return launched("project_point_forward");
}

int msplat_project_point_backward(const float *xyz, int64_t n, const float *intr, const float *extr,
                                  const float *depth, const float *dL_duv, const float *dL_ddepth, float *dL_dxyz,
                                  float *dL_dintr, float *dL_dextr, void *stream) {
  REQUIRE(n >= 0, "project_point_backward: n = %lld", (long long)n);
  hipStream_t s = as_stream(stream);
  const bool cam = dL_dintr || dL_dextr;
  launch_fill_zero_words({{dL_dintr, 4}, {dL_dextr, 12}}, s);
  if (n == 0) return launched("project_point_backward");
  REQUIRE(xyz && intr && extr && depth && dL_duv && dL_dxyz, "project_point_backward: null pointer");
  if (cam)
    project_bwd_kernel<true><<<grid_1d(n, kGeomThreads), kGeomThreads, 0, s>>>(
        xyz, n, intr, extr, depth, dL_duv, dL_ddepth, dL_dxyz, dL_dintr, dL_dextr);
  else
    project_bwd_kernel<false><<<grid_1d(n, kGeomThreads), kGeomThreads, 0, s>>>(
        xyz, n, intr, extr, depth, dL_duv, dL_ddepth, dL_dxyz, nullptr, nullptr);
  return launched("project_point_backward");
}

// ---------------------------------------------------------------- covariance

int msplat_compute_cov3d_forward(const float *scale, const float *rotate, const uint8_t *visible, int64_t n,
                                 float *cov3d, void *stream) {
  REQUIRE(n >= 0, "compute_cov3d_forward: n = %lld", (long long)n);
  if (n == 0) return MSPLAT_OK;
  REQUIRE(scale && rotate && cov3d, "compute_cov3d_forward: null pointer");
  cov3d_fwd_kernel<<<grid_1d(n, kGeomThreads), kGeomThreads, 0, as_stream(stream)>>>(scale, rotate, visible, n,
                                                                                     cov3d);
  return launched("compute_cov3d_forward");
}

int msplat_compute_cov3d_backward(const float *scale, const float *rotate, const uint8_t *visible,
                                  const float *dL_dcov3d, int64_t n, float *dL_dscale, float *dL_drotate,
                                  void *stream) {
  REQUIRE(n >= 0, "compute_cov3d_backward: n = %lld", (long long)n);
  if (n == 0) return MSPLAT_OK;
  REQUIRE(scale && rotate && dL_dcov3d && dL_dscale && dL_drotate, "compute_cov3d_backward: null pointer");
  cov3d_bwd_kernel<<<grid_1d(n, kGeomThreads), kGeomThreads, 0, as_stream(stream)>>>(
      scale, rotate, visible, dL_dcov3d, n, dL_dscale, dL_drotate);
  return launched("compute_cov3d_backward");
}

// ---------------------------------------------------------------- EWA

int msplat_ewa_project_forward(const float *xyz, const float *cov3d, const float *intr, const float *extr,
                               const float *uv, const uint8_t *visible, int64_t n, int W, int H, float *conic,
                               int32_t *radius, int32_t *tiles, void *stream) {
  REQUIRE(n >= 0, "ewa_project_forward: n = %lld", (long long)n);
  REQUIRE(valid_image(W, H), "ewa_project_forward: bad image size %dx%d", W, H);
  if (n == 0) return MSPLAT_OK;
  REQUIRE(xyz && cov3d && intr && extr && uv && conic && radius && tiles, "ewa_project_forward: null pointer");
  ewa_fwd_kernel<<<grid_1d(n, kGeomThreads), kGeomThreads, 0, as_stream(stream)>>>(
      xyz, cov3d, intr, extr, uv, visible, n, tiles_x(W), tiles_y(H), conic, radius, tiles);
  return launched("ewa_project_forward");
}

int msplat_ewa_project_backward(const float *xyz, const float *cov3d, const float *intr, const float *extr,
                                const int32_t *radius, const float *dL_dconic, int64_t n, float *dL_dxyz,
                                float *dL_dcov3d, float *dL_dintr, float *dL_dextr, void *stream) {
  REQUIRE(n >= 0, "ewa_project_backward: n = %lld", (long long)n);
  hipStream_t s = as_stream(stream);
  const bool cam = dL_dintr || dL_dextr;
  launch_fill_zero_words({{dL_dintr, 4}, {dL_dextr, 12}}, s);
  if (n == 0) return launched("ewa_project_backward");
  REQUIRE(xyz && cov3d && intr && extr && radius && dL_dconic && dL_dxyz && dL_dcov3d,
          "ewa_project_backward: null pointer");
  if (cam)
    ewa_bwd_kernel<true><<<grid_1d(n, kGeomThreads), kGeomThreads, 0, s>>>(
        xyz, cov3d, intr, extr, radius, dL_dconic, n, dL_dxyz, dL_dcov3d, dL_dintr, dL_dextr);
  else
    ewa_bwd_kernel<false><<<grid_1d(n, kGeomThreads), kGeomThreads, 0, s>>>(
        xyz, cov3d, intr, extr, radius, dL_dconic, n, dL_dxyz, dL_dcov3d, nullptr, nullptr);
  return launched("ewa_project_backward");
}

// ---------------------------------------------------------------- fused

int msplat_preprocess_forward_col(const float *xyz, const float *scale, const float *rotate, int64_t n,
                                  const float *intr, const float *extr, int W, int H, float nearest, float extent,
                                  float *uv, float *depth, float *conic, int32_t *radius, int32_t *tiles,
                                  float *depth_col, int ld_col, uint8_t *visible, void *stream) {
  REQUIRE(n >= 0, "preprocess_forward: n = %lld", (long long)n);
  REQUIRE(valid_image(W, H), "preprocess_forward: bad image size %dx%d", W, H);
  REQUIRE(!depth_col || ld_col >= 1, "preprocess_forward: depth column stride %d", ld_col);
  if (n == 0) return MSPLAT_OK;
  REQUIRE(xyz && scale && rotate && intr && extr && uv && depth && conic && radius && tiles,
          "preprocess_forward: null pointer");
  Timed timed_("preprocess_fwd", as_stream(stream));
  preprocess_fwd_kernel<<<grid_1d(n, kGeomThreads), kGeomThreads, 0, as_stream(stream)>>>(
      xyz, scale, rotate, n, intr, extr, W, H, tiles_x(W), tiles_y(H), nearest, extent, uv, depth, conic, radius,
      tiles, depth_col, ld_col, visible);
  return launched("preprocess_forward");
}

int msplat_preprocess_forward(const float *xyz, const float *scale, const float *rotate, int64_t n,
                              const float *intr, const float *extr, int W, int H, float nearest, float extent,
                              float *uv, float *depth, float *conic, int32_t *radius, int32_t *tiles,
                              void *stream) {
  return msplat_preprocess_forward_col(xyz, scale, rotate, n, intr, extr, W, H, nearest, extent, uv, depth, conic,
                                       radius, tiles, nullptr, 1, nullptr, stream);
}

int msplat_preprocess_backward_ld(const float *xyz, const float *scale, const float *rotate, int64_t n,
                                  const float *intr, const float *extr, const float *depth, const int32_t *radius,
                                  const float *dL_duv, const float *dL_ddepth, int ld_ddepth, const float *dL_dconic,
                                  float *dL_dxyz, float *dL_dscale, float *dL_drotate, float *dL_dintr,
                                  float *dL_dextr, void *stream) {
  REQUIRE(n >= 0, "preprocess_backward: n = %lld", (long long)n);
  REQUIRE(!dL_ddepth || ld_ddepth >= 1, "preprocess_backward: dL_ddepth stride %d", ld_ddepth);
  hipStream_t s = as_stream(stream);
  const bool cam = dL_dintr || dL_dextr;
  launch_fill_zero_words({{dL_dintr, 4}, {dL_dextr, 12}}, s);
  if (n == 0) return launched("preprocess_backward");
  REQUIRE(xyz && scale && rotate && intr && extr && depth && radius && dL_duv && dL_dconic && dL_dxyz &&
              dL_dscale && dL_drotate,
          "preprocess_backward: null pointer");
  Timed timed_("preprocess_bwd", s);
  if (cam)
    preprocess_bwd_kernel<true><<<cam_grid(n), kGeomThreads, 0, s>>>(
        xyz, scale, rotate, n, intr, extr, depth, radius, dL_duv, dL_ddepth, dL_dconic, dL_dxyz, dL_dscale,
        dL_drotate, dL_dintr, dL_dextr, ld_ddepth);
  else
    preprocess_bwd_kernel<false><<<grid_1d(n, kGeomThreads), kGeomThreads, 0, s>>>(
        xyz, scale, rotate, n, intr, extr, depth, radius, dL_duv, dL_ddepth, dL_dconic, dL_dxyz, dL_dscale,
        dL_drotate, nullptr, nullptr, ld_ddepth);
  return launched("preprocess_backward");
}

int msplat_preprocess_backward(const float *xyz, const float *scale, const float *rotate, int64_t n,
                               const float *intr, const float *extr, const float *depth, const int32_t *radius,
                               const float *dL_duv, const float *dL_ddepth, const float *dL_dconic, float *dL_dxyz,
                               float *dL_dscale, float *dL_drotate, float *dL_dintr, float *dL_dextr,
                               void *stream) {
  return msplat_preprocess_backward_ld(xyz, scale, rotate, n, intr, extr, depth, radius, dL_duv, dL_ddepth, 1,
                                       dL_dconic, dL_dxyz, dL_dscale, dL_drotate, dL_dintr, dL_dextr, stream);
}

// ---------------------------------------------------------------- SH

int msplat_compute_sh_forward(const float *shs, const float *dirs, const uint8_t *visible, int64_t n, int C, int D,
                              float *value, void *stream) {
  REQUIRE(n >= 0 && C >= 0, "compute_sh_forward: n = %lld, C = %d", (long long)n, C);
  const int L = sh_degree(D);
  REQUIRE(L >= 0, "compute_sh_forward: D = %d is not (deg+1)^2 with deg <= 10", D);
  if (n == 0 || C == 0) return MSPLAT_OK;
  REQUIRE(shs && dirs && value, "compute_sh_forward: null pointer");
  hipStream_t s = as_stream(stream);
#define CALL(LL) sh_fwd_launch<LL>(n, C, shs, ShDirArray{dirs, nullptr}, visible, value, C, s)
  MSPLAT_SH_DISPATCH(L, CALL)
#undef CALL
  return launched("compute_sh_forward");
}

int msplat_compute_sh_backward(const float *shs, const float *dirs, const uint8_t *visible,
                               const float *dL_dvalue, int64_t n, int C, int D, float *dL_dshs, float *dL_ddirs,
                               void *stream) {
  REQUIRE(n >= 0 && C >= 0, "compute_sh_backward: n = %lld, C = %d", (long long)n, C);
  const int L = sh_degree(D);
  REQUIRE(L >= 0, "compute_sh_backward: D = %d is not (deg+1)^2 with deg <= 10", D);
  if (n == 0) return MSPLAT_OK;
  hipStream_t s = as_stream(stream);
  if (C == 0) {
    launch_fill_zero(dL_ddirs, 3 * n, s);
    return launched("compute_sh_backward");
  }
  REQUIRE(shs && dirs && dL_dvalue && dL_dshs && dL_ddirs, "compute_sh_backward: null pointer");
#define CALL(LL) sh_bwd_launch<LL>(n, C, shs, ShDirArray{dirs, dL_ddirs}, visible, dL_dvalue, C, dL_dshs, s)
  MSPLAT_SH_DISPATCH(L, CALL)
#undef CALL
  return launched("compute_sh_backward");
}

int msplat_compute_sh_points_forward(const float *shs, const float *xyz, const float *campos, const uint8_t *visible,
                                     int64_t n, int C, int D, float *value, int ld_value, void *stream) {
  REQUIRE(n >= 0 && C >= 0, "compute_sh_points_forward: n = %lld, C = %d", (long long)n, C);
  const int L = sh_degree(D);
  REQUIRE(L >= 0, "compute_sh_points_forward: D = %d is not (deg+1)^2 with deg <= 10", D);
  if (n == 0 || C == 0) return MSPLAT_OK;
  REQUIRE(shs && xyz && campos && value, "compute_sh_points_forward: null pointer");
  REQUIRE(ld_value >= C, "compute_sh_points_forward: ld_value = %d < C = %d", ld_value, C);
  hipStream_t s = as_stream(stream);
  const ShDirFromPoints dir{xyz, campos, nullptr};
#define CALL(LL) sh_fwd_launch<LL>(n, C, shs, dir, visible, value, ld_value, s)
  MSPLAT_SH_DISPATCH(L, CALL)
#undef CALL
  return launched("compute_sh_points_forward");
}

}  // extern "C"

namespace {
int compute_sh_points_backward_impl(const char *name, const float *shs, const float *xyz, const float *campos,
                                    const uint8_t *visible, const float *dL_dvalue, int ld_dvalue, int64_t n, int C,
                                    int D, float *dL_dshs, float *dL_dxyz, bool add, void *stream) {
  REQUIRE(n >= 0 && C >= 0, "%s: n = %lld, C = %d", name, (long long)n, C);
  const int L = sh_degree(D);
  REQUIRE(L >= 0, "%s: D = %d is not (deg+1)^2 with deg <= 10", name, D);
  if (n == 0) return MSPLAT_OK;
  hipStream_t s = as_stream(stream);
  if (C == 0) {
    if (!add) launch_fill_zero(dL_dxyz, 3 * n, s);
    return launched(name);
  }
  REQUIRE(shs && xyz && campos && dL_dvalue && dL_dshs && dL_dxyz, "%s: null pointer", name);
  REQUIRE(ld_dvalue >= C, "%s: ld_dvalue = %d < C = %d", name, ld_dvalue, C);
  const ShDirFromPoints dir{xyz, campos, dL_dxyz, add};
#define CALL(LL) sh_bwd_launch<LL>(n, C, shs, dir, visible, dL_dvalue, ld_dvalue, dL_dshs, s)
  MSPLAT_SH_DISPATCH(L, CALL)
#undef CALL
  return launched(name);
}
}  // namespace

extern "C" {

int msplat_compute_sh_points_backward(const float *shs, const float *xyz, const float *campos,
                                      const uint8_t *visible, const float *dL_dvalue, int ld_dvalue, int64_t n,
                                      int C, int D, float *dL_dshs, float *dL_dxyz, void *stream) {
  return compute_sh_points_backward_impl("compute_sh_points_backward", shs, xyz, campos, visible, dL_dvalue,
                                         ld_dvalue, n, C, D, dL_dshs, dL_dxyz, false, stream);
}

int msplat_compute_sh_points_backward_add(const float *shs, const float *xyz, const float *campos,
                                          const uint8_t *visible, const float *dL_dvalue, int ld_dvalue, int64_t n,
                                          int C, int D, float *dL_dshs, float *dL_dxyz, void *stream) {
  return compute_sh_points_backward_impl("compute_sh_points_backward_add", shs, xyz, campos, visible, dL_dvalue,
                                         ld_dvalue, n, C, D, dL_dshs, dL_dxyz, true, stream);
}

// ---------------------------------------------------------------- sort

size_t msplat_sort_prepare_workspace_bytes(int64_t n, int W, int H) {
  if (!valid_image(W, H)) return 0;
  return prep_layout(n < 0 ? 0 : n, W, H).total;
}

int msplat_sort_prepare(const float *uv, const float *depth, const int32_t *radius, const int32_t *tiles, int64_t n,
                        int W, int H, void *workspace, size_t workspace_bytes, int64_t *d_count, void *stream) {
  REQUIRE(n >= 0 && n <= int64_t(INT32_MAX), "sort_prepare: n = %lld", (long long)n);
  REQUIRE(valid_image(W, H), "sort_prepare: bad image size %dx%d", W, H);
  REQUIRE(d_count, "sort_prepare: null count pointer");
  const PrepLayout L = prep_layout(n, W, H);
  if (workspace_bytes < L.total)
    return fail(MSPLAT_ERR_WORKSPACE, "sort_prepare: workspace %zu < %zu bytes", workspace_bytes, L.total);
  REQUIRE(workspace, "sort_prepare: null workspace");
  hipStream_t s = as_stream(stream);
  int32_t *flag = at<int32_t>(workspace, L.flag);
  uint32_t *error = sortk::radix_error_word(prep_radix(workspace, L));
  launch_fill_zero_words({{flag, 1}, {error, 1}}, s);
  if (n > 0) {
    REQUIRE(uv && depth && radius, "sort_prepare: null pointer");
    sortk::prepare_kernel<<<grid_1d(n, sortk::kThreads), sortk::kThreads, 0, s>>>(
        uv, depth, radius, tiles, n, tiles_x(W), tiles_y(H), at<int32_t>(workspace, L.area),
        at<uint32_t>(workspace, L.ka), at<uint32_t>(workspace, L.va), flag);
    depth_order(workspace, L, n, d_count, s);
  }
  sortk::finish_count_kernel<<<1, 1, 0, s>>>(flag, error, d_count, n);
  return launched("sort_prepare");
}

size_t msplat_sort_emit_workspace_bytes(int64_t n, int64_t k, int W, int H) {
  (void)n;
  if (!valid_image(W, H)) return 0;
  return emit_layout(k < 0 ? 0 : k, W, H).total;
}

int msplat_sort_emit(const float *uv, const float *depth, const int32_t *radius, int64_t n, int W, int H,
                     const void *prepared, int64_t k, void *workspace, size_t workspace_bytes, int32_t *idx_sorted,
                     int32_t *tile_range, int64_t *key_sorted, void *stream) {
  REQUIRE(n >= 0 && n <= int64_t(INT32_MAX), "sort_emit: n = %lld", (long long)n);
  REQUIRE(valid_image(W, H), "sort_emit: bad image size %dx%d", W, H);
  if (k < 0) return fail(MSPLAT_ERR_ARGUMENT, "sort_emit: tile counts disagree with the tile rectangles");
  if (k > int64_t(INT32_MAX))
    return fail(MSPLAT_ERR_CAPACITY, "sort_emit: %lld intersections exceed int32 indexing", (long long)k);
  REQUIRE(tile_range, "sort_emit: null tile_range");
  const EmitLayout L = emit_layout(k, W, H);
  if (workspace_bytes < L.total)
    return fail(MSPLAT_ERR_WORKSPACE, "sort_emit: workspace %zu < %zu bytes", workspace_bytes, L.total);
  hipStream_t s = as_stream(stream);
  const int gx = tiles_x(W), gy = tiles_y(H), T = gx * gy;
  REQUIRE(k == 0 || (uv && depth && radius && prepared && workspace && idx_sorted), "sort_emit: null pointer");
  uint32_t *error = k > 0 ? sortk::radix_error_word(emit_radix(workspace, L)) : nullptr;
  // step 1's error word (its depth sort): set when the caller went on after a count of -2
  const uint32_t *prep_error =
      k > 0 ? sortk::radix_error_word(prep_radix(const_cast<void *>(prepared), prep_layout(n, W, H))) : nullptr;
  launch_fill_zero_words({{tile_range, 2 * int64_t(T)}, {error, 1}}, s);
  if (k == 0) return launched("sort_emit");
  const uint32_t *sk = emit_by_class(uv, depth, radius, n, one_view(W, H, 1, n), 1, prepared, k, nullptr, workspace,
                                     L, reinterpret_cast<uint32_t *>(idx_sorted), s);
  sortk::range_vec_kernel<<<grid_1d((k + 3) / 4, sortk::kThreads), sortk::kThreads, 0, s>>>(sk, k, nullptr,
                                                                                           tile_range, T, error,
                                                                                           prep_error);
  if (key_sorted)
    sortk::sorted_keys_kernel<<<grid_1d(k, sortk::kThreads), sortk::kThreads, 0, s>>>(sk, idx_sorted, depth, k,
                                                                                      n, key_sorted);
  return launched("sort_emit");
}

// ---------------------------------------------------------------- cell binning

size_t msplat_bin_prepare_workspace_bytes(int64_t n, int W, int H, int cell) {
  if (!valid_image(W, H) || !valid_cell(cell)) return 0;
  return prep_layout(n < 0 ? 0 : n, W, H).total;
}

int msplat_bin_prepare(const float *uv, const float *depth, const int32_t *radius, int64_t n, int W, int H, int cell,
                       uint32_t *rect, void *workspace, size_t workspace_bytes, int64_t *d_count, void *stream) {
  return msplat_bin_views_prepare(uv, depth, radius, n, 1, &W, &H, nullptr, nullptr, cell, rect, workspace,
                                  workspace_bytes, d_count, stream);
}

size_t msplat_bin_emit_workspace_bytes(int64_t n, int64_t count, int W, int H, int cell) {
  (void)n;
  if (!valid_image(W, H) || !valid_cell(cell)) return 0;
  return emit_layout(count < 0 ? 0 : count, W, H).total;
}

int msplat_bin_emit(const float *uv, const float *depth, const int32_t *radius, int64_t n, int W, int H, int cell,
                    const void *prepared, int64_t count, const int64_t *d_count, void *workspace,
                    size_t workspace_bytes, int32_t *cell_ids, int32_t *cell_range, void *stream) {
  return msplat_bin_views_emit(uv, depth, radius, n, 1, &W, &H, nullptr, nullptr, cell, prepared, count, d_count,
                               workspace, workspace_bytes, cell_ids, cell_range, stream);
}

size_t msplat_bin_views_prepare_workspace_bytes(int64_t rows) {
  if (rows < 0) return 0;
  return prep_layout(rows, 1, 1).total;
}

// Step 1 of the cell binning in two halves (stages bit 1: the counting pass,
// bit 2: the depth order), so that the host can read the count while the GPU
// sorts.
int bin_views_prepare_stages(int stages, const float *uv, const float *depth, const int32_t *radius, int64_t n,
                             int B, const int *W, const int *H, const int64_t *first, const int64_t *count,
                             int cell, uint32_t *rect, void *workspace, size_t workspace_bytes, int64_t *d_count,
                             void *stream) {
  REQUIRE(n >= 0 && B >= 1, "bin_prepare: n = %lld, views = %d", (long long)n, B);
  REQUIRE(valid_cell(cell), "bin_prepare: cell = %d tiles", cell);
  Views vw;
  const std::string why = make_views(B, W, H, cell, first, count, n, vw);
  REQUIRE(why.empty(), "bin_prepare: %s", why.c_str());
  REQUIRE(d_count, "bin_prepare: null count pointer");
  const int64_t m = vw.vbase[B];  // per-view rows
  const PrepLayout L = prep_layout(m, 1, 1);
  if (workspace_bytes < L.total)
    return fail(MSPLAT_ERR_WORKSPACE, "bin_prepare: workspace %zu < %zu bytes", workspace_bytes, L.total);
  REQUIRE(workspace, "bin_prepare: null workspace");
  if (m > 0) {
    REQUIRE(uv && depth && radius && rect, "bin_prepare: null pointer");
    REQUIRE((reinterpret_cast<uintptr_t>(rect) & 7) == 0, "bin_prepare: rect must be 8-byte aligned");
  }
  hipStream_t s = as_stream(stream);
  const sortk::RadixBuffers rb = prep_radix(workspace, L);
  int64_t *nlisted = at<int64_t>(workspace, L.nlisted);
  uint32_t *ka = at<uint32_t>(workspace, L.ka), *va = at<uint32_t>(workspace, L.va);
  if (stages & 1) {
    // one launch zeroes the depth sort's digit totals, tickets and error
    // word, the look-back words, and the counts (which stay 0 when m = 0)
    uint64_t *bstatus = at<uint64_t>(workspace, L.bstatus);
    uint32_t *ctotal = fused_emit(vw) ? at<uint32_t>(workspace, L.ctotal) : nullptr;
    // a one-launch-per-pass depth sort gets its digit totals (in kCtSlices
    // slices) and zeroed look-back words from the counting pass: no histogram
    // launch
    uint32_t *dtotal = sortk::takes_onesweep(m, 4, true) ? at<uint32_t>(workspace, L.dtotal) : nullptr;
    launch_fill_zero_words({{rb.ghist, sortk::ghist_words()},
                            {bstatus, 2 * (sortk::prep_blocks(m) + 1)},
                            {nlisted, 2},
                            {d_count, 4},
                            {ctotal, int64_t(sortk::kCtSlices) * sortk::kFusedCells},
                            {dtotal, int64_t(sortk::kCtSlices) * sortk::kMaxPasses * 256}},
                           s);
    Timed timed_("bin_prepare", s);
    if (m > 0)
      sortk::bin_prepare_kernel<<<unsigned(sortk::prep_blocks(m)), sortk::kThreads, 0, s>>>(
          uv, depth, radius, m, vw, cell, reinterpret_cast<uint2 *>(rect), ka, va, bstatus, nlisted, d_count,
          sortk::radix_error_word(rb), ctotal, dtotal, dtotal ? rb.status : nullptr,
          dtotal ? int64_t(sortk::kMaxPasses) * 256 * sortk::radix_tiles(m) : 0);
  }
  if ((stages & 2) && m > 0) {
    // the listed rows (first *nlisted pairs) by depth; the last pass carries
    // their cell counts and rectangles into depth order
    const sortk::RadixGather gather{reinterpret_cast<const uint2 *>(rect), cell, at<int32_t>(workspace, L.area_s),
                                    at<uint2>(workspace, L.rect_s)};
    uint32_t *kb = at<uint32_t>(workspace, L.kb), *vb = at<uint32_t>(workspace, L.vb);
    Timed timed_("depth_sort", s);
    // a -> b -> a -> b -> a; digit totals from the counting pass
    sortk::radix_sort(ka, va, kb, vb, m, 32, 4, rb, s, nlisted, nullptr, gather, true,
                      sortk::takes_onesweep(m, 4, true) ? at<uint32_t>(workspace, L.dtotal) : nullptr,
                      sortk::kCtSlices);
  }
  return launched("bin_prepare");
}

int msplat_bin_views_prepare(const float *uv, const float *depth, const int32_t *radius, int64_t n, int B,
                             const int *W, const int *H, const int64_t *first, const int64_t *count, int cell,
                             uint32_t *rect, void *workspace, size_t workspace_bytes, int64_t *d_count,
                             void *stream) {
  return bin_views_prepare_stages(3, uv, depth, radius, n, B, W, H, first, count, cell, rect, workspace,
                                  workspace_bytes, d_count, stream);
}

int msplat_bin_views_count(const float *uv, const float *depth, const int32_t *radius, int64_t n, int B,
                           const int *W, const int *H, const int64_t *first, const int64_t *count, int cell,
                           uint32_t *rect, void *workspace, size_t workspace_bytes, int64_t *d_count,
                           void *stream) {
  return bin_views_prepare_stages(1, uv, depth, radius, n, B, W, H, first, count, cell, rect, workspace,
                                  workspace_bytes, d_count, stream);
}

int msplat_bin_views_order(const float *uv, const float *depth, const int32_t *radius, int64_t n, int B,
                           const int *W, const int *H, const int64_t *first, const int64_t *count, int cell,
                           uint32_t *rect, void *workspace, size_t workspace_bytes, int64_t *d_count,
                           void *stream) {
  return bin_views_prepare_stages(2, uv, depth, radius, n, B, W, H, first, count, cell, rect, workspace,
                                  workspace_bytes, d_count, stream);
}

size_t msplat_bin_views_emit_workspace_bytes(int64_t count) {
  return emit_layout(count < 0 ? 0 : count, 1, 1).total;
}

int msplat_bin_views_emit(const float *uv, const float *depth, const int32_t *radius, int64_t n, int B,
                          const int *W, const int *H, const int64_t *first, const int64_t *vcount, int cell,
                          const void *prepared, int64_t count, const int64_t *d_count, void *workspace,
                          size_t workspace_bytes, int32_t *cell_ids, int32_t *cell_range, void *stream) {
  REQUIRE(n >= 0 && B >= 1, "bin_emit: n = %lld, views = %d", (long long)n, B);
  REQUIRE(valid_cell(cell), "bin_emit: cell = %d tiles", cell);
  Views vw;
  const std::string why = make_views(B, W, H, cell, first, vcount, n, vw);
  REQUIRE(why.empty(), "bin_emit: %s", why.c_str());
  REQUIRE(count >= 0, "bin_emit: count = %lld", (long long)count);
  if (count > int64_t(INT32_MAX))
    return fail(MSPLAT_ERR_CAPACITY, "bin_emit: %lld cell entries exceed int32 indexing", (long long)count);
  REQUIRE(cell_range, "bin_emit: null cell_range");
  const EmitLayout L = emit_layout(count, 1, 1);
  if (workspace_bytes < L.total)
    return fail(MSPLAT_ERR_WORKSPACE, "bin_emit: workspace %zu < %zu bytes", workspace_bytes, L.total);
  REQUIRE(count == 0 || (uv && depth && radius && prepared && workspace && cell_ids), "bin_emit: null pointer");
  hipStream_t s = as_stream(stream);
  // the counting pass's and the depth sort's error word (step 1 zeroed it)
  const PrepLayout P = prep_layout(vw.vbase[B], 1, 1);
  const uint32_t *prep_error =
      prepared ? sortk::radix_error_word(prep_radix(const_cast<void *>(prepared), P)) : nullptr;
  if (count > 0 && count < (int64_t(1) << 30) && fused_emit(vw)) {
    // straight into the lists: one launch zeroes the ranges, the look-back
    // words and ticket, and the error word; one emits
    uint32_t *cst = at<uint32_t>(workspace, L.cstatus);
    const int64_t nblk = emit_blocks(count);
    uint32_t *error = sortk::radix_error_word(emit_radix(workspace, L));
    launch_fill_zero_words(
        {{cell_range, 2 * int64_t(vw.cell0[B])}, {cst, nblk * sortk::kFusedCells + 1}, {error, 1}}, s);
    const int64_t rows = vw.vbase[B] < count ? vw.vbase[B] : count;  // every listed row has a cell
    Timed timed_("emit", s);
    sortk::emit_cells_kernel<<<grid_1d(rows, sortk::kEmitThreads), sortk::kEmitThreads, 0, s>>>(
        at<uint32_t>(prepared, P.va), at<int32_t>(prepared, P.area_s), at<uint2>(prepared, P.rect_s), vw.vbase[B],
        at<int64_t>(prepared, P.nlisted), count, vw, cell, vw.cell0[B], at<uint32_t>(prepared, P.ctotal),
        cst + nblk * sortk::kFusedCells, cst, cell_ids, cell_range, error, prep_error);
    return launched("bin_emit");
  }
  // one launch zeroes the ranges, the class sort's digit totals and the
  // emission's look-back words
  launch_fill_zero_words({{cell_range, 2 * int64_t(vw.cell0[B])},
                          {count > 0 ? at<uint32_t>(workspace, L.ghist) : nullptr, sortk::ghist_words()},
                          {count > 0 ? at<uint64_t>(workspace, L.estatus) : nullptr, 2 * (emit_blocks(count) + 1)}},
                         s);
  if (count == 0) {
    // no entries: the range kernel only folds the error words
    sortk::range_vec_kernel<<<1, sortk::kThreads, 0, s>>>(nullptr, 0, nullptr, nullptr, 0, prep_error, nullptr);
    return launched("bin_emit");
  }
  const uint32_t *sk = emit_by_class(uv, depth, radius, vw.vbase[B], vw, cell, prepared, count, d_count, workspace,
                                     L, reinterpret_cast<uint32_t *>(cell_ids), s, true);
  sortk::range_vec_kernel<<<grid_1d((count + 3) / 4, sortk::kThreads), sortk::kThreads, 0, s>>>(
      sk, count, d_count, cell_range, vw.cell0[B], prep_error, sortk::radix_error_word(emit_radix(workspace, L)));
  return launched("bin_emit");
}

int msplat_gaussian_keys(const float *uv, const float *depth, const int32_t *radius, const int32_t *offsets,
                         int64_t n, int W, int H, int64_t k, int64_t *keys, int32_t *idx, void *stream) {
  REQUIRE(n >= 0 && k >= 0, "gaussian_keys: n = %lld, k = %lld", (long long)n, (long long)k);
  REQUIRE(valid_image(W, H), "gaussian_keys: bad image size %dx%d", W, H);
  hipStream_t s = as_stream(stream);
  if (k > 0) {
    REQUIRE(keys && idx, "gaussian_keys: null output");
    launch_fill_zero_words({{keys, 2 * k}, {idx, k}}, s);
  }
  if (n == 0 || k == 0) return launched("gaussian_keys");
  REQUIRE(uv && depth && radius && offsets, "gaussian_keys: null pointer");
  sortk::ref_keys_kernel<<<grid_1d(n, sortk::kThreads), sortk::kThreads, 0, s>>>(
      uv, depth, radius, offsets, n, tiles_x(W), tiles_y(H), k, keys, idx);
  return launched("gaussian_keys");
}

int msplat_tile_range_from_keys(const int64_t *key_sorted, int64_t k, int W, int H, int32_t *tile_range,
                                void *stream) {
  REQUIRE(k >= 0, "tile_range_from_keys: k = %lld", (long long)k);
  REQUIRE(valid_image(W, H), "tile_range_from_keys: bad image size %dx%d", W, H);
  REQUIRE(tile_range, "tile_range_from_keys: null tile_range");
  hipStream_t s = as_stream(stream);
  const int T = tiles_x(W) * tiles_y(H);
  launch_fill_zero(tile_range, 2 * int64_t(T), s);
  if (k > 0) {
    REQUIRE(key_sorted, "tile_range_from_keys: null keys");
    sortk::ref_range_kernel<<<grid_1d(k, sortk::kThreads), sortk::kThreads, 0, s>>>(key_sorted, k, T, tile_range);
  }
  return launched("tile_range_from_keys");
}

// ---------------------------------------------------------------- blending

int msplat_alpha_blending_forward(const float *uv, const float *conic, const float *opacity, const float *feature,
                                  int64_t n, int C, const int32_t *idx_sorted, const int32_t *tile_range, float bg,
                                  int W, int H, float *rendered, float *final_T, int32_t *ncontrib, void *stream) {
  REQUIRE(n >= 0 && n <= int64_t(INT32_MAX) && C >= 0, "alpha_blending_forward: n = %lld, C = %d", (long long)n,
          C);
  REQUIRE(valid_image(W, H), "alpha_blending_forward: bad image size %dx%d", W, H);
  REQUIRE(tile_range && final_T && ncontrib && (C == 0 || rendered), "alpha_blending_forward: null output");
  REQUIRE(n == 0 || (uv && conic && opacity && (C == 0 || feature)), "alpha_blending_forward: null input");
  const BlendIn in = blend_inputs(uv, conic, opacity, feature, C, idx_sorted, tile_range, nullptr, 0, 0, bg,
                                  one_view(W, H, 1, n), 0);
  blend_forward<false>(in, 0, rendered, final_T, ncontrib, nullptr, as_stream(stream));
  return launched("alpha_blending_forward");
}

int msplat_alpha_blending_backward(const float *uv, const float *conic, const float *opacity, const float *feature,
                                   int64_t n, int C, const int32_t *idx_sorted, const int32_t *tile_range, float bg,
                                   int W, int H, const float *final_T, const int32_t *ncontrib,
                                   const float *dL_drendered, float *dL_duv, float *dL_dconic, float *dL_dopacity,
                                   float *dL_dfeature, void *stream) {
  REQUIRE(n >= 0 && n <= int64_t(INT32_MAX) && C >= 0, "alpha_blending_backward: n = %lld, C = %d", (long long)n,
          C);
  REQUIRE(valid_image(W, H), "alpha_blending_backward: bad image size %dx%d", W, H);
  if (n == 0) return MSPLAT_OK;
  REQUIRE(uv && conic && opacity && idx_sorted && tile_range && final_T && ncontrib && dL_duv && dL_dconic &&
              dL_dopacity && (C == 0 || (feature && dL_drendered && dL_dfeature)),
          "alpha_blending_backward: null pointer");
  const BlendIn in = blend_inputs(uv, conic, opacity, feature, C, idx_sorted, tile_range, nullptr, 0, 0, bg,
                                  one_view(W, H, 1, n), 0);
  blend_backward<false>(in, 0, n, n, n, final_T, ncontrib, nullptr, dL_drendered, dL_duv, dL_dconic, dL_dopacity,
                        dL_dfeature, as_stream(stream));
  return launched("alpha_blending_backward");
}

// ---------------------------------------------------------------- rasterizer (cell lists)

int msplat_rasterize_forward(const float *uv, const float *conic, const float *opacity, const float *feature,
                             int64_t n, int C, const uint32_t *rect, const int32_t *cell_ids,
                             const int32_t *cell_range, int cell, float bg, int W, int H, int ppl, int cull,
                             float *rendered, float *final_T, int32_t *ncontrib, int32_t *walk, void *stream) {
  return msplat_rasterize_views_forward(uv, conic, opacity, feature, n, C, 0, 1, &W, &H, nullptr, nullptr, rect,
                                        cell_ids, cell_range, cell, bg, ppl, cull, rendered, final_T, ncontrib, walk,
                                        stream);
}

int msplat_rasterize_backward(const float *uv, const float *conic, const float *opacity, const float *feature,
                              int64_t n, int C, const uint32_t *rect, const int32_t *cell_ids,
                              const int32_t *cell_range, int cell, float bg, int W, int H, int ppl, int cull,
                              const float *final_T, const int32_t *ncontrib, const int32_t *walk,
                              const float *dL_drendered, float *dL_duv, float *dL_dconic, float *dL_dopacity,
                              float *dL_dfeature, void *stream) {
  return msplat_rasterize_views_backward(uv, conic, opacity, feature, n, C, 0, 1, &W, &H, rect, cell_ids,
                                         cell_range, cell, bg, ppl, cull, final_T, ncontrib, walk, dL_drendered,
                                         dL_duv, dL_dconic, dL_dopacity, dL_dfeature, stream);
}

}  // extern "C"

namespace {
// The cell-list forward of every entry point; target / part / loss: the
// fused squared-error loss (optional; part holds l2_slots * l2_passes floats).
int rasterize_forward_impl(const float *uv, const float *conic, const float *opacity, const float *feature,
                           int64_t n, int C, int feat_per_view, int B, const int *W, const int *H,
                           const int64_t *first, const int64_t *count, const uint32_t *rect,
                           const int32_t *cell_ids, const int32_t *cell_range, int cell, float bg, int ppl, int cull,
                           float *rendered, float *final_T, int32_t *ncontrib, int32_t *walk, const float *target,
                           float *part, float *loss, void *stream) {
  REQUIRE(n >= 0 && C >= 0 && B >= 1, "rasterize_forward: n = %lld, C = %d", (long long)n, C);
  REQUIRE(valid_cell(cell), "rasterize_forward: cell = %d tiles", cell);
  Views vw;
  const std::string why = make_views(B, W, H, cell, first, count, n, vw);
  REQUIRE(why.empty(), "rasterize_forward: %s", why.c_str());
  REQUIRE(cell_range && final_T && ncontrib && walk && (C == 0 || rendered), "rasterize_forward: null output");
  REQUIRE(n == 0 || (uv && conic && opacity && rect && cell_ids && (C == 0 || feature)),
          "rasterize_forward: null input");
  REQUIRE((reinterpret_cast<uintptr_t>(rect) & 7) == 0, "rasterize_forward: rect must be 8-byte aligned");
  BlendIn in = blend_inputs(uv, conic, opacity, feature, C, cell_ids, cell_range, rect, cell, cull, bg, vw,
                            feat_per_view != 0);
  if (target) {
    REQUIRE(part && loss && C > 0, "rasterize_forward: the fused loss needs a workspace, a loss output and channels");
    in.l2_target = target;
    in.l2_part = part;
    in.l2_slots = l2_slots(in.ntiles);
  }
  blend_forward<true>(in, ppl, rendered, final_T, ncontrib, walk, as_stream(stream), loss);
  return launched("rasterize_forward");
}
}  // namespace

extern "C" {

int msplat_rasterize_views_forward(const float *uv, const float *conic, const float *opacity, const float *feature,
                                   int64_t n, int C, int feat_per_view, int B, const int *W, const int *H,
                                   const int64_t *first, const int64_t *count, const uint32_t *rect,
                                   const int32_t *cell_ids, const int32_t *cell_range, int cell, float bg, int ppl,
                                   int cull, float *rendered, float *final_T, int32_t *ncontrib, int32_t *walk,
                                   void *stream) {
  return rasterize_forward_impl(uv, conic, opacity, feature, n, C, feat_per_view, B, W, H, first, count, rect,
                                cell_ids, cell_range, cell, bg, ppl, cull, rendered, final_T, ncontrib, walk, nullptr,
                                nullptr, nullptr, stream);
}

size_t msplat_rasterize_l2_workspace_bytes(int W, int H, int C) {
  if (!valid_image(W, H) || C < 0) return 0;
  return l2_workspace(int64_t(tiles_x(W)) * tiles_y(H), C);
}

int msplat_rasterize_forward_l2(const float *uv, const float *conic, const float *opacity, const float *feature,
                                int64_t n, int C, const uint32_t *rect, const int32_t *cell_ids,
                                const int32_t *cell_range, int cell, float bg, int W, int H, int ppl, int cull,
                                const float *target, void *workspace, size_t workspace_bytes, float *rendered,
                                float *final_T, int32_t *ncontrib, int32_t *walk, float *loss, void *stream) {
  REQUIRE(target && loss, "rasterize_forward_l2: null target or loss");
  const size_t need = msplat_rasterize_l2_workspace_bytes(W, H, C);
  if (workspace_bytes < need)
    return fail(MSPLAT_ERR_WORKSPACE, "rasterize_forward_l2: workspace %zu < %zu bytes", workspace_bytes, need);
  REQUIRE(workspace, "rasterize_forward_l2: null workspace");
  return rasterize_forward_impl(uv, conic, opacity, feature, n, C, 0, 1, &W, &H, nullptr, nullptr, rect, cell_ids,
                                cell_range, cell, bg, ppl, cull, rendered, final_T, ncontrib, walk, target,
                                static_cast<float *>(workspace), loss, stream);
}

}  // extern "C"

// ---- the fused-loss forward without a workspace argument -------------------
// For callers that cannot allocate (graph capture, serving loops that forbid
// hipMalloc): the loss partials live in this file-scope device array, one
// copy per GPU, held until the process ends.  Host code never passes the
// array's name to a kernel (on the host it is a stand-in, not a device
// address): the entry point's first call on a device looks up that device's
// copy with hipGetSymbolAddress and keeps the address.  Every call on a GPU
// shares its copy; the forward zeroes what must start at zero on each call,
// and calls that overlap on one GPU race.
namespace {
constexpr size_t kNoAllocScratchBytes = size_t(4) << 20;
constexpr int kNoAllocDevices = 64;
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

std::mutex g_scratch_mu;
void *g_scratch_ptr[kNoAllocDevices] = {};

// This device's copy of `scratch` (looked up once per device), or null with
// the error recorded.
void *scratch_ptr(int *status) {
  int dev = -1;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess || dev < 0 || dev >= kNoAllocDevices) {
    *status = fail(MSPLAT_ERR_LAUNCH, "no-allocation scratch: hipGetDevice: %s (device %d)", hipGetErrorString(e),
                   dev);
    return nullptr;
  }
  std::lock_guard<std::mutex> g(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void *p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
      *status = fail(MSPLAT_ERR_LAUNCH, "no-allocation scratch: hipGetSymbolAddress: %s", hipGetErrorString(e));
      return nullptr;
    }
    g_scratch_ptr[dev] = p;
  }
  return g_scratch_ptr[dev];
}
}  // namespace

extern "C" {

size_t msplat_noalloc_scratch_bytes(void) { return kNoAllocScratchBytes; }

int msplat_rasterize_forward_l2_noalloc(const float *uv, const float *conic, const float *opacity,
                                        const float *feature, int64_t n, int C, const uint32_t *rect,
                                        const int32_t *cell_ids, const int32_t *cell_range, int cell, float bg, int W,
                                        int H, int ppl, int cull, const float *target, float *rendered,
                                        float *final_T, int32_t *ncontrib, int32_t *walk, float *loss,
                                        void *stream) {
  REQUIRE(target && loss, "rasterize_forward_l2_noalloc: null target or loss");
  REQUIRE(valid_image(W, H) && C >= 0, "rasterize_forward_l2_noalloc: image %dx%d, C = %d", W, H, C);
  const size_t need = msplat_rasterize_l2_workspace_bytes(W, H, C);
  if (need > kNoAllocScratchBytes)
    return fail(MSPLAT_ERR_CAPACITY,
                "rasterize_forward_l2_noalloc: %dx%d with %d channels needs %zu bytes of scratch, more than the "
                "%zu this entry point holds; use msplat_rasterize_forward_l2 with a workspace",
                W, H, C, need, kNoAllocScratchBytes);
  int status = MSPLAT_OK;
  void *ws = scratch_ptr(&status);
  if (!ws) return status;
  return rasterize_forward_impl(uv, conic, opacity, feature, n, C, 0, 1, &W, &H, nullptr, nullptr, rect, cell_ids,
                                cell_range, cell, bg, ppl, cull, rendered, final_T, ncontrib, walk, target,
                                static_cast<float *>(ws), loss, stream);
}

}  // extern "C"

namespace {
// The cell-list adjoint of every entry point; rendered / target / dL_dloss:
// the fused loss's gradient (optional; dL_drendered may then be NULL).
int rasterize_backward_impl(const float *uv, const float *conic, const float *opacity, const float *feature,
                            int64_t n, int C, int feat_per_view, int B, const int *W, const int *H,
                            const uint32_t *rect, const int32_t *cell_ids, const int32_t *cell_range, int cell,
                            float bg, int ppl, int cull, const float *final_T, const int32_t *ncontrib,
                            const int32_t *walk, const float *dL_drendered, const float *rendered,
                            const float *target, const float *dL_dloss, float *dL_duv, float *dL_dconic,
                            float *dL_dopacity, float *dL_dfeature, void *stream) {
  REQUIRE(n >= 0 && C >= 0 && B >= 1, "rasterize_backward: n = %lld, C = %d", (long long)n, C);
  REQUIRE(valid_cell(cell), "rasterize_backward: cell = %d tiles", cell);
  Views vw;
  const std::string why = make_views(B, W, H, cell, nullptr, nullptr, n, vw);
  REQUIRE(why.empty(), "rasterize_backward: %s", why.c_str());
  if (n == 0) return MSPLAT_OK;
  const bool l2 = dL_dloss != nullptr;
  REQUIRE(!l2 || (rendered && target), "rasterize_backward: the fused loss needs the rendered and target images");
  REQUIRE(uv && conic && opacity && rect && cell_ids && cell_range && final_T && ncontrib && walk && dL_duv &&
              dL_dconic && dL_dopacity && (C == 0 || (feature && (dL_drendered || l2) && dL_dfeature)),
          "rasterize_backward: null pointer");
  REQUIRE((reinterpret_cast<uintptr_t>(rect) & 7) == 0, "rasterize_backward: rect must be 8-byte aligned");
  BlendIn in = blend_inputs(uv, conic, opacity, feature, C, cell_ids, cell_range, rect, cell, cull, bg, vw,
                            feat_per_view != 0);
  if (l2) {
    in.l2_rendered = rendered;
    in.l2_target = target;
    in.l2_dL = dL_dloss;
  }
  blend_backward<true>(in, ppl, vw.vbase[B], n, feat_per_view ? vw.vbase[B] : n, final_T, ncontrib, walk,
                       dL_drendered, dL_duv, dL_dconic, dL_dopacity, dL_dfeature, as_stream(stream));
  return launched("rasterize_backward");
}
}  // namespace

extern "C" {

int msplat_rasterize_views_backward(const float *uv, const float *conic, const float *opacity, const float *feature,
                                    int64_t n, int C, int feat_per_view, int B, const int *W, const int *H,
                                    const uint32_t *rect, const int32_t *cell_ids, const int32_t *cell_range,
                                    int cell, float bg, int ppl, int cull, const float *final_T,
                                    const int32_t *ncontrib, const int32_t *walk, const float *dL_drendered,
                                    float *dL_duv, float *dL_dconic, float *dL_dopacity, float *dL_dfeature,
                                    void *stream) {
  return rasterize_backward_impl(uv, conic, opacity, feature, n, C, feat_per_view, B, W, H, rect, cell_ids,
                                 cell_range, cell, bg, ppl, cull, final_T, ncontrib, walk, dL_drendered, nullptr,
                                 nullptr, nullptr, dL_duv, dL_dconic, dL_dopacity, dL_dfeature, stream);
}

int msplat_rasterize_backward_l2(const float *uv, const float *conic, const float *opacity, const float *feature,
                                 int64_t n, int C, const uint32_t *rect, const int32_t *cell_ids,
                                 const int32_t *cell_range, int cell, float bg, int W, int H, int ppl, int cull,
                                 const float *final_T, const int32_t *ncontrib, const int32_t *walk,
                                 const float *dL_drendered, const float *rendered, const float *target,
                                 const float *dL_dloss, float *dL_duv, float *dL_dconic, float *dL_dopacity,
                                 float *dL_dfeature, void *stream) {
  REQUIRE(dL_dloss, "rasterize_backward_l2: null dL_dloss");
  return rasterize_backward_impl(uv, conic, opacity, feature, n, C, 0, 1, &W, &H, rect, cell_ids, cell_range, cell,
                                 bg, ppl, cull, final_T, ncontrib, walk, dL_drendered, rendered, target, dL_dloss,
                                 dL_duv, dL_dconic, dL_dopacity, dL_dfeature, stream);
}

// ---------------------------------------------------------------- many views: preprocessing

int msplat_preprocess_views_forward(const float *xyz, const float *scale, const float *rotate, int64_t n, int B,
                                    const float *intr, const float *extr, const int *W, const int *H,
                                    const int64_t *first, const int64_t *count, float nearest, float extent,
                                    float *uv, float *depth, float *conic, int32_t *radius, int32_t *tiles,
                                    void *stream) {
  REQUIRE(n >= 0 && B >= 1, "preprocess_forward: n = %lld, views = %d", (long long)n, B);
  Views vw;
  const std::string why = make_views(B, W, H, 1, first, count, n, vw);
  REQUIRE(why.empty(), "preprocess_forward: %s", why.c_str());
  int64_t widest = 0;
  for (int b = 0; b < B; ++b) widest = imax64(widest, vw.vbase[b + 1] - vw.vbase[b]);
  if (widest == 0) return MSPLAT_OK;
  REQUIRE(xyz && scale && rotate && intr && extr && uv && depth && conic && radius && tiles,
          "preprocess_forward: null pointer");
  Timed timed_("preprocess_fwd", as_stream(stream));
  preprocess_fwd_views_kernel<<<dim3(grid_1d(widest, kGeomThreads), unsigned(B)), kGeomThreads, 0,
                                as_stream(stream)>>>(xyz, scale, rotate, intr, extr, vw, nearest, extent, uv, depth,
                                                     conic, radius, tiles);
  return launched("preprocess_forward");
}

int msplat_preprocess_views_backward(const float *xyz, const float *scale, const float *rotate, int64_t n, int B,
                                     const float *intr, const float *extr, const float *depth,
                                     const int32_t *radius, const float *dL_duv, const float *dL_ddepth,
                                     const float *dL_dconic, float *dL_dxyz, float *dL_dscale, float *dL_drotate,
                                     float *dL_dintr, float *dL_dextr, void *stream) {
  REQUIRE(n >= 0 && B >= 1 && B <= kMaxViews && n * B <= int64_t(INT32_MAX),
          "preprocess_backward: n = %lld, views = %d", (long long)n, B);
  hipStream_t s = as_stream(stream);
  const bool cam = dL_dintr || dL_dextr;
  launch_fill_zero_words({{dL_dintr, 4 * int64_t(B)}, {dL_dextr, 12 * int64_t(B)}}, s);
  if (n == 0) return launched("preprocess_backward");
  REQUIRE(xyz && scale && rotate && intr && extr && depth && radius && dL_duv && dL_dconic && dL_dxyz &&
              dL_dscale && dL_drotate,
          "preprocess_backward: null pointer");
  Timed timed_("preprocess_bwd", s);
  if (cam)
    preprocess_bwd_views_kernel<true><<<cam_grid(n), kGeomThreads, 0, s>>>(
        xyz, scale, rotate, n, intr, extr, B, depth, radius, dL_duv, dL_ddepth, dL_dconic, dL_dxyz, dL_dscale,
        dL_drotate, dL_dintr, dL_dextr);
  else
    preprocess_bwd_views_kernel<false><<<grid_1d(n, kGeomThreads), kGeomThreads, 0, s>>>(
        xyz, scale, rotate, n, intr, extr, B, depth, radius, dL_duv, dL_ddepth, dL_dconic, dL_dxyz, dL_dscale,
        dL_drotate, nullptr, nullptr);
  return launched("preprocess_backward");
}

}  // extern "C"
