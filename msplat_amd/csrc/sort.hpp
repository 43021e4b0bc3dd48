// Tile binning and sorting of (tile, Gaussian) intersections.
//
// The reference (msplat/sort_gaussian.py:42-52, msplat/src/sort_gaussian.cu)
// writes one 64-bit key (tile << 32 | int32 depth bits, sign-extended) per
// intersection in Gaussian order and stably sorts all K keys as signed int64
// (torch.sort, 8 radix passes over 64-bit keys and indices).  Sorting by that
// key is the same as sorting by (class, depth bits as unsigned, Gaussian id)
// where class = 0 for keys whose depth sign bit is set (the sign extension
// turns their tile into -1, so they sort first) and tile + 1 otherwise.
//
// We therefore
//   1. stably sort the n Gaussians by their 32-bit depth pattern (4 passes
//      over n, not K),
//   2. scan their tile counts in that order and emit the (class, id) pairs,
//      so the emitted sequence is already ordered by (depth bits, id),
//   3. stably radix-sort the K pairs by class alone: ceil(log2(T+1)) bits,
//      two 7-bit passes at 1080p,
// which yields the reference's order exactly, with 8-byte pairs instead of
// 16-byte (key, index) pairs and 2 passes over K instead of 8.
#pragma once

#include "common.hpp"

namespace msplat {
namespace sortk {

constexpr int kThreads = 256;
constexpr int kItems = 16;
constexpr int kRadixTile = kThreads * kItems;  // keys per radix workgroup
// A scatter pass over few tiles (a few hundred thousand keys: less than
// the chip holds at once) is a chain of latencies, and 16 short waves of 4
// keys per lane rank a tile sooner than 4 waves of 16.
constexpr int kScatterWide = 1024;
// Onesweep (one launch per pass, tiles chained by a decoupled look-back) up
// to this many tiles of capacity, multi-launch passes (histogram, scan,
// scatter) beyond.  When every tile starts at once, a tile walks back over
// the aggregates of all tiles started with it before it finds an inclusive
// prefix, so look-back reads grow with the square of the tiles in flight:
// from about 1.3 M keys on, the passes without look-back are faster.  A
// device length is usually well below its capacity, so a sort with one
// takes onesweep up to twice this capacity.
constexpr int64_t kOnesweepTiles = 320;
constexpr int kScanItems = 8;
constexpr int kScanTile = kThreads * kScanItems;  // elements per scan workgroup

MSPLAT_HD int64_t radix_tiles(int64_t len) { return (len + kRadixTile - 1) / kRadixTile; }
MSPLAT_HD int64_t scan_tiles(int64_t len) { return (len + kScanTile - 1) / kScanTile; }

MSPLAT_DEV uint32_t fbits(float v) { return __float_as_uint(v); }

// Length of a sequence whose buffers hold `cap` elements: *dlen (clamped to
// cap) when the length lives on the device, else cap.  Grids are sized for
// cap, so a launch needs no host read of the length.
MSPLAT_DEV int64_t dev_len(const int64_t *dlen, int64_t cap) {
  if (!dlen) return cap;
  const int64_t v = *dlen;
  return v < 0 ? 0 : (v < cap ? v : cap);
}

// ---- device-wide exclusive scan (reduce, scan of partials, apply) ------------

// Value at position i of the scanned sequence: in[perm[i]] (in[i] without
// perm); a permuted index past the cap elements of `in` (only an order left
// undefined by an expired wait holds one) reads 0.
template <typename In>
MSPLAT_DEV int64_t scan_load(const In *in, const uint32_t *perm, int64_t i, int64_t cap) {
  if (!perm) return static_cast<int64_t>(in[i]);
  const uint32_t p = perm[i];
  return int64_t(p) < cap ? static_cast<int64_t>(in[p]) : 0;
}

MSPLAT_DEV int64_t wave_incl_scan(int64_t v) {
  const int lane = __lane_id();
#pragma unroll
  for (int off = 1; off < kWave; off <<= 1) {
    const int64_t o = __shfl_up(v, off, kWave);
    if (lane >= off) v += o;
  }
  return v;
}

// Exclusive scan of one value per thread over a workgroup of NT threads;
// returns the workgroup total through *total.
template <int NT = kThreads>
MSPLAT_DEV int64_t block_excl_scan(int64_t v, int64_t *total) {
  __shared__ int64_t wsum[NT / kWave];
  const int lane = threadIdx.x % kWave, wave = threadIdx.x / kWave;
  const int64_t incl = wave_incl_scan(v);
  if (lane == kWave - 1) wsum[wave] = incl;
  __syncthreads();
  int64_t before = 0, all = 0;
#pragma unroll
  for (int w = 0; w < NT / kWave; ++w) {
    before += (w < wave) ? wsum[w] : 0;
    all += wsum[w];
  }
  __syncthreads();
  *total = all;
  return before + incl - v;
}

// Elements at or past *dlen (when given) count as 0 and are not loaded.
template <typename In>
__global__ __launch_bounds__(kThreads) void scan_reduce_kernel(const In *__restrict__ in,
                                                               const uint32_t *__restrict__ perm, int64_t cap,
                                                               const int64_t *__restrict__ dlen,
                                                               int64_t *__restrict__ sums) {
  const int64_t len = dev_len(dlen, cap);
  if (int64_t(blockIdx.x) * kScanTile >= len) {  // uniform: nothing to add
    if (threadIdx.x == 0) sums[blockIdx.x] = 0;
    return;
  }
  const int64_t base = int64_t(blockIdx.x) * kScanTile + int64_t(threadIdx.x) * kScanItems;
  int64_t s = 0;
#pragma unroll
  for (int j = 0; j < kScanItems; ++j)
    if (base + j < len) s += scan_load(in, perm, base + j, cap);
  int64_t total;
  block_excl_scan(s, &total);
  if (threadIdx.x == 0) sums[blockIdx.x] = total;
}

// One workgroup scans the partial sums in place and stores the grand total.
__global__ __launch_bounds__(kThreads) void scan_partials_kernel(int64_t *__restrict__ sums, int64_t nb,
                                                                 int64_t *__restrict__ total_out) {
  int64_t carry = 0;
  for (int64_t base = 0; base < nb; base += kThreads) {
    const int64_t i = base + threadIdx.x;
    const int64_t v = i < nb ? sums[i] : 0;
    int64_t total;
    const int64_t ex = block_excl_scan(v, &total);
    if (i < nb) sums[i] = carry + ex;
    carry += total;
  }
  if (threadIdx.x == 0 && total_out) *total_out = carry;
}

// Writes the first *dlen (or cap) outputs.
template <typename In, typename Out>
__global__ __launch_bounds__(kThreads) void scan_apply_kernel(const In *__restrict__ in,
                                                              const uint32_t *__restrict__ perm, int64_t cap,
                                                              const int64_t *__restrict__ dlen,
                                                              const int64_t *__restrict__ sums,
                                                              Out *__restrict__ out) {
  const int64_t len = dev_len(dlen, cap);
  if (int64_t(blockIdx.x) * kScanTile >= len) return;  // uniform
  const int64_t base = int64_t(blockIdx.x) * kScanTile + int64_t(threadIdx.x) * kScanItems;
  int64_t vals[kScanItems];
  int64_t s = 0;
#pragma unroll
  for (int j = 0; j < kScanItems; ++j) {
    vals[j] = (base + j < len) ? scan_load(in, perm, base + j, cap) : 0;
    s += vals[j];
  }
  int64_t total;
  int64_t run = block_excl_scan(s, &total) + sums[blockIdx.x];
#pragma unroll
  for (int j = 0; j < kScanItems; ++j) {
    if (base + j < len) out[base + j] = static_cast<Out>(run);
    run += vals[j];
  }
}

// sums must hold scan_tiles(len) entries.  With dlen, only the first *dlen
// (<= len) elements are scanned (the rest count as 0) and written.
template <typename In, typename Out>
inline void exclusive_scan(const In *in, const uint32_t *perm, int64_t len, int64_t *sums, Out *out,
                           int64_t *total, hipStream_t s, const int64_t *dlen = nullptr) {
  const int64_t nb = scan_tiles(len);
  if (nb == 0) return;
  scan_reduce_kernel<In><<<dim3(unsigned(nb)), dim3(kThreads), 0, s>>>(in, perm, len, dlen, sums);
  scan_partials_kernel<<<dim3(1), dim3(kThreads), 0, s>>>(sums, nb, total);
  scan_apply_kernel<In, Out><<<dim3(unsigned(nb)), dim3(kThreads), 0, s>>>(in, perm, len, dlen, sums, out);
}

// ---- stable LSD radix sort of (uint32 key, uint32 value) pairs --------------

// Per-workgroup digit histogram, stored digit-major (hist[d * tiles + t]) so
// that an exclusive scan of the flat array gives every (digit, workgroup) its
// first output slot in stable order.
__global__ __launch_bounds__(kThreads) void radix_hist_kernel(const uint32_t *__restrict__ keys, int64_t cap,
                                                              const int64_t *__restrict__ dlen, int shift,
                                                              int bits, uint32_t *__restrict__ hist,
                                                              int64_t ntiles) {
  __shared__ uint32_t h[256];
  const int64_t len = dev_len(dlen, cap);  // tiles past it store zero counts
  h[threadIdx.x] = 0;
  __syncthreads();
  const uint32_t mask = (1u << bits) - 1u;
  const int64_t base = int64_t(blockIdx.x) * kRadixTile;
#pragma unroll 4
  for (int it = 0; it < kItems; ++it) {
    const int64_t i = base + it * kThreads + threadIdx.x;
    if (i < len) atomicAdd(&h[(keys[i] >> shift) & mask], 1u);
  }
  __syncthreads();
  if (threadIdx.x < (1 << bits)) hist[int64_t(threadIdx.x) * ntiles + blockIdx.x] = h[threadIdx.x];
}

// Stable scatter of one 4096-key tile in three steps:
//  1. each wave ranks its own contiguous 1024 keys (16 rounds of 64) among
//     equal digits with ballot matching and per-wave running counters in LDS
//     (wave-synchronous: no workgroup barrier inside the loop);
//  2. one exchange turns per-wave counts into the tile's digit-sorted order
//     and the keys are reordered through LDS;
//  3. the tile is written out digit run by digit run, so consecutive lanes
//     store consecutive addresses (coalesced), at offsets from the global scan.
//
// Where each (digit, tile) run starts comes either from a scanned per-tile
// histogram (kOnesweep = false: histogram, scan and scatter are separate
// launches) or, in one launch per pass (kOnesweep = true), from the digit
// totals of the whole pass plus a decoupled look-back over the earlier tiles:
// workgroups take tiles in dispatch order from a ticket counter, publish
// their per-digit counts, then add up the earlier tiles' published counts
// until one of them carries its inclusive prefix.  A tile publishes its counts
// before it looks back, so every wait ends; each wait is also bounded and sets
// *error if it ever runs out.
// Optional records carried into the sorted order by a pass (cell binning):
// orect[p] = rect[value at p], the packed tile rectangle of that row, and
// oarea[p] = the number of cells of S x S tiles it touches (null: none).
struct RadixGather {
  const uint2 *rect = nullptr;
  int S = 1;
  int32_t *oarea = nullptr;
  uint2 *orect = nullptr;
};

// Cells of S x S tiles touched by a non-empty packed tile rectangle.
MSPLAT_HD int rect_cells(uint2 r, int S) {
  const int x0 = int(r.x & 0xffffu), y0 = int(r.x >> 16), x1 = int(r.y & 0xffffu), y1 = int(r.y >> 16);
  return ((x1 + S - 1) / S - x0 / S) * ((y1 + S - 1) / S - y0 / S);
}

constexpr int kMaxPasses = 4;
constexpr int kMaxTotalSlices = 8;  // slices of a onesweep sort's digit totals
constexpr uint32_t kFlagAgg = 1u << 30, kFlagIncl = 2u << 30, kValueMask = (1u << 30) - 1u;
constexpr int64_t kLookbackSpins = int64_t(1) << 16;
constexpr int kLookbackWidth = 2;

// ---- expired waits: see wait_expired in common.hpp ------------------------------
// One thread: out[0] = wait_expired, then clear it when asked.
__global__ void wait_status_kernel(int64_t *__restrict__ out, int clear) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    if (out) out[0] = int64_t(atomicOr(&wait_expired, 0u));
    if (clear) atomicAnd(&wait_expired, 0u);
  }
}

__global__ void set_force_expiry_kernel(uint32_t on) {
  if (threadIdx.x == 0 && blockIdx.x == 0) atomicExch(&force_expiry, on);
}

// Decoupled look-back for one digit (or class) d of tile `tile`: publishes
// the tile's count `tot` in status[tile * 256 + d] and returns the digit's
// total over all earlier tiles, reading kLookbackWidth earlier tiles' words
// per round (their loads' latencies overlap) until one carries an inclusive
// prefix.  A tile publishes before it looks back and looks back only at
// earlier tiles, which are running or done, so every wait ends; waits are
// also bounded, and an expired one sets *error (the prefix is then short).
MSPLAT_DEV uint32_t digit_lookback(uint32_t *status, int64_t tile, int d, uint32_t tot, uint32_t *error) {
  uint32_t *st = status + tile * 256;
  uint32_t prefix = 0;
  if (tile == 0) {
    __hip_atomic_store(st + d, kFlagIncl | tot, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    return 0u;
  }
  __hip_atomic_store(st + d, kFlagAgg | tot, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  bool found = false;
  if (expiry_forced()) {  // tests: give up without looking
    atomicOr(error, 1u);
    found = true;
  }
  for (int64_t t = tile - 1; t >= 0 && !found; t -= kLookbackWidth) {
    uint32_t w[kLookbackWidth];
#pragma unroll
    for (int q = 0; q < kLookbackWidth; ++q)
      w[q] = t - q >= 0 ? __hip_atomic_load(status + (t - q) * 256 + d, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)
                        : kFlagIncl;
#pragma unroll
    for (int q = 0; q < kLookbackWidth; ++q) {
      if (found) break;
      if (t - q < 0) {  // before the first tile: nothing more to add
        found = true;
        break;
      }
      int64_t spins = 0;
      while ((w[q] & ~kValueMask) == 0u && ++spins < kLookbackSpins)
        w[q] = __hip_atomic_load(status + (t - q) * 256 + d, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      if ((w[q] & ~kValueMask) == 0u) {  // never published: give up, flag it
        atomicOr(error, 1u);
        found = true;
        break;
      }
      prefix += w[q] & kValueMask;
      if ((w[q] & ~kValueMask) == kFlagIncl) found = true;
    }
  }
  __hip_atomic_store(st + d, kFlagIncl | (prefix + tot), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  return prefix;
}

template <bool kOnesweep, int kScatterThreads>
__global__ __launch_bounds__(kScatterThreads) void radix_scatter_tile_kernel(
    const uint32_t *__restrict__ kin, const uint32_t *__restrict__ vin, int64_t cap,
    const int64_t *__restrict__ dlen, int shift, int bits, const uint32_t *__restrict__ offsets, int nsl,
    int64_t ntiles, uint32_t *__restrict__ kout, uint32_t *__restrict__ vout, uint32_t *__restrict__ ticket,
    uint32_t *__restrict__ status, uint32_t *__restrict__ error, RadixGather gather) {
  const int64_t len = dev_len(dlen, cap);
  int64_t tile = blockIdx.x;
  // Workgroups past the data leave before taking a ticket: the grid is sized
  // for the capacity, and every ticket is an atomic on one shared word.
  // Exactly ceil(len / tile) workgroups take tickets, so every ticket names a
  // tile with data, and no later tile waits on a workgroup that left.
  if (tile * kRadixTile >= len) return;
  constexpr int kScatterItems = kRadixTile / kScatterThreads;  // keys per lane
  constexpr int kWaves = kScatterThreads / kWave;
  constexpr int kPerWave = kRadixTile / kWaves;  // contiguous keys per wave
  __shared__ uint32_t s_key[kRadixTile];
  __shared__ uint32_t s_val[kRadixTile];
  __shared__ uint32_t s_run[kWaves][256];  // per-wave running digit counts, then wave offsets
  __shared__ uint32_t s_lstart[256];       // tile-local first slot of each digit
  __shared__ uint32_t s_gstart[256];       // global first slot of each digit for this tile
  __shared__ uint32_t s_ticket;
  const int tid = threadIdx.x, lane = tid % kWave, wave = tid / kWave;
  const uint32_t mask = (1u << bits) - 1u;
  const int ndig = 1 << bits;
  // onesweep: the pass's digit totals (the sum of nsl slices kMaxPasses x 256
  // words apart) do not depend on the tile; their loads stay in flight until
  // the look-back (the barriers before it order LDS only)
  uint32_t part[kMaxTotalSlices];  // added up where the look-back needs them
#pragma unroll
  for (int k = 0; k < kMaxTotalSlices; ++k)
    part[k] = (kOnesweep && tid < ndig && k < nsl) ? offsets[k * (kMaxPasses * 256) + tid] : 0u;
  if (kOnesweep && tid == 0) s_ticket = atomicAdd(ticket, 1u);
  for (int j = tid; j < kWaves * 256; j += kScatterThreads) s_run[j / 256][j % 256] = 0u;
  if (!kOnesweep && tid < ndig) s_gstart[tid] = offsets[int64_t(tid) * ntiles + tile];
  lds_barrier();
  if constexpr (kOnesweep) tile = s_ticket;

  const int64_t base = tile * kRadixTile + int64_t(wave) * kPerWave;
  uint32_t key[kScatterItems], val[kScatterItems], rnk[kScatterItems];
  // All loads first: the ranking rounds below are separated by LDS fences,
  // which would otherwise hold each round's load until the previous round's
  // LDS traffic completes (one memory latency per round, in a row).
#pragma unroll
  for (int it = 0; it < kScatterItems; ++it) {
    const int64_t i = base + it * kWave + lane;
    key[it] = i < len ? kin[i] : 0u;
    val[it] = i < len ? vin[i] : 0u;
  }
  // Rank in two sweeps.  First the lanes holding an equal digit in each
  // round of 64 keys, by bit-wise ballots (the rounds are independent, so
  // their ballot chains overlap).  Then, round by round, the lowest lane of
  // each digit group adds the group's size to the wave's running count of
  // that digit with an LDS atomic whose return value -- the count of that
  // digit in earlier rounds -- it hands to its group.  A wave's LDS
  // operations execute in program order and each digit has one leader per
  // round, so ranks follow key order (stable), with no fence or wait between
  // rounds.
  uint32_t dig[kScatterItems];
  uint64_t grp[kScatterItems];
#pragma unroll
  for (int it = 0; it < kScatterItems; ++it) {
    const bool valid = base + it * kWave + lane < len;
    const uint32_t d = (key[it] >> shift) & mask;
    uint64_t peers = __ballot(valid);
#pragma unroll
    for (int b = 0; b < 8; ++b) {
      if (b < bits) {  // uniform
        const bool bit = (d >> b) & 1u;
        const uint64_t bb = __ballot(bit);
        peers &= bit ? bb : ~bb;
      }
    }
    dig[it] = d;
    grp[it] = valid ? peers : 0ull;
  }
#pragma unroll
  for (int it = 0; it < kScatterItems; ++it) {
    const uint64_t peers = grp[it];
    const uint32_t below =
        __builtin_amdgcn_mbcnt_hi(uint32_t(peers >> 32), __builtin_amdgcn_mbcnt_lo(uint32_t(peers), 0u));
    uint32_t before = 0u;
    if (peers != 0ull && below == 0u) before = atomicAdd(&s_run[wave][dig[it]], uint32_t(__popcll(peers)));
    const int leader = peers != 0ull ? __ffsll((unsigned long long)peers) - 1 : lane;
    before = __shfl(before, leader, kWave);
    rnk[it] = peers != 0ull ? before + below : 0xffffffffu;
  }
  __syncthreads();
  // digit totals -> tile-local starts (exclusive scan over digits) and wave offsets
  uint32_t tot = 0;
  if (tid < ndig) {
#pragma unroll
    for (int w = 0; w < kWaves; ++w) tot += s_run[w][tid];
  }
  int64_t dummy;
  if constexpr (kOnesweep) {
    // this tile's first slot per digit = digits below it in the whole pass
    // (offsets holds the pass's digit totals) + this digit in earlier tiles
    uint32_t gtot = 0u;
#pragma unroll
    for (int k = 0; k < kMaxTotalSlices; ++k) gtot += part[k];
    const uint32_t doff = uint32_t(block_excl_scan<kScatterThreads>(int64_t(gtot), &dummy));
    if (tid < ndig) s_gstart[tid] = doff + digit_lookback(status, tile, tid, tot, error);
  }
  const uint32_t lstart = uint32_t(block_excl_scan<kScatterThreads>(tid < ndig ? int64_t(tot) : 0, &dummy));
  if (tid < ndig) {
    s_lstart[tid] = lstart;
    uint32_t off = lstart;
#pragma unroll
    for (int w = 0; w < kWaves; ++w) {
      const uint32_t c = s_run[w][tid];
      s_run[w][tid] = off;
      off += c;
    }
  }
  __syncthreads();
#pragma unroll
  for (int it = 0; it < kScatterItems; ++it) {
    if (rnk[it] != 0xffffffffu) {
      const uint32_t d = (key[it] >> shift) & mask;
      const uint32_t slot = s_run[wave][d] + rnk[it];
      s_key[slot] = key[it];
      s_val[slot] = val[it];
    }
  }
  __syncthreads();
  const int64_t tbase = tile * kRadixTile;
  const int cnt = len - tbase < kRadixTile ? int(len - tbase) : kRadixTile;
  for (int e = tid; e < cnt; e += kScatterThreads) {
    const uint32_t k = s_key[e];
    const uint32_t d = (k >> shift) & mask;
    const uint32_t pos = s_gstart[d] + (uint32_t(e) - s_lstart[d]);
    const uint32_t v = s_val[e];
    kout[pos] = k;
    vout[pos] = v;
    if (gather.oarea) {  // uniform: per-value records, moved into the sorted order
      const uint2 r = v < cap ? gather.rect[v] : uint2{0u, 0u};  // v >= cap: an order left undefined
      gather.orect[pos] = r;
      gather.oarea[pos] = rect_cells(r, gather.S);
    }
  }
}

// Workspace of the radix sorts.  hist / sums: the multi-launch pass;
// ghist (kMaxPasses x 256 digit totals, then kMaxPasses tickets and an error
// word) and status (kMaxPasses x tiles x 256 look-back words): onesweep.
struct RadixBuffers {
  uint32_t *hist;    // 256 * radix_tiles(len)
  int64_t *sums;     // scan_tiles(256 * radix_tiles(len))
  uint32_t *ghist;   // kMaxPasses * 256 + kMaxPasses + 1
  uint32_t *status;  // kMaxPasses * 256 * radix_tiles(len)
};

MSPLAT_HD int64_t ghist_words() { return int64_t(kMaxPasses) * 256 + kMaxPasses + 1; }
// The error word at the end of ghist: set when a bounded wait expires.
inline uint32_t *radix_error_word(const RadixBuffers &rb) { return rb.ghist + kMaxPasses * 256 + kMaxPasses; }
MSPLAT_HD int64_t status_words(int64_t len) { return int64_t(kMaxPasses) * 256 * radix_tiles(len); }

// Digit totals of every pass in one read of the keys (pass p covers bits
// [p db, min(p db + db, bits))), and the look-back words cleared.  ghist must
// be zero on entry.
__global__ __launch_bounds__(kThreads) void onesweep_hist_kernel(const uint32_t *__restrict__ keys, int64_t cap,
                                                                 const int64_t *__restrict__ dlen, int npass,
                                                                 int db, int bits, uint32_t *__restrict__ ghist,
                                                                 uint32_t *__restrict__ status, int64_t nstatus) {
  __shared__ uint32_t h[kMaxPasses][256];
  const int tid = threadIdx.x;
#pragma unroll
  for (int p = 0; p < kMaxPasses; ++p) h[p][tid] = 0u;
  __syncthreads();
  const int64_t len = dev_len(dlen, cap);
  const int64_t stride = int64_t(gridDim.x) * kThreads;
  // chunks of kRadixTile keys, all of a chunk's loads in flight together
  for (int64_t c0 = int64_t(blockIdx.x) * kRadixTile; c0 < len; c0 += int64_t(gridDim.x) * kRadixTile) {
    uint32_t k[kItems];
#pragma unroll
    for (int it = 0; it < kItems; ++it) {
      const int64_t i = c0 + it * kThreads + tid;
      k[it] = i < len ? keys[i] : 0u;
    }
#pragma unroll
    for (int it = 0; it < kItems; ++it) {
      if (c0 + it * kThreads + tid >= len) break;
      for (int p = 0; p < npass; ++p) {
        const int sh = p * db;
        const int pb = bits - sh < db ? bits - sh : db;
        atomicAdd(&h[p][(k[it] >> sh) & ((1u << pb) - 1u)], 1u);
      }
    }
  }
  __syncthreads();
  for (int p = 0; p < npass; ++p)
    if (h[p][tid]) atomicAdd(&ghist[p * 256 + tid], h[p][tid]);
  for (int64_t j = int64_t(blockIdx.x) * kThreads + tid; j < nstatus; j += stride) status[j] = 0u;
}

// One stable pass over bits [shift, shift + bits) of the first `cap` pairs,
// or of the first *dlen (<= cap) when dlen is given: histogram, scan and
// scatter launches.
inline void radix_pass(const uint32_t *kin, const uint32_t *vin, uint32_t *kout, uint32_t *vout, int64_t cap,
                       int shift, int bits, const RadixBuffers &rb, hipStream_t s,
                       const int64_t *dlen = nullptr, RadixGather gather = {}) {
  const int64_t nt = radix_tiles(cap);
  if (nt == 0) return;
  radix_hist_kernel<<<dim3(unsigned(nt)), dim3(kThreads), 0, s>>>(kin, cap, dlen, shift, bits, rb.hist, nt);
  const int64_t hlen = int64_t(1 << bits) * nt;
  exclusive_scan<uint32_t, uint32_t>(rb.hist, nullptr, hlen, rb.sums, rb.hist, nullptr, s);
  radix_scatter_tile_kernel<false, kScatterWide><<<dim3(unsigned(nt)), dim3(kScatterWide), 0, s>>>(
      kin, vin, cap, dlen, shift, bits, rb.hist, 1, nt, kout, vout, nullptr, nullptr, nullptr, gather);
}

// Stable LSD sort of `cap` (or *dlen) pairs over `bits` key bits in npass
// passes of db bits (db <= 8, npass <= kMaxPasses), ping-ponging a -> b -> a
// ...; returns which buffer holds the result (0: a, 1: b).  Onesweep for at
// most kOnesweepTiles tiles of capacity (twice that with dlen) below 2^30 keys
// (the look-back words carry 30-bit counts); other inputs take the
// multi-launch passes.  vlast, when given, receives the values of the last
// pass instead of the other buffer.  ghist_zeroed: the caller has zeroed
// rb.ghist (with other buffers, in one launch).
inline bool takes_onesweep(int64_t cap, int npass, bool has_dlen) {
  return cap < (int64_t(1) << 30) && npass <= kMaxPasses && radix_tiles(cap) <= (has_dlen ? 2 : 1) * kOnesweepTiles;
}

// totals: the caller has stored the onesweep digit totals there, in nslices
// slices of kMaxPasses x 256 words that add up to them, and zeroed the
// look-back words: no histogram launch (ignored when the sort takes the
// multi-launch passes).
inline int radix_sort(uint32_t *ka, uint32_t *va, uint32_t *kb, uint32_t *vb, int64_t cap, int bits, int npass,
                      const RadixBuffers &rb, hipStream_t s, const int64_t *dlen = nullptr,
                      uint32_t *vlast = nullptr, RadixGather gather = {}, bool ghist_zeroed = false,
                      const uint32_t *totals = nullptr, int nslices = 1) {
  const int db = (bits + npass - 1) / npass;
  const int64_t nt = radix_tiles(cap);
  const bool onesweep = takes_onesweep(cap, npass, dlen != nullptr);
  if (nt > 0 && onesweep && !totals) {
    if (!ghist_zeroed) launch_fill_zero(rb.ghist, ghist_words(), s);
    const int64_t nst = int64_t(npass) * 256 * nt;
    const unsigned hb = unsigned(imin(int(nt), 1024));
    onesweep_hist_kernel<<<dim3(hb), dim3(kThreads), 0, s>>>(ka, cap, dlen, npass, db, bits, rb.ghist, rb.status,
                                                            nst);
  }
  uint32_t *sk = ka, *sv = va;
  int in_b = 0;
  for (int p = 0; p < npass; ++p) {
    const int shift = p * db;
    const int pb = bits - shift < db ? bits - shift : db;
    uint32_t *dk = in_b ? ka : kb;
    uint32_t *dv = (p == npass - 1 && vlast) ? vlast : (in_b ? va : vb);
    const RadixGather g = p == npass - 1 ? gather : RadixGather{};
    if (nt > 0) {
      if (onesweep)
        radix_scatter_tile_kernel<true, kScatterWide><<<dim3(unsigned(nt)), dim3(kScatterWide), 0, s>>>(
            sk, sv, cap, dlen, shift, pb, (totals ? totals : rb.ghist) + p * 256, totals ? nslices : 1, nt, dk, dv,
            rb.ghist + kMaxPasses * 256 + p, rb.status + int64_t(p) * 256 * nt,
            rb.ghist + kMaxPasses * 256 + kMaxPasses, g);
      else
        radix_pass(sk, sv, dk, dv, cap, shift, pb, rb, s, dlen, g);
    }
    sk = dk;
    sv = dv;
    in_b ^= 1;
  }
  return in_b;
}

// ---- sort-specific kernels ---------------------------------------------------

// Tile count per Gaussian from its rectangle, depth bits as the first sort
// key, identity as the value; flags any disagreement with caller-given tiles.
__global__ __launch_bounds__(kThreads) void prepare_kernel(const float *__restrict__ uv,
                                                           const float *__restrict__ depth,
                                                           const int32_t *__restrict__ radius,
                                                           const int32_t *__restrict__ tiles, int64_t n, int gx,
                                                           int gy, int32_t *__restrict__ area,
                                                           uint32_t *__restrict__ keys,
                                                           uint32_t *__restrict__ vals, int32_t *__restrict__ flag) {
  const int64_t i = int64_t(blockIdx.x) * kThreads + threadIdx.x;
  if (i >= n) return;
  const int r = radius[i];
  int a = 0;
  if (r > 0) a = tile_rect(uv[2 * i], uv[2 * i + 1], r, gx, gy).area();
  area[i] = a;
  keys[i] = fbits(depth[i]);
  vals[i] = uint32_t(i);
  if (tiles && tiles[i] != a) atomicOr(flag, 1);
}

// ---- single-pass scans with a decoupled look-back (cell binning) -----------
// Workgroup b publishes its aggregate in status[b], then one wave adds up its
// predecessors' words, 64 at a time, back to the nearest one that carries an
// inclusive prefix, and b publishes its own inclusive prefix.  A status word
// holds a flag (2 bits) and two 31-bit sums (saturating).  Workgroups are
// dispatched in index order and never wait on a later one, so every
// predecessor is running or done and every wait ends; waits are also bounded
// (kLookbackSpins polls) and report an expired one.  No ticket counter: one
// atomic per workgroup on a shared word would serialise thousands of
// workgroups.  status must be zero before the launch.
constexpr uint64_t kLbAgg = 1ull << 62, kLbIncl = 2ull << 62, kLbFlags = 3ull << 62;
constexpr int64_t kLbField = (int64_t(1) << 31) - 1;

MSPLAT_DEV uint64_t lb_word(uint64_t flag, int64_t hi, int64_t lo) {
  const uint64_t h = uint64_t(hi < kLbField ? hi : kLbField), l = uint64_t(lo < kLbField ? lo : kLbField);
  return flag | (h << 31) | l;
}

// Lanes below this one in a ballot mask.
MSPLAT_DEV uint32_t mask_rank_lo(uint64_t m) {
  return __builtin_amdgcn_mbcnt_hi(uint32_t(m >> 32), __builtin_amdgcn_mbcnt_lo(uint32_t(m), 0u));
}

MSPLAT_DEV int64_t wave_sum_i64(int64_t v) {
#pragma unroll
  for (int off = kWave / 2; off >= 1; off >>= 1) v += __shfl_xor(v, off, kWave);
  return v;
}

// Called by one whole wave of workgroup b >= 1: the sums of both fields over
// workgroups [0, b); false if a wait expired.
MSPLAT_DEV bool lookback_pair(const uint64_t *status, int64_t b, int64_t &hi, int64_t &lo) {
  const int lane = __lane_id();
  hi = 0;
  lo = 0;
  if (expiry_forced()) return false;  // tests: give up without looking
  int64_t spins = 0;
  for (int64_t p = b - 1;; p -= kWave) {
    const int64_t q = p - lane;  // before workgroup 0: an inclusive zero
    uint64_t w = q >= 0 ? __hip_atomic_load(status + q, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) : kLbIncl;
    while (__ballot((w & kLbFlags) == 0ull) != 0ull) {
      if (++spins >= kLookbackSpins) return false;
      if ((w & kLbFlags) == 0ull) w = __hip_atomic_load(status + q, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    const uint64_t incl = __ballot((w & kLbFlags) == kLbIncl);
    const int stop = incl ? __ffsll((unsigned long long)incl) - 1 : kWave - 1;  // nearest inclusive word
    const bool take = lane <= stop;
    hi += wave_sum_i64(take ? int64_t((w >> 31) & uint64_t(kLbField)) : 0);
    lo += wave_sum_i64(take ? int64_t(w & uint64_t(kLbField)) : 0);
    if (incl) return true;
  }
}

// lookback_pair by a whole workgroup of NT threads (every thread calls it:
// it has barriers): in a round, wave k reads the words of workgroups
// r - 64 k - lane (r: the round's most recent), each wave sums down to its
// nearest inclusive word, and the waves' sums are added in order through LDS
// until one of them met an inclusive word; a round covers NT earlier
// workgroups instead of 64.  Every thread returns the same sums.
template <int NT>
MSPLAT_DEV bool lookback_pair_wg(const uint64_t *status, int64_t b, int64_t &hi, int64_t &lo) {
  constexpr int kW = NT / kWave;
  __shared__ int64_t s_h[kW], s_l[kW];
  __shared__ uint32_t s_f[kW];
  const int lane = threadIdx.x % kWave, wave = threadIdx.x / kWave;
  hi = 0;
  lo = 0;
  if (expiry_forced()) return false;  // tests: give up without looking
  for (int64_t r = b - 1;; r -= NT) {
    const int64_t q = r - int64_t(wave) * kWave - lane;  // before workgroup 0: an inclusive zero
    uint64_t w = q >= 0 ? __hip_atomic_load(status + q, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) : kLbIncl;
    bool expired = false;
    int64_t spins = 0;
    while (__ballot((w & kLbFlags) == 0ull) != 0ull) {
      if (++spins >= kLookbackSpins) {
        expired = true;
        break;
      }
      if ((w & kLbFlags) == 0ull) w = __hip_atomic_load(status + q, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    const uint64_t incl = __ballot((w & kLbFlags) == kLbIncl);
    const int stop = incl ? __ffsll((unsigned long long)incl) - 1 : kWave - 1;  // nearest inclusive word
    const bool take = lane <= stop;
    const int64_t wh = wave_sum_i64(take ? int64_t((w >> 31) & uint64_t(kLbField)) : 0);
    const int64_t wl = wave_sum_i64(take ? int64_t(w & uint64_t(kLbField)) : 0);
    if (lane == 0) {
      s_h[wave] = wh;
      s_l[wave] = wl;
      s_f[wave] = (incl != 0ull ? 1u : 0u) | (expired ? 2u : 0u);
    }
    __syncthreads();
    bool fin = false, ok = true;
#pragma unroll
    for (int k = 0; k < kW; ++k) {
      if (fin) break;
      hi += s_h[k];
      lo += s_l[k];
      if (s_f[k] & 2u) ok = false;
      if (s_f[k]) fin = true;
    }
    __syncthreads();  // the slots are reused by the next round
    if (fin) return ok;
  }
}

// Publishes workgroup b's sums (nh, nl) and returns its exclusive prefixes
// (wave 0 calls it; lane 0 stores).  false if a wait expired.
MSPLAT_DEV bool lookback_publish(uint64_t *status, int64_t b, int64_t nh, int64_t nl, int64_t &ph, int64_t &pl) {
  const int lane = __lane_id();
  ph = 0;
  pl = 0;
  bool ok = true;
  if (b > 0) {
    if (lane == 0) __hip_atomic_store(status + b, lb_word(kLbAgg, nh, nl), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    ok = lookback_pair(status, b, ph, pl);
  }
  // an expired wait leaves the aggregate: later workgroups look past it
  if (lane == 0 && ok)
    __hip_atomic_store(status + b, lb_word(kLbIncl, ph + nh, pl + nl), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  return ok;
}

// Cell binning, first pass (one launch).  Per row (the per-view rows of the
// views in `v`): its packed tile rectangle, and the number of S x S-tile cells
// the rectangle touches (0 for a negative depth, whose reference key sorts
// into no tile, and for an empty rectangle).  The rows with cells ("listed")
// get their depth bits and row index written, in row order, to the first
// slots of keys / vals, at positions from a single-pass scan; the last
// workgroup stores the number of listed rows and of cell entries (-1 when a
// wait expired; 2^31 when the entries reach 2^31 - 1, which no int32 list
// holds).
constexpr int kPrepRows = 8;  // rows per thread of the counting pass (2048 per workgroup)

MSPLAT_HD int64_t prep_blocks(int64_t n) {
  const int64_t per = int64_t(kThreads) * kPrepRows;
  return n > 0 ? (n + per - 1) / per : 1;
}

// Workgroup w covers rows w * 2048 + it * 256 + tid (it < kPrepRows): the
// loads of all its rounds are in flight together, and a few hundred
// workgroups (all resident at once) keep the look-back short.  Ranks follow
// row order: round, then wave, then lane.
// With dhist, it also adds the 8-bit digit counts of the listed depth keys
// (4 passes x 256) to dhist -- the onesweep depth sort's digit totals, in
// kCtSlices slices (workgroup b adds to slice b mod kCtSlices: the top digits'
// few bins would otherwise take every workgroup's atomic on one address) --
// and zeroes that sort's ndstatus look-back words in dstatus.
// With ctotal (at most kFusedCells cells), it also adds every cell's entry
// count to ctotal[k * kFusedCells + cell] (zero before the launch), for the
// fused emission: kCtSlices slices, workgroup b adding to slice b mod
// kCtSlices, so that the workgroups' atomics on one cell spread over several
// addresses (the emission adds the slices up).
constexpr int kFusedCells = 256;
constexpr int kCtSlices = 8;
static_assert(kCtSlices <= kMaxTotalSlices, "the depth sort reads at most kMaxTotalSlices slices of digit totals");

__global__ __launch_bounds__(kThreads) void bin_prepare_kernel(
    const float *__restrict__ uv, const float *__restrict__ depth, const int32_t *__restrict__ radius, int64_t n,
    Views v, int S, uint2 *__restrict__ rect, uint32_t *__restrict__ keys, uint32_t *__restrict__ vals,
    uint64_t *__restrict__ status, int64_t *__restrict__ nlisted, int64_t *__restrict__ count,
    uint32_t *__restrict__ error, uint32_t *__restrict__ ctotal, uint32_t *__restrict__ dhist,
    uint32_t *__restrict__ dstatus, int64_t ndstatus) {
  static_assert(kThreads == kFusedCells, "one thread per cell adds the cell counts");
  static_assert(kThreads == 256, "one thread per digit adds the digit totals");
  constexpr int kWaves = kThreads / kWave;
  __shared__ int64_t s_base;
  __shared__ uint32_t s_cnt[kPrepRows][kWaves];
  __shared__ int64_t s_area[kWaves];
  __shared__ uint32_t s_ct[kFusedCells];
  __shared__ uint32_t s_dh[kMaxPasses][256];
  __shared__ uint32_t s_ticket;
  // Workgroups take their rows in ticket order (status[nblocks] is the
  // ticket): a workgroup's look-back waits only on workgroups that have
  // started, whichever XCD runs them and whatever else shares the GPU.
  if (threadIdx.x == 0) s_ticket = atomicAdd(reinterpret_cast<uint32_t *>(status + gridDim.x), 1u);
  if (ctotal) s_ct[threadIdx.x] = 0u;  // uniform
  if (dhist) {  // uniform
#pragma unroll
    for (int p = 0; p < kMaxPasses; ++p) s_dh[p][threadIdx.x] = 0u;
    // the depth sort's look-back words, zeroed for its passes
    for (int64_t j = int64_t(blockIdx.x) * kThreads + threadIdx.x; j < ndstatus; j += int64_t(gridDim.x) * kThreads)
      dstatus[j] = 0u;
  }
  __syncthreads();
  const int64_t blk = s_ticket;
  const int lane = threadIdx.x % kWave, wave = threadIdx.x / kWave;
  const int64_t base = blk * kThreads * kPrepRows + threadIdx.x;
  int a[kPrepRows];
  uint32_t db[kPrepRows];
  uint32_t cr[kPrepRows];  // fused emission: the row's cells (first cell, width - 1, height - 1, ncx - 1)
  uint64_t m[kPrepRows];
  int64_t my_area = 0;
  // Every row's inputs first, their loads in flight together; then the
  // rectangles.  A workgroup's rows nearly always lie in one view, whose
  // fields are then read once as uniform values rather than per lane from
  // the view table (a dependent load per row).
  int rr[kPrepRows];
  float dd[kPrepRows], ux[kPrepRows], uy[kPrepRows];
#pragma unroll
  for (int it = 0; it < kPrepRows; ++it) {
    const int64_t i = base + int64_t(it) * kThreads;
    const bool in = i < n;
    rr[it] = in ? radius[i] : 0;
    dd[it] = in ? depth[i] : 0.f;
    ux[it] = in ? uv[2 * i] : 0.f;
    uy[it] = in ? uv[2 * i + 1] : 0.f;
  }
  const int64_t first = blk * kThreads * kPrepRows;
  const int64_t last_row = (first + kThreads * kPrepRows < n ? first + kThreads * kPrepRows : n) - 1;
  const int vb0 = __builtin_amdgcn_readfirstlane(view_of_row(v, first));
  const int vb1 = __builtin_amdgcn_readfirstlane(view_of_row(v, last_row));
  // x / S for the tile coordinates (0 <= x <= 4096 + S): a float quotient
  // (off by at most one) and one exact correction, in place of an integer
  // division sequence per use
  const float inv_s = 1.0f / float(S);
  auto div_s = [&](int x) {
    int q = int(float(x) * inv_s);
    const int rem = x - q * S;
    q += int(rem >= S) - int(rem < 0);
    return q;
  };
  struct ViewGrid {
    int gx, gy, cell0, ncx;
  };
  auto grid_of = [&](int b) { return ViewGrid{int(v.gx[b]), int(v.gy[b]), v.cell0[b], int(v.ncx[b])}; };
  const ViewGrid g0 = grid_of(vb0);
  auto rows = [&](auto one_view) {
#pragma unroll
    for (int it = 0; it < kPrepRows; ++it) {
      const int64_t i = base + int64_t(it) * kThreads;
      a[it] = 0;
      db[it] = 0u;
      cr[it] = 0u;
      if (i < n) {
        const ViewGrid vg = decltype(one_view)::value ? g0 : grid_of(view_of_row(v, i));
        const int r = rr[it];
        db[it] = fbits(dd[it]);
        TileRect t{0, 0, 0, 0};
        if (r > 0 && !(db[it] >> 31)) {
          t = tile_rect(ux[it], uy[it], r, vg.gx, vg.gy);
          if (t.area() > 0) {
            // the cells touched, as rect_cells(pack_rect(t), S) counts them
            const int cx0 = div_s(t.x0), cx1 = div_s(t.x1 + S - 1), cy0 = div_s(t.y0), cy1 = div_s(t.y1 + S - 1);
            a[it] = (cx1 - cx0) * (cy1 - cy0);
            if (ctotal) {  // the rectangle of cells, packed; counted after the loop
              cr[it] = uint32_t(vg.cell0 + cy0 * vg.ncx + cx0) | (uint32_t(cx1 - cx0 - 1) << 8) |
                       (uint32_t(cy1 - cy0 - 1) << 16) | (uint32_t(vg.ncx - 1) << 24);
            }
          }
        }
        rect[i] = pack_rect(t);
      }
      m[it] = __ballot(a[it] > 0);
      my_area += a[it];
      if (lane == 0) s_cnt[it][wave] = uint32_t(__popcll(m[it]));
    }
  };
  if (vb0 == vb1)  // uniform
    rows(std::true_type{});
  else
    rows(std::false_type{});
  my_area = wave_sum_i64(my_area);
  if (lane == 0) s_area[wave] = my_area;
  __syncthreads();
  int64_t nl = 0, ne = 0;
  uint32_t pre[kPrepRows];  // listed rows of this workgroup before (round it, this wave)
#pragma unroll
  for (int it = 0; it < kPrepRows; ++it) {
    pre[it] = uint32_t(nl);
#pragma unroll
    for (int w = 0; w < kWaves; ++w) {
      pre[it] += w < wave ? s_cnt[it][w] : 0u;
      nl += s_cnt[it][w];
    }
  }
#pragma unroll
  for (int w = 0; w < kWaves; ++w) ne += s_area[w];
  // this workgroup's counts are published first, so that later workgroups'
  // look-backs find them while this one adds up its rows' cells
  const int64_t b = blk;
  if (b > 0 && threadIdx.x == 0)
    __hip_atomic_store(status + b, lb_word(kLbAgg, nl, ne), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  // the rows' cells
  if (ctotal) {  // uniform
#pragma unroll
    for (int it = 0; it < kPrepRows; ++it) {
      if (a[it] <= 0) continue;
      const int c0 = int(cr[it] & 0xffu), w = int((cr[it] >> 8) & 0xffu) + 1, h = int((cr[it] >> 16) & 0xffu) + 1;
      const int ncx = int(cr[it] >> 24) + 1;
      for (int y = 0; y < h; ++y)
        for (int x = 0; x < w; ++x) atomicAdd(&s_ct[c0 + y * ncx + x], 1u);
    }
  }
  if (dhist) {  // uniform: the listed depth keys' 8-bit digits, for the depth sort's passes
#pragma unroll
    for (int it = 0; it < kPrepRows; ++it) {
      if (a[it] <= 0) continue;
#pragma unroll
      for (int p = 0; p < kMaxPasses; ++p) atomicAdd(&s_dh[p][(db[it] >> (8 * p)) & 0xffu], 1u);
    }
  }
  {
    // the whole workgroup looks back (489 workgroups at 1 M rows: two rounds)
    int64_t pl = 0, pe = 0;
    bool ok = true;
    if (b > 0) ok = lookback_pair_wg<kThreads>(status, b, pl, pe);
    // an expired wait leaves the aggregate: later workgroups look past it
    if (threadIdx.x == 0 && ok)
      __hip_atomic_store(status + b, lb_word(kLbIncl, pl + nl, pe + ne), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (threadIdx.x == 0) {
      s_base = pl;
      if (!ok) atomicOr(error, 1u);
      if (blk == int64_t(gridDim.x) - 1) {
        *nlisted = pl + nl;
        const int64_t e = pe + ne;
        count[0] = ok ? (e >= kLbField ? kLbField + 1 : e) : -1;
        count[1] = int64_t(expired_state());  // an earlier binning's expired wait
      }
    }
  }
  __syncthreads();
#pragma unroll
  for (int it = 0; it < kPrepRows; ++it) {
    if (a[it] > 0) {
      const int64_t j = s_base + pre[it] + mask_rank_lo(m[it]);
      keys[j] = db[it];
      vals[j] = uint32_t(base + int64_t(it) * kThreads);
    }
  }
  if (ctotal && s_ct[threadIdx.x])  // s_ct is final: the barrier above
    atomicAdd(&ctotal[(blockIdx.x % kCtSlices) * kFusedCells + threadIdx.x], s_ct[threadIdx.x]);
  if (dhist) {  // uniform; s_dh is final: the barrier above
#pragma unroll
    for (int p = 0; p < kMaxPasses; ++p)
      if (s_dh[p][threadIdx.x])
        atomicAdd(&dhist[(blockIdx.x % kCtSlices) * (kMaxPasses * 256) + p * 256 + threadIdx.x], s_dh[p][threadIdx.x]);
  }
}

// Cell binning, emission (one launch): row i of the depth order writes its
// Gaussian's cells, row-major, at slots o(i)..., o = the exclusive scan of
// the rows' cell counts in depth order, computed here by a single-pass scan
// (status: one word per workgroup, zero before the launch).  Slots past the
// capacity are not written; workgroups past nstatus hold no slot below it
// (every listed row has a cell, so o(i) >= i) and leave at once.
// 1024-thread workgroups: a quarter of the look-back chain of 256-thread
// ones (same single-view step; the 16-view step 0.5 % shorter).
constexpr int kEmitThreads = 1024;

__global__ __launch_bounds__(kEmitThreads) void emit_direct_kernel(
    const uint32_t *__restrict__ order, const int32_t *__restrict__ area_s, const uint2 *__restrict__ rect_s,
    int64_t n, const int64_t *__restrict__ nrows, int64_t cap, const int64_t *__restrict__ dlen, Views v, int S,
    uint32_t *__restrict__ cls, uint32_t *__restrict__ gid, uint64_t *__restrict__ status, int64_t nstatus,
    uint32_t *__restrict__ error, const uint32_t *__restrict__ upstream) {
  // Row i of the depth order: its Gaussian row order[i], cell count area_s[i]
  // and tile rectangle rect_s[i] (both carried into depth order by the depth
  // sort's last pass, so every read here is coalesced).  Listed rows have
  // positive depth, so every entry has a cell class.
  __shared__ int64_t s_base;
  __shared__ uint32_t s_ticket;
  // An expired wait of the counting pass or the depth sort (upstream) left the
  // depth order undefined: nothing here may be read.  The lists stay
  // unwritten; the range kernel flags them.
  if (*upstream) return;  // uniform: the word is final before this launch
  const int64_t rows = dev_len(nrows, n);  // rows past the listed ones have nothing to emit
  // Workgroups past the rows leave before taking a ticket (status[nstatus]):
  // the others take their rows in ticket order, so a look-back waits only on
  // workgroups that have started.
  if (int64_t(blockIdx.x) * kEmitThreads >= rows || int64_t(blockIdx.x) >= nstatus) return;  // uniform
  if (threadIdx.x == 0) s_ticket = atomicAdd(reinterpret_cast<uint32_t *>(status + nstatus), 1u);
  __syncthreads();
  const int64_t blk = s_ticket;
  const int64_t i = blk * kEmitThreads + threadIdx.x;
  const int lane = threadIdx.x % kWave;
  const int a = i < rows ? area_s[i] : 0;
  int64_t tot;
  const int64_t ex = block_excl_scan<kEmitThreads>(a, &tot);
  if (threadIdx.x < kWave) {
    int64_t unused, pe;
    const bool ok = lookback_publish(status, blk, 0, tot, unused, pe);
    if (threadIdx.x == 0) {
      s_base = pe;
      if (!ok) atomicOr(error, 1u);
    }
  }
  __syncthreads();
  uint32_t g = 0u, base = 0u;
  int w = 1, ncx = 1;
  const int64_t o = s_base + ex;
  const int64_t k = dev_len(dlen, cap);
  if (a > 0) {
    g = order[i];
    const int b = view_of_row(v, g);
    const uint2 rp = rect_s[i];
    const int x0 = int(rp.x & 0xffffu), y0 = int(rp.x >> 16), x1 = int(rp.y & 0xffffu);
    const int cx0 = x0 / S;
    w = imax(1, (x1 + S - 1) / S - cx0);
    ncx = v.ncx[b];
    base = uint32_t(v.cell0[b] + (y0 / S) * ncx + cx0) + 1u;
  }
  // A Gaussian's entries: the cells of its box, row by row, from offset o.
  // Small boxes are written by their own thread; large ones by the whole
  // wave, one at a time (a thread alone would hold up its wave and the grid).
  constexpr int kSmall = 16;
  if (a > 0 && a <= kSmall) {
    int q = 0, r = 0;
    for (int e = 0; e < a && o + e < k; ++e) {
      cls[o + e] = base ? base + uint32_t(q * ncx + r) : 0u;
      gid[o + e] = g;
      if (++r == w) {
        r = 0;
        ++q;
      }
    }
  }
  uint64_t big = __ballot(a > kSmall);
  while (big) {
    const int src = __ffsll((unsigned long long)big) - 1;
    big &= big - 1;
    const int sa = __shfl(a, src, kWave), sw = __shfl(w, src, kWave), sn = __shfl(ncx, src, kWave);
    const uint32_t sg = uint32_t(__shfl(int(g), src, kWave)), sb = uint32_t(__shfl(int(base), src, kWave));
    const int64_t so = (int64_t(uint32_t(__shfl(int(uint32_t(o >> 32)), src, kWave))) << 32) |
                       int64_t(uint32_t(__shfl(int(uint32_t(o)), src, kWave)));
    for (int e = lane; e < sa && so + e < k; e += kWave) {
      const int q = e / sw, r = e - q * sw;
      cls[so + e] = sb ? sb + uint32_t(q * sn + r) : 0u;
      gid[so + e] = sg;
    }
  }
}

// Cell binning, emission straight into the cell lists (one launch), when
// there are at most kFusedCells cells: row i of the depth order writes its
// Gaussian into each of its cells' lists at
//   start[c] + (entries of c from earlier workgroups) + (from earlier waves
//   of this workgroup) + (from earlier lanes of this wave),
// the depth-ordered (stable) position the emission + class sort would give it
// (a row holds a cell once).  A wave records which of its lanes hold each
// cell as a 64-bit mask in LDS; the workgroup's per-cell counts are summed
// over its waves and looked back over the earlier workgroups one cell per
// thread (workgroups take tiles in ticket order, so every predecessor is
// running or done); start[c] is the exclusive scan of the totals the
// counting pass added up.  The first workgroup writes the cell ranges
// (clipped to the capacity; entries at or past it are not written).
// Cells [lo, hi) of a 256-cell set as 4 words of 64 bits, or-ed into bits
// (words at or past nwords are left alone).
MSPLAT_DEV void cell_range_bits(int lo, int hi, int nwords, uint64_t (&bits)[kFusedCells / 64]) {
#pragma unroll
  for (int k = 0; k < kFusedCells / 64; ++k) {
    if (k >= nwords) break;  // uniform
    const int a = imax(lo, 64 * k), b = imin(hi, 64 * k + 64);
    if (a < b) {
      const int len = b - a;
      bits[k] |= (len == 64 ? ~0ull : (1ull << len) - 1ull) << (a - 64 * k);
    }
  }
}

MSPLAT_DEV uint64_t wave_or_u64(uint64_t v) {
#pragma unroll
  for (int off = kWave / 2; off >= 1; off >>= 1) v |= __shfl_xor(v, off, kWave);
  return v;
}

// At most 72 SGPRs (78 with VCC and the other reserved registers): two of
// these 16-wave workgroups then fit on a CU (measured: 78 in all fit two,
// 86 only one), so the 282 workgroups of the 1080p frame all start at once
// instead of 26 of them waiting for others to finish.
__global__ __launch_bounds__(kEmitThreads) __attribute__((amdgpu_num_sgpr(72))) void emit_cells_kernel(
    const uint32_t *__restrict__ order, const int32_t *__restrict__ area_s, const uint2 *__restrict__ rect_s,
    int64_t n, const int64_t *__restrict__ nrows, int64_t cap, Views v, int S, int ncls,
    const uint32_t *__restrict__ ctotal, uint32_t *__restrict__ ticket, uint32_t *__restrict__ status,
    int32_t *__restrict__ cell_ids, int32_t *__restrict__ cell_range, uint32_t *__restrict__ error,
    const uint32_t *__restrict__ upstream) {
  constexpr int kWaves = kEmitThreads / kWave;
  constexpr int kWords = kFusedCells / 64;
  // Rows of more than kSmall cells, and every row of a wave whose rows hold
  // more than kHeavy cells in all (the near, large Gaussians at the front of
  // the depth order), are marked by cell, not by row: the wave walks the union
  // of their cells, gathers each cell's ballot into lane (cell mod 64) and
  // adds them with one LDS atomic per 64 cells, where a per-row walk would
  // issue an atomic per row and cell, most of them on the same few words.
  // (The 1080p frame's first workgroups took 13 us to mark their rows by
  // row, the others about 3 us.)
  constexpr int kSmall = 16;
  constexpr int64_t kHeavy = 3 * kWave;
  __shared__ uint64_t s_mask[kWaves][kFusedCells];
  __shared__ uint32_t s_wbase[kWaves][kFusedCells];
  __shared__ uint32_t s_gbase[kFusedCells];
  __shared__ uint32_t s_tile;
  // An expired wait of the counting pass or the depth sort left the depth
  // order undefined: nothing is read, the lists stay empty (zero ranges),
  // and the device's expired-wait word makes the blending write NaN.
  if (*upstream) {  // uniform: the word is final before this launch
    if (blockIdx.x == 0 && threadIdx.x == 0) atomicOr(&wait_expired, 1u);
    return;
  }
  const int64_t rows = dev_len(nrows, n);
  if (int64_t(blockIdx.x) * kEmitThreads >= rows) return;  // uniform; no ticket taken
  const int tid = threadIdx.x, lane = tid % kWave, wave = tid / kWave;
  if (tid == 0) s_tile = atomicAdd(ticket, 1u);
  for (int j = tid; j < kWaves * kFusedCells; j += kEmitThreads) s_mask[j / kFusedCells][j % kFusedCells] = 0ull;
  __syncthreads();
  const int64_t blk = s_tile;
  const int64_t i = blk * kEmitThreads + tid;
  const int a = i < rows ? area_s[i] : 0;
  uint32_t g = 0u;
  int w = 1, h = 0, ncx = 1, base = 0;  // the row's cells: base + y ncx + x, x < w, y < h
  if (a > 0) {
    g = order[i];
    const int b = view_of_row(v, g);
    const uint2 rp = rect_s[i];
    const int x0 = int(rp.x & 0xffffu), y0 = int(rp.x >> 16), x1 = int(rp.y & 0xffffu);
    const int cx0 = x0 / S;
    w = imax(1, (x1 + S - 1) / S - cx0);
    h = a / w;  // a = w h
    ncx = v.ncx[b];
    base = v.cell0[b] + (y0 / S) * ncx + cx0;
  }
  const bool heavy = __builtin_amdgcn_readfirstlane(uint32_t(wave_sum_i64(a) > kHeavy)) != 0u;
  const bool big = heavy ? a > 0 : a > kSmall;
  const uint64_t below = (1ull << lane) - 1ull;
  // which lanes of the wave hold each cell: small rows one atomic per cell
  if (a > 0 && !big)
    for (int y = 0; y < h; ++y)
      for (int x = 0; x < w; ++x) {
        const int c = base + y * ncx + x;
        if (c < ncls) atomicOr(&s_mask[wave][c], 1ull << lane);
      }
  // large rows by cell: their cells as bit sets, and the wave's union
  uint64_t bits[kWords], uni[kWords];
#pragma unroll
  for (int k = 0; k < kWords; ++k) bits[k] = 0ull;
  const bool any_big = __ballot(big) != 0ull;
  const int nwords = (ncls + 63) / 64;
  if (any_big) {  // uniform
    if (big)
      for (int y = 0; y < h; ++y) cell_range_bits(base + y * ncx, base + y * ncx + w, nwords, bits);
#pragma unroll
    for (int k = 0; k < kWords; ++k) {
      uni[k] = 0ull;
      if (k >= nwords) continue;  // uniform
      const int vc = imin(imax(ncls - 64 * k, 0), 64);  // valid cells of word k
      const uint64_t valid = vc == 64 ? ~0ull : (1ull << vc) - 1ull;
      const uint64_t u = wave_or_u64(bits[k]) & valid;
      uni[k] = uniform_u64(u);
      uint32_t mlo = 0u, mhi = 0u;  // lane j: the ballot of cell 64 k + j
      for (uint64_t r = uni[k]; r; r &= r - 1) {
        const int bit = __ffsll((unsigned long long)r) - 1;
        const uint64_t m = __ballot((bits[k] >> bit) & 1ull);
        mlo = lane == bit ? uint32_t(m) : mlo;
        mhi = lane == bit ? uint32_t(m >> 32) : mhi;
      }
      if ((uni[k] >> lane) & 1ull) atomicOr(&s_mask[wave][64 * k + lane], (uint64_t(mhi) << 32) | mlo);
    }
  }
  __syncthreads();
  // per cell: the waves' offsets and the workgroup's count
  uint32_t tot = 0u;
  if (tid < ncls) {
#pragma unroll
    for (int wv = 0; wv < kWaves; ++wv) {
      s_wbase[wv][tid] = tot;
      tot += uint32_t(__popcll(s_mask[wv][tid]));
    }
  }
  int64_t dummy;
  int64_t ct = 0;  // entries of cell tid, over the counting pass's slices
  if (tid < ncls)
#pragma unroll
    for (int k = 0; k < kCtSlices; ++k) ct += ctotal[k * kFusedCells + tid];
  const int64_t start = block_excl_scan<kEmitThreads>(ct, &dummy);
  if (tid < ncls) {
    s_gbase[tid] = uint32_t(start) + digit_lookback(status, blk, tid, tot, error);
    if (blk == 0) {
      const int64_t e = start + ct;
      cell_range[2 * tid] = int32_t(start < cap ? start : cap);
      cell_range[2 * tid + 1] = int32_t(e < cap ? e : cap);
    }
  }
  __syncthreads();
  // an expired look-back (this workgroup's, or an earlier one's): the lists are undefined
  if (tid == 0 && __hip_atomic_load(error, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) atomicOr(&wait_expired, 1u);
  auto slot = [&](int c) {
    return int64_t(s_gbase[c]) + s_wbase[wave][c] + __popcll(s_mask[wave][c] & below);
  };
  if (a > 0 && !big)
    for (int y = 0; y < h; ++y)
      for (int x = 0; x < w; ++x) {
        const int c = base + y * ncx + x;
        if (c >= ncls) continue;
        const int64_t pos = slot(c);
        if (pos < cap) cell_ids[pos] = int32_t(g);
      }
  if (any_big) {  // uniform: the large rows, cell by cell of the union
#pragma unroll
    for (int k = 0; k < kWords; ++k) {
      if (uni[k] == 0ull) continue;  // uniform
      // lane j holds cell 64 k + j's first slot for this wave and its mask
      const int cj = 64 * k + lane;
      const bool on = (uni[k] >> lane) & 1ull;
      const uint32_t sb = on ? s_gbase[cj] + s_wbase[wave][cj] : 0u;
      const uint64_t sm = on ? s_mask[wave][cj] : 0ull;
      for (uint64_t r = uni[k]; r; r &= r - 1) {
        const int bit = __ffsll((unsigned long long)r) - 1;
        const uint32_t base_c = uint32_t(__builtin_amdgcn_readlane(sb, bit));
        const uint64_t m = lane_u64(sm, bit);
        if ((bits[k] >> bit) & 1ull) {
          const int64_t pos = int64_t(base_c) + __popcll(m & below);
          if (pos < cap) cell_ids[pos] = int32_t(g);
        }
      }
    }
  }
}

// tile_range[c-1] = [first, last+1) of class c in the sorted classes; class 0
// (negative depth) has no tile.  Each thread reads 4 consecutive classes with
// one 16-byte load and compares them with their neighbours.  The caller
// zero-fills tile_range (unset tiles stay (0, 0)).
// It is a sort's last launch: its first thread also folds the sort's error
// words e0, e1 (either may be NULL) into wait_expired.  Classes above
// nranges are not written (lists left undefined by an expired wait may hold
// any class).
__global__ __launch_bounds__(kThreads) void range_vec_kernel(const uint32_t *__restrict__ cls, int64_t cap,
                                                             const int64_t *__restrict__ dlen,
                                                             int32_t *__restrict__ range, int64_t nranges,
                                                             const uint32_t *__restrict__ e0,
                                                             const uint32_t *__restrict__ e1) {
  if (blockIdx.x == 0 && threadIdx.x == 0 && ((e0 && *e0) || (e1 && *e1))) atomicOr(&wait_expired, 1u);
  const int64_t k = dev_len(dlen, cap);
  const int64_t i0 = (int64_t(blockIdx.x) * kThreads + threadIdx.x) * 4;
  if (i0 >= k) return;
  uint32_t c[6];  // c[0] = class at i0 - 1, c[1..4] = i0..i0+3, c[5] = i0 + 4 (0xffffffff outside)
  c[0] = i0 > 0 ? cls[i0 - 1] : 0xffffffffu;
  if (i0 + 4 <= k) {
    const uint4 v = *reinterpret_cast<const uint4 *>(cls + i0);
    c[1] = v.x;
    c[2] = v.y;
    c[3] = v.z;
    c[4] = v.w;
  } else {
#pragma unroll
    for (int j = 0; j < 4; ++j) c[1 + j] = i0 + j < k ? cls[i0 + j] : 0xffffffffu;
  }
  c[5] = i0 + 4 < k ? cls[i0 + 4] : 0xffffffffu;
#pragma unroll
  for (int j = 1; j <= 4; ++j) {
    const uint32_t cj = c[j];
    if (cj == 0u || int64_t(cj) > nranges) continue;
    const int64_t i = i0 + j - 1;
    if (c[j - 1] != cj) range[2 * (int64_t(cj) - 1)] = int32_t(i);
    if (c[j + 1] != cj) range[2 * (int64_t(cj) - 1) + 1] = int32_t(i + 1);
  }
}

// The sort's count: 0 for no Gaussians, -1 when `tiles` disagrees with the
// rectangles, -2 when a wait of the depth sort expired.
__global__ void finish_count_kernel(const int32_t *__restrict__ flag, const uint32_t *__restrict__ error,
                                    int64_t *__restrict__ count, int64_t n) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    if (n == 0) *count = 0;
    if (*flag) *count = -1;
    if (*error) *count = -2;
  }
}

// ---- emission with equal output chunks per workgroup --------------------------
// The K output slots are cut into chunks of kEmitChunk; workgroup c writes
// chunk c.  Depth order clusters the large near splats, so splitting by
// output slots (not by Gaussians) is what balances the work.

constexpr int kEmitChunk = 4096;

MSPLAT_HD int64_t emit_chunks(int64_t k) { return (k + kEmitChunk - 1) / kEmitChunk; }

// owner[c] = depth-ordered position of the Gaussian whose run holds slot c * kEmitChunk.
__global__ __launch_bounds__(kThreads) void emit_owner_kernel(const int32_t *__restrict__ area,
                                                              const uint32_t *__restrict__ order,
                                                              const int64_t *__restrict__ offs, int64_t n,
                                                              int64_t nchunks, int64_t *__restrict__ owner,
                                                              const uint32_t *__restrict__ upstream) {
  const int64_t i = int64_t(blockIdx.x) * kThreads + threadIdx.x;
  if (i >= n || *upstream) return;  // upstream: the depth sort's error word (its order is then undefined)
  const int a = area[order[i]];
  if (a == 0) return;
  const int64_t o = offs[i];
  for (int64_t c = (o + kEmitChunk - 1) / kEmitChunk; c * kEmitChunk < o + a && c < nchunks; ++c) owner[c] = i;
}

// Slots are written p = tid, tid + 256, ... (coalesced).  The workgroup stages
// 256 consecutive Gaussians (starting at the chunk's owner) at a time: start
// offsets relative to the chunk, box and class; a thread finds the Gaussian
// holding its first slot by binary search and later ones by walking forward.
// The box is the tile rectangle (S = 1, classes are tiles + 1, ncx = gx) or
// the cells of S x S tiles it touches (classes are cells + 1, ncx cells per
// row).
__global__ __launch_bounds__(kThreads) void emit_chunk_kernel(
    const float *__restrict__ uv, const float *__restrict__ depth, const int32_t *__restrict__ radius,
    const uint32_t *__restrict__ order, const int64_t *__restrict__ offs, const int64_t *__restrict__ owner,
    int64_t n, int64_t cap, const int64_t *__restrict__ dlen, Views v, int S,
    uint32_t *__restrict__ cls, uint32_t *__restrict__ gid, const uint32_t *__restrict__ upstream) {
  const int64_t k = dev_len(dlen, cap);  // slots past it (or past cap) are not written
  if (int64_t(blockIdx.x) * kEmitChunk >= k || *upstream) return;  // upstream: as emit_owner_kernel
  __shared__ int32_t s_rel[kThreads + 1];
  __shared__ int32_t s_w[kThreads], s_ncx[kThreads];
  __shared__ uint32_t s_g[kThreads], s_c[kThreads];  // id, class of the box's first cell (0: no tile)
  const int tid = threadIdx.x;
  const int64_t c0 = int64_t(blockIdx.x) * kEmitChunk;
  const int span = int(k - c0 < kEmitChunk ? k - c0 : kEmitChunk);
  auto rel_of = [&](int64_t off) {  // offset relative to the chunk, clamped into int32
    const int64_t r = off - c0;
    return int32_t(r < -(int64_t(1) << 30) ? -(1 << 30) : (r > (int64_t(1) << 30) ? (1 << 30) : r));
  };
  int64_t gbase = owner[blockIdx.x];
  if (gbase < 0 || gbase >= n) gbase = 0;  // every chunk has an owner; never read outside [0, n)
  for (;;) {
    const int64_t i = gbase + tid;
    int32_t rel = 1 << 30;
    if (i < n) {
      const uint32_t g = order[i];
      const int r = radius[g];
      rel = rel_of(offs[i]);
      s_g[tid] = g;
      if (r > 0) {
        const int b = view_of_row(v, g);
        const TileRect t = tile_rect(uv[2 * g], uv[2 * g + 1], r, v.gx[b], v.gy[b]);
        const int x0 = t.x0 / S, x1 = (t.x1 + S - 1) / S;  // cell columns (tiles when S = 1)
        s_w[tid] = imax(1, x1 - x0);
        s_ncx[tid] = v.ncx[b];
        // depth sign bit set: the reference key's upper word is all ones
        s_c[tid] = (fbits(depth[g]) >> 31) ? 0u : uint32_t(v.cell0[b] + (t.y0 / S) * v.ncx[b] + x0) + 1u;
      }
    }
    s_rel[tid] = rel;
    if (tid == 0) s_rel[kThreads] = gbase + kThreads < n ? rel_of(offs[gbase + kThreads]) : rel_of(k);
    __syncthreads();
    const int lo_slot = imax(0, s_rel[0]);
    const int hi_slot = imin(span, s_rel[kThreads]);
    int p = lo_slot + tid;
    if (p < hi_slot) {
      int lo = 0, hi = kThreads - 1;  // largest j with s_rel[j] <= p (skips empty runs)
      while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (s_rel[mid] <= p) lo = mid; else hi = mid - 1;
      }
      for (; p < hi_slot; p += kThreads) {
        while (lo + 1 < kThreads && s_rel[lo + 1] <= p) ++lo;
        const int t = p - s_rel[lo];
        const int w = s_w[lo];
        // t / w through the fp32 reciprocal, corrected to the exact quotient
        int q = int(float(t) * __builtin_amdgcn_rcpf(float(w)));
        int rr = t - q * w;
        if (rr < 0) { --q; rr += w; } else if (rr >= w) { ++q; rr -= w; }
        cls[c0 + p] = s_c[lo] ? s_c[lo] + uint32_t(q * s_ncx[lo] + rr) : 0u;
        gid[c0 + p] = s_g[lo];
      }
    }
    if (s_rel[kThreads] >= span) break;  // uniform: every thread read the same value
    gbase += kThreads;
    __syncthreads();
  }
}

// The reference's 64-bit keys, rebuilt from (class, id) after sorting.
__global__ __launch_bounds__(kThreads) void sorted_keys_kernel(const uint32_t *__restrict__ cls,
                                                               const int32_t *__restrict__ idx,
                                                               const float *__restrict__ depth, int64_t k,
                                                               int64_t n, int64_t *__restrict__ keys) {
  const int64_t i = int64_t(blockIdx.x) * kThreads + threadIdx.x;
  if (i >= k) return;
  const uint32_t c = cls[i];
  const uint32_t g = uint32_t(idx[i]);  // >= n only in lists left undefined by an expired wait
  const uint32_t b = int64_t(g) < n ? fbits(depth[g]) : 0u;
  keys[i] = c == 0 ? int64_t(int32_t(b)) : ((int64_t(c) - 1) << 32) | int64_t(b);
}

// Reference-order keygen (sort_gaussian.cu:17-43): Gaussian i writes its keys
// from offsets[i-1] on; writes stay inside [0, k).
__global__ __launch_bounds__(kThreads) void ref_keys_kernel(const float *__restrict__ uv,
                                                            const float *__restrict__ depth,
                                                            const int32_t *__restrict__ radius,
                                                            const int32_t *__restrict__ offsets, int64_t n, int gx,
                                                            int gy, int64_t k, int64_t *__restrict__ keys,
                                                            int32_t *__restrict__ idx) {
  const int64_t i = int64_t(blockIdx.x) * kThreads + threadIdx.x;
  if (i >= n || radius[i] <= 0) return;
  const TileRect t = tile_rect(uv[2 * i], uv[2 * i + 1], radius[i], gx, gy);
  int64_t cur = i == 0 ? 0 : offsets[i - 1];
  const int64_t dk = int64_t(int32_t(fbits(depth[i])));
  for (int y = t.y0; y < t.y1; ++y)
    for (int x = t.x0; x < t.x1; ++x, ++cur)
      if (cur >= 0 && cur < k) {
        keys[cur] = (int64_t(y * gx + x) << 32) | dk;
        idx[cur] = int32_t(i);
      }
}

// sort_gaussian.cu:45-71 with bounds on the tile id.
__global__ __launch_bounds__(kThreads) void ref_range_kernel(const int64_t *__restrict__ keys, int64_t k,
                                                             int ntiles, int32_t *__restrict__ range) {
  const int64_t i = int64_t(blockIdx.x) * kThreads + threadIdx.x;
  if (i >= k) return;
  const int cur = int(keys[i] >> 32);
  const bool cur_ok = cur >= 0 && cur < ntiles;
  if (i == 0 && cur_ok) range[2 * cur] = 0;
  if (i == k - 1 && cur_ok) range[2 * cur + 1] = int32_t(k);
  if (i == 0) return;
  const int prev = int(keys[i - 1] >> 32);
  if (prev != cur) {
    if (prev >= 0 && prev < ntiles) range[2 * prev + 1] = int32_t(i);
    if (cur_ok) range[2 * cur] = int32_t(i);
  }
}

}  // namespace sortk
}  // namespace msplat
