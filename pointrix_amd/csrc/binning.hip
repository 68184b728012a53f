// Tile binning (msplat.sort_gaussian, pointrix/model/renderer/msplat.py:133-135).
//
// Contract: every splat with radius > 0 emits one instance per tile of its
// rectangle; instances are ordered by the 64-bit key (tile << 32 | f32 bits of
// depth), ties by splat index; tile_range[t] = [first, last + 1).
//
// Design: the low 32 key bits (depth) are equal for all instances of a splat,
// so the LSD radix sort of the 64-bit key runs its four depth passes over the
// Nv visible splats instead of the M instances (M ~ 10-20 x Nv), then emits
// instances in depth order and finishes with ceil(log2(T)/8) stable passes on
// the tile bits over M.  Result is identical to the full 64-bit sort.
//
// Radix passes: digits of at most 8 bits (the pass count's bits split evenly),
// 4096 items per 256-thread block.  Ranking inside a block uses wave64
// ballots (match mask of lanes with equal digits) and per-wave LDS counters,
// which keeps the sort stable; the block is reordered by digit in LDS before
// it is written, so writes go out as runs.  Element counts live on the
// device so no pass needs a host sync.
#include <mutex>

#include "gsr_internal.h"

using namespace gsr;

namespace {

constexpr int kThreads = 256;
constexpr int kItems = 16;
constexpr int kTileItems = kThreads * kItems;  // 4096
constexpr int kRadixBits = 8;
constexpr int kRadix = 1 << kRadixBits;

__device__ __forceinline__ uint32_t count_of(const uint32_t *n_ptr, uint32_t cap) {
  return n_ptr ? min(*n_ptr, cap) : cap;
}

// Logical block index of the radix kernels.  Workgroups go to the 8 XCDs
// round-robin (workgroup i to XCD i % 8), each XCD with its own L2.  Blocks
// whose writes share cache lines -- neighbouring columns of a digit's
// histogram row, the line where one block's output run of a digit ends and
// the next block's begins -- get consecutive logical indices on one XCD, so
// their partial-line writes merge in one L2 instead of leaving eight L2s as
// separate partial writes.  Only the first na blocks hold items (grids are
// sized for a capacity, the item count is on the device): those are spread
// over all eight XCDs, the rest keep their index.  A bijection of
// [0, gridDim.x); the histogram and scatter kernels of a pass compute it from
// the same count.
__device__ __forceinline__ uint32_t logical_block(uint32_t na) {
  const uint32_t b = blockIdx.x;
  if (b >= na) return b;
  const uint32_t q = na >> 3, r = na & 7u, x = b & 7u, k = b >> 3;
  return x * q + min(x, r) + k;
}
__device__ __forceinline__ uint32_t blocks_of(uint32_t n) {
  return (n + kTileItems - 1) / kTileItems;
}

// The wave's index in its block as a scalar: the compiler cannot tell that
// threadIdx.x >> 6 is the same on every lane of a wave, and would otherwise
// keep it and every address and loop bound derived from it in vector
// registers (loop control then runs on the VALU with exec-mask updates).
__device__ __forceinline__ int wave_index() {
  return __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
}

// Inclusive prefix sum over the 64 lanes of a wave in registers (DPP, no LDS
// round trip; every lane of the wave active): row_shr 1, 2, 4, 8 scan each
// row of 16 lanes (a lane whose source is outside its row adds the 0 of
// `old`), row_bcast:15 on rows 1 and 3 adds lane 15's and lane 47's sums,
// row_bcast:31 on rows 2 and 3 adds lane 31's.
__device__ __forceinline__ uint32_t wave_inclusive_scan(uint32_t v) {
  int x = (int)v;
  x += __builtin_amdgcn_update_dpp(0, x, 0x111, 0xf, 0xf, false);  // row_shr:1
  x += __builtin_amdgcn_update_dpp(0, x, 0x112, 0xf, 0xf, false);  // row_shr:2
  x += __builtin_amdgcn_update_dpp(0, x, 0x114, 0xf, 0xf, false);  // row_shr:4
  x += __builtin_amdgcn_update_dpp(0, x, 0x118, 0xf, 0xf, false);  // row_shr:8
  x += __builtin_amdgcn_update_dpp(0, x, 0x142, 0xa, 0xf, false);  // row_bcast:15
  x += __builtin_amdgcn_update_dpp(0, x, 0x143, 0xc, 0xf, false);  // row_bcast:31
  return (uint32_t)x;
}
__device__ __forceinline__ uint32_t wave_last(uint32_t x) {
  return (uint32_t)__builtin_amdgcn_readlane((int)x, 63);
}

// Exclusive scan of one value per thread over a 256-thread block.
__device__ __forceinline__ uint32_t block_exclusive_scan(uint32_t v, uint32_t &total) {
  __shared__ uint32_t s_w[kThreads / 64];
  const int lane = lane_id(), w = wave_index();
  const uint32_t inc = wave_inclusive_scan(v);
  if (lane == 63) s_w[w] = inc;
  __syncthreads();
  uint32_t base = 0;
  total = 0;
#pragma unroll
  for (int k = 0; k < kThreads / 64; ++k) {
    const uint32_t s = s_w[k];
    if (k < w) base += s;
    total += s;
  }
  __syncthreads();
  return base + inc - v;
}

// --------------------------------------------------------------------------
// Generic two-kernel exclusive scan (block sums / downsweep) of uint32_t
// values, or of uint2 values (two independent sums in one pass).
// --------------------------------------------------------------------------
__device__ __forceinline__ uint32_t vadd(uint32_t a, uint32_t b) { return a + b; }
__device__ __forceinline__ uint2 vadd(uint2 a, uint2 b) {
  return make_uint2(a.x + b.x, a.y + b.y);
}
__device__ __forceinline__ uint32_t vsub(uint32_t a, uint32_t b) { return a - b; }
__device__ __forceinline__ uint2 vsub(uint2 a, uint2 b) {
  return make_uint2(a.x - b.x, a.y - b.y);
}
__device__ __forceinline__ uint32_t vscan(uint32_t x) { return wave_inclusive_scan(x); }
__device__ __forceinline__ uint2 vscan(uint2 x) {
  return make_uint2(wave_inclusive_scan(x.x), wave_inclusive_scan(x.y));
}
template <class V>
__device__ __forceinline__ V vzero() {
  return V{};
}

// block_exclusive_scan over V
template <class V>
__device__ __forceinline__ V block_exclusive_scan_v(V v, V &total) {
  __shared__ V s_wv[kThreads / 64];
  const int lane = lane_id(), w = wave_index();
  const V inc = vscan(v);
  if (lane == 63) s_wv[w] = inc;
  __syncthreads();
  V base = vzero<V>();
  total = vzero<V>();
#pragma unroll
  for (int k = 0; k < kThreads / 64; ++k) {
    const V s = s_wv[k];
    if (k < w) base = vadd(base, s);
    total = vadd(total, s);
  }
  __syncthreads();
  return vadd(base, vsub(inc, v));
}

template <class Load>
__global__ __launch_bounds__(kThreads) void scan_reduce_kernel(
    Load ld, const uint32_t *n_ptr, uint32_t cap, typename Load::value_type *partials) {
  using V = typename Load::value_type;
  const uint32_t n = count_of(n_ptr, cap);
  const uint32_t base = blockIdx.x * kTileItems;
  V s = vzero<V>();
  if (base < n) {
#pragma unroll 4
    for (int k = 0; k < kItems; ++k) {
      const uint32_t i = base + k * kThreads + threadIdx.x;
      if (i < n) s = vadd(s, ld(i));
    }
  }
  V total;
  block_exclusive_scan_v(s, total);
  if (threadIdx.x == 0) partials[blockIdx.x] = total;
}

// What block 0 of a downsweep does with the grand total (all threads of the
// block call it).
struct TotNone {
  template <class V>
  __device__ void operator()(V) const {}
};
struct TotU32 {
  uint32_t *out;
  __device__ void operator()(uint32_t t) const {
    if (threadIdx.x == 0) *out = t;
  }
};
// the compaction: the total (visible splats) and the plain sum of aux[0, naux)
// (the count blocks' instance totals: M)
// (the count also goes straight to host_out, a device-visible pinned host
// word, when given: no copy is queued behind the kernel)
struct TotCompact {
  uint32_t *out;
  const uint32_t *aux;
  uint32_t naux;
  uint32_t *aux_out;
  uint32_t *host_out = nullptr;
  uint32_t *high = nullptr;  // receives max(*high, the sum)
  __device__ void operator()(uint32_t t) const {
    uint32_t a = 0;
#pragma unroll 4
    for (uint32_t i = threadIdx.x; i < naux; i += kThreads) a += aux[i];
    uint32_t sum;
    block_exclusive_scan(a, sum);
    if (threadIdx.x == 0) {
      *out = t;
      *aux_out = sum;
      if (host_out != nullptr) *host_out = sum;
      if (high != nullptr) *high = max(*high, sum);
    }
  }
};
// two sums: .x -> out_x, .y -> out_y
struct TotPair {
  uint32_t *out_x, *out_y;
  __device__ void operator()(uint2 t) const {
    if (threadIdx.x == 0) {
      *out_x = t.x;
      *out_y = t.y;
    }
  }
};

// Downsweep: thread t owns items [16t, 16t + 16) of the block for the serial
// part of the scan, but items are loaded and stored in coalesced order
// (item r * 256 + t by thread t) through an LDS transpose (row padded to 17
// values: conflict-free both ways).  The block's offset is the sum of the
// block sums before it (read here: no separate scan of the block sums);
// block 0 sums them all for the grand total, also when the scan is empty.
template <class Load, class Store, class Tot>
__global__ __launch_bounds__(kThreads) void scan_down_kernel(
    Load ld, Store st, const uint32_t *n_ptr, uint32_t cap,
    const typename Load::value_type *__restrict__ partials, uint32_t nb, Tot tot) {
  using V = typename Load::value_type;
  constexpr int kRow = kItems + 1;
  __shared__ V s_v[kThreads * kRow];
  __shared__ V s_e[kThreads * kRow];
  const uint32_t n = count_of(n_ptr, cap);
  const uint32_t base = blockIdx.x * kTileItems;
  if (base >= n && blockIdx.x != 0) return;  // uniform over the block
  const uint32_t lim = blockIdx.x == 0 ? nb : blockIdx.x;
  V pre = vzero<V>();
#pragma unroll 8
  for (uint32_t k = threadIdx.x; k < lim; k += kThreads) pre = vadd(pre, partials[k]);
  V before;
  block_exclusive_scan_v(pre, before);
  if (blockIdx.x == 0) {
    tot(before);
    before = vzero<V>();
  }
  if (base >= n) return;  // block 0 of an empty scan
  typename Store::Pre spre[kItems];
#pragma unroll
  for (int r = 0; r < kItems; ++r) {
    const uint32_t idx = r * kThreads + threadIdx.x;
    const uint32_t i = base + idx;
    s_v[(idx / kItems) * kRow + idx % kItems] = i < n ? ld(i) : vzero<V>();
    if (i < n) spre[r] = st.pre(i);
  }
  __syncthreads();
  V v[kItems];
  V s = vzero<V>();
#pragma unroll
  for (int k = 0; k < kItems; ++k) {
    v[k] = s_v[threadIdx.x * kRow + k];
    s = vadd(s, v[k]);
  }
  V total;
  V run = vadd(before, block_exclusive_scan_v(s, total));
#pragma unroll
  for (int k = 0; k < kItems; ++k) {
    s_e[threadIdx.x * kRow + k] = run;
    run = vadd(run, v[k]);
  }
  __syncthreads();
#pragma unroll
  for (int r = 0; r < kItems; ++r) {
    const uint32_t idx = r * kThreads + threadIdx.x;
    const uint32_t i = base + idx;
    const uint32_t slot = (idx / kItems) * kRow + idx % kItems;
    if (i < n) st(i, s_e[slot], s_v[slot], spre[r]);
  }
}

struct LoadU32 {
  using value_type = uint32_t;
  const uint32_t *p;
  __device__ uint32_t operator()(uint32_t i) const { return p[i]; }
};
// Store functors of the downsweep: pre(i) is requested with the scanned
// value (the loads of a store that gathers are then in flight with the
// scan's own), operator() stores item i with its exclusive prefix.
struct NoPre {};
struct StoreU32 {
  using Pre = NoPre;
  uint32_t *p;
  __device__ NoPre pre(uint32_t) const { return NoPre{}; }
  __device__ void operator()(uint32_t i, uint32_t excl, uint32_t, NoPre) const { p[i] = excl; }
};

// --------------------------------------------------------------------------
// Radix sort passes (32-bit key, 32-bit value, stable)
// --------------------------------------------------------------------------
// A pass sorts by the digit (key >> shift) & mask, mask = 2^bits - 1 with
// bits <= 8; a sort over nbits key bits uses ceil(nbits / 8) passes of equal
// width (12 tile bits: two passes of 6), which keeps every pass's buckets as
// few and as long as the pass count allows.
// Tile counts gathered by the histogram of the last of two tile passes (its
// input is sorted by the low digit, the key bits below `shift`): per block b,
// the items of the block's first low digit (slot 0) and of the next one
// (slot 1) counted per high digit into tcnt[b][slot][digit]; items of later low
// digits (a block spanning more than two) add to over[tile] with a global
// atomic.  tile_count_kernel then sums each tile's entries over the blocks.
struct TileCount {
  uint32_t *tcnt;  // [nb][2][kRadix]; nullptr: not gathered
  uint32_t *over;  // [T], zero on entry
};

template <int kIt = kItems>
__global__ __launch_bounds__(kThreads) void radix_hist_kernel(const uint32_t *__restrict__ keys,
                                                              const uint32_t *n_ptr,
                                                              uint32_t cap, int shift,
                                                              uint32_t mask,
                                                              uint32_t *__restrict__ hist,
                                                              uint32_t nb, TileCount tc) {
  constexpr uint32_t kBlk = kThreads * kIt;
  __shared__ uint32_t s_h[kRadix];
  __shared__ uint32_t s_t[2][kRadix];
  s_h[threadIdx.x] = 0;
  s_t[0][threadIdx.x] = 0;
  s_t[1][threadIdx.x] = 0;
  __syncthreads();
  const uint32_t n = count_of(n_ptr, cap);
  const uint32_t lb = logical_block((n + kBlk - 1) / kBlk);
  const uint32_t base = lb * kBlk;
  if (base >= n) return;  // no items: no column (the digit scan covers the others)
  const bool tiles = tc.tcnt != nullptr;
  const uint32_t lmask = (1u << shift) - 1u;
  const uint32_t lf = (tiles && base < n) ? (keys[base] & lmask) : 0u;
  // a block of one low digit (nearly all of them) has its slot-0 counts in
  // s_h already; only blocks that span several count per slot
  const bool split =
      tiles && base < n && (keys[min(base + kBlk, n) - 1] & lmask) != lf;
  auto count = [&](uint32_t k) {
    const uint32_t d = (k >> shift) & mask;
    atomicAdd(&s_h[d], 1u);
    if (split) {
      const uint32_t slot = (k & lmask) - lf;
      if (slot < 2u)
        atomicAdd(&s_t[slot][d], 1u);
      else
        atomicAdd(&tc.over[k], 1u);
    }
  };
  if (base + kBlk <= n) {
    // full block: 16-byte loads, all in flight before the LDS counts
    // (keys and the block base are 16-byte aligned)
    uint4 v[kIt / 4];
#pragma unroll
    for (int k = 0; k < kIt / 4; ++k)
      v[k] = reinterpret_cast<const uint4 *>(keys + base)[k * kThreads + threadIdx.x];
#pragma unroll
    for (int k = 0; k < kIt / 4; ++k) {
      count(v[k].x);
      count(v[k].y);
      count(v[k].z);
      count(v[k].w);
    }
  } else if (base < n) {
#pragma unroll 4
    for (int k = 0; k < kIt; ++k) {
      const uint32_t i = base + k * kThreads + threadIdx.x;
      if (i < n) count(keys[i]);
    }
  }
  __syncthreads();
  hist[threadIdx.x * nb + lb] = s_h[threadIdx.x];
  if (tiles) {
    tc.tcnt[(2 * (size_t)lb) * kRadix + threadIdx.x] =
        split ? s_t[0][threadIdx.x] : s_h[threadIdx.x];
    tc.tcnt[(2 * (size_t)lb + 1) * kRadix + threadIdx.x] =
        split ? s_t[1][threadIdx.x] : 0u;
  }
}

// counts[t] for tile t = (high digit h, low digit l).  The first pass's digit
// totals give the low digits' starts P_l in its output, so block b (items
// [4096 b, 4096 b + 4096)) starts in digit l exactly when ceil(P_l / 4096) <= b
// < ceil(P_{l+1} / 4096): the blocks holding digit l in slot 1 (first digit
// l - 1) and slot 0 (first digit l) are two known runs, and count[t] is the sum
// of their entries plus over[t].  One workgroup per low digit; thread (g, h),
// h the high digit (hbits bits) and g one of the 256 >> hbits groups, sums
// every group-th block of the runs (independent loads, eight in flight); the
// groups are added in LDS.
__global__ __launch_bounds__(kThreads) void tile_count_kernel(
    const uint32_t *__restrict__ tcnt, const uint32_t *__restrict__ totals1, int lbits,
    const uint32_t *__restrict__ over, int hbits, int T, uint32_t *__restrict__ counts) {
  __shared__ uint32_t s_c[kThreads];
  __shared__ uint32_t s_b[3];
  const uint32_t l = blockIdx.x;
  {
    // P_{l-1}, P_l, P_{l+1}: prefix sums of the first pass's totals
    const uint32_t d = threadIdx.x;
    const uint32_t v = d < (1u << lbits) ? totals1[d] : 0u;
    uint32_t tot;
    const uint32_t excl = block_exclusive_scan(v, tot);
    const uint32_t blk = (excl + kTileItems - 1) / kTileItems;  // ceil(P_d / 4096)
    if (l > 0u && d == l - 1u) s_b[0] = blk;
    if (d == l) s_b[1] = blk;
    if (d == l + 1u) s_b[2] = blk;
    if (l == 0u && d == 0u) s_b[0] = blk;
    if (l + 1u == (1u << lbits) && d == 0u) s_b[2] = (tot + kTileItems - 1) / kTileItems;
  }
  __syncthreads();
  const uint32_t h = threadIdx.x & ((1u << hbits) - 1u), grp = threadIdx.x >> hbits;
  const uint32_t ngrp = (uint32_t)kThreads >> hbits;
  const uint32_t b0 = s_b[0], b1 = s_b[1], b2 = s_b[2];
  uint32_t c = 0;
  for (uint32_t b = b0 + grp; b < b2; b += 8 * ngrp) {
    uint32_t v[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) {
      const uint32_t bb = b + k * ngrp;
      // blocks [b0, b1) start in digit l - 1 (slot 1), [b1, b2) in digit l
      v[k] = bb < b2 ? tcnt[(2 * (size_t)bb + (bb < b1 ? 1u : 0u)) * kRadix + h] : 0u;
    }
#pragma unroll
    for (int k = 0; k < 8; ++k) c += v[k];
  }
  s_c[threadIdx.x] = c;
  __syncthreads();
  const int t = (int)((h << lbits) | l);
  if (grp == 0 && t < T) {
    for (uint32_t q = 1; q < ngrp; ++q) c += s_c[(q << hbits) | h];
    counts[t] = c + over[t];
  }
}

// tile_range[t] = [start, start + counts[t]) with start the exclusive prefix of
// the counts, (0, 0) for an empty tile; one workgroup, 4096 tiles per round.
__global__ __launch_bounds__(1024) void ranges_from_counts_kernel(
    const uint32_t *__restrict__ counts, int T, int32_t *__restrict__ tile_range) {
  __shared__ uint32_t s_w[16];
  __shared__ uint32_t s_carry;
  const int lane = lane_id(), w = wave_index();
  if (threadIdx.x == 0) s_carry = 0;
  __syncthreads();
  for (int r0 = 0; r0 < T; r0 += 4096) {
    const int t0 = r0 + 4 * threadIdx.x;
    uint32_t v[4], sum = 0;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      v[j] = t0 + j < T ? counts[t0 + j] : 0u;
      sum += v[j];
    }
    const uint32_t inc = wave_inclusive_scan(sum);
    if (lane == 63) s_w[w] = inc;
    __syncthreads();
    uint32_t run = s_carry + inc - sum, tot = 0;
    for (int k = 0; k < 16; ++k) {
      if (k < w) run += s_w[k];
      tot += s_w[k];
    }
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      if (t0 + j < T) {
        tile_range[2 * (t0 + j)] = v[j] ? (int32_t)run : 0;
        tile_range[2 * (t0 + j) + 1] = v[j] ? (int32_t)(run + v[j]) : 0;
      }
      run += v[j];
    }
    __syncthreads();
    if (threadIdx.x == 0) s_carry += tot;
    __syncthreads();
  }
}

// Histogram columns a digit scan covers: of the nb a grid has, those of the
// blocks that hold items, ceil(min(*p, cap) / div) + add (p == nullptr: all);
// the blocks past them write no column.
struct Cols {
  const uint32_t *p = nullptr;
  uint32_t cap = 0, div = kTileItems, add = 0;
  __device__ uint32_t of(uint32_t nb) const {
    return p ? min(nb, (min(*p, cap) + div - 1) / div + add) : nb;
  }
};

// Per digit d (one block each): hist[d][b] <- sum of hist[d][b'] over
// b' < b (exclusive, in place) and totals[d] <- the digit's total, over the
// columns cols covers (rows of nb).  Thread t owns a contiguous run of blocks;
// run sums are scanned across the block.
__global__ __launch_bounds__(kThreads) void digit_scan_kernel(uint32_t *__restrict__ hist,
                                                              uint32_t nb,
                                                              uint32_t *__restrict__ totals,
                                                              Cols cols) {
  uint32_t *row = hist + (size_t)blockIdx.x * nb;
  const uint32_t nc = cols.of(nb);
  const uint32_t per = (nc + kThreads - 1) / kThreads;
  const uint32_t b0 = min(threadIdx.x * per, nc), b1 = min(b0 + per, nc);
  constexpr uint32_t kHeld = 32;
  if (per <= kHeld) {
    // the run held in registers: one read (all loads in flight), one write
    uint32_t v[kHeld], sum = 0;
#pragma unroll
    for (uint32_t k = 0; k < kHeld; ++k) {
      v[k] = b0 + k < b1 ? row[b0 + k] : 0u;
      sum += v[k];
    }
    uint32_t total;
    uint32_t run = block_exclusive_scan(sum, total);
#pragma unroll
    for (uint32_t k = 0; k < kHeld; ++k) {
      if (b0 + k < b1) row[b0 + k] = run;
      run += v[k];
    }
    if (threadIdx.x == 0) totals[blockIdx.x] = total;
    return;
  }
  // longer runs are read in groups of 8 independent loads (one latency per group)
  uint32_t sum = 0;
  for (uint32_t b = b0; b < b1; b += 8) {
    uint32_t v[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) v[k] = b + k < b1 ? row[b + k] : 0u;
#pragma unroll
    for (int k = 0; k < 8; ++k) sum += v[k];
  }
  uint32_t total;
  uint32_t run = block_exclusive_scan(sum, total);
  for (uint32_t b = b0; b < b1; b += 8) {
    uint32_t v[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) v[k] = b + k < b1 ? row[b + k] : 0u;
#pragma unroll
    for (int k = 0; k < 8; ++k) {
      if (b + k < b1) row[b + k] = run;
      run += v[k];
    }
  }
  if (threadIdx.x == 0) totals[blockIdx.x] = total;
}

// Stable scatter of one block of kTileItems items.  Each wave ranks its
// items per digit with ballots (match mask of equal digits, rank = earlier
// lanes of the mask plus the wave's running count); the block then places
// every item at its digit-sorted position in LDS and writes the block out in
// LDS order, so consecutive lanes write consecutive addresses of a digit's
// output run instead of scattering single words.
// Items of a radix block from arrays in memory.
struct ListItems {
  static constexpr bool kPackedItems = false;
  const uint32_t *kin;
  const int32_t *vin;
  // rows wbase + 64 r + lane of this wave (every load issued before the
  // ranking, whose wave barriers would otherwise expose one latency per row)
  template <int kIt>
  __device__ __forceinline__ void load(uint32_t, uint32_t, uint32_t wbase, uint32_t n,
                                       uint32_t (&key)[kIt], int32_t (&val)[kIt],
                                       uint32_t *, int32_t *, uint32_t *) const {
    const int lane = lane_id();
#pragma unroll
    for (int r = 0; r < kIt; ++r) {
      const uint32_t i = wbase + r * 64 + lane;
      const bool valid = i < n;
      key[r] = valid ? kin[i] : 0u;
      val[r] = valid ? vin[i] : 0;
    }
  }
};

// What a scatter pass writes besides the sorted values (all optional):
// kgather: the key output receives kgather[value] (a per-item array carried
// into the sorted order) instead of the key; rank_out: each value's sorted
// position (the inverse permutation); pin / pout: an 8-byte payload per item
// (pin in the input order) moved with it; kout_area: the key output receives
// the payload's packed rectangle area (w * h of (x0 | y0 << 16, w | h << 16)).
// binfo (row binning's pass B): block b covers items [binfo[b].x, binfo[b].y)
// of row ty = binfo[b].w & 255, whose blocks are blk0 = binfo[b].z onwards
// (binfo[b].w >> 8 of them); digits are the row's columns, each placed from
// its tile's start (the row's first instance plus the earlier columns'
// counts, both from the scanned histogram), and the row's first block writes
// tile_range for the row's gx tiles.
struct ScatterExtras {
  const uint32_t *kgather = nullptr;
  uint32_t *rank_out = nullptr;
  const uint2 *pin = nullptr;
  uint2 *pout = nullptr;
  bool kout_area = false;
  const uint4 *binfo = nullptr;
  const uint32_t *row_blocks = nullptr;  // with binfo: the number of row blocks
  int32_t *tile_range = nullptr;
  int gx = 0;
  // e_out (row binning's pass A): each item's first instance in the sorted
  // order, E = the digit's (row's) first instance + the widths (key >> 16) of
  // the digit's items before it: whist = the scanned width histogram rows,
  // wtot = the width totals per digit
  uint32_t *e_out = nullptr;
  const uint32_t *whist = nullptr;
  const uint32_t *wtot = nullptr;
  uint32_t *e_own = nullptr;  // with e_out: e_own[k] = the item holding instance 1024 k
  uint32_t e_own_cap = 0;
};

// kPay: the pass carries an 8-byte payload (ex.pin / ex.pout), loaded with
// the items and staged through LDS beside them; its LDS array exists only
// then, so the other passes keep their occupancy.
// kBits: the digit bits the ranking ballots (bits >= kBits are zero in every
// digit of the pass; their ballots would keep the match masks as they are).
template <class Items, bool kPay = false, int kIt = kItems, int kBits = kRadixBits>
__global__ __launch_bounds__(kThreads) void radix_scatter_kernel(
    Items items, uint32_t *__restrict__ kout, int32_t *__restrict__ vout, const uint32_t *n_ptr,
    uint32_t cap, int shift, int bits, const uint32_t *__restrict__ hist_scanned, uint32_t nb,
    const uint32_t *__restrict__ totals, ScatterExtras ex) {
  static_assert(kRadix == kThreads, "one digit per thread");
  constexpr uint32_t kBlk = kThreads * kIt;
  __shared__ uint32_t s_cnt[kThreads / 64][kRadix];  // per-wave counts, then LDS offsets
  __shared__ int32_t s_delta[kRadix];                // global start - LDS start per digit
  // packed items (Items::kPackedItems): the key word carries the value in its
  // low 24 bits, and no value array is staged
  constexpr bool kPk = Items::kPackedItems;
  __shared__ uint32_t s_key[kBlk];
  __shared__ int32_t s_val[kPk ? 1 : kBlk];
  __shared__ uint2 s_pay[kPay ? kBlk : 1];
  __shared__ uint32_t s_flag[kThreads];
  uint32_t n, base, lb;
  uint4 bi = make_uint4(0u, 0u, 0u, 0u);
  if (ex.binfo != nullptr) {
    lb = logical_block(*ex.row_blocks);
    bi = ex.binfo[lb];
    base = bi.x;
    n = bi.y;
  } else {
    n = count_of(n_ptr, cap);
    lb = logical_block((n + kBlk - 1) / kBlk);
    base = lb * kBlk;
  }
  if (base >= n) return;  // uniform over the block
  const uint32_t mask = (1u << bits) - 1u;
  const int w = wave_index(), lane = lane_id();
#pragma unroll
  for (int k = 0; k < kThreads / 64; ++k) s_cnt[k][threadIdx.x] = 0;
  const uint64_t lt = lanemask_lt();
  uint32_t key[kIt];
  int32_t val[kIt];
  uint32_t rank[kIt];
  const uint32_t wbase = base + w * (kBlk / (kThreads / 64));
  // this thread's digit total and block offset, requested with the items
  const bool my_digit = (int)threadIdx.x < (1 << bits);
  const uint32_t my_total = my_digit ? totals[threadIdx.x] : 0u;
  const uint32_t my_hist = hist_scanned[threadIdx.x * nb + lb];
  // row blocks: this column's count before the row's first block and after its last
  uint32_t row_p0 = 0, row_p1 = 0;
  if (ex.binfo != nullptr && my_digit) {
    row_p0 = hist_scanned[threadIdx.x * nb + bi.z];
    row_p1 = hist_scanned[threadIdx.x * nb + bi.z + (bi.w >> 8)];
  }
  // E: this digit's widths in earlier blocks and in all blocks
  uint32_t my_w = 0, my_wt = 0;
  if (ex.e_out != nullptr && my_digit) {
    my_w = ex.whist[threadIdx.x * nb + lb];
    my_wt = ex.wtot[threadIdx.x];
  }
  items.template load<kIt>(lb, base, wbase, n, key, val, s_key, s_val, s_flag);
  uint2 pay[kPay ? kIt : 1];
  if constexpr (kPay) {
#pragma unroll
    for (int r = 0; r < kIt; ++r) {
      const uint32_t i = wbase + r * 64 + lane;
      pay[r] = i < n ? ex.pin[i] : make_uint2(0u, 0u);
    }
  }
  __syncthreads();
#pragma unroll
  for (int r = 0; r < kIt; ++r) {
    const uint32_t i = wbase + r * 64 + lane;
    const bool valid = i < n;
    const uint32_t d = (key[r] >> shift) & mask;
    // m = the valid lanes whose digit equals this lane's: one ballot per
    // digit bit, all kBits of them with no branch (the bits at and above
    // `bits` are zero in every digit, so their ballots keep m as it is)
    // (per bit: v_bfe_i32, v_cmp into the ballot, two xor and the and; the
    // empty asm keeps the compare on clr instead of a shift of the digit)
    const uint64_t vm = __ballot(valid);
    uint32_t mlo = (uint32_t)vm, mhi = (uint32_t)(vm >> 32);
    const uint32_t nd = ~d;
#pragma unroll
    for (int b = 0; b < kBits; ++b) {
      uint32_t clr = (uint32_t)__builtin_amdgcn_sbfe((int)nd, b, 1);  // 0: bit b set, ~0: clear
      asm("" : "+v"(clr));
      const uint64_t bal = __ballot(clr == 0u);
      mlo &= (uint32_t)bal ^ clr;
      mhi &= (uint32_t)(bal >> 32) ^ clr;
    }
    const uint64_t m = ((uint64_t)mhi << 32) | mlo;
    const uint32_t before = (uint32_t)__popcll(m & lt);
    const uint32_t prior = s_cnt[w][d];  // d is in range on every lane
    __builtin_amdgcn_wave_barrier();
    // the group's last lane stores the digit's count after this row
    if (valid && (m >> lane) == 1ull) s_cnt[w][d] = prior + before + 1u;
    __builtin_amdgcn_wave_barrier();
    rank[r] = prior + before;
  }
  __syncthreads();
  {
    const int d = threadIdx.x;
    uint32_t run = 0;
#pragma unroll
    for (int k = 0; k < kThreads / 64; ++k) {
      const uint32_t t = s_cnt[k][d];
      s_cnt[k][d] = run;
      run += t;
    }
    uint32_t total;
    const uint32_t lstart = block_exclusive_scan(run, total);
#pragma unroll
    for (int k = 0; k < kThreads / 64; ++k) s_cnt[k][d] += lstart;
    if (ex.binfo == nullptr) {
      // global start of digit d: all smaller digits, then d in earlier blocks
      const uint32_t dstart = block_exclusive_scan(my_total, total);
      s_delta[d] = (int32_t)(dstart + my_hist - lstart);
    } else {
      // tile (ty, d): the row's first instance, the row's smaller columns,
      // then d in the row's earlier blocks
      const uint32_t blk0 = bi.z, ty = bi.w & 0xFFu;
      const uint32_t c = row_p1 - row_p0;
      const uint32_t e0 = base - (lb - blk0) * kBlk;
      const uint32_t tstart = e0 + block_exclusive_scan(c, total);
      s_delta[d] = (int32_t)(tstart + (my_hist - row_p0) - lstart);
      if (lb == blk0 && d < ex.gx) {
        const int64_t t = (int64_t)ty * ex.gx + d;
        ex.tile_range[2 * t] = c ? (int32_t)tstart : 0;
        ex.tile_range[2 * t + 1] = c ? (int32_t)(tstart + c) : 0;
      }
    }
  }
  __syncthreads();
#pragma unroll
  for (int r = 0; r < kIt; ++r) {
    const uint32_t i = wbase + r * 64 + lane;
    if (i < n) {
      const uint32_t lp = s_cnt[w][(key[r] >> shift) & mask] + rank[r];
      s_key[lp] = key[r];
      if constexpr (!kPk) s_val[lp] = val[r];
      if constexpr (kPay) s_pay[lp] = pay[r];
    }
  }
  __syncthreads();
  const uint32_t cnt = min((uint32_t)kBlk, n - base);
#pragma unroll 4
  for (int k = 0; k < kIt; ++k) {
    const uint32_t i = k * kThreads + threadIdx.x;
    if (i < cnt) {
      const uint32_t kk = s_key[i];
      const uint32_t pos = i + (uint32_t)s_delta[(kk >> shift) & mask];
      const int32_t vv = kPk ? (int32_t)(kk & 0xFFFFFFu) : s_val[i];
      uint2 pl = make_uint2(0u, 0u);
      if constexpr (kPay) {
        pl = s_pay[i];
        ex.pout[pos] = pl;
      }
      if (kout != nullptr)
        kout[pos] = ex.kout_area ? (pl.y & 0xFFFFu) * (pl.y >> 16)
                                 : (ex.kgather != nullptr ? ex.kgather[vv] : kk);
      vout[pos] = vv;
      if (ex.rank_out != nullptr) ex.rank_out[vv] = pos;
    }
  }
  if constexpr (!kPk) {
    if (ex.e_out != nullptr) {  // uniform
      // the block's items in LDS (digit) order: wave w scans the widths of
      // items [1024 w, 1024 w + 1024) in rows of 64 (lane = item, no bank
      // conflicts) into each item's exclusive width prefix within the wave,
      // kept in s_val (read above, free now); s_woff = the waves' offsets
      __shared__ uint32_t s_woff[kThreads / 64];
      __syncthreads();
      const int wv = wave_index(), ln = lane_id();
      uint32_t carry = 0;
#pragma unroll 4
      for (int r = 0; r < kIt; ++r) {
        const uint32_t idx = (uint32_t)wv * (kBlk / 4) + r * 64 + ln;
        const uint32_t wd = idx < cnt ? s_key[idx] >> 16 : 0u;
        const uint32_t inc = wave_inclusive_scan(wd);
        if (idx < cnt) s_val[idx] = (int32_t)(carry + inc - wd);
        carry += wave_last(inc);
      }
      if (ln == 0) s_woff[wv] = carry;
      __syncthreads();
      if (threadIdx.x == 0) {
        uint32_t run = 0;
        for (int k = 0; k < kThreads / 64; ++k) {
          const uint32_t t = s_woff[k];
          s_woff[k] = run;
          run += t;
        }
      }
      __syncthreads();
      uint32_t tot;
      {
        // digit d's first instance in this block: the digit's first instance
        // overall (prefix of the totals) + its widths in earlier blocks, less
        // the width prefix at the digit's first LDS slot (s_cnt[0][d]; a
        // digit without items here is never looked up)
        const int d = threadIdx.x;
        const uint32_t dstart = block_exclusive_scan(my_wt, tot);
        const uint32_t ls = s_cnt[0][d];
        const uint32_t gls =
            ls < cnt ? (uint32_t)s_val[ls] + s_woff[ls / (kBlk / 4)] : 0u;
        s_flag[d] = dstart + my_w - gls;
      }
      __syncthreads();
#pragma unroll 4
      for (int k = 0; k < kIt; ++k) {
        const uint32_t i = k * kThreads + threadIdx.x;
        if (i < cnt) {
          const uint32_t kk = s_key[i];
          const uint32_t d = (kk >> shift) & mask;
          const uint32_t pos = i + (uint32_t)s_delta[d];
          const uint32_t e = s_flag[d] + (uint32_t)s_val[i] + s_woff[i / (kBlk / 4)];
          ex.e_out[pos] = e;
          const uint32_t kq = (e + 1023u) >> 10;  // an item holds < 1024 instances
          if ((kq << 10) < e + (kk >> 16) && kq < ex.e_own_cap) ex.e_own[kq] = pos;
        }
      }
    }
  }
}

// --------------------------------------------------------------------------
// Binning kernels
// --------------------------------------------------------------------------
// Tile grid of the items being binned.  One view: (gx, gy) and image size
// (W, H) with vi == nullptr.  Several views: item j belongs to view j / npv,
// whose grid, size and first global tile come from the view table vi (int32
// [V][8]: W, H, gx, gy, tile_off, pix_off, -, -); emitted tile ids are global
// (tile_off + local).  With conic != nullptr ("tight" binning) a splat is
// emitted only into the tiles of its rectangle that tile_reachable accepts
// (opacity of item j at opac[j % npv], or opac[j] for one view).
struct Grid {
  int gx, gy, tw, th;
  const int32_t *vi;
  uint32_t npv;
  int W, H;
  const float *conic;
  const float *opac;
};

struct ItemGrid {
  int gx, gy, toff, W, H;
};

__device__ __forceinline__ ItemGrid item_grid(const Grid &g, uint32_t j) {
  if (g.vi == nullptr) return ItemGrid{g.gx, g.gy, 0, g.W, g.H};
  const int32_t *row = g.vi + 8 * (j / g.npv);
  return ItemGrid{row[2], row[3], row[4], row[0], row[1]};
}

// ---- tight binning: a wave works through the flattened (splat, tile)
// candidates of its 64 splats, 64 candidates at a time.  Lane l holds splat l
// (rectangle, test parameters) and the exclusive prefix P_l of the
// rectangle areas; candidate c belongs to the last lane with P <= c (found by
// a 6-step binary search over shuffles), whose parameters are then shuffled
// in.  Candidates are ordered by splat, then row-major over the rectangle.
struct TLane {
  TightSplat t;
  int x0, y0, w, area, W, H;
  float inv_w;
};

__device__ __forceinline__ TLane tight_lane(const Grid &g, const float *uv,
                                            const int32_t *radius, uint32_t j, bool valid) {
  TLane L;
  L.t = TightSplat{0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
  L.x0 = L.y0 = L.area = L.W = L.H = 0;
  L.w = 1;
  L.inv_w = 1.0f;
  if (!valid) return L;
  const int r = radius[j];
  if (r <= 0) return L;
  const ItemGrid ig = item_grid(g, j);
  const float u = uv[2 * (int64_t)j], v = uv[2 * (int64_t)j + 1];
  const Rect q = tile_rect(u, v, r, ig.gx, ig.gy, g.tw, g.th);
  const uint32_t oi = g.vi == nullptr ? j : j % g.npv;
  bool reach;
  L.t = tight_setup(u, v, g.conic[3 * (int64_t)j], g.conic[3 * (int64_t)j + 1],
                    g.conic[3 * (int64_t)j + 2], g.opac[oi], reach);
  L.x0 = q.x0;
  L.y0 = q.y0;
  L.w = max(q.x1 - q.x0, 1);
  L.inv_w = 1.0f / (float)L.w;
  L.area = reach ? (q.x1 - q.x0) * (q.y1 - q.y0) : 0;
  L.W = ig.W;
  L.H = ig.H;
  return L;
}

__device__ __forceinline__ float shfl_f(float x, int src) { return __shfl(x, src, 64); }
__device__ __forceinline__ int shfl_i(int x, int src) { return __shfl(x, src, 64); }

// Owner lane of candidate c: the last lane with P <= c (P nondecreasing).
__device__ __forceinline__ int tight_owner(int P, uint32_t c) {
  int lo = 0;
#pragma unroll
  for (int step = 32; step >= 1; step >>= 1) {
    const int mid = lo + step;
    if ((uint32_t)shfl_i(P, mid) <= c) lo = mid;
  }
  return lo;
}

// Tile of candidate k of a rectangle of width w: row = trunc((k + 0.5) / w)
// in float (exact for every k of a tile grid), column = the rest.
__device__ __forceinline__ void cand_tile(int k, int w, float inv_w, int x0, int y0, int &tx,
                                          int &ty) {
  const int row = (int)(((float)k + 0.5f) * inv_w);
  tx = x0 + (k - row * w);
  ty = y0 + row;
}

// The tile test of candidate (owner lane o, tile (tx, ty)); every lane takes
// part in the shuffles.
__device__ __forceinline__ bool owner_test(const TLane &L, int o, int tx, int ty, int tw,
                                           int th) {
  TightSplat t;
  t.u = shfl_f(L.t.u, o);
  t.v = shfl_f(L.t.v, o);
  t.a = shfl_f(L.t.a, o);
  t.b = shfl_f(L.t.b, o);
  t.c = shfl_f(L.t.c, o);
  t.rbc = shfl_f(L.t.rbc, o);
  t.rba = shfl_f(L.t.rba, o);
  t.tau = shfl_f(L.t.tau, o);
  const int W = shfl_i(L.W, o), H = shfl_i(L.H, o);
  const int px0 = tx * tw, py0 = ty * th;
  const int px1 = min(px0 + tw, W) - 1, py1 = min(py0 + th, H) - 1;
  return tile_reachable(t, (float)px0, (float)px1, (float)py0, (float)py1);
}

// exclusive prefix of v over the wave (all lanes active); *total = the
// wave's sum
__device__ __forceinline__ int wave_excl_scan(int v, int &total) {
  const uint32_t inc = wave_inclusive_scan((uint32_t)v);
  total = (int)wave_last(inc);
  return (int)inc - v;
}

// Per splat: tile count of its rectangle, and the rectangle packed as
// (x0 | y0 << 16, w | h << 16) for the emission (which then gathers 8 bytes
// per splat instead of its position and radius).
// Instance total of the block's splats -> bsum[blockIdx.x] (their sum is M,
// known before the depth sort).
__device__ __forceinline__ void store_block_sum(uint32_t c, uint32_t *__restrict__ bsum) {
  uint32_t total;
  block_exclusive_scan(c, total);
  if (threadIdx.x == 0) bsum[blockIdx.x] = total;
}

// Exact counts, 4096 splats per block (16 per thread, every load issued
// first): per splat its tile rectangle packed as (x0 | y0 << 16, w | h << 16)
// (the tile count is w h); per block the instance total (bsum, summed into M)
// and the number of splats with instances (vsum, the block sums of the
// compaction scan, which then needs no reduction kernel).
__global__ __launch_bounds__(kThreads) void count_kernel(const float *__restrict__ uv,
                                                         const int32_t *__restrict__ radius,
                                                         uint32_t n, Grid g,
                                                         uint2 *__restrict__ rects,
                                                         uint32_t *__restrict__ bsum,
                                                         uint32_t *__restrict__ vsum) {
  const uint32_t base = blockIdx.x * kTileItems;
  int r[kItems];
  float2 p[kItems];
#pragma unroll
  for (int k = 0; k < kItems; ++k) {
    const uint32_t i = base + k * kThreads + threadIdx.x;
    r[k] = i < n ? radius[i] : 0;
    p[k] = i < n ? ld_f2(uv + 2 * (int64_t)i) : make_float2(0.f, 0.f);
  }
  uint2 sums = make_uint2(0u, 0u);
#pragma unroll
  for (int k = 0; k < kItems; ++k) {
    const uint32_t i = base + k * kThreads + threadIdx.x;
    if (i < n) {
      uint32_t c = 0;
      uint2 packed = make_uint2(0u, 0u);
      if (r[k] > 0) {
        const ItemGrid ig = item_grid(g, i);
        const Rect q = tile_rect(p[k].x, p[k].y, r[k], ig.gx, ig.gy, g.tw, g.th);
        const uint32_t w = (uint32_t)(q.x1 - q.x0), h = (uint32_t)(q.y1 - q.y0);
        c = w * h;
        packed = make_uint2((uint32_t)q.x0 | ((uint32_t)q.y0 << 16), w | (h << 16));
      }
      rects[i] = packed;
      sums.x += c;
      sums.y += c > 0u ? 1u : 0u;
    }
  }
  uint2 tot;
  block_exclusive_scan_v(sums, tot);
  if (threadIdx.x == 0) {
    bsum[blockIdx.x] = tot.x;
    vsum[blockIdx.x] = tot.y;
  }
}

// Tight counts: per splat the number of rectangle tiles tile_reachable
// accepts (flattened per wave; lane j adds the passes that fall in its own
// candidate range of each 64-candidate chunk).
__global__ __launch_bounds__(kThreads) void count_tight_kernel(const float *__restrict__ uv,
                                                               const int32_t *__restrict__ radius,
                                                               uint32_t n, Grid g,
                                                               uint32_t *__restrict__ cnt,
                                                               uint64_t *__restrict__ masks,
                                                               uint32_t *__restrict__ bsum) {
  const uint32_t wave_base = (blockIdx.x * kThreads + threadIdx.x) & ~63u;
  uint32_t mine = 0;
  if (wave_base < n) {  // uniform per wave
    const int lane = lane_id();
    const uint32_t i = wave_base + lane;
    const TLane L = tight_lane(g, uv, radius, i, i < n);
    int total;
    const int P = wave_excl_scan(L.area, total);
    uint64_t mbits = 0;
    for (int c0 = 0; c0 < total; c0 += 64) {
      const uint32_t c = (uint32_t)(c0 + lane);
      const int o = tight_owner(P, c);
      const int Po = shfl_i(P, o), w = shfl_i(L.w, o), x0 = shfl_i(L.x0, o), y0 = shfl_i(L.y0, o);
      const float inv_w = shfl_f(L.inv_w, o);
      int tx, ty;
      cand_tile((int)c - Po, w, inv_w, x0, y0, tx, ty);
      const bool test = owner_test(L, o, tx, ty, g.tw, g.th);
      const uint64_t bal = __ballot(c < (uint32_t)total && test);
      const int lo = max(P, c0) - c0, hi = min(P + L.area, c0 + 64) - c0;
      if (hi > lo) {
        const uint64_t up = hi >= 64 ? ~0ull : ((1ull << hi) - 1ull);
        const uint64_t piece = (bal & up) >> lo;
        mine += (uint32_t)__popcll(piece);
        mbits |= piece << (max(P, c0) - P);  // only kept for areas <= 64
      }
    }
    if (i < n) {
      cnt[i] = mine;
      if (L.area > 0 && L.area <= 64) masks[i] = mbits;
    }
  }
  store_block_sum(mine, bsum);
}

struct LoadFlag {
  using value_type = uint32_t;
  const uint32_t *cnt;
  __device__ uint32_t operator()(uint32_t i) const { return cnt[i] > 0 ? 1u : 0u; }
};
// exact binning: a splat is listed when its packed rectangle has tiles
struct LoadRectFlag {
  using value_type = uint32_t;
  const uint2 *rects;
  __device__ uint32_t operator()(uint32_t i) const {
    const uint32_t wh = rects[i].y;
    return (wh & 0xFFFFu) * (wh >> 16) > 0u ? 1u : 0u;
  }
};
// Compaction of the splats that touch a tile: depth bits as keys, the splat
// index as values, and (exact binning) the packed rectangle as the depth
// sort's payload.
struct CompactPre {
  float depth;
  uint2 rect;
};
struct StoreCompact {
  using Pre = CompactPre;
  const float *depth;
  uint32_t *keys;
  int32_t *vals;
  const uint2 *rin;  // nullptr: no payload
  uint2 *rout;
  // depth and rectangle of every splat (visible ones share lines with the
  // others at 1.3 of 3 visible, so reading all costs about the same traffic)
  __device__ CompactPre pre(uint32_t i) const {
    return CompactPre{depth[i], rin != nullptr ? rin[i] : make_uint2(0u, 0u)};
  }
  __device__ void operator()(uint32_t i, uint32_t excl, uint32_t flag, CompactPre p) const {
    if (flag) {
      keys[excl] = __float_as_uint(p.depth);
      vals[excl] = (int32_t)i;
      if (rin != nullptr) rout[excl] = p.rect;
    }
  }
};

// Emission, wave-cooperative over instances: lane l of a wave loads
// depth-sorted splat (wave base + l) -- its tile rectangle as the count kernel
// packed it, one 8-byte gather -- and the wave then writes the instance range
// of its 64 splats 64 consecutive instances at a time (coalesced stores; a
// round serves as many small splats as start in it, where a round per splat
// would leave most lanes idle).  The owner of instance c0 + l is found without
// a search: the splat holding c0 is the last lane whose offset is <= c0, and
// each later owner starts at a chunk position that its lane flags in LDS, so
// the owner is that splat plus the number of flagged positions in [1, l].
// Row/column of instance k within the w-wide rectangle: q = trunc((k + 0.5)
// / w) in float is exact while k < 2^22 / w (k is never within 0.5/w of a
// multiple of w), far beyond the largest tile grid.
// The wave's 64 depth-sorted splats, one per lane (as the count kernel
// packed their rectangles), and the instance range [r0, r1) they own (r1
// truncated to the instance capacity: a caller that sized the instance
// buffers below the device count -- graph replays -- gets an in-bounds prefix).
struct EmitWave {
  int id, toff;
  uint32_t xy, wgx, off, end;
  float inv_w;
  uint32_t r0, r1;
};

__device__ __forceinline__ EmitWave emit_wave_load(const int32_t *__restrict__ sorted_ids,
                                                   const uint32_t *__restrict__ offsets,
                                                   const uint2 *__restrict__ rects,
                                                   uint32_t nv, const Grid &g,
                                                   uint32_t wave_base, uint32_t cap_inst) {
  EmitWave e{0, 0, 0u, 1u, 0xFFFFFFFFu, 0u, 1.0f, 0u, 0u};
  const uint32_t s = wave_base + (uint32_t)lane_id();
  if (s < nv) {
    e.id = sorted_ids[s];
    const ItemGrid ig = item_grid(g, (uint32_t)e.id);
    e.toff = ig.toff;
    const uint2 r = rects[e.id];
    e.xy = r.x;  // x0 | y0 << 16
    const uint32_t rw = r.y & 0xFFFFu, rh = r.y >> 16;
    const uint32_t w = rw > 1u ? rw : 1u;
    e.wgx = w | ((uint32_t)ig.gx << 16);  // the tile grid is < 65536 per side
    e.inv_w = 1.0f / (float)w;
    e.off = offsets[s];
    e.end = e.off + rw * rh;
  }
  const int nsplat = min(64, (int)(nv - wave_base));
  e.r0 = (uint32_t)__builtin_amdgcn_readlane((int)e.off, 0);
  e.r1 = min((uint32_t)__builtin_amdgcn_readlane((int)e.end, nsplat - 1), cap_inst);
  return e;
}

// Instances [cb, ce) of the wave's range, 64 consecutive ones per round:
// every lane of the wave calls f(q, tile, id, valid, single) once per round,
// q = c0 + lane, valid = q < ce; single = the whole round lies inside one
// splat (every lane valid, the 64 tiles distinct).
template <class F>
__device__ __forceinline__ void gen_rounds(const EmitWave &e, uint32_t cb, uint32_t ce,
                                           uint32_t *flag, F &f) {
  const int lane = lane_id();
  const uint64_t upto = lane == 63 ? ~0ull : ((2ull << lane) - 1ull);
  for (uint32_t c0 = cb; c0 < ce; c0 += 64) {
    const int j0 = 63 - __clzll(__ballot(e.off <= c0));  // splat holding instance c0
    const uint32_t jend = min((uint32_t)__builtin_amdgcn_readlane((int)e.end, j0), ce);
    if (jend >= c0 + 64u) {
      // whole rounds inside splat j0 (uniform): no owner search
      const uint32_t joff = (uint32_t)__builtin_amdgcn_readlane((int)e.off, j0);
      const uint32_t jxy = (uint32_t)__builtin_amdgcn_readlane((int)e.xy, j0);
      const uint32_t jwgx = (uint32_t)__builtin_amdgcn_readlane((int)e.wgx, j0);
      const int jtoff = __builtin_amdgcn_readlane(e.toff, j0);
      const int jid = __builtin_amdgcn_readlane(e.id, j0);
      const float jinv = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(e.inv_w), j0));
      const int w = (int)(jwgx & 0xFFFFu), gx = (int)(jwgx >> 16);
      const int x0 = (int)(jxy & 0xFFFFu), y0 = (int)(jxy >> 16);
      for (; c0 + 64u <= jend; c0 += 64) {
        const uint32_t q = c0 + (uint32_t)lane;
        const int k = (int)(q - joff);
        const int row = (int)(((float)k + 0.5f) * jinv);
        const int col = k - row * w;
        f(q, (uint32_t)(jtoff + (y0 + row) * gx + (x0 + col)), jid, true, true);
      }
      c0 -= 64;  // the loop increment moves to the first round not yet written
      continue;
    }
    flag[lane] = 0u;
    __builtin_amdgcn_wave_barrier();
    const uint32_t p = e.off - c0;  // round position of this splat's first instance
    if (p >= 1u && p < 64u) flag[p] = 1u;
    __builtin_amdgcn_wave_barrier();
    const uint64_t starts = __ballot(flag[lane] != 0u);
    __builtin_amdgcn_wave_barrier();
    const int o = j0 + (int)__popcll(starts & upto);
    const uint32_t ooff = (uint32_t)shfl_i((int)e.off, o);
    const uint32_t oxy = (uint32_t)shfl_i((int)e.xy, o), owgx = (uint32_t)shfl_i((int)e.wgx, o);
    const int otoff = shfl_i(e.toff, o), oid = shfl_i(e.id, o);
    const float oinv = shfl_f(e.inv_w, o);
    const uint32_t q = c0 + (uint32_t)lane;
    const bool valid = q < ce;
    const int k = valid ? (int)(q - ooff) : 0;
    const int w = (int)(owgx & 0xFFFFu), gx = (int)(owgx >> 16);
    const int row = (int)(((float)k + 0.5f) * oinv);
    const int col = k - row * w;
    const uint32_t tile =
        (uint32_t)(otoff + ((int)(oxy >> 16) + row) * gx + ((int)(oxy & 0xFFFFu) + col));
    f(q, valid ? tile : 0u, oid, valid, false);
  }
}

// The emission's writer: instance q -> (tile key, splat id) at q.
struct EmitStore {
  uint32_t *tkeys;
  int32_t *tvals;
  __device__ __forceinline__ void operator()(uint32_t q, uint32_t tile, int id, bool valid,
                                             bool) const {
    if (valid) {
      tkeys[q] = tile;
      tvals[q] = id;
    }
  }
};

__device__ __forceinline__ void emit_rounds(const EmitWave &e, uint32_t cb, uint32_t ce,
                                            uint32_t *flag, uint32_t *__restrict__ tkeys,
                                            int32_t *__restrict__ tvals) {
  EmitStore st{tkeys, tvals};
  gen_rounds(e, cb, ce, flag, st);
}

// A wave writes at most kEmitSpan instances of its range; the rest (the waves
// of large near splats) is queued as items of kEmitChunk instances that
// emit_rest_kernel spreads over the chip, one atomic per such wave.
constexpr uint32_t kEmitSpan = 8192;
constexpr uint32_t kEmitChunk = 4096;
// items[] bound: a wave queues items only beyond kEmitSpan instances, so at
// most M / kEmitSpan waves queue, each at most (its instances) / kEmitChunk + 1
inline uint64_t emit_items_cap(uint64_t m) { return m / kEmitChunk + m / kEmitSpan + 1; }

__global__ __launch_bounds__(kThreads) void emit_kernel(
    const int32_t *__restrict__ sorted_ids, const uint32_t *__restrict__ offsets,
    const uint2 *__restrict__ rects, const uint32_t *nv_ptr, uint32_t cap, Grid g,
    uint32_t cap_inst, uint32_t *__restrict__ tkeys, int32_t *__restrict__ tvals,
    uint2 *__restrict__ items, uint32_t *__restrict__ n_items, uint32_t items_cap) {
  __shared__ uint32_t s_flag[kThreads];
  const uint32_t nv = count_of(nv_ptr, cap);
  const uint32_t wave_base = (blockIdx.x * kThreads + threadIdx.x) & ~63u;
  if (wave_base >= nv) return;  // uniform per wave
  const EmitWave e = emit_wave_load(sorted_ids, offsets, rects, nv, g, wave_base, cap_inst);
  // r1 < r0 when the capacity truncates before this wave's range: nothing to do
  const uint32_t len = e.r1 > e.r0 ? e.r1 - e.r0 : 0u;
  const uint32_t ce = e.r0 + min(len, kEmitSpan);
  if (ce < e.r1) {  // uniform: queue the rest
    const uint32_t nit = (e.r1 - ce + kEmitChunk - 1) / kEmitChunk;
    uint32_t base = 0;
    if (lane_id() == 0) base = atomicAdd(n_items, nit);
    base = (uint32_t)__builtin_amdgcn_readfirstlane((int)base);
    for (uint32_t t = (uint32_t)lane_id(); t < nit; t += 64)
      if (base + t < items_cap) items[base + t] = make_uint2(wave_base, ce + t * kEmitChunk);
  }
  emit_rounds(e, e.r0, ce, s_flag + (threadIdx.x & ~63u), tkeys, tvals);
}

// The queued rest: wave v takes item v (a wave's 64 splats and the first
// instance of a kEmitChunk range of theirs).
__global__ __launch_bounds__(kThreads) void emit_rest_kernel(
    const int32_t *__restrict__ sorted_ids, const uint32_t *__restrict__ offsets,
    const uint2 *__restrict__ rects, const uint32_t *nv_ptr, uint32_t cap, Grid g,
    uint32_t cap_inst, const uint2 *__restrict__ items, const uint32_t *__restrict__ n_items,
    uint32_t items_cap, uint32_t *__restrict__ tkeys, int32_t *__restrict__ tvals) {
  __shared__ uint32_t s_flag[kThreads];
  const uint32_t v = (blockIdx.x * kThreads + threadIdx.x) >> 6;
  if (v >= min(*n_items, items_cap)) return;  // uniform per wave
  const uint2 item = items[v];
  const uint32_t nv = count_of(nv_ptr, cap);
  const EmitWave e = emit_wave_load(sorted_ids, offsets, rects, nv, g, item.x, cap_inst);
  const uint32_t cb = item.y, ce = min(item.y + kEmitChunk, e.r1);
  emit_rounds(e, cb, ce, s_flag + (threadIdx.x & ~63u), tkeys, tvals);
}

// ---- the first tile pass generates its items ---------------------------
// The first LSD pass over the tile bits does not read an emitted instance
// list: each 4096-instance radix block is generated where it is used, from
// the depth-ordered splats (id, instance offset and packed rectangle, the
// rectangles put in depth order by the depth sort's last pass), exactly as
// the emission would have written it.  A block's wave w owns instances
// [base + 1024 w, base + 1024 (w + 1)), i.e. the rows the radix kernels give
// it, and starts at the splat holding its first instance (a 64-way search
// over the offsets in the histogram kernel, kept for the scatter).  This
// removes the emitted list's write and its re-read by the first pass
// (8 + 12 bytes per instance).
struct SplatStream {
  static constexpr bool kOneRow = false;
  static constexpr bool kLimit = false;
  const int32_t *ids;   // depth-sorted splat ids
  const uint32_t *offs; // instance offset of each (strictly increasing)
  const uint2 *rects;   // packed rectangle of each, in depth order
};

// Group of 64 depth-sorted splats [sb, sb + 64) as emit_wave_load gives it,
// the rectangle read in depth order (no gather through the id); a group at
// or past nv (a prefetch beyond the last) comes back empty.
__device__ __forceinline__ EmitWave stream_load(const SplatStream &st, uint32_t nv, const Grid &g,
                                                uint32_t sb) {
  EmitWave e{0, 0, 0u, 1u, 0xFFFFFFFFu, 0u, 1.0f, 0u, 0u};
  if (sb >= nv) return e;  // uniform per wave
  const uint32_t s = sb + (uint32_t)lane_id();
  if (s < nv) {
    e.id = st.ids[s];
    const uint2 r = st.rects[s];
    e.off = st.offs[s];
    const ItemGrid ig = item_grid(g, (uint32_t)e.id);
    e.toff = ig.toff;
    e.xy = r.x;
    const uint32_t rw = r.y & 0xFFFFu, rh = r.y >> 16;
    const uint32_t w = rw > 1u ? rw : 1u;
    e.wgx = w | ((uint32_t)ig.gx << 16);
    e.inv_w = 1.0f / (float)w;
    e.end = e.off + rw * rh;
  }
  const int nsplat = min(64, (int)(nv - sb));
  e.r0 = (uint32_t)__builtin_amdgcn_readlane((int)e.off, 0);
  e.r1 = (uint32_t)__builtin_amdgcn_readlane((int)e.end, nsplat - 1);
  return e;
}

// ---- row binning: segments (one per splat and rectangle row) ---------------
// A grid of at most 256 x 256 tiles is binned in two generated passes instead
// of two passes over the instances.  Pass A sorts the segments -- one per
// rectangle row of each depth-ordered splat, key ty | x0 << 8 | w << 16, value
// the splat id -- stably by the row ty; pass B expands every segment into its
// w tiles and sorts them stably by the column tx, each radix block inside one
// row, so the row-major tile order and the depth order inside a tile follow.
// Pass A moves S ~ M / 4 items and also writes each segment's first instance
// E (its histogram counts the widths per row beside the segments), and pass B
// writes the final ids and the tile ranges: no pass reads or writes the
// instance keys.
//
// Both passes reuse the generator of the first tile pass: an element with n
// items is a one-row rectangle of width n (inv_w = 1 / n, row 0), and the
// generated "tile" of item k is toff + x0 + k (grid width 0).
//
// Pass A's elements are the depth-ordered splats: item k = rectangle row
// y0 + k, key toff + y0 + k with toff = x0 << 8 | w << 16 (the row keeps the
// low 8 bits).
struct RowStream {
  static constexpr bool kOneRow = true;
  static constexpr bool kLimit = true;
  const int32_t *ids;    // depth-sorted splat ids
  const uint32_t *offs;  // segment offset of each (strictly increasing)
  const uint2 *rects;    // packed rectangle of each, in depth order
  // pass A's histogram computes the segment limit (seg_limit) from these
  const uint32_t *ioffs;  // instance offset of each splat
  uint32_t *sc;           // scalars: [0] Nv, [1] M, [6] S in; [2..5] written
  uint32_t m;             // instance capacity
  const uint32_t *own;    // own[k]: the splat holding segment 1024 k
};

__device__ __forceinline__ EmitWave stream_load(const RowStream &st, uint32_t nv, const Grid &,
                                                uint32_t sb) {
  EmitWave e{0, 0, 0u, 1u, 0xFFFFFFFFu, 0u, 1.0f, 0u, 0u};
  if (sb >= nv) return e;  // uniform per wave
  const uint32_t s = sb + (uint32_t)lane_id();
  if (s < nv) {
    e.id = st.ids[s];
    const uint2 r = st.rects[s];
    e.off = st.offs[s];
    const uint32_t x0 = r.x & 0xFFFFu, y0 = r.x >> 16;
    const uint32_t w = r.y & 0xFFFFu, h = r.y >> 16;
    const uint32_t hh = h > 1u ? h : 1u;
    e.toff = (int)((x0 << 8) | (w << 16));
    e.xy = y0;
    e.wgx = hh;  // grid width 0
    e.inv_w = 1.0f / (float)hh;
    e.end = e.off + h;
  }
  const int nsplat = min(64, (int)(nv - sb));
  e.r0 = (uint32_t)__builtin_amdgcn_readlane((int)e.off, 0);
  e.r1 = (uint32_t)__builtin_amdgcn_readlane((int)e.end, nsplat - 1);
  return e;
}

// Pass B's elements are pass A's output in order: segment s (key, splat id)
// starting at instance E[s] (the exclusive prefix of the widths in that
// order, written by pass A's scatter); item k = column x0 + k.
struct SegStream {
  static constexpr bool kOneRow = true;
  static constexpr bool kLimit = false;
  const int32_t *ids;     // splat id of each segment
  const uint32_t *offs;   // E: first instance of each (strictly increasing)
  const uint32_t *keys;   // ty | x0 << 8 | w << 16
  // pass B's histogram builds the row table (row_table) from these
  const uint32_t *seg_per_row;  // pass A's digit totals
  const uint32_t *sc;           // scalars: [2] segments kept, [5] instances kept
  int gx, gy;
  int32_t *tile_range;          // the tiles of rows without instances get (0, 0)
  uint32_t *nblocks;            // receives the number of row blocks
  const uint32_t *own;          // own[k]: the segment holding instance 1024 k
};

__device__ __forceinline__ EmitWave stream_load(const SegStream &st, uint32_t nv, const Grid &,
                                                uint32_t sb) {
  EmitWave e{0, 0, 0u, 1u, 0xFFFFFFFFu, 0u, 1.0f, 0u, 0u};
  if (sb >= nv) return e;  // uniform per wave
  const uint32_t s = sb + (uint32_t)lane_id();
  if (s < nv) {
    e.id = st.ids[s];
    const uint32_t k = st.keys[s];
    e.off = st.offs[s];
    const uint32_t w = k >> 16;
    const uint32_t ww = w > 1u ? w : 1u;
    e.xy = (k >> 8) & 0xFFu;
    e.wgx = ww;  // grid width 0
    e.inv_w = 1.0f / (float)ww;
    e.end = e.off + w;
  }
  const int nsplat = min(64, (int)(nv - sb));
  e.r0 = (uint32_t)__builtin_amdgcn_readlane((int)e.off, 0);
  e.r1 = (uint32_t)__builtin_amdgcn_readlane((int)e.end, nsplat - 1);
  return e;
}

// Depth-order index of the splat holding instance q: the last s < nv with
// offs[s] <= q (offsets strictly increase: listed splats own >= 1 instance);
// a 64-way search, one probe per lane per step.
// The same search over [lo, hi), given offs[lo] <= q.
__device__ __forceinline__ uint32_t find_owner_in(const uint32_t *__restrict__ offs, uint32_t lo,
                                                  uint32_t hi, uint32_t q) {
  const uint32_t lane = (uint32_t)lane_id();
  while (hi - lo > 1u) {
    const uint32_t step = (hi - lo + 63u) / 64u;
    const uint32_t idx = lo + lane * step;
    const uint64_t b = __ballot(idx < hi && offs[idx] <= q);
    lo += (uint32_t)(63 - __clzll(b)) * step;  // b != 0: offs[lo] <= q
    hi = min(lo + step, hi);
  }
  return lo;
}
__device__ __forceinline__ uint32_t find_owner(const uint32_t *__restrict__ offs, uint32_t nv,
                                               uint32_t q) {
  return find_owner_in(offs, 0u, nv, q);
}

// Row binning's owner tables: own[k] = the element holding item 1024 k (an
// element holds fewer than 1024 items, so each writes at most one entry).
// Pass A's waves start at multiples of 1024 and read their first element
// there; pass B's waves start anywhere and search between two entries.
__device__ __forceinline__ uint32_t stream_owner(const SplatStream &st, uint32_t nv, uint32_t q) {
  return find_owner(st.offs, nv, q);
}
__device__ __forceinline__ uint32_t stream_owner(const RowStream &st, uint32_t, uint32_t q) {
  return st.own[q >> 10];  // q is a multiple of 1024
}
__device__ __forceinline__ uint32_t stream_owner(const SegStream &st, uint32_t nv, uint32_t q) {
  const uint32_t k = q >> 10;
  const uint32_t lo = st.own[k];
  const uint32_t hi = ((k + 1u) << 10) < st.sc[5] ? st.own[k + 1u] + 1u : nv;
  return find_owner_in(st.offs, lo, hi, q);
}

// Instances [q0, q1) in order, from the splat s0 that holds q0, over
// consecutive groups of 64 splats; the next group is requested before the
// current one is walked.
// gen_rounds for streams of one-row elements (RowStream, SegStream): the key
// of item q of element j is tb_j + q with tb_j = toff + x0 - off, so a round
// shuffles two values from each item's owner (tb, id) and computes no row.
template <class F>
__device__ __forceinline__ void gen_rounds1(const EmitWave &e, uint32_t cb, uint32_t ce,
                                            uint32_t *flag, F &f) {
  const int lane = lane_id();
  const uint64_t upto = lane == 63 ? ~0ull : ((2ull << lane) - 1ull);
  const uint32_t tb = (uint32_t)e.toff + (e.xy & 0xFFFFu) - e.off;
  for (uint32_t c0 = cb; c0 < ce; c0 += 64) {
    const int j0 = 63 - __clzll(__ballot(e.off <= c0));  // element holding item c0
    const uint32_t jend = min((uint32_t)__builtin_amdgcn_readlane((int)e.end, j0), ce);
    if (jend >= c0 + 64u) {
      // whole rounds inside element j0 (uniform): no owner search
      const uint32_t jtb = (uint32_t)__builtin_amdgcn_readlane((int)tb, j0);
      const int jid = __builtin_amdgcn_readlane(e.id, j0);
      for (; c0 + 64u <= jend; c0 += 64) {
        const uint32_t q = c0 + (uint32_t)lane;
        f(q, jtb + q, jid, true, true);
      }
      c0 -= 64;  // the loop increment moves to the first round not yet written
      continue;
    }
    flag[lane] = 0u;
    __builtin_amdgcn_wave_barrier();
    const uint32_t p = e.off - c0;  // round position of this element's first item
    if (p >= 1u && p < 64u) flag[p] = 1u;
    __builtin_amdgcn_wave_barrier();
    const uint64_t starts = __ballot(flag[lane] != 0u);
    __builtin_amdgcn_wave_barrier();
    const int o = j0 + (int)__popcll(starts & upto);
    const uint32_t otb = (uint32_t)shfl_i((int)tb, o);
    const int oid = shfl_i(e.id, o);
    const uint32_t q = c0 + (uint32_t)lane;
    const bool valid = q < ce;
    f(q, valid ? otb + q : 0u, oid, valid, false);
  }
}

template <class Stream, class F>
__device__ __forceinline__ void gen_range(const Stream &st, uint32_t nv, const Grid &g,
                                          uint32_t s0, uint32_t q0, uint32_t q1, uint32_t *flag,
                                          F &f) {
  EmitWave cur = stream_load(st, nv, g, s0);
  uint32_t c0 = q0;
  for (uint32_t sb = s0; c0 < q1; sb += 64) {
    const EmitWave nxt = stream_load(st, nv, g, sb + 64);
    const uint32_t ce = min(q1, cur.r1);
    if constexpr (Stream::kOneRow)
      gen_rounds1(cur, c0, ce, flag, f);
    else
      gen_rounds(cur, c0, ce, flag, f);
    c0 = ce;
    cur = nxt;
  }
}

// kFix: item qfix gets the width wfix in its key (pass A: the segment that an
// instance capacity below the count cuts short).  kPacked: one word per item,
// id | tile << 24 (pass B: tile = the column, ids below 2^24), no values.
template <bool kFix = false, bool kPacked = false>
struct GenStage {
  uint32_t *key;
  int32_t *val;
  uint32_t base;
  uint32_t qfix, wfix;
  __device__ __forceinline__ void operator()(uint32_t q, uint32_t tile, int id, bool valid,
                                             bool) const {
    if (valid) {
      if constexpr (kFix) {
        if (q == qfix) tile = (tile & 0xFFFFu) | (wfix << 16);
      }
      if constexpr (kPacked) {
        key[q - base] = (uint32_t)id | (tile << 24);
      } else {
        key[q - base] = tile;
        val[q - base] = id;
      }
    }
  }
};

// Segments of the first Mc = min(M, m) instances of the depth-first sequence
// (an instance capacity m below the count M keeps that prefix), from sc[0] =
// Nv, sc[1] = M, sc[6] = S (all segments): {segments kept, the last kept one
// if the capacity cuts it short (else 0xFFFFFFFF), its width, Mc}.  Called by
// a whole wave (the search is wave-wide).
__device__ __forceinline__ uint4 seg_limit(const uint32_t *__restrict__ offs,
                                           const uint32_t *__restrict__ soffs,
                                           const uint2 *__restrict__ rects, uint32_t cap_n,
                                           uint32_t m, const uint32_t *sc) {
  const uint32_t nv = min(sc[0], cap_n), M = sc[1];
  const uint32_t mc = min(M, m);
  uint4 r = make_uint4(sc[6], 0xFFFFFFFFu, 0u, mc);
  if (M > m) {  // uniform
    r.x = 0u;
    if (mc > 0u && nv > 0u) {
      const uint32_t q = mc - 1u;  // the last kept instance
      const uint32_t s = find_owner(offs, nv, q);
      const uint32_t w = max(rects[s].y & 0xFFFFu, 1u);
      const uint32_t k = q - offs[s];
      const uint32_t row = k / w;
      r.x = soffs[s] + row + 1u;
      r.y = r.x - 1u;
      r.z = k - row * w + 1u;
    }
  }
  return r;
}

// Pass B's row table, one row per thread of the block: row ty's segments are
// [seg0, seg1) of pass A's output (prefix of the digit totals), its instances
// [E[seg0], E[seg1]) (Mc past the last segment), cut into ceil(count / 4096)
// blocks numbered row after row from blk0.  -> {e0, e1, blk0, block count};
// nblocks = the blocks of all rows.
__device__ __forceinline__ uint4 row_table(const SegStream &st, uint32_t &nblocks) {
  const int ty = threadIdx.x;
  const uint32_t seff = st.sc[2], mc = st.sc[5];
  const uint32_t cnt = ty < st.gy ? st.seg_per_row[ty] : 0u;
  uint32_t total;
  const uint32_t seg0 = block_exclusive_scan(cnt, total);
  const uint32_t seg1 = seg0 + cnt;
  const uint32_t e0 = seg0 < seff ? st.offs[seg0] : mc;
  const uint32_t e1 = seg1 < seff ? st.offs[seg1] : mc;
  const uint32_t nbr = (e1 - e0 + kTileItems - 1) / kTileItems;
  const uint32_t blk0 = block_exclusive_scan(nbr, nblocks);
  return make_uint4(e0, e1, blk0, nbr);
}

// Histogram of the first tile pass (digit = tile & mask, D = mask + 1 a power
// of two <= 256) per radix block, as radix_hist_kernel would count the emitted
// list, without generating the instances: a splat's instances in one row of
// its rectangle are the consecutive tiles t0 .. t0 + len - 1, whose digits are
// len / D full cycles plus a run of len % D digits from t0 & mask, wrapping.
// Each lane takes one splat of the wave's current 64 and adds, for each of its
// rectangle rows inside the wave's instance range (the first and last rows may
// be partial), the full cycles to one counter and the run to a difference
// array over the digits (+1 at its start, -1 past its end, split at the wrap);
// a block scan over the digits then gives the counts.  first[4 b + w] = the
// stream index of the element holding wave w's first item.
// The same kernel counts row binning's passes (Stream = RowStream, SegStream).
// Pass A (Stream::kLimit): the items are the segments kept (seg_limit, which
// every block computes -- two reads unless the capacity cuts -- and block 0
// stores in sc[2..5] for the later kernels).
// kRows: the blocks are pass B's, each inside one tile row: block b is the
// row's (b - blk0)-th run of 4096 instances (row_table, which every block
// builds); binfo[b] = {first, end, blk0, ty | count << 8} ({0, 0, 0, 0}: no
// row, an empty block).  The last block, never a row's, clears the ranges of
// the rows without instances.
template <class Stream, bool kRows = false>
__global__ __launch_bounds__(kThreads) void gen_hist_kernel(
    Stream st, const uint32_t *nv_ptr, uint32_t cap_n, const uint32_t *m_ptr,
    uint32_t cap_m, Grid g, uint32_t mask, uint32_t *__restrict__ hist, uint32_t nb,
    uint32_t *__restrict__ first, uint4 *__restrict__ binfo) {
  constexpr bool kW = Stream::kLimit;  // pass A: also the widths per row
  __shared__ int32_t s_d[kRadix + 1];
  __shared__ int32_t s_dw[kW ? kRadix + 1 : 1];
  __shared__ uint32_t s_full, s_wfull;
  __shared__ uint4 s_rng;
  s_d[threadIdx.x] = 0;
  if constexpr (kW) s_dw[threadIdx.x] = 0;
  if (threadIdx.x == 0) {
    s_d[kRadix] = 0;
    if constexpr (kW) s_dw[kRadix] = 0;
    s_full = 0;
    s_wfull = 0;
    s_rng = make_uint4(0u, 0u, 0u, 0u);
  }
  __syncthreads();
  uint32_t base, end, lb, rows_end = 0;
  if constexpr (Stream::kLimit) {
    if (threadIdx.x < 64) {  // wave 0
      const uint4 lim = seg_limit(st.ioffs, st.offs, st.rects, cap_n, st.m, st.sc);
      if (threadIdx.x == 0) {
        s_rng = lim;
        if (blockIdx.x == 0) {
          st.sc[2] = lim.x;
          st.sc[3] = lim.y;
          st.sc[4] = lim.z;
          st.sc[5] = lim.w;
        }
      }
    }
    __syncthreads();
    end = s_rng.x;
    lb = logical_block(blocks_of(end));
    base = lb * kTileItems;
  } else if constexpr (kRows) {
    uint32_t nrb;
    const uint4 r = row_table(st, nrb);
    lb = logical_block(nrb);
    rows_end = nrb;
    if (blockIdx.x == 0 && threadIdx.x == 0) st.nblocks[0] = nrb;
    const uint32_t b = lb;
    if (r.w > 0u && b >= r.z && b - r.z < r.w) {  // one row matches at most
      const uint32_t lo = r.x + (b - r.z) * kTileItems;
      s_rng = make_uint4(lo, min(lo + kTileItems, r.y), r.z, threadIdx.x | (r.w << 8));
    }
    if (blockIdx.x == gridDim.x - 1 && (int)threadIdx.x < st.gy && r.w == 0u)
      for (int x = 0; x < st.gx; ++x) {
        const int64_t t = (int64_t)threadIdx.x * st.gx + x;
        st.tile_range[2 * t] = 0;
        st.tile_range[2 * t + 1] = 0;
      }
    __syncthreads();
    const uint4 bi = s_rng;
    if (threadIdx.x == 0) binfo[b] = bi;
    base = bi.x;
    end = bi.y;
  } else {
    end = count_of(m_ptr, cap_m);
    lb = logical_block(blocks_of(end));
    base = lb * kTileItems;
  }
  const uint32_t nv = count_of(nv_ptr, cap_n);
  const int w = wave_index();
  const uint32_t D = mask + 1u;
  const uint32_t dbits = (uint32_t)__popc(mask);
  constexpr uint32_t kWaveItems = kTileItems / (kThreads / 64);
  const uint32_t q0 = base + (uint32_t)w * kWaveItems;
  if (q0 < end) {  // uniform per wave
    const uint32_t q1 = min(q0 + kWaveItems, end);
    const uint32_t s0 = stream_owner(st, nv, q0);
    if (lane_id() == 0) first[lb * (kThreads / 64) + w] = s0;
    EmitWave cur = stream_load(st, nv, g, s0);
    for (uint32_t sb = s0;; sb += 64) {
      // the next group is requested before this one is counted
      const EmitWave nxt = stream_load(st, nv, g, sb + 64);
      // this lane's splat: its instances inside [q0, q1), relative to its
      // first instance, row by row of its rectangle
      const uint32_t a = max(cur.off, q0), b = min(cur.end, q1);
      if constexpr (Stream::kOneRow) {
        // one row (segments, or a splat's rows): items a .. b - 1 are the
        // digits d0 .. d1 - 1 with d1 <= the grid side <= D (no full cycle,
        // no wrap; a run of all D digits ends at the extra slot s_d[D])
        if (cur.off != 0xFFFFFFFFu && a < b) {
          const uint32_t d0 = (cur.xy & 0xFFFFu) + (a - cur.off), d1 = d0 + (b - a);
          atomicAdd(&s_d[d0], 1);
          atomicAdd(&s_d[d1], -1);
          if constexpr (kW) {
            const int32_t wt = (int32_t)((uint32_t)cur.toff >> 16);
            atomicAdd(&s_dw[d0], wt);
            atomicAdd(&s_dw[d1], -wt);
            const uint32_t qf = s_rng.y;
            if (qf >= a && qf < b) {
              const uint32_t dq = d0 + (qf - a);
              atomicAdd(&s_dw[dq], (int32_t)s_rng.z - wt);
              atomicAdd(&s_dw[dq + 1], wt - (int32_t)s_rng.z);
            }
          }
        }
      } else if (cur.off != 0xFFFFFFFFu && a < b) {
        const int rw = (int)(cur.wgx & 0xFFFFu), gx = (int)(cur.wgx >> 16);
        const int x0 = (int)(cur.xy & 0xFFFFu), y0 = (int)(cur.xy >> 16);
        const int k0 = (int)(a - cur.off), k1 = (int)(b - cur.off) - 1;
        // row of instance k: trunc((k + 0.5) / w) in float, exact here
        const int r0 = (int)(((float)k0 + 0.5f) * cur.inv_w);
        const int r1 = (int)(((float)k1 + 0.5f) * cur.inv_w);
        const int c0 = k0 - r0 * rw, c1 = k1 - r1 * rw;
        for (int r = r0; r <= r1; ++r) {
          const int cs = r == r0 ? c0 : 0, ce = r == r1 ? c1 : rw - 1;
          const uint32_t len = (uint32_t)(ce - cs + 1);
          const uint32_t t0 = (uint32_t)(cur.toff + (y0 + r) * gx + x0 + cs);
          const uint32_t full = len >> dbits, rem = len & mask;
          if (full) atomicAdd(&s_full, full);
          if (rem) {
            const uint32_t d0 = t0 & mask, d1 = d0 + rem;
            atomicAdd(&s_d[d0], 1);
            if (d1 <= D) {
              atomicAdd(&s_d[d1], -1);
            } else {  // wraps past the last digit
              atomicAdd(&s_d[0], 1);
              atomicAdd(&s_d[d1 - D], -1);
            }
          }
          if constexpr (kW) {
            // the same run weighted by the splat's tile width (rows: no wrap,
            // a splat covers at most the D >= gy rows once)
            const int32_t wt = (int32_t)((uint32_t)cur.toff >> 16);
            if (full) atomicAdd(&s_wfull, (uint32_t)wt);
            if (rem) {
              const uint32_t d0 = t0 & mask;
              atomicAdd(&s_dw[d0], wt);
              atomicAdd(&s_dw[d0 + rem], -wt);
            }
            // the segment a capacity cuts short has width wfix, not wt
            const uint32_t qf = s_rng.y;
            if (qf >= a && qf < b) {
              const uint32_t dq = (t0 + (qf - cur.off - (uint32_t)k0)) & mask;
              atomicAdd(&s_dw[dq], (int32_t)s_rng.z - wt);
              atomicAdd(&s_dw[dq + 1], wt - (int32_t)s_rng.z);
            }
          }
        }
      }
      if (cur.r1 >= q1 || sb + 64 >= nv) break;  // uniform
      cur = nxt;
    }
  }
  __syncthreads();
  // blocks without items write no column, except in row mode the one right
  // after the last row block (zeros: it closes the row counts)
  if (kRows ? base >= end && lb != rows_end : base >= end) return;  // uniform
  // counts = the cycles every digit got + the prefix of the difference array
  const int32_t v = threadIdx.x < D ? s_d[threadIdx.x] : 0;
  uint32_t total;
  const uint32_t excl = block_exclusive_scan((uint32_t)v, total);
  hist[threadIdx.x * nb + lb] = threadIdx.x < D ? s_full + excl + (uint32_t)v : 0u;
  if constexpr (kW) {
    // rows D .. 2D - 1: the widths (instances) per row
    const int32_t vw = threadIdx.x < D ? s_dw[threadIdx.x] : 0;
    const uint32_t ew = block_exclusive_scan((uint32_t)vw, total);
    if (threadIdx.x < D) hist[(D + threadIdx.x) * nb + lb] = s_wfull + ew + (uint32_t)vw;
  }
}

// Items of a radix block generated from a stream (the first tile pass: the
// depth-ordered splats; row binning: RowStream, SegStream): the wave writes
// its 1024 items (key, id) into its quarter of the block's LDS staging
// arrays, then reads its rows back.  kFix: fix[0], fix[1] = the item and
// width GenStage corrects (pass A).
template <class Stream, bool kFix = false, bool kPacked = false>
struct GenItems {
  static constexpr bool kPackedItems = kPacked;
  Stream st;
  const uint32_t *nv_ptr;
  uint32_t cap_n;
  Grid g;
  const uint32_t *first;
  const uint32_t *fix;
  template <int kIt>
  __device__ __forceinline__ void load(uint32_t block, uint32_t base, uint32_t wbase, uint32_t n,
                                       uint32_t (&key)[kIt], int32_t (&val)[kIt],
                                       uint32_t *s_key, int32_t *s_val, uint32_t *s_flag) const {
    static_assert(kIt == kItems, "generated blocks hold 4096 items");
    const int w = wave_index(), lane = lane_id();
    constexpr uint32_t kWaveItems = kTileItems / (kThreads / 64);
    if (wbase < n) {  // uniform per wave
      GenStage<kFix, kPacked> f{s_key, s_val, base, 0xFFFFFFFFu, 0u};
      if constexpr (kFix) {
        f.qfix = fix[0];
        f.wfix = fix[1];
      }
      gen_range(st, count_of(nv_ptr, cap_n), g, first[block * (kThreads / 64) + w], wbase,
                min(wbase + kWaveItems, n), s_flag + 64 * w, f);
    }
    // the rows were written by this wave (any lane): LDS is in order per wave
    __builtin_amdgcn_wave_barrier();
#pragma unroll
    for (int r = 0; r < kItems; ++r) {
      const uint32_t i = wbase + r * 64 + lane;
      const bool valid = i < n;
      key[r] = valid ? s_key[i - base] : 0u;
      if constexpr (!kPacked) val[r] = valid ? s_val[i - base] : 0;
    }
  }
};

// Tight emission in splat index order, as the tight count walked them: the
// depth order would put the large near splats, with a reachability test per
// candidate tile, into the same few waves.  A wave's 64 splats' candidates
// are flattened 64 at a time exactly as in count_tight_kernel; a candidate
// that passes goes to its splat's depth-order offset (offsets[rank[i]]) plus
// the passes of that splat before it (earlier rounds: the splat lane's running
// count; this round: the ballot from the splat's first lane in the round).
// count_tight accepted exactly the same candidates, so every splat fills
// [offset, offset + count).
__global__ __launch_bounds__(kThreads) void emit_tight_kernel(
    const float *__restrict__ uv, const int32_t *__restrict__ radius,
    const uint32_t *__restrict__ cnt, const uint32_t *__restrict__ rank,
    const uint32_t *__restrict__ offsets, const uint64_t *__restrict__ masks, uint32_t n, Grid g,
    uint32_t cap_inst, uint32_t *__restrict__ tkeys, int32_t *__restrict__ tvals) {
  const uint32_t wave_base = (blockIdx.x * kThreads + threadIdx.x) & ~63u;
  if (wave_base >= n) return;  // uniform per wave
  const int lane = lane_id();
  const uint64_t lt = lanemask_lt();
  const uint32_t i = wave_base + lane;
  const bool valid = i < n;
  const TLane L = tight_lane(g, uv, radius, i, valid);
  int gxi = 0, toff = 0;
  uint64_t mine = 0;
  uint32_t base = 0;
  if (valid && L.area > 0) {
    const ItemGrid ig = item_grid(g, i);
    gxi = ig.gx;
    toff = ig.toff;
    if (L.area <= 64) mine = masks[i];
    if (cnt[i] > 0) base = offsets[rank[i]];  // splats with no accepted tile are unlisted
  }
  const int mlo = (int)(uint32_t)mine, mhi = (int)(uint32_t)(mine >> 32);
  int total;
  const int P = wave_excl_scan(L.area, total);
  uint32_t done = 0;  // this lane's splat: accepted candidates so far
  for (int c0 = 0; c0 < total; c0 += 64) {
    const uint32_t c = (uint32_t)(c0 + lane);
    const bool in = c < (uint32_t)total;
    const int o = tight_owner(P, c);
    const int Po = shfl_i(P, o), w = shfl_i(L.w, o), x0 = shfl_i(L.x0, o), y0 = shfl_i(L.y0, o);
    const int area_o = shfl_i(L.area, o);
    const float inv_w = shfl_f(L.inv_w, o);
    const uint64_t mo = (uint64_t)(uint32_t)shfl_i(mlo, o) |
                        ((uint64_t)(uint32_t)shfl_i(mhi, o) << 32);
    int tx, ty;
    const int k = (int)c - Po;
    cand_tile(k, w, inv_w, x0, y0, tx, ty);
    // areas <= 64: the count kernel's recorded decisions; larger: the test
    bool pass = in && area_o <= 64 && ((mo >> (k & 63)) & 1ull);
    if (__ballot(in && area_o > 64) != 0ull) {  // uniform
      const bool test = owner_test(L, o, tx, ty, g.tw, g.th);
      pass = pass || (in && area_o > 64 && test);
    }
    const int ogx = shfl_i(gxi, o), otoff = shfl_i(toff, o);
    const uint32_t obase = (uint32_t)shfl_i((int)(base + done), o);
    const uint64_t bal = __ballot(pass);
    if (pass) {
      const int first = max(Po - c0, 0);  // the owner's first lane in this round
      const uint32_t pos = obase + (uint32_t)__popcll(bal & lt & (~0ull << first));
      if (pos < cap_inst) {  // capacity guard as in emit_kernel
        tkeys[pos] = (uint32_t)(otoff + ty * ogx + tx);
        tvals[pos] = (int32_t)(wave_base + (uint32_t)o);
      }
    }
    const int lo = max(P, c0) - c0, hi = min(P + L.area, c0 + 64) - c0;
    if (hi > lo) {
      const uint64_t up = hi >= 64 ? ~0ull : ((1ull << hi) - 1ull);
      done += (uint32_t)__popcll((bal & up) >> lo);
    }
  }
}

__global__ __launch_bounds__(kThreads) void fill_kernel(int32_t *__restrict__ p, int64_t n,
                                                        int32_t v) {
  const int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x;
  if (i < n) p[i] = v;
}

// Zero-fill of p[0, n) and of q[0, nq) in one launch.
__global__ __launch_bounds__(kThreads) void zero2_kernel(int32_t *__restrict__ p, int64_t n,
                                                         uint32_t *__restrict__ q, int64_t nq) {
  const int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x;
  if (i < n)
    p[i] = 0;
  else if (i - n < nq)
    q[i - n] = 0u;
}

// tile_range[t] = [first, last + 1) of tile t in the sorted keys: a key that
// differs from its predecessor starts a range, one that differs from its
// successor ends one.  Four consecutive keys per thread (16-byte loads; the
// key arrays are 256-byte aligned).
__global__ __launch_bounds__(kThreads) void ranges_kernel(const uint32_t *__restrict__ tkeys,
                                                           const uint32_t *m_ptr, uint32_t cap,
                                                           int32_t *__restrict__ ranges) {
  const uint32_t q = blockIdx.x * kThreads + threadIdx.x;
  const uint32_t M = count_of(m_ptr, cap);
  const uint32_t m0 = 4 * q;
  if (m0 >= M) return;
  uint32_t k[6];  // k[0] = key m0 - 1, k[1..4] = keys m0..m0+3, k[5] = key m0 + 4
  if (m0 + 4 <= M) {
    const uint4 v = reinterpret_cast<const uint4 *>(tkeys)[q];
    k[1] = v.x;
    k[2] = v.y;
    k[3] = v.z;
    k[4] = v.w;
  } else {
#pragma unroll
    for (int j = 0; j < 4; ++j) k[1 + j] = m0 + j < M ? tkeys[m0 + j] : 0u;
  }
  k[0] = m0 > 0 ? tkeys[m0 - 1] : ~k[1];
  k[5] = m0 + 4 < M ? tkeys[m0 + 4] : ~k[4];
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    const uint32_t m = m0 + j;
    if (m < M) {
      const uint32_t t = k[1 + j];
      if (k[j] != t) ranges[2 * t] = (int32_t)m;
      if (m == M - 1 || k[2 + j] != t) ranges[2 * t + 1] = (int32_t)(m + 1);
    }
  }
}

// ---- row binning: the small kernels -----------------------------------------
// Instance and segment count (rectangle area and rows) of each depth-ordered
// splat, scanned together into the instance and the segment offsets.
struct LoadAreaH {
  using value_type = uint2;
  const uint2 *rects;
  __device__ uint2 operator()(uint32_t i) const {
    const uint32_t wh = rects[i].y;
    return make_uint2((wh & 0xFFFFu) * (wh >> 16), wh >> 16);
  }
};
struct StoreOffs2 {
  using Pre = NoPre;
  uint32_t *offs, *soffs;
  uint32_t *own;     // own[k]: the splat holding segment 1024 k (k < own_cap)
  uint32_t own_cap;
  __device__ NoPre pre(uint32_t) const { return NoPre{}; }
  __device__ void operator()(uint32_t i, uint2 excl, uint2 v, NoPre) const {
    offs[i] = excl.x;
    soffs[i] = excl.y;
    const uint32_t k = (excl.y + 1023u) >> 10;  // first multiple of 1024 at or after
    if ((k << 10) < excl.y + v.y && k < own_cap) own[k] = i;
  }
};

// --------------------------------------------------------------------------
// Host-side plumbing
// --------------------------------------------------------------------------
inline size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }
inline uint32_t nblk(uint64_t n) { return (uint32_t)((n + kTileItems - 1) / kTileItems); }
// radix blocks of kThreads * it items
inline uint32_t nblk_it(uint64_t n, int it) {
  const uint64_t b = (uint64_t)kThreads * it;
  return (uint32_t)((n + b - 1) / b);
}
// The depth sort (a few million keys at most) runs blocks of 1024 keys: four
// times the blocks of the tile passes' 4096, so ~5 per CU hide each other's
// load, ranking and scatter latencies (one 4096-key block per CU leaves them
// exposed).  A digit's output run per block is then short, but neighbouring
// blocks run on one XCD (logical_block) and their partial lines merge in its
// L2.  (Measured before the XCD-aware order, 1024-key blocks were slower.)
constexpr int kDepthItems = 4;

// Workspace carve-up for a capacity of `cap` items of 32-bit keys + values.
struct SortBufs {
  uint32_t *keyA, *keyB;
  int32_t *valA, *valB;
  uint32_t *hist;      // [256][nb]
  uint32_t *totals;    // [256] per-digit totals of a pass
  uint32_t *partials;  // data scan partials
};

// hist holds the histograms of radix blocks of kThreads * hist_it items;
// partials the block sums of 4096-item scan blocks
inline size_t sortbufs_bytes(uint64_t cap, int hist_it = kItems) {
  const uint64_t nb = nblk(cap) ? nblk(cap) : 1;
  const uint64_t nh = nblk_it(cap, hist_it) ? nblk_it(cap, hist_it) : 1;
  const uint64_t hn = (uint64_t)kRadix * nh;
  return 4 * align256(cap * 4) + align256(hn * 4) + align256(kRadix * 4) +
         align256((nb + 1) * 4);
}

inline char *carve(char *p, size_t bytes, void **out) {
  *out = p;
  return p + align256(bytes);
}

inline char *carve_sortbufs(char *p, uint64_t cap, SortBufs &b, int hist_it = kItems) {
  const uint64_t nb = nblk(cap) ? nblk(cap) : 1;
  const uint64_t nh = nblk_it(cap, hist_it) ? nblk_it(cap, hist_it) : 1;
  const uint64_t hn = (uint64_t)kRadix * nh;
  p = carve(p, cap * 4, (void **)&b.keyA);
  p = carve(p, cap * 4, (void **)&b.keyB);
  p = carve(p, cap * 4, (void **)&b.valA);
  p = carve(p, cap * 4, (void **)&b.valB);
  p = carve(p, hn * 4, (void **)&b.hist);
  p = carve(p, kRadix * 4, (void **)&b.totals);
  p = carve(p, (nb + 1) * 4, (void **)&b.partials);
  return p;
}

template <class Load, class Store, class Tot>
void run_scan(Load ld, Store st, const uint32_t *n_ptr, uint32_t cap,
              typename Load::value_type *partials, Tot tot, hipStream_t s) {
  const uint32_t nb = nblk(cap) ? nblk(cap) : 1;
  scan_reduce_kernel<<<nb, kThreads, 0, s>>>(ld, n_ptr, cap, partials);
  scan_down_kernel<<<nb, kThreads, 0, s>>>(ld, st, n_ptr, cap, partials, nb, tot);
}

inline int radix_passes(int nbits) { return (nbits + kRadixBits - 1) / kRadixBits; }

// Stable LSD passes over bits [0, nbits) of 32-bit keys.  Data starts in
// (k_cur, v_cur) and the final result is returned through the pointers.
// Passes [p0, passes) run (p0 > 0: the earlier passes were done elsewhere and
// their result is in (k_cur, v_cur)).  An 8-byte payload per item (pay_cur,
// ping-ponged with pay_alt; nullptr for none) moves with the items.  The last
// pass can write, instead of the keys, gathered keys (final_gather[value]) or
// the payload's rectangle area (final_area), and each value's sorted position
// (final_rank); final_tc gathers its tile counts.
struct RadixOpts {
  bool final_keys = true;
  const uint32_t *final_gather = nullptr;
  uint32_t *final_rank = nullptr;
  uint2 *pay_cur = nullptr, *pay_alt = nullptr;
  bool final_area = false;
  int p0 = 0;
  TileCount final_tc = TileCount{};
  uint32_t *totals0 = nullptr;  // the first pass's digit totals, kept apart
};

template <int kIt = kItems>
void run_radix(uint32_t *&k_cur, int32_t *&v_cur, uint32_t *k_alt, int32_t *v_alt,
               const uint32_t *n_ptr, uint32_t cap, int nbits, SortBufs &b, hipStream_t s,
               RadixOpts o = RadixOpts{}) {
  const uint32_t nb = nblk_it(cap, kIt) ? nblk_it(cap, kIt) : 1;
  const int passes = radix_passes(nbits);
  const int per = passes ? (nbits + passes - 1) / passes : 0;
  for (int p = o.p0; p < passes; ++p) {
    const int shift = p * per;
    const int bits = min(per, nbits - shift);
    const uint32_t mask = (1u << bits) - 1u;
    const bool last = p == passes - 1;
    uint32_t *totals = (p == 0 && o.totals0 != nullptr) ? o.totals0 : b.totals;
    radix_hist_kernel<kIt><<<nb, kThreads, 0, s>>>(k_cur, n_ptr, cap, shift, mask, b.hist, nb,
                                                   last ? o.final_tc : TileCount{});
    digit_scan_kernel<<<1u << bits, kThreads, 0, s>>>(b.hist, nb, totals,
                                                      Cols{n_ptr, cap, kThreads * kIt, 0});
    ScatterExtras ex;
    ex.pin = o.pay_cur;
    ex.pout = o.pay_alt;
    if (last) {
      ex.kgather = o.final_gather;
      ex.rank_out = o.final_rank;
      ex.kout_area = o.final_area;
    }
    const bool keys_out = !last || o.final_keys || o.final_gather != nullptr || o.final_area;
    if (ex.pin != nullptr)
      radix_scatter_kernel<ListItems, true, kIt><<<nb, kThreads, 0, s>>>(
          ListItems{k_cur, v_cur}, keys_out ? k_alt : nullptr, v_alt, n_ptr, cap, shift, bits,
          b.hist, nb, totals, ex);
    else
      radix_scatter_kernel<ListItems, false, kIt><<<nb, kThreads, 0, s>>>(
          ListItems{k_cur, v_cur}, keys_out ? k_alt : nullptr, v_alt, n_ptr, cap, shift, bits,
          b.hist, nb, totals, ex);
    uint32_t *kt = k_cur;
    k_cur = k_alt;
    k_alt = kt;
    int32_t *vt = v_cur;
    v_cur = v_alt;
    v_alt = vt;
    uint2 *pt = o.pay_cur;
    o.pay_cur = o.pay_alt;
    o.pay_alt = pt;
  }
}

int tile_bits(int T) {
  int b = 0;
  while ((1 << b) < T) ++b;
  return b;
}

// Tile grids of at most kCountTiles tiles (two tile passes or fewer) get their
// tile ranges from counts gathered by the last pass's histogram, so that pass
// writes no keys and no kernel re-reads them.
constexpr int kCountTiles = 1 << 16;

// Exact binning of one view: mode 2 (default) bins grids of at most 256 x 256
// tiles by rows (two generated passes, the second writing ids and ranges) and
// larger grids as mode 1; mode 1 generates the first tile pass's items from
// the depth-ordered splats (gen_hist_kernel, radix_scatter_kernel<GenItems>);
// mode 0 writes the instance list and list passes sort it.
int g_binning_mode = 2;
constexpr int kRowGrid = 256;  // row binning: columns and rows per grid at most
// pass B's launch: at most ceil(m / 4096) + (rows - 1) blocks hold instances
// (each row adds at most one partial block), plus one empty block whose zero
// histogram column closes the scanned histogram
inline uint32_t row_blocks(uint64_t m) { return nblk(m) + kRowGrid + 1; }
// entries of pass B's owner table (one per 1024 of the m instances)
inline uint64_t own_e_cap(uint64_t m) { return m / 1024 + 2; }

// Row binning's grids: exact, one view, at most 256 x 256 tiles, splat ids
// below 2^24 (pass B packs them with the column into one word).  sort_count
// writes the segment offsets for every exact binning, so a mode change
// between count and emit stays consistent.
bool row_binning(const Grid &gr, int T, int64_t n) {
  return g_binning_mode == 2 && gr.conic == nullptr && gr.vi == nullptr && T > 1 &&
         gr.gx <= kRowGrid && gr.gy <= kRowGrid && n < (int64_t(1) << 24);
}

// Gaussian-level workspace: cnt, offsets, scalars + sort buffers.
struct GaussWs {
  // scalars[0] = Nv, scalars[1] = M; row binning: [2..5] (seg_limit), [6] = S,
  // [7] = pass B's row blocks
  uint32_t *cnt, *offs, *scalars;
  uint32_t *soffs;                 // row binning: segment offsets, in depth order
  uint32_t *bsum;                  // instances per count block
  uint64_t *masks;                 // tight binning: accepted tiles of rectangles <= 64
  uint2 *rects;                    // exact binning: packed tile rectangles
  uint2 *rects2;                   // their ping-pong partner in the depth sort
  uint2 *partials2;                // block sums of the two-value offset scan
  uint32_t *own1k;                 // row binning: the splat holding segment 1024 k
  SortBufs sb;
};
// entries of own1k: S <= 256 Nv segments (rows binning's grids have <= 256 rows)
inline uint64_t own1k_cap(uint64_t c) { return c / 4 + 2; }

inline size_t gauss_ws_bytes(uint64_t n) {
  const uint64_t c = n ? n : 1;
  return 3 * align256(c * 4) + align256(64) + align256((c / kThreads + 1) * 4) +
         3 * align256(c * 8) + align256((nblk(c) + 1) * 8) + align256(own1k_cap(c) * 4) +
         sortbufs_bytes(c, kDepthItems);
}

inline void carve_gauss(void *ws, uint64_t n, GaussWs &g) {
  const uint64_t c = n ? n : 1;
  char *p = (char *)ws;
  p = carve(p, c * 4, (void **)&g.cnt);
  p = carve(p, c * 4, (void **)&g.offs);
  p = carve(p, 64, (void **)&g.scalars);
  p = carve(p, c * 4, (void **)&g.soffs);
  p = carve(p, (c / kThreads + 1) * 4, (void **)&g.bsum);
  p = carve(p, c * 8, (void **)&g.masks);
  p = carve(p, c * 8, (void **)&g.rects);
  p = carve(p, c * 8, (void **)&g.rects2);
  p = carve(p, (nblk(c) + 1) * 8, (void **)&g.partials2);
  p = carve(p, own1k_cap(c) * 4, (void **)&g.own1k);
  carve_sortbufs(p, c, g.sb, kDepthItems);
}

}  // namespace

extern "C" {

size_t gsr_sort_ws_gauss_bytes(int64_t n) { return gauss_ws_bytes((uint64_t)(n > 0 ? n : 1)); }

size_t gsr_sort_ws_inst_bytes(int64_t m) {
  const uint64_t c = (uint64_t)(m > 0 ? m : 1);
  const uint64_t nb = nblk(c);
  const uint64_t nbr = row_blocks(c);
  return sortbufs_bytes(c) + align256(64) + align256(emit_items_cap(c) * 8) +
         align256(nb * (kThreads / 64) * 4) + align256(nb * 2 * kRadix * 4) +
         align256(kRadix * 4) + 2 * align256((uint64_t)kCountTiles * 4) +
         // row binning: pass B's first[], histogram, totals, binfo, owner table
         align256(nbr * (kThreads / 64) * 4) + align256(nbr * kRadix * 4) +
         align256(kRadix * 4) + align256(nbr * 16) + align256(own_e_cap(c) * 4);
}

int gsr_set_binning_mode(int mode) {
  const int prev = g_binning_mode;
  if (mode >= 0 && mode <= 2) g_binning_mode = mode;
  return prev;
}

}  // extern "C"

namespace {

// M to the host as soon as it is known (m_host, pinned; event recorded after
// it), ahead of the depth sort that is queued behind it.  written: the kernel
// that sums M already stored it through the host word's device address.
int copy_count_to_host(const int32_t *m_out, int32_t *m_host, void *event, hipStream_t s,
                       bool written = false) {
  if (m_host == nullptr) return GSR_OK;
  if (!written)
    GSR_REQUIRE(hipMemcpyAsync(m_host, m_out, sizeof(int32_t), hipMemcpyDeviceToHost, s) ==
                    hipSuccess,
                "sort_count: instance count copy failed");
  if (event != nullptr)
    GSR_REQUIRE(hipEventRecord((hipEvent_t)event, s) == hipSuccess,
                "sort_count: event record failed");
  return GSR_OK;
}

// Device address of a pinned host word (nullptr if it has none: the count is
// then copied).
uint32_t *host_word_on_device(int32_t *m_host) {
  if (m_host == nullptr) return nullptr;
  void *d = nullptr;
  if (hipHostGetDevicePointer(&d, m_host, 0) != hipSuccess) {
    (void)hipGetLastError();  // clear it: the copy path is taken instead
    return nullptr;
  }
  return static_cast<uint32_t *>(d);
}

// Scratch of gsr_sort_gaussian_noalloc: one file-scope device array, which
// every GPU the library runs on holds (its own copy) until the process ends.
// Host code reaches it only through the address hipGetSymbolAddress returns
// for the current device (taken on the first call on that device and kept);
// the array's name is not a device address on the host.
constexpr size_t kScratchBytes = size_t(64) << 20;
__device__ __align__(256) unsigned char scratch[kScratchBytes];
constexpr int kMaxDevices = 64;
char *g_scratch_ptr[kMaxDevices] = {};
std::mutex g_scratch_mu;

int scratch_of_current_device(char **out) {
  int dev = -1;
  GSR_REQUIRE(hipGetDevice(&dev) == hipSuccess && dev >= 0 && dev < kMaxDevices,
              "sort_gaussian_noalloc: no current device");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (g_scratch_ptr[dev] == nullptr) {
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || p == nullptr) {
      gsr_set_error("sort_gaussian_noalloc: hipGetSymbolAddress: %s", hipGetErrorString(e));
      return e != hipSuccess ? -(int)e : GSR_ERR_ARG;
    }
    g_scratch_ptr[dev] = static_cast<char *>(p);
  }
  *out = g_scratch_ptr[dev];
  return GSR_OK;
}

int sort_count_impl(const float *uv, const float *depth, const int32_t *radius, int64_t n,
                    const Grid &gr, void *ws_gauss, int32_t *m_out, hipStream_t s,
                    int32_t *m_host = nullptr, void *event = nullptr, int32_t *m_high = nullptr) {
  GaussWs g;
  carve_gauss(ws_gauss, (uint64_t)n, g);
  const uint32_t cap = (uint32_t)(n > 0 ? n : 1);
  if (n == 0) {
    fill_kernel<<<1, kThreads, 0, s>>>((int32_t *)g.scalars, 2, 0);
    fill_kernel<<<1, kThreads, 0, s>>>(m_out, 1, 0);
    const int rc = copy_count_to_host(m_out, m_host, event, s);
    return rc != GSR_OK ? rc : gsr_launch_status("sort_count");
  }
  // compact visible splats with their depth bits as keys -> Nv in scalars[0];
  // the same launch sums the count blocks' totals -> M in m_out (and in m_host)
  const bool exact = gr.conic == nullptr;
  uint32_t *m_host_dev = host_word_on_device(m_host);
  if (exact) {
    // the count kernel writes the compaction's block sums itself
    const uint32_t nb = nblk((uint64_t)n);
    count_kernel<<<nb, kThreads, 0, s>>>(uv, radius, (uint32_t)n, gr, g.rects, g.bsum,
                                         g.sb.partials);
    scan_down_kernel<<<nb, kThreads, 0, s>>>(
        LoadRectFlag{g.rects},
        StoreCompact{depth, g.sb.keyA, g.sb.valA, g.rects, reinterpret_cast<uint2 *>(g.masks)},
        nullptr, (uint32_t)n, g.sb.partials, nb,
        TotCompact{g.scalars, g.bsum, nb, (uint32_t *)m_out, m_host_dev, (uint32_t *)m_high});
  } else {
    const uint32_t ncb = (uint32_t)gsr_blocks(n, kThreads);
    count_tight_kernel<<<ncb, kThreads, 0, s>>>(uv, radius, (uint32_t)n, gr, g.cnt, g.masks,
                                                g.bsum);
    run_scan(LoadFlag{g.cnt}, StoreCompact{depth, g.sb.keyA, g.sb.valA, nullptr, nullptr},
             nullptr, (uint32_t)n, g.sb.partials,
             TotCompact{g.scalars, g.bsum, ncb, (uint32_t *)m_out, m_host_dev, (uint32_t *)m_high},
             s);
  }
  {
    const int rc = copy_count_to_host(m_out, m_host, event, s, m_host_dev != nullptr);
    if (rc != GSR_OK) return rc;
  }
  // depth sort of the visible splats (4 passes over 32 bits); for tight
  // binning the last pass writes each splat's instance count, in depth order,
  // where the keys would go
  uint32_t *kc = g.sb.keyA;
  int32_t *vc = g.sb.valA;
  RadixOpts o;
  o.final_keys = false;
  if (exact) {
    // the packed rectangles travel with the splats (payload, from the masks
    // array -- only tight binning uses it otherwise -- and back after the four
    // passes), so the offset scan and the generated passes of sort_emit read
    // them in depth order (the counts are their areas)
    o.pay_cur = reinterpret_cast<uint2 *>(g.masks);
    o.pay_alt = g.rects2;
  } else {
    // tight binning: counts gathered by id, and each splat's depth rank (its
    // emission walks splats in index order), kept in the unused rectangle array
    o.final_gather = g.cnt;
    o.final_rank = reinterpret_cast<uint32_t *>(g.rects);
  }
  run_radix<kDepthItems>(kc, vc, g.sb.keyB, g.sb.valB, g.scalars, cap, 32, g.sb, s, o);
  // offsets of each splat's instances, in depth order -> M in scalars[1]; exact
  // binning also those of its segments (rectangle rows) -> S in scalars[6]
  if (exact)
    run_scan(LoadAreaH{reinterpret_cast<const uint2 *>(g.masks)},
             StoreOffs2{g.offs, g.soffs, g.own1k, (uint32_t)own1k_cap((uint64_t)cap)}, g.scalars,
             cap, g.partials2, TotPair{g.scalars + 1, g.scalars + 6}, s);
  else
    run_scan(LoadU32{kc}, StoreU32{g.offs}, g.scalars, cap, g.sb.partials,
             TotU32{g.scalars + 1}, s);
  return gsr_launch_status("sort_count");
}

// A generated pass's scatter with its ranking cut to the digit width (6, 7
// or 8 bits: the tile rows and columns of a grid up to 64, 128 or 256 wide).
template <class Items>
void gen_scatter(int bits, uint32_t nblocks, hipStream_t s, const Items &it, uint32_t *kout,
                 int32_t *vout, const uint32_t *n_ptr, uint32_t cap, int shift,
                 const uint32_t *hist, uint32_t nb, const uint32_t *totals,
                 const ScatterExtras &ex) {
  if (bits <= 6)
    radix_scatter_kernel<Items, false, kItems, 6><<<nblocks, kThreads, 0, s>>>(
        it, kout, vout, n_ptr, cap, shift, bits, hist, nb, totals, ex);
  else if (bits == 7)
    radix_scatter_kernel<Items, false, kItems, 7><<<nblocks, kThreads, 0, s>>>(
        it, kout, vout, n_ptr, cap, shift, bits, hist, nb, totals, ex);
  else
    radix_scatter_kernel<Items, false, kItems, 8><<<nblocks, kThreads, 0, s>>>(
        it, kout, vout, n_ptr, cap, shift, bits, hist, nb, totals, ex);
}

int sort_emit_impl(const float *uv, const int32_t *radius, int64_t n, const Grid &gr, int T,
                   void *ws_gauss, int64_t m, void *ws_inst, int32_t *ids_out,
                   int32_t *tile_range, hipStream_t s) {
  if (m == 0 || n == 0) {
    fill_kernel<<<gsr_blocks(2 * (int64_t)T, kThreads), kThreads, 0, s>>>(tile_range,
                                                                          2 * (int64_t)T, 0);
    return gsr_launch_status("sort_emit");
  }
  GaussWs g;
  carve_gauss(ws_gauss, (uint64_t)n, g);
  SortBufs ib;
  char *p = carve_sortbufs((char *)ws_inst, (uint64_t)m, ib);
  uint32_t *n_items, *first, *totals1, *over, *counts;
  uint32_t *tcnt;
  uint2 *items;
  const uint64_t items_cap = emit_items_cap((uint64_t)m);
  const uint32_t nb = nblk((uint64_t)m);
  p = carve(p, 64, (void **)&n_items);
  p = carve(p, items_cap * 8, (void **)&items);
  p = carve(p, (uint64_t)nb * (kThreads / 64) * 4, (void **)&first);
  p = carve(p, (uint64_t)nb * 2 * kRadix * 4, (void **)&tcnt);
  p = carve(p, (uint64_t)kRadix * 4, (void **)&totals1);
  p = carve(p, (uint64_t)kCountTiles * 4, (void **)&over);
  p = carve(p, (uint64_t)kCountTiles * 4, (void **)&counts);
  if (row_binning(gr, T, n)) {
    const uint32_t nbB = row_blocks((uint64_t)m);
    uint32_t *firstB, *histB, *totalsB;
    uint4 *binfo;
    p = carve(p, (uint64_t)nbB * (kThreads / 64) * 4, (void **)&firstB);
    uint32_t *ownE;
    p = carve(p, (uint64_t)nbB * kRadix * 4, (void **)&histB);
    p = carve(p, (uint64_t)kRadix * 4, (void **)&totalsB);
    p = carve(p, (uint64_t)nbB * 16, (void **)&binfo);
    carve(p, own_e_cap((uint64_t)m) * 4, (void **)&ownE);
    const uint32_t mu = (uint32_t)m;
    const int bitsA = tile_bits(gr.gy), bitsB = tile_bits(gr.gx);
    const uint2 *rects = reinterpret_cast<const uint2 *>(g.masks);
    // pass A: segments of the depth-ordered splats, stably by row; its
    // histogram rows [0, D) count segments and rows [D, 2D) their widths (in
    // tcnt, 2 x 256 rows), the digit totals go to counts ([0, D) segments per
    // row, [D, 2D) instances per row), and its scatter writes each segment's
    // first instance E
    const uint32_t DA = 1u << bitsA;
    uint32_t *histA = tcnt, *totA = counts;
    const RowStream rs{g.sb.valA, g.soffs, rects, g.offs, g.scalars, mu, g.own1k};
    gen_hist_kernel<RowStream><<<nb, kThreads, 0, s>>>(rs, g.scalars, (uint32_t)n, nullptr, mu,
                                                       gr, DA - 1u, histA, nb, first, nullptr);
    digit_scan_kernel<<<2 * DA, kThreads, 0, s>>>(histA, nb, totA, Cols{g.scalars + 2, mu});
    ScatterExtras exA;
    exA.e_out = ib.keyB;
    exA.whist = histA + (size_t)DA * nb;
    exA.wtot = totA + DA;
    exA.e_own = ownE;
    exA.e_own_cap = (uint32_t)own_e_cap((uint64_t)m);
    gen_scatter(bitsA, nb, s,
                GenItems<RowStream, true>{rs, g.scalars, (uint32_t)n, gr, first, g.scalars + 3},
                ib.keyA, ib.valA, g.scalars + 2, mu, 0, histA, nb, totA, exA);
    // pass B: each row's instances, stably by column, straight to ids_out
    const SegStream ss{ib.valA, ib.keyB, ib.keyA,     totA,          g.scalars,
                       gr.gx,   gr.gy,   tile_range, g.scalars + 7, ownE};
    gen_hist_kernel<SegStream, true><<<nbB, kThreads, 0, s>>>(
        ss, g.scalars + 2, mu, g.scalars + 5, mu, gr, (1u << bitsB) - 1u, histB, nbB, firstB,
        binfo);
    digit_scan_kernel<<<1u << bitsB, kThreads, 0, s>>>(histB, nbB, totalsB,
                                                       Cols{g.scalars + 7, nbB, 1, 1});
    ScatterExtras ex;
    ex.binfo = binfo;
    ex.row_blocks = g.scalars + 7;
    ex.tile_range = tile_range;
    ex.gx = gr.gx;
    gen_scatter(bitsB, nbB, s,
                GenItems<SegStream, false, true>{ss, g.scalars + 2, mu, gr, firstB, nullptr},
                nullptr, ids_out, g.scalars + 5, mu, 24, histB, nbB, totalsB, ex);
    return gsr_launch_status("sort_emit");
  }
  const int nbits = tile_bits(T);
  const int passes = radix_passes(nbits);
  // ranges from counts: one pass (its digit totals are the tile counts) or two
  // (counts gathered by the last histogram); more passes, or one tile (no
  // pass), take them from the sorted keys
  const bool by_counts = T <= kCountTiles && (passes == 1 || passes == 2);
  const bool by_keys = !by_counts;
  // the emission's queue counter starts at zero; so do tile_range when it is
  // filled from the keys, and the spill counts of the last histogram
  zero2_kernel<<<gsr_blocks((by_keys ? 2 * (int64_t)T : 0) + 1, kThreads), kThreads, 0, s>>>(
      tile_range, by_keys ? 2 * (int64_t)T : 0, n_items, 1);
  if (passes == 2 && by_counts)
    fill_kernel<<<gsr_blocks(T, kThreads), kThreads, 0, s>>>(reinterpret_cast<int32_t *>(over),
                                                            T, 0);
  const int per = passes ? (nbits + passes - 1) / passes : 0;
  const TileCount tc = (passes == 2 && by_counts) ? TileCount{tcnt, over} : TileCount{};
  auto ranges = [&](uint32_t *keys_final) {
    if (by_keys) {
      ranges_kernel<<<gsr_blocks(((int64_t)m + 3) / 4, kThreads), kThreads, 0, s>>>(
          keys_final, g.scalars + 1, (uint32_t)m, tile_range);
      return;
    }
    const uint32_t *cnt = totals1;  // one pass: its digit totals
    if (passes == 2) {
      tile_count_kernel<<<1u << per, kThreads, 0, s>>>(tcnt, totals1, per, over, nbits - per,
                                                        T, counts);
      cnt = counts;
    }
    ranges_from_counts_kernel<<<1, 1024, 0, s>>>(cnt, T, tile_range);
  };
  if (gr.conic == nullptr && passes >= 1 && g_binning_mode >= 1) {
    // first tile pass over generated items (sort_count's last depth pass put
    // the rectangles in depth order into the masks array), then the rest as
    // list passes; values alternate so that the last pass writes ids_out
    int32_t *v_pass0 = (passes % 2 == 1) ? ids_out : ib.valA;
    int32_t *v_pass1 = (passes % 2 == 1) ? ib.valA : ids_out;
    const int bits0 = min(per, nbits);
    const SplatStream st{g.sb.valA, g.offs, reinterpret_cast<const uint2 *>(g.masks)};
    gen_hist_kernel<SplatStream><<<nb, kThreads, 0, s>>>(st, g.scalars, (uint32_t)n,
                                                         g.scalars + 1, (uint32_t)m, gr,
                                                         (1u << bits0) - 1u, ib.hist, nb, first,
                                                         nullptr);
    digit_scan_kernel<<<1u << bits0, kThreads, 0, s>>>(ib.hist, nb, totals1,
                                                       Cols{g.scalars + 1, (uint32_t)m});
    radix_scatter_kernel<<<nb, kThreads, 0, s>>>(
        GenItems<SplatStream>{st, g.scalars, (uint32_t)n, gr, first, nullptr},
        (passes == 1 && by_counts) ? nullptr : ib.keyA, v_pass0, g.scalars + 1, (uint32_t)m, 0,
        bits0, ib.hist, nb, totals1, ScatterExtras{});
    uint32_t *kc = ib.keyA;
    int32_t *vc = v_pass0;
    RadixOpts o;
    o.final_keys = by_keys;
    o.p0 = 1;
    o.final_tc = tc;
    run_radix(kc, vc, ib.keyB, v_pass1, g.scalars + 1, (uint32_t)m, nbits, ib, s, o);
    ranges(kc);
    return gsr_launch_status("sort_emit");
  }
  // depth-sorted splat ids are where sort_count's 4 passes left them: keyA/valA
  int32_t *v_first = (passes % 2 == 0) ? ids_out : ib.valA;
  int32_t *v_second = (passes % 2 == 0) ? ib.valA : ids_out;
  if (gr.conic == nullptr) {
    emit_kernel<<<gsr_blocks(n, kThreads), kThreads, 0, s>>>(
        g.sb.valA, g.offs, g.rects, g.scalars, (uint32_t)n, gr, (uint32_t)m, ib.keyA, v_first,
        items, n_items, (uint32_t)items_cap);
    emit_rest_kernel<<<gsr_blocks((int64_t)items_cap * 64, kThreads), kThreads, 0, s>>>(
        g.sb.valA, g.offs, g.rects, g.scalars, (uint32_t)n, gr, (uint32_t)m, items, n_items,
        (uint32_t)items_cap, ib.keyA, v_first);
  } else
    emit_tight_kernel<<<gsr_blocks(n, kThreads), kThreads, 0, s>>>(
        uv, radius, g.cnt, reinterpret_cast<const uint32_t *>(g.rects), g.offs, g.masks,
        (uint32_t)n, gr, (uint32_t)m, ib.keyA, v_first);
  uint32_t *kc = ib.keyA;
  int32_t *vc = v_first;
  RadixOpts o;
  o.final_keys = by_keys;
  o.final_tc = tc;
  o.totals0 = totals1;
  run_radix(kc, vc, ib.keyB, v_second, g.scalars + 1, (uint32_t)m, nbits, ib, s, o);
  ranges(kc);
  return gsr_launch_status("sort_emit");
}

}  // namespace

extern "C" {

int gsr_sort_count_tight(const float *uv, const float *depth, const int32_t *radius,
                         const float *conic, const float *opacity, int64_t n, int W, int H,
                         int tile_w, int tile_h, void *ws_gauss, int32_t *m_out, void *stream) {
  GSR_REQUIRE(n >= 0 && n < (int64_t)INT32_MAX && W > 0 && H > 0 && tile_w > 0 && tile_h > 0 &&
                  (W + tile_w - 1) / tile_w < 65536 && (H + tile_h - 1) / tile_h < 65536,
              "sort_count: bad sizes");
  GSR_REQUIRE((conic == nullptr) == (opacity == nullptr), "sort_count: conic without opacity");
  const Grid gr{(W + tile_w - 1) / tile_w, (H + tile_h - 1) / tile_h, tile_w, tile_h, nullptr, 0,
                W, H, conic, opacity};
  return sort_count_impl(uv, depth, radius, n, gr, ws_gauss, m_out, (hipStream_t)stream);
}

int gsr_sort_count_async(const float *uv, const float *depth, const int32_t *radius,
                         const float *conic, const float *opacity, int64_t n, int W, int H,
                         int tile_w, int tile_h, void *ws_gauss, int32_t *m_out,
                         int32_t *m_host, void *event, void *stream) {
  GSR_REQUIRE(n >= 0 && n < (int64_t)INT32_MAX && W > 0 && H > 0 && tile_w > 0 && tile_h > 0 &&
                  (W + tile_w - 1) / tile_w < 65536 && (H + tile_h - 1) / tile_h < 65536,
              "sort_count: bad sizes");
  GSR_REQUIRE((conic == nullptr) == (opacity == nullptr), "sort_count: conic without opacity");
  GSR_REQUIRE(m_host != nullptr, "sort_count_async: no host count buffer");
  const Grid gr{(W + tile_w - 1) / tile_w, (H + tile_h - 1) / tile_h, tile_w, tile_h, nullptr, 0,
                W, H, conic, opacity};
  return sort_count_impl(uv, depth, radius, n, gr, ws_gauss, m_out, (hipStream_t)stream, m_host,
                         event);
}

int gsr_sort_emit_tight(const float *uv, const int32_t *radius, const float *conic,
                        const float *opacity, int64_t n, int W, int H, int tile_w, int tile_h,
                        void *ws_gauss, int64_t m, void *ws_inst, int32_t *ids_out,
                        int32_t *tile_range, void *stream) {
  GSR_REQUIRE(n >= 0 && m >= 0 && m < (int64_t)INT32_MAX && W > 0 && H > 0 && tile_w > 0 &&
                  tile_h > 0,
              "sort_emit: bad sizes");
  GSR_REQUIRE((conic == nullptr) == (opacity == nullptr), "sort_emit: conic without opacity");
  const int gx = (W + tile_w - 1) / tile_w, gy = (H + tile_h - 1) / tile_h;
  const Grid gr{gx, gy, tile_w, tile_h, nullptr, 0, W, H, conic, opacity};
  return sort_emit_impl(uv, radius, n, gr, gx * gy, ws_gauss, m, ws_inst, ids_out, tile_range,
                        (hipStream_t)stream);
}

int gsr_sort_count_high(const float *uv, const float *depth, const int32_t *radius,
                        const float *conic, const float *opacity, int64_t n, int W, int H,
                        int tile_w, int tile_h, void *ws_gauss, int32_t *m_out, int32_t *m_high,
                        void *stream) {
  GSR_REQUIRE(n >= 0 && n < (int64_t)INT32_MAX && W > 0 && H > 0 && tile_w > 0 && tile_h > 0 &&
                  (W + tile_w - 1) / tile_w < 65536 && (H + tile_h - 1) / tile_h < 65536,
              "sort_count: bad sizes");
  GSR_REQUIRE((conic == nullptr) == (opacity == nullptr), "sort_count: conic without opacity");
  GSR_REQUIRE(m_high != nullptr, "sort_count_high: no high-water word");
  const Grid gr{(W + tile_w - 1) / tile_w, (H + tile_h - 1) / tile_h, tile_w, tile_h, nullptr, 0,
                W, H, conic, opacity};
  return sort_count_impl(uv, depth, radius, n, gr, ws_gauss, m_out, (hipStream_t)stream, nullptr,
                         nullptr, m_high);
}

int gsr_sort_count(const float *uv, const float *depth, const int32_t *radius, int64_t n,
                   int W, int H, int tile_w, int tile_h, void *ws_gauss, int32_t *m_out,
                   void *stream) {
  return gsr_sort_count_tight(uv, depth, radius, nullptr, nullptr, n, W, H, tile_w, tile_h,
                              ws_gauss, m_out, stream);
}

int gsr_sort_emit(const float *uv, const int32_t *radius, int64_t n, int W, int H, int tile_w,
                  int tile_h, void *ws_gauss, int64_t m, void *ws_inst, int32_t *ids_out,
                  int32_t *tile_range, void *stream) {
  return gsr_sort_emit_tight(uv, radius, nullptr, nullptr, n, W, H, tile_w, tile_h, ws_gauss, m,
                             ws_inst, ids_out, tile_range, stream);
}

int gsr_sort_gaussian_noalloc(const float *uv, const float *depth, const int32_t *radius,
                              int64_t n, int W, int H, int tile_w, int tile_h, int64_t m_cap,
                              int32_t *m_out, int32_t *ids_out, int32_t *tile_range,
                              void *stream) {
  GSR_REQUIRE(n >= 0 && n < (int64_t)INT32_MAX && m_cap >= 0 && m_cap < (int64_t)INT32_MAX &&
                  W > 0 && H > 0 && tile_w > 0 && tile_h > 0 &&
                  (W + tile_w - 1) / tile_w < 65536 && (H + tile_h - 1) / tile_h < 65536,
              "sort_gaussian_noalloc: bad sizes");
  const size_t g_bytes = align256(gauss_ws_bytes((uint64_t)(n > 0 ? n : 1)));
  const size_t i_bytes = gsr_sort_ws_inst_bytes(m_cap);
  GSR_REQUIRE(g_bytes + i_bytes <= kScratchBytes,
              "sort_gaussian_noalloc: %lld splats with an instance capacity of %lld need %zu "
              "bytes of scratch, more than its %zu",
              (long long)n, (long long)m_cap, g_bytes + i_bytes, kScratchBytes);
  char *base = nullptr;
  {
    const int rc = scratch_of_current_device(&base);
    if (rc != GSR_OK) return rc;
  }
  const int gx = (W + tile_w - 1) / tile_w, gy = (H + tile_h - 1) / tile_h;
  const Grid gr{gx, gy, tile_w, tile_h, nullptr, 0, W, H, nullptr, nullptr};
  hipStream_t s = (hipStream_t)stream;
  const int rc = sort_count_impl(uv, depth, radius, n, gr, base, m_out, s);
  if (rc != GSR_OK) return rc;
  return sort_emit_impl(uv, radius, n, gr, gx * gy, base, m_cap, base + g_bytes, ids_out,
                        tile_range, s);
}

int gsr_mv_sort_count(const float *uv, const float *depth, const int32_t *radius,
                      const float *conic, const float *opacity, int64_t n_per_view, int V,
                      const int32_t *view_int, int tile_w, int tile_h, void *ws_gauss,
                      int32_t *m_out, void *stream) {
  const int64_t n = n_per_view * (int64_t)V;
  GSR_REQUIRE(V >= 1 && n_per_view >= 0 && n < (int64_t)INT32_MAX && tile_w > 0 && tile_h > 0,
              "mv_sort_count: bad sizes");
  GSR_REQUIRE((conic == nullptr) == (opacity == nullptr), "mv_sort_count: conic without opacity");
  const Grid gr{0, 0, tile_w, tile_h, view_int, (uint32_t)(n_per_view > 0 ? n_per_view : 1),
                0, 0, conic, opacity};
  return sort_count_impl(uv, depth, radius, n, gr, ws_gauss, m_out, (hipStream_t)stream);
}

int gsr_mv_sort_count_async(const float *uv, const float *depth, const int32_t *radius,
                            const float *conic, const float *opacity, int64_t n_per_view, int V,
                            const int32_t *view_int, int tile_w, int tile_h, void *ws_gauss,
                            int32_t *m_out, int32_t *m_host, void *event, void *stream) {
  const int64_t n = n_per_view * (int64_t)V;
  GSR_REQUIRE(V >= 1 && n_per_view >= 0 && n < (int64_t)INT32_MAX && tile_w > 0 && tile_h > 0,
              "mv_sort_count: bad sizes");
  GSR_REQUIRE((conic == nullptr) == (opacity == nullptr), "mv_sort_count: conic without opacity");
  GSR_REQUIRE(m_host != nullptr, "mv_sort_count_async: no host count buffer");
  const Grid gr{0, 0, tile_w, tile_h, view_int, (uint32_t)(n_per_view > 0 ? n_per_view : 1),
                0, 0, conic, opacity};
  return sort_count_impl(uv, depth, radius, n, gr, ws_gauss, m_out, (hipStream_t)stream, m_host,
                         event);
}

int gsr_mv_sort_emit(const float *uv, const int32_t *radius, const float *conic,
                     const float *opacity, int64_t n_per_view, int V, const int32_t *view_int,
                     int total_tiles, int tile_w, int tile_h, void *ws_gauss, int64_t m,
                     void *ws_inst, int32_t *ids_out, int32_t *tile_range, void *stream) {
  const int64_t n = n_per_view * (int64_t)V;
  GSR_REQUIRE(V >= 1 && n_per_view >= 0 && n < (int64_t)INT32_MAX && m >= 0 &&
                  m < (int64_t)INT32_MAX && total_tiles > 0,
              "mv_sort_emit: bad sizes");
  GSR_REQUIRE((conic == nullptr) == (opacity == nullptr), "mv_sort_emit: conic without opacity");
  const Grid gr{0, 0, tile_w, tile_h, view_int, (uint32_t)(n_per_view > 0 ? n_per_view : 1),
                0, 0, conic, opacity};
  return sort_emit_impl(uv, radius, n, gr, total_tiles, ws_gauss, m, ws_inst, ids_out,
                        tile_range, (hipStream_t)stream);
}

}  // extern "C"
