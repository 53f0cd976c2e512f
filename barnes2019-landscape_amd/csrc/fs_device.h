// fs_device.h -- device code of the FastScape library: constants, lane and
// wave helpers, and every kernel.  Included by each HIP translation unit of
// the library; everything here has internal linkage.
#ifndef FS_DEVICE_H
#define FS_DEVICE_H

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstddef>
#include <cstdint>

#include "fastscape.h"
#include "fs_internal.h"
#include "fs_receiver.h"

namespace {

constexpr int kWave = 64;
constexpr uint8_t kNoFlow = 0xFF;
constexpr int kWalkCap = 64;           // hops walked per cell in k_topology
constexpr int kJumpPasses = 12;        // kWalkCap * 2^10 > FS_MAX_LEVELS
constexpr int kScanBlock = 256;
constexpr int kRootItems = 16;         // cells per thread in k_roots
constexpr int kRootTile = kScanBlock * kRootItems;
constexpr int kBfsItems = 8;           // stack entries per thread in k_bfs
constexpr int kBfsTile = kScanBlock * kBfsItems;
constexpr int kSmallBlock = 1024;
constexpr int kSmallItems = 4;
constexpr int kSmallLevel = kSmallBlock * kSmallItems;  // one-workgroup levels
constexpr int kNewtonCap = 100;
constexpr int kHdrInts = 4;            // maxdepth, n_unfinished, err, spare
constexpr int kHdrCounts = 1024;       // level counts fetched with the header
constexpr int kTopoBlocks = 2048;
constexpr int kJumpBlocks = 1024;

// device error bits (hdr[2])
constexpr int kErrLookback = 1, kErrNewton = 2, kErrIndex = 4, kErrDepth = 8;

// ---------------------------------------------------------------------------
// neighbourhood encodings
// ---------------------------------------------------------------------------
// Receiver direction r = 0..7 is W, NW, N, NE, E, SE, S, SW (RB:79-82); its
// offset is nshift[r] = {-1, -W-1, -W, -W+1, 1, W+1, W, W-1} (RB:137).
__device__ __forceinline__ int dir_offset(int r, int W) {
  constexpr unsigned kDx = (0u << 0) | (0u << 2) | (1u << 4) | (2u << 6) |
                           (2u << 8) | (2u << 10) | (1u << 12) | (0u << 14);
  constexpr unsigned kDy = (1u << 0) | (0u << 2) | (0u << 4) | (0u << 6) |
                           (1u << 8) | (2u << 10) | (2u << 12) | (2u << 14);
  const int dx = static_cast<int>((kDx >> (2 * r)) & 3u) - 1;
  const int dy = static_cast<int>((kDy >> (2 * r)) & 3u) - 1;
  return dy * W + dx;
}

// Donor bitmask slot j = 0..7 is the neighbour at offset
// {-W-1, -W, -W+1, -1, +1, W-1, W, W+1}[j]: ascending cell index, so the set
// bits in slot order list the donors in the order RB's push builds them
// (RB:194-202).  The neighbour in slot j is a donor iff its receiver
// direction is kSlotDir[j].
__device__ __forceinline__ int slot_offset(int j, int W) {
  constexpr unsigned kSx = (0u << 0) | (1u << 2) | (2u << 4) | (0u << 6) |
                           (2u << 8) | (0u << 10) | (1u << 12) | (2u << 14);
  constexpr unsigned kSy = (0u << 0) | (0u << 2) | (0u << 4) | (1u << 6) |
                           (1u << 8) | (2u << 10) | (2u << 12) | (2u << 14);
  const int dx = static_cast<int>((kSx >> (2 * j)) & 3u) - 1;
  const int dy = static_cast<int>((kSy >> (2 * j)) & 3u) - 1;
  return dy * W + dx;
}

// ---------------------------------------------------------------------------
// scans
// ---------------------------------------------------------------------------
// Exclusive prefix sum over the block; *total gets the block sum.  sh holds
// BLOCK/64 + 1 words.  Ends with a barrier, so sh may be reused at once.
template <int BLOCK>
__device__ __forceinline__ uint32_t block_excl_scan(uint32_t v, uint32_t *total,
                                                    uint32_t *sh) {
  const int lane = threadIdx.x & (kWave - 1), wid = threadIdx.x / kWave;
  uint32_t x = v;
#pragma unroll
  for (int d = 1; d < kWave; d <<= 1) {
    const uint32_t y = __shfl_up(x, d);
    if (lane >= d) x += y;
  }
  if (lane == kWave - 1) sh[wid] = x;
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t run = 0;
    for (int w = 0; w < BLOCK / kWave; w++) {
      const uint32_t t = sh[w];
      sh[w] = run;
      run += t;
    }
    sh[BLOCK / kWave] = run;
  }
  __syncthreads();
  const uint32_t ex = sh[wid] + x - v;
  *total = sh[BLOCK / kWave];
  __syncthreads();
  return ex;
}

// Tile status word: [epoch:30 | flag:2 | value:32].  flag 1 = tile aggregate,
// 2 = inclusive prefix.  A word whose epoch is not the launch's epoch is
// "not yet published", so the array never needs clearing.
__device__ __forceinline__ uint64_t status_word(uint32_t epoch, uint32_t flag, uint32_t v) {
  return (static_cast<uint64_t>(epoch & 0x3fffffffu) << 34) |
         (static_cast<uint64_t>(flag) << 32) | v;
}

// Decoupled look-back (Merrill & Garland, "Single-pass Parallel Prefix Scan
// with Decoupled Look-back", 2016), run by one full wave.  Publishes the
// tile's aggregate, sums predecessors' aggregates 64 at a time until an
// inclusive prefix is found, publishes its own inclusive prefix and returns
// the tile's exclusive prefix in every lane.  Tiles are numbered by an atomic
// ticket, so every predecessor is resident and publishes its aggregate
// without waiting; the wait is still bounded and reported in *err.
__device__ uint32_t lookback(uint64_t *status, uint32_t tile, uint32_t epoch,
                             uint32_t agg, int32_t *err) {
  const int lane = threadIdx.x & (kWave - 1);
  epoch &= 0x3fffffffu;
  if (tile == 0) {
    if (lane == 0)
      __hip_atomic_store(&status[0], status_word(epoch, 2, agg), __ATOMIC_RELAXED,
                         __HIP_MEMORY_SCOPE_AGENT);
    return 0;
  }
  if (lane == 0)
    __hip_atomic_store(&status[tile], status_word(epoch, 1, agg), __ATOMIC_RELAXED,
                       __HIP_MEMORY_SCOPE_AGENT);
  uint32_t excl = 0;
  long long base = static_cast<long long>(tile) - 1;
  int spins = 0;
  while (true) {
    const long long idx = base - lane;
    uint64_t w = status_word(epoch, 2, 0);
    if (idx >= 0)
      w = __hip_atomic_load(&status[idx], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    const uint32_t flag = static_cast<uint32_t>(w >> 32) & 3u;
    const bool valid = (static_cast<uint32_t>(w >> 34) == epoch) && flag != 0;
    const unsigned long long inval = __ballot(!valid);
    const unsigned long long pmask = __ballot(valid && flag == 2);
    const int first_p = pmask ? __builtin_ctzll(pmask) : kWave;
    const int first_inv = inval ? __builtin_ctzll(inval) : kWave;
    if (first_inv < first_p) {
      if (++spins > (1 << 22)) {
        if (lane == 0) atomicOr(err, kErrLookback);
        break;
      }
      __builtin_amdgcn_s_sleep(1);
      continue;
    }
    uint32_t v = (lane <= first_p) ? static_cast<uint32_t>(w) : 0u;
#pragma unroll
    for (int o = kWave / 2; o > 0; o >>= 1) v += __shfl_xor(v, o);
    excl += v;
    if (first_p < kWave) break;
    base -= kWave;
  }
  if (lane == 0)
    __hip_atomic_store(&status[tile], status_word(epoch, 2, excl + agg),
                       __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  return excl;
}

// ---------------------------------------------------------------------------
// kernels
// ---------------------------------------------------------------------------
__global__ void k_fill_counts(int32_t *hdr, int n_counts) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i == 0) {
    hdr[0] = 0;   // maxdepth
    hdr[1] = 0;   // n_unfinished
    hdr[3] = 0;   // tiles whose levels exceed the tile halo
  }
  if (i < n_counts) hdr[kHdrInts + i] = 0;
}

// ComputeReceivers (RB:160-182).  Each thread owns one column of a strip of
// kRecvRows interior rows and slides a 3x3 window down it, so every
// elevation is fetched from HBM about once (neighbouring columns come from
// L1).  fs::receiver_fast picks the reference's direction -- steepest
// (hc - h[n]) / dr[n] over n = 0..7 with a strict '>' from 0, ties to the
// lowest n, flat cells and pits NO_FLOW -- with two IEEE divisions instead of
// eight (see fs_receiver.h).
constexpr int kRecvRows = 16;
__device__ __forceinline__ void receivers_strip(const double *__restrict__ h,
                                                uint8_t *__restrict__ rcv, int W, int H, int x,
                                                int strip, double d_card, double d_diag) {
  const int y0 = strip * kRecvRows + 2;
  const int y1 = min(y0 + kRecvRows, H - 2);
  const double *p = h + static_cast<size_t>(y0 - 1) * W + x;
  double a0 = p[-1], a1 = p[0], a2 = p[1];        // row y-1
  p += W;
  double b0 = p[-1], b1 = p[0], b2 = p[1];        // row y
  for (int y = y0; y < y1; y++) {
    p += W;
    const double c0 = p[-1], c1 = p[0], c2 = p[1];  // row y+1
    // neighbours in direction order W, NW, N, NE, E, SE, S, SW (RB:79-82)
    const double nb[8] = {b0, a0, a1, a2, b2, c2, c1, c0};
    rcv[static_cast<size_t>(y) * W + x] =
        static_cast<uint8_t>(fs::receiver_fast(b1, nb, d_card, d_diag));
    a0 = b0; a1 = b1; a2 = b2;
    b0 = c0; b1 = c1; b2 = c2;
  }
}

__global__ __launch_bounds__(256) void k_receivers(const double *__restrict__ h,
                                                   uint8_t *__restrict__ rcv, int W, int H,
                                                   double d_card, double d_diag) {
  const int x = blockIdx.x * 256 + threadIdx.x + 2;
  if (x >= W - 2) return;
  const int nstrips = (H - 4 + kRecvRows - 1) / kRecvRows;
  for (int strip = blockIdx.y; strip < nstrips; strip += gridDim.y)
    receivers_strip(h, rcv, W, H, x, strip, d_card, d_diag);
}

// Donor bitmask and level of every cell.  Rows are strided over the grid so
// each block folds its level histogram in LDS before touching HBM.  The
// level is found by walking receivers (all reads hit L1/L2: the receiver
// field is one byte per cell and the walk moves one cell per hop); walks
// longer than kWalkCap hops record (cell reached, hops) for k_jump.
__global__ __launch_bounds__(256) void k_topology(
    const uint8_t *__restrict__ rcv, uint8_t *__restrict__ dmask,
    uint16_t *__restrict__ depth, unsigned long long *__restrict__ jump,
    int32_t *__restrict__ list, int32_t *__restrict__ hdr, int W, int H) {
  __shared__ int hist[kWalkCap + 1];
  __shared__ int smax;
  for (int i = threadIdx.x; i <= kWalkCap; i += blockDim.x) hist[i] = 0;
  if (threadIdx.x == 0) smax = 0;
  __syncthreads();
  int32_t *counts = hdr + kHdrInts;
  int local_max = 0;
  for (int y = blockIdx.x; y < H; y += gridDim.x) {
    const bool yin = y > 0 && y < H - 1;
    for (int x = threadIdx.x; x < W; x += blockDim.x) {
      const int c = y * W + x;
      uint32_t m = 0;
      if (yin && x > 0 && x < W - 1) {
        const uint8_t *up = rcv + c - W;
        const uint8_t *dn = rcv + c + W;
        m |= (up[-1] == 5) ? 0x01u : 0u;
        m |= (up[0] == 6) ? 0x02u : 0u;
        m |= (up[1] == 7) ? 0x04u : 0u;
        m |= (rcv[c - 1] == 4) ? 0x08u : 0u;
        m |= (rcv[c + 1] == 0) ? 0x10u : 0u;
        m |= (dn[-1] == 3) ? 0x20u : 0u;
        m |= (dn[0] == 2) ? 0x40u : 0u;
        m |= (dn[1] == 1) ? 0x80u : 0u;
      }
      dmask[c] = static_cast<uint8_t>(m);
      int cur = c, d = 0;
      int r = rcv[c];
      while (r < 8 && d < kWalkCap) {
        cur += dir_offset(r, W);
        d++;
        r = rcv[cur];
      }
      if (r >= 8) {
        depth[c] = static_cast<uint16_t>(d);
        atomicAdd(&hist[d], 1);
        local_max = max(local_max, d);
      } else {
        depth[c] = 0xFFFF;
        jump[c] = (static_cast<unsigned long long>(d) << 32) | static_cast<uint32_t>(cur);
        const int slot = atomicAdd(&hdr[1], 1);
        list[slot] = c;
      }
    }
  }
  atomicMax(&smax, local_max);
  __syncthreads();
  for (int i = threadIdx.x; i <= kWalkCap; i += blockDim.x)
    if (hist[i]) atomicAdd(&counts[i], hist[i]);
  if (threadIdx.x == 0) atomicMax(&hdr[0], smax);
}

// One pointer-jumping pass over the cells k_topology left unfinished: a cell
// whose recorded ancestor has a level takes level = hops + that level;
// otherwise it jumps to its ancestor's ancestor.  (ancestor, hops) pairs are
// read and written as one 64-bit word, so every pair a thread sees is
// consistent and the update is safe in place.
__global__ __launch_bounds__(256) void k_jump(const int32_t *__restrict__ list,
                                              uint16_t *depth, unsigned long long *jump,
                                              int32_t *hdr) {
  const int n = hdr[1];
  int32_t *counts = hdr + kHdrInts;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const int c = list[i];
    if (__hip_atomic_load(&depth[c], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0xFFFF)
      continue;
    const unsigned long long j =
        __hip_atomic_load(&jump[c], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    const int a = static_cast<int>(j & 0xffffffffu);
    const unsigned d = static_cast<unsigned>(j >> 32);
    const unsigned da = __hip_atomic_load(&depth[a], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (da != 0xFFFF) {
      const unsigned nd = d + da;
      if (nd > FS_MAX_LEVELS) {
        atomicOr(&hdr[2], kErrDepth);
        continue;
      }
      __hip_atomic_store(&depth[c], static_cast<uint16_t>(nd), __ATOMIC_RELAXED,
                         __HIP_MEMORY_SCOPE_AGENT);
      atomicAdd(&counts[nd], 1);
      atomicMax(&hdr[0], static_cast<int>(nd));
    } else {
      const unsigned long long ja =
          __hip_atomic_load(&jump[a], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      const unsigned nd = d + static_cast<unsigned>(ja >> 32);
      if (nd > FS_MAX_LEVELS + 1u) {
        atomicOr(&hdr[2], kErrDepth);
        continue;
      }
      __hip_atomic_store(&jump[c],
                         (static_cast<unsigned long long>(nd) << 32) | (ja & 0xffffffffu),
                         __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
  }
}

// Level 0 of GenerateOrder (RB:223-226): every NO_FLOW cell, ring 0
// included, in ascending index.  One tile = 4096 cells, 16 per thread read
// as one 16-byte load.
__global__ __launch_bounds__(kScanBlock) void k_roots(
    const uint8_t *__restrict__ rcv, int32_t *__restrict__ stack, uint64_t *status,
    unsigned long long *ticket, unsigned long long tbase, uint32_t epoch, int32_t N,
    int32_t *hdr) {
  __shared__ uint32_t sh[kScanBlock / kWave + 1];
  __shared__ uint32_t s_tile, s_pre;
  if (threadIdx.x == 0) s_tile = static_cast<uint32_t>(atomicAdd(ticket, 1ull) - tbase);
  __syncthreads();
  const uint32_t tile = s_tile;
  const long long base = static_cast<long long>(tile) * kRootTile +
                         static_cast<long long>(threadIdx.x) * kRootItems;
  uint32_t flags = 0;
  if (base + kRootItems <= N) {
    const uint4 v = *reinterpret_cast<const uint4 *>(rcv + base);
    const uint32_t w4[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
    for (int k = 0; k < 4; k++)
#pragma unroll
      for (int b = 0; b < 4; b++)
        if (((w4[k] >> (8 * b)) & 0xFFu) == kNoFlow) flags |= 1u << (4 * k + b);
  } else {
    for (int i = 0; i < kRootItems; i++)
      if (base + i < N && rcv[base + i] == kNoFlow) flags |= 1u << i;
  }
  uint32_t total;
  const uint32_t ex = block_excl_scan<kScanBlock>(__popc(flags), &total, sh);
  if (threadIdx.x < kWave) {
    const uint32_t pre = lookback(status, tile, epoch, total, &hdr[2]);
    if (threadIdx.x == 0) s_pre = pre;
  }
  __syncthreads();
  long long pos = static_cast<long long>(s_pre) + ex;
  while (flags) {
    const int i = __builtin_ctz(flags);
    flags &= flags - 1;
    if (pos < N) stack[pos] = static_cast<int32_t>(base + i);
    pos++;
  }
}

// One BFS level of GenerateOrder (RB:234-247): stack[ls, le) -> the donor
// lists of those cells, in order, written from stack[le].
__global__ __launch_bounds__(kScanBlock) void k_bfs(
    int32_t *__restrict__ stack, const uint8_t *__restrict__ dmask, int32_t ls, int32_t le,
    int W, int32_t N, uint64_t *status, unsigned long long *ticket,
    unsigned long long tbase, uint32_t epoch, int32_t *hdr) {
  __shared__ uint32_t sh[kScanBlock / kWave + 1];
  __shared__ uint32_t s_tile, s_pre;
  if (threadIdx.x == 0) s_tile = static_cast<uint32_t>(atomicAdd(ticket, 1ull) - tbase);
  __syncthreads();
  const uint32_t tile = s_tile;
  const long long p0 = static_cast<long long>(ls) + static_cast<long long>(tile) * kBfsTile +
                       static_cast<long long>(threadIdx.x) * kBfsItems;
  int32_t cell[kBfsItems];
  unsigned long long masks = 0;
  uint32_t cnt = 0;
#pragma unroll
  for (int i = 0; i < kBfsItems; i++) {
    const long long p = p0 + i;
    int32_t c = 0;
    uint32_t m = 0;
    if (p < le) {
      c = stack[p];
      if (static_cast<uint32_t>(c) < static_cast<uint32_t>(N)) m = dmask[c];
      else atomicOr(&hdr[2], kErrIndex);
    }
    cell[i] = c;
    masks |= static_cast<unsigned long long>(m) << (8 * i);
    cnt += __popc(m);
  }
  uint32_t total;
  const uint32_t ex = block_excl_scan<kScanBlock>(cnt, &total, sh);
  if (threadIdx.x < kWave) {
    const uint32_t pre = lookback(status, tile, epoch, total, &hdr[2]);
    if (threadIdx.x == 0) s_pre = pre;
  }
  __syncthreads();
  long long pos = static_cast<long long>(le) + s_pre + ex;
#pragma unroll
  for (int i = 0; i < kBfsItems; i++) {
    uint32_t m = static_cast<uint32_t>(masks >> (8 * i)) & 0xFFu;
    while (m) {
      const int j = __builtin_ctz(m);
      m &= m - 1;
      if (pos < N) stack[pos] = cell[i] + slot_offset(j, W);
      pos++;
    }
  }
}

// Levels [L0, L1) of the BFS in one workgroup, for runs of small levels.
__global__ __launch_bounds__(kSmallBlock) void k_bfs_small(
    int32_t *__restrict__ stack, const uint8_t *__restrict__ dmask,
    const int32_t *__restrict__ levels, int L0, int L1, int W, int32_t N, int32_t *hdr) {
  __shared__ uint32_t sh[kSmallBlock / kWave + 1];
  for (int L = L0; L < L1; L++) {
    const int32_t ls = levels[L], le = levels[L + 1];
    long long out = le;
    for (long long b = ls; b < le; b += kSmallLevel) {
      const long long p0 = b + static_cast<long long>(threadIdx.x) * kSmallItems;
      int32_t cell[kSmallItems];
      uint32_t masks = 0, cnt = 0;
#pragma unroll
      for (int i = 0; i < kSmallItems; i++) {
        const long long p = p0 + i;
        int32_t c = 0;
        uint32_t m = 0;
        if (p < le) {
          c = stack[p];
          if (static_cast<uint32_t>(c) < static_cast<uint32_t>(N)) m = dmask[c];
          else atomicOr(&hdr[2], kErrIndex);
        }
        cell[i] = c;
        masks |= m << (8 * i);
        cnt += __popc(m);
      }
      uint32_t total;
      const uint32_t ex = block_excl_scan<kSmallBlock>(cnt, &total, sh);
      long long pos = out + ex;
#pragma unroll
      for (int i = 0; i < kSmallItems; i++) {
        uint32_t m = (masks >> (8 * i)) & 0xFFu;
        while (m) {
          const int j = __builtin_ctz(m);
          m &= m - 1;
          if (pos < N) stack[pos] = cell[i] + slot_offset(j, W);
          pos++;
        }
      }
      out += total;
    }
    __syncthreads();
  }
}

// Upstream cell count of one cell: 1 + the counts of its donors, which sit
// one level higher and are already final.  accum = count * cell_area exactly
// (RB:262-279: every partial sum there is an integer multiple of cell_area).
__device__ __forceinline__ void flowacc_cell(int32_t c, const uint8_t *__restrict__ dmask,
                                             uint32_t *cnt, int W, int32_t N, int32_t *hdr) {
  if (static_cast<uint32_t>(c) >= static_cast<uint32_t>(N)) {
    atomicOr(&hdr[2], kErrIndex);
    return;
  }
  uint32_t m = dmask[c];
  uint32_t s = 1;
  while (m) {
    const int j = __builtin_ctz(m);
    m &= m - 1;
    s += cnt[c + slot_offset(j, W)];
  }
  cnt[c] = s;
}

__global__ __launch_bounds__(256) void k_flowacc(const int32_t *__restrict__ stack,
                                                 const uint8_t *__restrict__ dmask,
                                                 uint32_t *cnt, int32_t ls, int32_t le, int W,
                                                 int32_t N, int32_t *hdr) {
  const long long p = static_cast<long long>(ls) + blockIdx.x * 256ll + threadIdx.x;
  if (p >= le) return;
  flowacc_cell(stack[p], dmask, cnt, W, N, hdr);
}

// Levels L_hi down to L_lo (inclusive) in one workgroup.
__global__ __launch_bounds__(kSmallBlock) void k_flowacc_small(
    const int32_t *__restrict__ stack, const uint8_t *__restrict__ dmask, uint32_t *cnt,
    const int32_t *__restrict__ levels, int L_hi, int L_lo, int W, int32_t N, int32_t *hdr) {
  for (int L = L_hi; L >= L_lo; L--) {
    const int32_t ls = levels[L], le = levels[L + 1];
    for (long long p = ls + threadIdx.x; p < le; p += kSmallBlock)
      flowacc_cell(stack[p], dmask, cnt, W, N, hdr);
    __syncthreads();
  }
}

// AddUplift (RB:285-296) for the interior outlets (level-0 cells away from
// rings 0 and 1).  Interior cells with a receiver are uplifted in k_erode.
__global__ __launch_bounds__(256) void k_uplift_roots(const int32_t *__restrict__ stack,
                                                      int32_t n0, double *h, int W, int H,
                                                      double ud, int32_t N, int32_t *hdr) {
  const int p = blockIdx.x * 256 + threadIdx.x;
  if (p >= n0) return;
  const int32_t c = stack[p];
  if (static_cast<uint32_t>(c) >= static_cast<uint32_t>(N)) {
    atomicOr(&hdr[2], kErrIndex);
    return;
  }
  const int y = c / W, x = c - y * W;
  if (x >= 2 && x < W - 2 && y >= 2 && y < H - 2) h[c] += ud;
}

// Plain 16-byte-per-thread copy, used only to calibrate hardware traffic
// counters against a kernel whose HBM bytes are known exactly.
__global__ __launch_bounds__(256) void k_copy_calibration(const uint4 *__restrict__ src,
                                                          uint4 *__restrict__ dst, size_t n) {
  for (size_t i = blockIdx.x * 256ull + threadIdx.x; i < n; i += gridDim.x * 256ull)
    dst[i] = src[i];
}

struct ErodeConst {
  double kdt;       // keq*dt
  double lenpow0;   // pow(dx, neq)        cardinal receiver
  double lenpow1;   // pow(dx*SQRT2, neq)  diagonal receiver
  double ud;        // ueq*dt
  double neq;
  double tol;
};

// Per-model constants of the erosion (RB:313-315) and the pow mode of neq.
inline ErodeConst make_erode_const(const fs_params &p, int *nmode) {
  ErodeConst k;
  k.kdt = p.keq * p.dt;
  k.lenpow0 = std::pow(p.dx, p.neq);
  k.lenpow1 = std::pow(p.dx * fs::kSqrt2, p.neq);
  k.ud = p.ueq * p.dt;
  k.neq = p.neq;
  k.tol = p.tol;
  *nmode = (p.neq == 1.0) ? 1 : (p.neq == 2.0) ? 2 : 0;
  return k;
}

// The implicit update of one cell (RB:313-326), operation for operation.
// NMODE 1: n = 1, where libm returns pow(x,1) = x and pow(x,0) = 1 exactly;
// NMODE 2: n = 2 with x*x for pow(x,2); NMODE 0: any n through the device
// pow.  fact = (K*dt)*A^m / dr^n, h0 is the uplifted elevation h + U*dt
// (RB:294) and hn the receiver's final elevation.
template <int NMODE>
__device__ __forceinline__ double newton_update(double h0, double hn, double fact,
                                                const ErodeConst &k, int32_t *hdr) {
  double hnew = h0, hp = h0, diff = 2.0 * k.tol;
  int it = 0;
  while (fabs(diff) > k.tol) {
    const double x = hnew - hn;
    double p1, p0;
    if (NMODE == 1) {
      p1 = x;
      p0 = 1.0;
    } else if (NMODE == 2) {
      p1 = x * x;
      p0 = x;
    } else {
      p1 = pow(x, k.neq);
      p0 = pow(x, k.neq - 1.0);
    }
    hnew = hnew - ((hnew - h0) + fact * p1) / (1.0 + (fact * k.neq) * p0);
    diff = hnew - hp;
    hp = hnew;
    if (++it >= kNewtonCap) {
      atomicOr(&hdr[2], kErrNewton);
      break;
    }
  }
  return hnew;
}

// Erode for one stack entry (RB:306-328): h[c] from h[c] + U*dt and the
// receiver's final elevation (its level was processed earlier).
template <int NMODE>
__device__ __forceinline__ void erode_cell(int32_t c, const uint8_t *__restrict__ rcv,
                                           const uint32_t *__restrict__ cnt,
                                           const double *__restrict__ ftab, double *h,
                                           int W, int32_t N, const ErodeConst &k,
                                           int32_t *hdr) {
  if (static_cast<uint32_t>(c) >= static_cast<uint32_t>(N)) {
    atomicOr(&hdr[2], kErrIndex);
    return;
  }
  const int r = rcv[c];
  if (r >= 8) return;
  const int32_t n = c + dir_offset(r, W);
  uint32_t a = cnt[c];
  if (a > static_cast<uint32_t>(N)) {
    atomicOr(&hdr[2], kErrIndex);
    a = static_cast<uint32_t>(N);
  }
  h[c] = newton_update<NMODE>(h[c] + k.ud, h[n], ftab[2 * a + (r & 1)], k, hdr);
}

template <int NMODE>
__global__ __launch_bounds__(256) void k_erode(const int32_t *__restrict__ stack,
                                               const uint8_t *__restrict__ rcv,
                                               const uint32_t *__restrict__ cnt,
                                               const double *__restrict__ ftab, double *h,
                                               int32_t ls, int32_t le, int W, int32_t N,
                                               ErodeConst k, int32_t *hdr) {
  const long long p = static_cast<long long>(ls) + blockIdx.x * 256ll + threadIdx.x;
  if (p >= le) return;
  erode_cell<NMODE>(stack[p], rcv, cnt, ftab, h, W, N, k, hdr);
}

template <int NMODE>
__global__ __launch_bounds__(kSmallBlock) void k_erode_small(
    const int32_t *__restrict__ stack, const uint8_t *__restrict__ rcv,
    const uint32_t *__restrict__ cnt, const double *__restrict__ ftab, double *h,
    const int32_t *__restrict__ levels, int L_lo, int L_hi, int W, int32_t N, ErodeConst k,
    int32_t *hdr) {
  for (int L = L_lo; L <= L_hi; L++) {
    const int32_t ls = levels[L], le = levels[L + 1];
    for (long long p = ls + threadIdx.x; p < le; p += kSmallBlock)
      erode_cell<NMODE>(stack[p], rcv, cnt, ftab, h, W, N, k, hdr);
    __syncthreads();
  }
}

// ---------------------------------------------------------------------------
// tile path
// ---------------------------------------------------------------------------
// A cell's receiver chain and its whole upstream tree lie within D cells
// (Chebyshev distance) of it, D being the deepest level of the grid.  So
// when D <= kHalo, a kTile x kTile tile plus a kHalo-wide halo holds
// everything the tile's cells depend on, and accumulation and erosion can
// run level by level inside LDS with one HBM read of the region and one
// write of the tile.  Every grid cell is a tile cell of exactly one tile.
constexpr int kTile = 64;
constexpr int kHalo = 16;
constexpr int kReg = kTile + 2 * kHalo;       // region side
constexpr int kRegCells = kReg * kReg;
constexpr int kAccBlock = 1024;
constexpr int kEroBlock = 1024;
constexpr uint8_t kOutside = 0xFD;            // region cell beyond the grid
constexpr uint8_t kEscaped = 0xFF;            // chain leaves the region
constexpr int kFactCache = 256;               // erosion factors cached in LDS

// Region-local neighbour step in direction r; false if it leaves the region.
__device__ __forceinline__ bool region_step(int r, int &lx, int &ly) {
  constexpr unsigned kDx = (0u << 0) | (0u << 2) | (1u << 4) | (2u << 6) |
                           (2u << 8) | (2u << 10) | (1u << 12) | (0u << 14);
  constexpr unsigned kDy = (1u << 0) | (0u << 2) | (0u << 4) | (0u << 6) |
                           (1u << 8) | (2u << 10) | (2u << 12) | (2u << 14);
  lx += static_cast<int>((kDx >> (2 * r)) & 3u) - 1;
  ly += static_cast<int>((kDy >> (2 * r)) & 3u) - 1;
  return lx >= 0 && lx < kReg && ly >= 0 && ly < kReg;
}

// Workgroups are dealt to the 8 XCDs round-robin by block index; hand each
// XCD one contiguous run of tiles so that tiles sharing halos share an L2.
__device__ __forceinline__ int xcd_tile(int b, int ntiles) {
  const int q = ntiles / 8, rem = ntiles % 8;
  const int x = b % 8, k = b / 8;
  return x * q + min(x, rem) + k;
}

// Region-local index offset of direction r (0..7), from one packed constant.
__device__ __forceinline__ int region_offset(int r) {
  constexpr int8_t kOff[8] = {-1, -kReg - 1, -kReg, -kReg + 1, 1, kReg + 1, kReg, kReg - 1};
  constexpr uint64_t kPacked =
      static_cast<uint64_t>(static_cast<uint8_t>(kOff[0])) |
      static_cast<uint64_t>(static_cast<uint8_t>(kOff[1])) << 8 |
      static_cast<uint64_t>(static_cast<uint8_t>(kOff[2])) << 16 |
      static_cast<uint64_t>(static_cast<uint8_t>(kOff[3])) << 24 |
      static_cast<uint64_t>(static_cast<uint8_t>(kOff[4])) << 32 |
      static_cast<uint64_t>(static_cast<uint8_t>(kOff[5])) << 40 |
      static_cast<uint64_t>(static_cast<uint8_t>(kOff[6])) << 48 |
      static_cast<uint64_t>(static_cast<uint8_t>(kOff[7])) << 56;
  return static_cast<int8_t>(static_cast<uint8_t>(kPacked >> (8 * r)));
}

// Pass 1: upstream cell counts (ComputeFlowAcc, RB:262-279) and levels
// (receiver hops to an outlet, the level GenerateOrder assigns, RB:213-254)
// of the tile's cells.  Every region cell walks its receiver chain once and
// adds 1 to each cell it passes, itself included, so a cell's count is the
// number of region cells draining through it -- its whole upstream area when
// every level is at most kHalo.  A tile cell deeper than that (or whose chain
// leaves the region) raises hdr[3]; the host then runs the level path.
// Halo cells the erosion pass solves per tile at most: those on the receiver
// chains of the tile's cells (mean 60-66, at most 92 per tile of 5120 over
// 120 steps at the benchmark preset).
constexpr int kHaloCap = 256;

__device__ __forceinline__ bool in_tile(int lx, int ly) {
  return lx >= kHalo && lx < kHalo + kTile && ly >= kHalo && ly < kHalo + kTile;
}

// Accumulation word of a region cell in LDS: bits 0-13 its count, bit 23
// set for a tile cell, bits 24-31 the signed region-index offset of its
// receiver or a stop code.  One atomic add both counts a walker's visit and
// tells it where to go next.
constexpr uint32_t kAccCount = 0x3FFFu;
constexpr uint32_t kAccTile = 1u << 23;
constexpr int kStopRoot = -128;      // an outlet (NO_FLOW)
constexpr int kStopLeaves = -127;    // the receiver lies outside the region
constexpr int kStopOutside = -126;   // beyond the grid
constexpr int kStopMax = -126;       // codes <= kStopMax end a walk
static_assert(kRegCells <= static_cast<int>(kAccCount) && kReg + 1 < -kStopMax,
              "accumulation word layout");

__device__ __forceinline__ int acc_code(uint32_t w) { return static_cast<int32_t>(w) >> 24; }

// Load the receivers of the region into accumulation words with zero counts
// (every load of a thread is issued before the first LDS store).
template <int BLOCK>
__device__ __forceinline__ void load_region_words(const uint8_t *__restrict__ rcv,
                                                  uint32_t *w_s, int rx0, int ry0, int W,
                                                  int H) {
  constexpr int kPer = kRegCells / BLOCK;
  static_assert(kRegCells % BLOCK == 0, "region must split evenly over the block");
  // a thread's cells are threadIdx.x + k * BLOCK: step the region column and
  // row instead of dividing
  constexpr int kStepY = BLOCK / kReg, kStepX = BLOCK % kReg;
  const int lx0 = threadIdx.x % kReg, ly0 = threadIdx.x / kReg;
  const bool inside = rx0 >= 0 && ry0 >= 0 && rx0 + kReg <= W && ry0 + kReg <= H;
  uint8_t rv[kPer];
  {
    int lx = lx0, ly = ly0;
#pragma unroll
    for (int k = 0; k < kPer; k++) {
      const int gx = rx0 + lx, gy = ry0 + ly;
      const bool in = inside || (gx >= 0 && gx < W && gy >= 0 && gy < H);
      rv[k] = in ? rcv[static_cast<size_t>(gy) * W + gx] : kOutside;
      lx += kStepX;
      ly += kStepY;
      if (lx >= kReg) {
        lx -= kReg;
        ly++;
      }
    }
  }
  int lx = lx0, ly = ly0;
#pragma unroll
  for (int k = 0; k < kPer; k++) {
    const int i = k * BLOCK + threadIdx.x;
    const int r = rv[k];
    int code;
    if (r == kOutside) {
      code = kStopOutside;
    } else if (r >= 8) {
      code = kStopRoot;
    } else {
      int nx = lx, ny = ly;
      code = region_step(r, nx, ny) ? region_offset(r) : kStopLeaves;
    }
    w_s[i] = static_cast<uint32_t>(code & 0xFF) << 24 | (in_tile(lx, ly) ? kAccTile : 0u);
    lx += kStepX;
    ly += kStepY;
    if (lx >= kReg) {
      lx -= kReg;
      ly++;
    }
  }
}

// Bytes of a 4-entry table, first entry lowest.
constexpr uint32_t pack4(int a, int b, int c, int d) {
  return static_cast<uint32_t>(a & 0xFF) | static_cast<uint32_t>(b & 0xFF) << 8 |
         static_cast<uint32_t>(c & 0xFF) << 16 | static_cast<uint32_t>(d & 0xFF) << 24;
}

// Same words as load_region_words, four cells per dword, for a region that
// lies inside the grid with 4-byte aligned rows.  v_perm_b32 looks bytes up
// in an 8-entry table (selector 0..7; 12 gives 0): one lookup turns four
// receiver directions into four offsets, four more flag the directions that
// leave the region through its first/last column and row.
template <int BLOCK>
__device__ __forceinline__ void load_region_words_x4(const uint8_t *__restrict__ rcv,
                                                     uint32_t *w_s, int rx0, int ry0, int W) {
  constexpr int kRowWords = kReg / 4;
  constexpr int kWords = kRegCells / 4;
  constexpr int kIters = (kWords + BLOCK - 1) / BLOCK;
  static_assert(kReg % 4 == 0 && kHalo % 4 == 0 && kTile % 4 == 0, "dword cells");
  // offsets of directions 0..3 (low table) and 4..7 (high table)
  constexpr uint32_t kOffLo = pack4(-1, -kReg - 1, -kReg, -kReg + 1);
  constexpr uint32_t kOffHi = pack4(1, kReg + 1, kReg, kReg - 1);
  // 0xFF for the directions that leave through each side
  constexpr uint32_t kWestLo = pack4(-1, -1, 0, 0), kWestHi = pack4(0, 0, 0, -1);
  constexpr uint32_t kEastLo = pack4(0, 0, 0, -1), kEastHi = pack4(-1, -1, 0, 0);
  constexpr uint32_t kNorthLo = pack4(0, -1, -1, -1), kNorthHi = 0u;
  constexpr uint32_t kSouthLo = 0u, kSouthHi = pack4(0, -1, -1, -1);
  const uint8_t *base = rcv + static_cast<size_t>(ry0) * W + rx0;
  uint32_t v[kIters];
#pragma unroll
  for (int k = 0; k < kIters; k++) {
    const int d = k * BLOCK + threadIdx.x;
    const int row = d / kRowWords, col = d % kRowWords;
    v[k] = d < kWords ? *reinterpret_cast<const uint32_t *>(base + static_cast<size_t>(row) * W +
                                                            4 * col)
                      : 0u;
  }
#pragma unroll
  for (int k = 0; k < kIters; k++) {
    const int d = k * BLOCK + threadIdx.x;
    if (d >= kWords) continue;
    const int row = d / kRowWords, col = d % kRowWords;
    const uint32_t x = v[k];
    // outlets are 0xFF bytes: m is 0xFF in their byte lanes
    const uint32_t rb = (x >> 7) & 0x01010101u;
    const uint32_t m = (rb << 8) - rb;
    const uint32_t sel = (x & 0x07070707u & ~m) | (m & 0x0C0C0C0Cu);
    uint32_t code = __builtin_amdgcn_perm(kOffHi, kOffLo, sel) | (rb << 7);
    uint32_t ex = 0;
    if (col == 0) ex |= __builtin_amdgcn_perm(kWestHi, kWestLo, sel) & 0x000000FFu;
    if (col == kRowWords - 1) ex |= __builtin_amdgcn_perm(kEastHi, kEastLo, sel) & 0xFF000000u;
    if (row == 0) ex |= __builtin_amdgcn_perm(kNorthHi, kNorthLo, sel);
    if (row == kReg - 1) ex |= __builtin_amdgcn_perm(kSouthHi, kSouthLo, sel);
    constexpr uint32_t kLeaves4 = 0x01010101u * static_cast<uint32_t>(kStopLeaves & 0xFF);
    code = (code & ~ex) | (ex & kLeaves4);
    const bool tile = static_cast<unsigned>(row - kHalo) < static_cast<unsigned>(kTile) &&
                      static_cast<unsigned>(col - kHalo / 4) < static_cast<unsigned>(kTile / 4);
    const uint32_t t = tile ? kAccTile : 0u;
    uint4 q;
    q.x = (code << 24) | t;
    q.y = ((code << 16) & 0xFF000000u) | t;
    q.z = ((code << 8) & 0xFF000000u) | t;
    q.w = (code & 0xFF000000u) | t;
    *reinterpret_cast<uint4 *>(&w_s[4 * d]) = q;
  }
}

// Mark (bitmask over region cells) the halo cells on the receiver chains of
// the tile's cells: walks start at tile cells whose receiver leaves the tile
// and stop at an outlet, the region edge, a tile cell, an already marked cell
// or after kHalo hops.  mark must be zero and w_s loaded before the call
// (counts may still be changing: only the top byte and the tile bit are read).
template <int BLOCK>
__device__ __forceinline__ void mark_halo_chains(const uint32_t *w_s, uint32_t *mark) {
  for (int it = 0; it < kTile * kTile / BLOCK; it++) {
    const int j = it * BLOCK + threadIdx.x;
    int c = (kHalo + j / kTile) * kReg + kHalo + j % kTile;
    int code = acc_code(w_s[c]);
    for (int hop = 0; hop <= kHalo && code > kStopMax; hop++) {
      c += code;
      const uint32_t wd = w_s[c];
      if (wd & kAccTile) break;
      const uint32_t bit = 1u << (c & 31);
      if (atomicOr(&mark[c >> 5], bit) & bit) break;   // marked by another walk
      code = acc_code(wd);
    }
  }
}

// Per-tile erosion plan, written by the accumulation pass for the erosion
// pass.  Slots: the tile's cells at their tile index, then the marked halo
// cells in region order.  rslot holds each slot's receiver slot with
// kPlanDiag set for a diagonal receiver, or kPlanRoot for an outlet.
constexpr int kTileCells = kTile * kTile;
constexpr int kSlots = kTileCells + kHaloCap;
constexpr uint16_t kPlanDiag = 0x8000u;
constexpr uint16_t kPlanRoot = 0x7FFFu;
constexpr uint16_t kPlanSlot = 0x7FFFu;       // slot bits of an rslot entry
static_assert(kSlots < kPlanRoot, "slot numbers must fit below kPlanRoot");
struct alignas(16) TilePlan {
  uint16_t rslot[kSlots];
  uint16_t hreg[kHaloCap];     // region index of each halo slot
  uint8_t hlvl[kHaloCap];      // level of each halo slot
  int32_t nhalo;
};

struct AccShared {
  alignas(16) uint32_t w_s[kRegCells];   // accumulation words
  uint32_t mark[kRegCells / 32];
  uint16_t pre[kRegCells / 32];   // marks before each mark word
  uint32_t scan[1];               // number of marks
  int over;
  int hist[kHalo + 1];
};

// Slot of region cell c in the plan: a tile cell, or a marked halo cell.
__device__ __forceinline__ int plan_slot(const uint32_t *mark, const uint16_t *pre, int c) {
  const int lx = c % kReg, ly = c / kReg;
  if (in_tile(lx, ly)) return (ly - kHalo) * kTile + (lx - kHalo);
  return kTileCells + pre[c >> 5] + __popc(mark[c >> 5] & ((1u << (c & 31)) - 1u));
}

// rslot entry of region cell c (a tile cell or a marked halo cell).
__device__ __forceinline__ uint16_t plan_receiver(const uint32_t *w_s, const uint32_t *mark,
                                                  const uint16_t *pre, int c) {
  const int code = acc_code(w_s[c]);
  if (code <= kStopMax) return kPlanRoot;   // an outlet (leaving the region makes the tile over)
  const int s = plan_slot(mark, pre, c + code);
  const bool diag = code != 1 && code != -1 && code != kReg && code != -kReg;
  return static_cast<uint16_t>(s | (diag ? kPlanDiag : 0u));
}

// Returns nonzero (block-uniform) when a tile cell is deeper than kHalo or
// the tile's chains need more than kHaloCap halo cells in the erosion pass;
// otherwise also writes the tile's erosion plan.  With level_hist, adds the
// tile cells' level histogram to it.
template <int BLOCK>
__device__ __forceinline__ int tile_acc_body(AccShared &sh, const uint8_t *__restrict__ rcv,
                                             uint32_t *__restrict__ cnt,
                                             uint8_t *__restrict__ lvl,
                                             TilePlan *__restrict__ plan, int W, int H, int t,
                                             int tiles_x, int32_t *level_hist) {
  static_assert(BLOCK == kAccBlock, "shared layout is sized for kAccBlock");
  uint32_t *w_s = sh.w_s;
  const int x0 = (t % tiles_x) * kTile, y0 = (t / tiles_x) * kTile;
  const int rx0 = x0 - kHalo, ry0 = y0 - kHalo;
  if (threadIdx.x == 0) sh.over = 0;
  if (threadIdx.x < kHalo + 1) sh.hist[threadIdx.x] = 0;
  if (threadIdx.x < kRegCells / 32) sh.mark[threadIdx.x] = 0;
  if (rx0 >= 0 && ry0 >= 0 && rx0 + kReg <= W && ry0 + kReg <= H && W % 4 == 0)
    load_region_words_x4<BLOCK>(rcv, w_s, rx0, ry0, W);
  else
    load_region_words<BLOCK>(rcv, w_s, rx0, ry0, W, H);
  __syncthreads();
  // Every region cell walks its receiver chain and adds 1 to each cell it
  // passes, itself included; the thread's chains advance together, one cell
  // per round.  A walker stops at an outlet (its level is the round it got
  // there), where its chain leaves the region, or after kHalo hops (kHalo + 1
  // rounds).  Walkers hold byte addresses into w_s.
  constexpr int kPer = kRegCells / BLOCK;
  static_assert(kPer <= 16, "walkers per thread");
  int xa[kPer], hops[kPer];
  uint32_t live = 0, rooted = 0;
#pragma unroll
  for (int j = 0; j < kPer; j++) {
    const int i = j * BLOCK + threadIdx.x;
    xa[j] = 4 * i;
    hops[j] = 0;
    if (acc_code(w_s[i]) != kStopOutside) live |= 1u << j;
  }
  for (int round = 0; round <= kHalo && live; round++) {
#pragma unroll
    for (int j = 0; j < kPer; j++) {
      if (!(live & (1u << j))) continue;
      uint32_t *w = reinterpret_cast<uint32_t *>(reinterpret_cast<char *>(w_s) + xa[j]);
      const int code = acc_code(atomicAdd(w, 1u));
      if (code > kStopMax) {
        xa[j] += 4 * code;
      } else {
        if (code == kStopRoot) {
          rooted |= 1u << j;
          hops[j] = round;
        }
        live &= ~(1u << j);
      }
    }
  }
  // the level of a cell is the length of its walk; only slots whose region
  // rows can meet the tile's rows hold tile cells
  int over = 0;
  constexpr int kJ0 = (kHalo * kReg) / BLOCK;
  constexpr int kJ1 = ((kHalo + kTile) * kReg - 1) / BLOCK;
#pragma unroll
  for (int j = kJ0; j <= kJ1; j++) {
    const int i = j * BLOCK + threadIdx.x;
    const int tx = i % kReg - kHalo, ty = i / kReg - kHalo;
    if (tx < 0 || tx >= kTile || ty < 0 || ty >= kTile) continue;
    if (acc_code(w_s[i]) == kStopOutside) continue;
    const size_t gi = static_cast<size_t>(y0 + ty) * W + (x0 + tx);
    if (rooted & (1u << j)) {
      lvl[gi] = static_cast<uint8_t>(hops[j]);
      if (level_hist) atomicAdd(&sh.hist[hops[j]], 1);
    } else {
      over = 1;
    }
  }
  if (over) sh.over = 1;
  mark_halo_chains<BLOCK>(w_s, sh.mark);
  __syncthreads();
  // halo slot bases: one wave scans the mark words, kMarkRun per lane
  constexpr int kMarkWords = kRegCells / 32;
  constexpr int kMarkRun = (kMarkWords + kWave - 1) / kWave;
  if (threadIdx.x < kWave) {
    const int w0 = threadIdx.x * kMarkRun;
    uint32_t v = 0;
#pragma unroll
    for (int k = 0; k < kMarkRun; k++)
      if (w0 + k < kMarkWords) v += __popc(sh.mark[w0 + k]);
    uint32_t x = v;
#pragma unroll
    for (int d = 1; d < kWave; d <<= 1) {
      const uint32_t y = __shfl_up(x, d);
      if (static_cast<int>(threadIdx.x) >= d) x += y;
    }
    uint32_t run = x - v;
#pragma unroll
    for (int k = 0; k < kMarkRun; k++)
      if (w0 + k < kMarkWords) {
        sh.pre[w0 + k] = static_cast<uint16_t>(run);
        run += __popc(sh.mark[w0 + k]);
      }
    if (threadIdx.x == kWave - 1) sh.scan[0] = x;
  }
  __syncthreads();
  const uint32_t nhalo = sh.scan[0];
  const int tile_over = sh.over || nhalo > static_cast<uint32_t>(kHaloCap);
  if (level_hist && threadIdx.x < kHalo + 1 && sh.hist[threadIdx.x])
    atomicAdd(&level_hist[threadIdx.x], sh.hist[threadIdx.x]);
  if (!tile_over) {
    // the plan: every needed cell's receiver slot; halo cells' region index
    // and level
#pragma unroll
    for (int j = 0; j < kPer; j++) {
      const int i = j * BLOCK + threadIdx.x;
      const uint32_t mw = sh.mark[i >> 5];
      if (!(mw & (1u << (i & 31)))) continue;
      const int s = kTileCells + sh.pre[i >> 5] + __popc(mw & ((1u << (i & 31)) - 1u));
      plan->rslot[s] = plan_receiver(w_s, sh.mark, sh.pre, i);
      plan->hreg[s - kTileCells] = static_cast<uint16_t>(i);
      plan->hlvl[s - kTileCells] = static_cast<uint8_t>(hops[j]);
    }
    for (int it = 0; it < kTileCells / BLOCK; it++) {
      const int tj = it * BLOCK + threadIdx.x;
      const int tx = tj % kTile, ty = tj / kTile;
      const int i = (ty + kHalo) * kReg + tx + kHalo;
      const int code = acc_code(w_s[i]);
      uint16_t rs = kPlanRoot;
      if (code > kStopMax) {
        // code = dy * kReg + dx with dx, dy in {-1, 0, 1}
        const int dy = (code > 1) - (code < -1);
        const int dx = code - dy * kReg;
        int sl;
        if (static_cast<unsigned>(tx + dx) < static_cast<unsigned>(kTile) &&
            static_cast<unsigned>(ty + dy) < static_cast<unsigned>(kTile)) {
          sl = tj + dy * kTile + dx;
        } else {
          const int nb = i + code;
          sl = kTileCells + sh.pre[nb >> 5] + __popc(sh.mark[nb >> 5] & ((1u << (nb & 31)) - 1u));
        }
        rs = static_cast<uint16_t>(sl | ((dx != 0 && dy != 0) ? kPlanDiag : 0u));
      }
      plan->rslot[tj] = rs;
    }
    if (threadIdx.x == 0) plan->nhalo = static_cast<int32_t>(nhalo);
  }
  for (int it = 0; it < kTileCells / BLOCK; it++) {
    const int i = it * BLOCK + threadIdx.x;
    const int tx = i % kTile, ty = i / kTile;
    const int gx = x0 + tx, gy = y0 + ty;
    if (gx >= W || gy >= H) continue;
    cnt[static_cast<size_t>(gy) * W + gx] = w_s[(ty + kHalo) * kReg + tx + kHalo] & kAccCount;
  }
  return tile_over;
}

__global__ __launch_bounds__(kAccBlock) void k_tile_acc(const uint8_t *__restrict__ rcv,
                                                        uint32_t *__restrict__ cnt,
                                                        uint8_t *__restrict__ lvl,
                                                        TilePlan *__restrict__ plan, int32_t *hdr,
                                                        int W, int H, int tiles_x, int ntiles) {
  __shared__ AccShared sh;
  const int t = xcd_tile(blockIdx.x, ntiles);
  const int over = tile_acc_body<kAccBlock>(sh, rcv, cnt, lvl, plan + t, W, H, t, tiles_x, nullptr);
  if (threadIdx.x == 0 && over) atomicAdd(&hdr[3], 1);
}

// Pass 2: uplift and erosion (AddUplift RB:285-296, Erode RB:306-328) of the
// tile's cells.  Counts and levels come from pass 1, and so does the tile's
// plan: which halo cells lie on the receiver chains of the tile's cells (at
// most kHaloCap; the chains lie in the region because every level is at most
// kHalo) and every needed cell's receiver slot.  Only the tile's cells and
// those halo cells are solved, in slots, so LDS holds 77 KB and two tiles
// share a CU.  The needed cells are bucketed by level into a list whose
// level segments are padded to whole 64-entry chunks; each wave takes chunks
// in increasing order and each lane solves its cell with the reference's
// Newton iteration as soon as its receiver (one level lower, hence in an
// earlier chunk) is marked done.  With one level per chunk no lane waits on a
// lane of its own wave, and the lowest chunk in progress never waits on an
// unfinished one, so the wavefront always advances; the wait is bounded all
// the same and reported in hdr[2].  Reads h_old, writes h_new: neighbouring
// tiles read this tile's cells as their halo.  Callers that run several
// tiles per workgroup put a barrier between calls.
constexpr uint8_t kDone = 0xFD;
constexpr int kListCap = kSlots + (kHalo + 1) * (kWave - 1);
constexpr int kEroFactCache = 64;             // erosion factors cached in LDS
struct EroShared {
  double hs[kSlots];
  uint32_t c_s[kSlots];
  double ft_s[2 * kEroFactCache];
  uint16_t list[kListCap];
  uint16_t rs_s[kSlots];        // receiver slot (plan rslot)
  uint16_t hreg[kHaloCap];      // region index of each halo slot
  uint8_t d_s[kSlots];          // level, then kDone once solved
  uint32_t scan[kEroBlock / kWave + 1];
  int wcount[kHalo + 1][kEroBlock / kWave];
  int ltotal[kHalo + 1], loff[kHalo + 2];
};

template <int NMODE>
__device__ __forceinline__ void tile_erode_body(EroShared &sh, const double *__restrict__ h_old,
                                                double *__restrict__ h_new,
                                                const TilePlan *__restrict__ plan,
                                                const uint8_t *__restrict__ lvl,
                                                const uint32_t *__restrict__ cnt,
                                                const double *__restrict__ ftab, int W, int H,
                                                int t, int tiles_x, const ErodeConst &k,
                                                int32_t *hdr) {
  constexpr int kWaves = kEroBlock / kWave;
  constexpr int kKeys = kHalo + 1;
  constexpr int kTilePer = kTileCells / kEroBlock;
  constexpr int kSlotPer = (kSlots + kEroBlock - 1) / kEroBlock;
  static_assert(kKeys <= 32, "level keys must fit the 5-bit match");
  static_assert(kHaloCap <= kEroBlock && 2 * kEroFactCache <= kEroBlock, "block shape");
  static_assert(kEroFactCache <= kFactCache, "factor table holds the cached counts");
  double *hs = sh.hs, *ft_s = sh.ft_s;
  uint32_t *c_s = sh.c_s;
  uint16_t *list = sh.list, *rs_s = sh.rs_s;
  uint8_t *d_s = sh.d_s;
  int(*wcount)[kWaves] = sh.wcount;
  int *ltotal = sh.ltotal, *loff = sh.loff;
  const int x0 = (t % tiles_x) * kTile, y0 = (t / tiles_x) * kTile;
  const int rx0 = x0 - kHalo, ry0 = y0 - kHalo;
  const int wave = threadIdx.x / kWave, lane = threadIdx.x % kWave;
  for (int i = threadIdx.x; i < kKeys * kWaves; i += kEroBlock) (&wcount[0][0])[i] = 0;
  if (threadIdx.x < 2 * kEroFactCache) ft_s[threadIdx.x] = ftab[threadIdx.x];
  // the plan and the tile's elevations, counts and levels; a halo cell's
  // elevation and count as soon as its region index arrives
  const int nhalo = plan->nhalo;
  const int nh = nhalo < 0 ? 0 : (nhalo > kHaloCap ? kHaloCap : nhalo);
  {
    uint8_t tl[kTilePer];
    double th[kTilePer];
    uint32_t tc[kTilePer];
    uint16_t rv[kSlotPer];
#pragma unroll
    for (int j = 0; j < kTilePer; j++) {
      const int tj = j * kEroBlock + threadIdx.x;
      const int gx = x0 + tj % kTile, gy = y0 + tj / kTile;
      const bool in = gx < W && gy < H;
      const size_t g = in ? static_cast<size_t>(gy) * W + gx : 0;
      th[j] = in ? h_old[g] : 0.0;
      tc[j] = in ? cnt[g] : 0u;
      tl[j] = in ? lvl[g] : kEscaped;
    }
#pragma unroll
    for (int q = 0; q < kSlotPer; q++) {
      const int sl = q * kEroBlock + threadIdx.x;
      rv[q] = sl < kTileCells + nh ? plan->rslot[sl] : kPlanRoot;
    }
    const bool halo = threadIdx.x < nh;
    const int hr = halo ? plan->hreg[threadIdx.x] : 0;
    const uint8_t hl = halo ? plan->hlvl[threadIdx.x] : 0;
    const size_t hg = static_cast<size_t>(ry0 + hr / kReg) * W + (rx0 + hr % kReg);
    const double hh = halo ? h_old[hg] : 0.0;
    const uint32_t hc = halo ? cnt[hg] : 0u;
#pragma unroll
    for (int j = 0; j < kTilePer; j++) {
      hs[j * kEroBlock + threadIdx.x] = th[j];
      c_s[j * kEroBlock + threadIdx.x] = tc[j];
      d_s[j * kEroBlock + threadIdx.x] = tl[j];
    }
#pragma unroll
    for (int q = 0; q < kSlotPer; q++) {
      const int sl = q * kEroBlock + threadIdx.x;
      if (sl < kSlots) rs_s[sl] = rv[q];
    }
    if (halo) {
      hs[kTileCells + threadIdx.x] = hh;
      c_s[kTileCells + threadIdx.x] = hc;
      d_s[kTileCells + threadIdx.x] = hl;
      sh.hreg[threadIdx.x] = static_cast<uint16_t>(hr);
    }
  }
  if (nh != nhalo) {   // excluded by pass 1
    if (threadIdx.x == 0) atomicOr(&hdr[2], kErrIndex);
    return;
  }
  __syncthreads();
  // bucket the slots by level: rank within (level, wave) by matching keys
  // with five ballots, no atomics; then level bases
  const int nslots = kTileCells + nh;
  const uint64_t lt = (lane == 0) ? 0ull : (~0ull >> (kWave - lane));
  int rank[kSlotPer];
#pragma unroll
  for (int q = 0; q < kSlotPer; q++) {
    rank[q] = -1;
    if (q * kEroBlock + wave * kWave >= nslots) continue;   // the whole wave is past the slots
    const int sl = q * kEroBlock + threadIdx.x;
    const int d = sl < nslots ? d_s[sl] : 0xFF;
    const bool valid = d < kKeys;
    uint64_t m = __ballot(valid);
#pragma unroll
    for (int b = 0; b < 5; b++) {
      // keep the lanes whose bit b equals this lane's: ~(ballot ^ all-ones-if-set)
      const uint64_t bb = __ballot((d >> b) & 1);
      const uint64_t mine = static_cast<uint64_t>(-static_cast<int64_t>((d >> b) & 1));
      m &= ~(bb ^ mine);
    }
    const int before = __popcll(m & lt);
    const int base = valid ? wcount[d][wave] : 0;
    rank[q] = valid ? base + before : -1;
    if (valid && before == 0) wcount[d][wave] = base + __popcll(m);
  }
  __syncthreads();
  if (threadIdx.x < kKeys) {
    int run = 0;
    for (int w = 0; w < kWaves; w++) {
      const int c = wcount[threadIdx.x][w];
      wcount[threadIdx.x][w] = run;
      run += c;
    }
    ltotal[threadIdx.x] = run;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    int run = 0;
    for (int L = 0; L < kKeys; L++) {
      loff[L] = run;
      run += (ltotal[L] + kWave - 1) / kWave * kWave;   // whole chunks per level
    }
    loff[kKeys] = run;
  }
  __syncthreads();
#pragma unroll
  for (int q = 0; q < kSlotPer; q++) {
    if (q * kEroBlock + wave * kWave >= nslots) continue;
    const int sl = q * kEroBlock + threadIdx.x;
    if (rank[q] < 0) continue;
    const int d = d_s[sl];
    list[loff[d] + wcount[d][wave] + rank[q]] = static_cast<uint16_t>(sl);
  }
  if (threadIdx.x < kKeys)
    for (int e = loff[threadIdx.x] + ltotal[threadIdx.x]; e < loff[threadIdx.x + 1]; e++)
      list[e] = 0xFFFF;
  __syncthreads();
  const int nchunks = loff[kKeys] / kWave, root_chunks = loff[1] / kWave;
  for (int ch = wave; ch < nchunks; ch += kWaves) {
    const int sl = list[ch * kWave + lane];
    if (sl == 0xFFFF) continue;
    if (ch < root_chunks) {
      // level 0: uplift of interior outlets; rings 0 and 1 stay fixed
      int gx, gy;
      if (sl < kTileCells) {
        gx = x0 + sl % kTile;
        gy = y0 + sl / kTile;
      } else {
        const int i = sh.hreg[sl - kTileCells];
        gx = rx0 + i % kReg;
        gy = ry0 + i / kReg;
      }
      if (gx >= 2 && gx < W - 2 && gy >= 2 && gy < H - 2) hs[sl] += k.ud;
    } else {
      const int rs = rs_s[sl];
      const int nb = rs & kPlanSlot;
      if (nb < nslots) {                      // always, for a cell above level 0
        // the cell's own inputs do not depend on its receiver: read them
        // before the acquire wait, which later loads may not pass
        const uint32_t a = c_s[sl];
        const int ti = 2 * static_cast<int>(a) + (rs >> 15);
        const double fact = a < kEroFactCache ? ft_s[ti] : ftab[ti];
        const double h0 = hs[sl] + k.ud;
        int spins = 0;
        while (__hip_atomic_load(&d_s[nb], __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_WORKGROUP) !=
               kDone) {
          if (++spins > (1 << 22)) {
            atomicOr(&hdr[2], kErrLookback);
            break;
          }
          __builtin_amdgcn_s_sleep(1);
        }
        hs[sl] = newton_update<NMODE>(h0, hs[nb], fact, k, hdr);
      } else {
        atomicOr(&hdr[2], kErrIndex);
      }
    }
    __hip_atomic_store(&d_s[sl], kDone, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_WORKGROUP);
  }
  __syncthreads();
#pragma unroll
  for (int j = 0; j < kTilePer; j++) {
    const int tj = j * kEroBlock + threadIdx.x;
    const int gx = x0 + tj % kTile, gy = y0 + tj / kTile;
    if (gx >= W || gy >= H) continue;
    h_new[static_cast<size_t>(gy) * W + gx] = hs[tj];
  }
}

template <int NMODE>
__global__ __launch_bounds__(kEroBlock) void k_tile_erode(
    const double *__restrict__ h_old, double *__restrict__ h_new,
    const TilePlan *__restrict__ plan, const uint8_t *__restrict__ lvl,
    const uint32_t *__restrict__ cnt, const double *__restrict__ ftab, int W, int H,
    int tiles_x, int ntiles, ErodeConst k, int32_t *hdr) {
  __shared__ EroShared sh;
  const int t = xcd_tile(blockIdx.x, ntiles);
  tile_erode_body<NMODE>(sh, h_old, h_new, plan + t, lvl, cnt, ftab, W, H, t, tiles_x, k, hdr);
}

}  // namespace

#endif
