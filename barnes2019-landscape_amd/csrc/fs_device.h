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
// After the level counts: one "cells still unfinished" flag per pointer-
// jumping pass, so the passes after the last needed one return at once.
constexpr int kHdrJump = kHdrInts + FS_MAX_LEVELS + 3;
constexpr int kHdrAlloc = kHdrJump + kJumpPasses + 1;   // header ints allocated
constexpr int kTopoBlocks = 2048;
constexpr int kJumpBlocks = 1024;
constexpr int kJumpHist = 2048;         // levels k_jump counts in LDS

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
  if (i <= kJumpPasses) hdr[kHdrJump + i] = 0;
}

// ComputeReceivers (RB:160-182).  Each thread owns one column of a run of
// interior rows (whole strips of kRecvRows) and slides a 3x3 window down it,
// so every elevation is fetched from HBM about once (neighbouring columns
// come from L1, and a run re-reads only its two halo rows).
// fs::receiver_try picks the reference's direction -- steepest
// (hc - h[n]) / dr[n] over n = 0..7 with a strict '>' from 0, ties to the
// lowest n, flat cells and pits NO_FLOW -- without dividing; the rare cells it
// leaves undecided (near-ties, extreme magnitudes) send the column's run
// through fs::receiver_fast again, which divides where it must, so the row
// loop has no branch.  Rows are read and written through buffer resources
// built per row in scalar registers: the lane's column is a constant offset,
// so the loop spends no vector instructions on addresses.
constexpr int kRecvRows = 16;
constexpr int kRecvStrips = 2;                  // strips per thread in k_receivers
constexpr int kRsrcWord3 = 0x00020000;          // raw (unformatted) buffer, gfx9
using v4u = unsigned __attribute__((ext_vector_type(4)));
using v2u = unsigned __attribute__((ext_vector_type(2)));

__device__ __forceinline__ __amdgpu_buffer_rsrc_t row_rsrc(const void *row, uint32_t bytes) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(row), 0, static_cast<int>(bytes),
                                           kRsrcWord3);
}

// columns x-1, x, x+1 of a row; xo = 8*(x-1) (a row of W doubles is below
// 2^32 bytes: W*H < 2^31 with H >= 5)
__device__ __forceinline__ void load_cols(__amdgpu_buffer_rsrc_t r, uint32_t xo, double &l,
                                          double &c, double &rr) {
  const v4u a = __builtin_amdgcn_raw_buffer_load_b128(r, static_cast<int>(xo), 0, 0);
  const v2u b = __builtin_amdgcn_raw_buffer_load_b64(r, static_cast<int>(xo) + 16, 0, 0);
  l = __builtin_bit_cast(double, v2u{a.x, a.y});
  c = __builtin_bit_cast(double, v2u{a.z, a.w});
  rr = __builtin_bit_cast(double, b);
}

// Rows y0 .. y0+n-1 (all interior: 2 <= y0, y0+n <= H-2) of column x.  The
// window holds rows y-1, y, y+1 and the next row's load is in flight while
// row y is decided (row y+2 <= H-1 always exists).
__device__ __forceinline__ void receivers_rows(const double *__restrict__ h,
                                               uint8_t *__restrict__ rcv, int W, int x, int y0,
                                               int n, double d_card, double d_diag) {
  const bool div_ok = fs::divisors_in_range(d_card, d_diag);
  const uint32_t rowb = static_cast<uint32_t>(W) * 8u;
  const uint32_t xo = static_cast<uint32_t>(x - 1) * 8u;
  const double *row = h + static_cast<size_t>(y0 - 1) * W;   // row y-1 of the first cell
  uint8_t *out = rcv + static_cast<size_t>(y0) * W;
  double a0, a1, a2, b0, b1, b2, c0, c1, c2;
  load_cols(row_rsrc(row, rowb), xo, a0, a1, a2);
  load_cols(row_rsrc(row + W, rowb), xo, b0, b1, b2);
  load_cols(row_rsrc(row + 2 * W, rowb), xo, c0, c1, c2);
  uint32_t undecided = 0;
  auto one_row = [&]() {
    double d0, d1, d2;
    load_cols(row_rsrc(row + 3 * W, rowb), xo, d0, d1, d2);
    // neighbours in direction order W, NW, N, NE, E, SE, S, SW (RB:79-82)
    const double nb[8] = {b0, a0, a1, a2, b2, c2, c1, c0};
    bool ok;
    const int r = fs::receiver_try(b1, nb, d_card, d_diag, div_ok, &ok);
    __builtin_amdgcn_raw_buffer_store_b8(static_cast<uint8_t>(r), row_rsrc(out, W), x, 0, 0);
    undecided |= ok ? 0u : 1u;
    // settle the flag here: left symbolic, every row's comparison masks stay
    // live in scalar registers until after the loop
    asm volatile("" : "+v"(undecided));
    row += W;
    out += W;
    a0 = b0; a1 = b1; a2 = b2;
    b0 = c0; b1 = c1; b2 = c2;
    c0 = d0; c1 = d1; c2 = d2;
  };
  int k = 0;
  for (; k + kRecvRows <= n; k += kRecvRows) {
#pragma unroll
    for (int u = 0; u < kRecvRows; u++) one_row();
  }
  for (; k < n; k++) one_row();
  if (undecided) {   // rare: the column's rows again, with the divisions where needed
    for (k = 0; k < n; k++) {
      const double *up = h + static_cast<size_t>(y0 + k - 1) * W + x;
      const double *mid = up + W, *dn = mid + W;
      const double nb[8] = {mid[-1], up[-1], up[0], up[1], mid[1], dn[1], dn[0], dn[-1]};
      rcv[static_cast<size_t>(y0 + k) * W + x] =
          static_cast<uint8_t>(fs::receiver_fast(mid[0], nb, d_card, d_diag, div_ok));
    }
  }
}

// Strips s0 .. s1-1 of kRecvRows interior rows each, as one run of rows.
__device__ __forceinline__ void receivers_strips(const double *__restrict__ h,
                                                 uint8_t *__restrict__ rcv, int W, int H, int x,
                                                 int s0, int s1, double d_card, double d_diag) {
  const int y0 = s0 * kRecvRows + 2;
  const int n = min(s1 * kRecvRows + 2, H - 2) - y0;
  if (n > 0) receivers_rows(h, rcv, W, x, y0, n, d_card, d_diag);
}

// Also clears the header words the tile path's accumulation reports into
// (hdr[0], hdr[1], hdr[3]; as k_fill_counts with no level counts).  Compiled
// for eight waves per SIMD (its waves wait on memory most of their cycles;
// 106 scalar registers held it to fewer).
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(8))) void k_receivers(const double *__restrict__ h,
                                                   uint8_t *__restrict__ rcv, int W, int H,
                                                   double d_card, double d_diag,
                                                   int32_t *__restrict__ hdr) {
  if (blockIdx.x == 0 && blockIdx.y == 0 && threadIdx.x == 0) {
    hdr[0] = 0;
    hdr[1] = 0;
    hdr[3] = 0;
  }
  const int x = blockIdx.x * 256 + threadIdx.x + 2;
  if (x >= W - 2) return;
  const int nstrips = (H - 4 + kRecvRows - 1) / kRecvRows;
  for (int s = blockIdx.y * kRecvStrips; s < nstrips; s += gridDim.y * kRecvStrips)
    receivers_strips(h, rcv, W, H, x, s, min(s + kRecvStrips, nstrips), d_card, d_diag);
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
// consistent and the update is safe in place.  A pass that leaves a cell
// unfinished raises the next pass's flag; a pass whose flag is down (the
// previous pass finished every cell) returns at once.
__global__ __launch_bounds__(256) void k_jump(const int32_t *__restrict__ list,
                                              uint16_t *depth, unsigned long long *jump,
                                              int32_t *hdr, int pass) {
  // level counts below kJumpHist and the deepest level are folded per block
  // in LDS: finished cells share a few hundred levels, so global atomics per
  // cell would queue on the same words
  __shared__ int hist[kJumpHist];
  __shared__ int smax, smore;
  if (pass > 0 &&
      __hip_atomic_load(&hdr[kHdrJump + pass], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == 0)
    return;
  for (int i = threadIdx.x; i < kJumpHist; i += blockDim.x) hist[i] = 0;
  if (threadIdx.x == 0) smax = smore = 0;
  __syncthreads();
  const int n = hdr[1];
  int32_t *counts = hdr + kHdrInts;
  int local_max = 0, more = 0;
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
      if (nd < static_cast<unsigned>(kJumpHist)) atomicAdd(&hist[nd], 1);
      else atomicAdd(&counts[nd], 1);
      local_max = max(local_max, static_cast<int>(nd));
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
      more = 1;
    }
  }
  if (local_max > 0) atomicMax(&smax, local_max);
  if (more) smore = 1;
  __syncthreads();
  const int top = min(smax, kJumpHist - 1);
  for (int i = threadIdx.x; i <= top; i += blockDim.x)
    if (hist[i]) atomicAdd(&counts[i], hist[i]);
  if (threadIdx.x == 0) {
    if (smax > 0) atomicMax(&hdr[0], smax);
    if (smore)
      __hip_atomic_store(&hdr[kHdrJump + pass + 1], 1, __ATOMIC_RELAXED,
                         __HIP_MEMORY_SCOPE_AGENT);
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

// Stack-order records of the level path, written as the breadth-first
// order is built: for stack position p, fdon[p] is where its donor list was
// appended and ndon[p] its length; for every appended donor, rdir is its
// receiver direction 0..7 (RB:79-82).  The donor in mask slot j drains in
// direction kSlotDir[j] (slot 0, the north-west neighbour, drains south-east
// and so on).
__device__ __forceinline__ int slot_dir(int j) {
  constexpr uint32_t kSlotDir = 5u | 6u << 3 | 7u << 6 | 4u << 9 | 0u << 12 | 3u << 15 |
                                2u << 18 | 1u << 21;
  return static_cast<int>((kSlotDir >> (3 * j)) & 7u);
}

__device__ __forceinline__ void append_donors(int32_t c, uint32_t m, int32_t p, long long pos,
                                              int32_t *__restrict__ stack,
                                              uint8_t *__restrict__ rdir,
                                              int32_t *__restrict__ fdon,
                                              uint8_t *__restrict__ ndon, int W, int32_t N) {
  fdon[p] = static_cast<int32_t>(pos < N ? pos : N);
  ndon[p] = static_cast<uint8_t>(__popc(m));
  while (m) {
    const int j = __builtin_ctz(m);
    m &= m - 1;
    if (pos < N) {
      stack[pos] = c + slot_offset(j, W);
      rdir[pos] = static_cast<uint8_t>(slot_dir(j));
    }
    pos++;
  }
}

// One BFS level of GenerateOrder (RB:234-247): stack[ls, le) -> the donor
// lists of those cells, in order, written from stack[le].
__global__ __launch_bounds__(kScanBlock) void k_bfs(
    int32_t *__restrict__ stack, const uint8_t *__restrict__ dmask, uint8_t *__restrict__ rdir,
    int32_t *__restrict__ fdon, uint8_t *__restrict__ ndon, int32_t ls, int32_t le, int W,
    int32_t N, uint64_t *status, unsigned long long *ticket, unsigned long long tbase,
    uint32_t epoch, int32_t *hdr) {
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
    const long long p = p0 + i;
    if (p >= le) break;
    const uint32_t m = static_cast<uint32_t>(masks >> (8 * i)) & 0xFFu;
    append_donors(cell[i], m, static_cast<int32_t>(p), pos, stack, rdir, fdon, ndon, W, N);
    pos += __popc(m);
  }
}

// Levels [L0, L1) of the BFS in one workgroup, for runs of small levels.
__global__ __launch_bounds__(kSmallBlock) void k_bfs_small(
    int32_t *__restrict__ stack, const uint8_t *__restrict__ dmask, uint8_t *__restrict__ rdir,
    int32_t *__restrict__ fdon, uint8_t *__restrict__ ndon, const int32_t *__restrict__ levels,
    int L0, int L1, int W, int32_t N, int32_t *hdr) {
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
        const long long p = p0 + i;
        if (p >= le) break;
        const uint32_t m = (masks >> (8 * i)) & 0xFFu;
        append_donors(cell[i], m, static_cast<int32_t>(p), pos, stack, rdir, fdon, ndon, W, N);
        pos += __popc(m);
      }
      out += total;
    }
    __syncthreads();
  }
}

// Upstream cell count of stack entry p: 1 + the counts of its donors, one
// level higher, already final and contiguous from fdon[p].  Counts are kept
// in stack order (cnt_s); accum = count * cell_area exactly (RB:262-279:
// every partial sum there is an integer multiple of cell_area).  Entries of
// the deepest level have no donors (the breadth-first pass never records
// them): top is true for those.
__device__ __forceinline__ void flowacc_entry(int32_t p, bool top,
                                              const int32_t *__restrict__ fdon,
                                              const uint8_t *__restrict__ ndon, uint32_t *cnt_s,
                                              int32_t N, int32_t *hdr) {
  if (top) {
    cnt_s[p] = 1;
    return;
  }
  const int32_t f = fdon[p];
  const int k = ndon[p];
  uint32_t s = 1;
  if (f > p && static_cast<long long>(f) + k <= N) {
    for (int i = 0; i < k; i++) s += cnt_s[f + i];
  } else if (k > 0) {
    atomicOr(&hdr[2], kErrIndex);
  }
  cnt_s[p] = s;
}

__global__ __launch_bounds__(256) void k_flowacc(const int32_t *__restrict__ fdon,
                                                 const uint8_t *__restrict__ ndon,
                                                 uint32_t *cnt_s, int32_t ls, int32_t le,
                                                 bool top, int32_t N, int32_t *hdr) {
  const long long p = static_cast<long long>(ls) + blockIdx.x * 256ll + threadIdx.x;
  if (p >= le) return;
  flowacc_entry(static_cast<int32_t>(p), top, fdon, ndon, cnt_s, N, hdr);
}

// Levels L_hi down to L_lo (inclusive) in one workgroup; L_top is the
// deepest level of the step.
__global__ __launch_bounds__(kSmallBlock) void k_flowacc_small(
    const int32_t *__restrict__ fdon, const uint8_t *__restrict__ ndon, uint32_t *cnt_s,
    const int32_t *__restrict__ levels, int L_hi, int L_lo, int L_top, int32_t N,
    int32_t *hdr) {
  for (int L = L_hi; L >= L_lo; L--) {
    const int32_t ls = levels[L], le = levels[L + 1];
    for (long long p = ls + threadIdx.x; p < le; p += kSmallBlock)
      flowacc_entry(static_cast<int32_t>(p), L == L_top, fdon, ndon, cnt_s, N, hdr);
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

// Erode for stack entry p (RB:306-328): the cell's new elevation from
// h + U*dt and its receiver's final elevation (one level lower, so already
// written; it sits next to the cell in the grid, usually in the same cache
// line).  The count and the receiver direction come from stack order.
template <int NMODE>
__device__ __forceinline__ void erode_entry(int32_t p, const int32_t *__restrict__ stack,
                                            const uint8_t *__restrict__ rdir,
                                            const uint32_t *__restrict__ cnt_s,
                                            const double *__restrict__ ftab, double *h, int W,
                                            int32_t N, const ErodeConst &k, int32_t *hdr) {
  const int32_t c = stack[p];
  const int r = rdir[p] & 7;
  uint32_t a = cnt_s[p];
  const int32_t nb = c + dir_offset(r, W);
  if (static_cast<uint32_t>(c) >= static_cast<uint32_t>(N) ||
      static_cast<uint32_t>(nb) >= static_cast<uint32_t>(N) || a > static_cast<uint32_t>(N)) {
    atomicOr(&hdr[2], kErrIndex);
    return;
  }
  h[c] = newton_update<NMODE>(h[c] + k.ud, h[nb], ftab[2 * a + (r & 1)], k, hdr);
}

template <int NMODE>
__global__ __launch_bounds__(256) void k_erode(const int32_t *__restrict__ stack,
                                               const uint8_t *__restrict__ rdir,
                                               const uint32_t *__restrict__ cnt_s,
                                               const double *__restrict__ ftab, double *h,
                                               int32_t ls, int32_t le, int W, int32_t N,
                                               ErodeConst k, int32_t *hdr) {
  const long long p = static_cast<long long>(ls) + blockIdx.x * 256ll + threadIdx.x;
  if (p >= le) return;
  erode_entry<NMODE>(static_cast<int32_t>(p), stack, rdir, cnt_s, ftab, h, W, N, k, hdr);
}

template <int NMODE>
__global__ __launch_bounds__(kSmallBlock) void k_erode_small(
    const int32_t *__restrict__ stack, const uint8_t *__restrict__ rdir,
    const uint32_t *__restrict__ cnt_s, const double *__restrict__ ftab, double *h,
    const int32_t *__restrict__ levels, int L_lo, int L_hi, int W, int32_t N, ErodeConst k,
    int32_t *hdr) {
  for (int L = L_lo; L <= L_hi; L++) {
    const int32_t ls = levels[L], le = levels[L + 1];
    for (long long p = ls + threadIdx.x; p < le; p += kSmallBlock)
      erode_entry<NMODE>(static_cast<int32_t>(p), stack, rdir, cnt_s, ftab, h, W, N, k, hdr);
    __syncthreads();
  }
}

// ---------------------------------------------------------------------------
// depression filling (an extension: the reference has none)
// ---------------------------------------------------------------------------
// The Priority-Flood+epsilon surface (Barnes, Lehman & Mulla 2014): the
// unique fixed point of F = max(h, nextafter(min over the D8 neighbours of F,
// +inf)) on interior cells, rings 0 and 1 held at h.  Starting from +inf the
// update only lowers F, so tiles relax it in LDS against their neighbours'
// last values (reads may see older, higher values: the iteration still
// decreases to the fixed point); the host relaunches until a launch changes
// nothing, which leaves every cell consistent with its neighbours' final
// values.
constexpr int kFillTile = 64;
constexpr int kFillReg = kFillTile + 2;
constexpr int kFillBlock = 256;
constexpr int kFillSweeps = 512;          // local sweeps per launch at most

// nextafter(x, +inf) for IEEE doubles.
__device__ __forceinline__ double next_up(double x) {
  if (!(x < __builtin_huge_val())) return x;          // +inf, NaN
  if (x == 0.0) return 0x1p-1074;                      // +-0 -> smallest subnormal
  const long long u = __double_as_longlong(x);
  return __longlong_as_double(x > 0.0 ? u + 1 : u - 1);
}

// F = h on rings 0 and 1, +inf inside.
__global__ __launch_bounds__(256) void k_fill_init(const double *__restrict__ h,
                                                   double *__restrict__ F, int W, int H) {
  const long long N = static_cast<long long>(W) * H;
  for (long long c = blockIdx.x * 256ll + threadIdx.x; c < N; c += gridDim.x * 256ll) {
    const int y = static_cast<int>(c / W), x = static_cast<int>(c - static_cast<long long>(y) * W);
    const bool inner = x >= 2 && x < W - 2 && y >= 2 && y < H - 2;
    F[c] = inner ? __builtin_huge_val() : h[c];
  }
}

__global__ __launch_bounds__(kFillBlock) void k_fill(const double *__restrict__ h, double *F,
                                                     int W, int H, int tiles_x,
                                                     int32_t *changed) {
  __shared__ double f_s[kFillReg * kFillReg];
  __shared__ int s_more, s_any;
  constexpr int kPer = kFillTile * kFillTile / kFillBlock;
  const int x0 = (blockIdx.x % tiles_x) * kFillTile, y0 = (blockIdx.x / tiles_x) * kFillTile;
  for (int i = threadIdx.x; i < kFillReg * kFillReg; i += kFillBlock) {
    const int gx = x0 - 1 + i % kFillReg, gy = y0 - 1 + i / kFillReg;
    f_s[i] = (gx >= 0 && gx < W && gy >= 0 && gy < H)
                 ? __hip_atomic_load(&F[static_cast<size_t>(gy) * W + gx], __ATOMIC_RELAXED,
                                     __HIP_MEMORY_SCOPE_AGENT)
                 : __builtin_huge_val();
  }
  double hv[kPer];
  bool inner[kPer];
#pragma unroll
  for (int k = 0; k < kPer; k++) {
    const int i = k * kFillBlock + threadIdx.x;
    const int gx = x0 + i % kFillTile, gy = y0 + i / kFillTile;
    inner[k] = gx >= 2 && gx < W - 2 && gy >= 2 && gy < H - 2;
    hv[k] = inner[k] ? h[static_cast<size_t>(gy) * W + gx] : 0.0;
  }
  if (threadIdx.x == 0) s_any = 0;
  __syncthreads();
  for (int sweep = 0; sweep < kFillSweeps; sweep++) {
    if (threadIdx.x == 0) s_more = 0;
    __syncthreads();
    int more = 0;
#pragma unroll
    for (int k = 0; k < kPer; k++) {
      if (!inner[k]) continue;
      const int i = k * kFillBlock + threadIdx.x;
      const int r = (i / kFillTile + 1) * kFillReg + i % kFillTile + 1;
      double m = f_s[r - kFillReg - 1];
      m = fmin(m, f_s[r - kFillReg]);
      m = fmin(m, f_s[r - kFillReg + 1]);
      m = fmin(m, f_s[r - 1]);
      m = fmin(m, f_s[r + 1]);
      m = fmin(m, f_s[r + kFillReg - 1]);
      m = fmin(m, f_s[r + kFillReg]);
      m = fmin(m, f_s[r + kFillReg + 1]);
      const double v = fmax(hv[k], next_up(m));
      if (v < f_s[r]) {
        f_s[r] = v;
        more = 1;
      }
    }
    if (more) s_more = 1;
    __syncthreads();
    const int go = s_more;
    if (go && threadIdx.x == 0) s_any = 1;
    __syncthreads();
    if (!go) break;
  }
  if (!s_any) return;
#pragma unroll
  for (int k = 0; k < kPer; k++) {
    if (!inner[k]) continue;
    const int i = k * kFillBlock + threadIdx.x;
    const int gx = x0 + i % kFillTile, gy = y0 + i / kFillTile;
    const int r = (i / kFillTile + 1) * kFillReg + i % kFillTile + 1;
    __hip_atomic_store(&F[static_cast<size_t>(gy) * W + gx], f_s[r], __ATOMIC_RELAXED,
                       __HIP_MEMORY_SCOPE_AGENT);
  }
  if (threadIdx.x == 0) atomicOr(changed, 1);
}

// ---------------------------------------------------------------------------
// phase timing of the tile kernels
// ---------------------------------------------------------------------------
// The tile path fuses the reference's stages (donors, order, accumulation,
// uplift, erosion; RB:367-372) into two kernels.  To report the reference's
// per-stage timers, the timed instantiations of the tile kernels (run on a
// sample of steps) let one workgroup in kPhaseSample add the wall-clock
// ticks of each of its phases (between barriers, never inside a loop:
// reading the clock is a slow scalar-memory access) to a row of `phase`
// (rows spread the atomics); the host splits the kernels' event times over
// the stages in proportion.
constexpr int kPhaseSample = 16;     // workgroups per sampled one in a timed launch
constexpr int kPhaseEvery = 32;      // tile steps per timed step
constexpr int kPhaseSlots = 32;
constexpr int kPhaseStride = 8;
enum PhaseId {
  kPhAccLoad = 0,    // receivers -> region words (the donor structure)   Donors
  kPhAccWalk,        // receiver-chain walks: counts and levels          FlowAcc
  kPhAccPlan,        // halo chains, slot scan, erosion plan             Order
  kPhAccStore,       // count and level stores                           FlowAcc
  kPhEroBucket,      // loads and level bucketing of the slots           Order
  kPhEroRoots,       // uplift of the outlets                            Uplift
  kPhEroLoop,        // the erosion wavefront                            Erosion
  kPhEroStore,       // elevation stores                                 Erosion
  kPhCount
};
static_assert(kPhCount <= kPhaseStride, "phase row");

// This workgroup's phase row, or null when it is not sampled.
__device__ __forceinline__ unsigned long long *phase_row(unsigned long long *phase) {
  if (!phase || blockIdx.x % kPhaseSample != 0) return nullptr;
  return phase + (blockIdx.x / kPhaseSample % kPhaseSlots) * kPhaseStride;
}

// ---------------------------------------------------------------------------
// tile path, 16-cell halo (fs_tile.h)
// ---------------------------------------------------------------------------
#define FS_TILE_HALO 16
#include "fs_tile.h"
#undef FS_TILE_HALO

}  // namespace

#endif
