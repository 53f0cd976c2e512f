// fs_deep.h -- the tile-graph path: a time-step for drainage of any depth,
// tile by tile with no halo.  Included once by fastscape_hip.hip, after
// fs_device.h.
//
// Within a 64x64 tile, each cell's receiver chain is followed only while it
// stays in the tile.  The chain ends at the cell's tail: a NO_FLOW cell of
// the tile (a root) or the last tile cell before the chain leaves the tile
// (an exit).  Exits are the nodes of a forest that spans the tiles: the
// parent of exit x is the tail of x's receiver when that tail is an exit.
//
//   k_deep_local  per tile: the in-tile walks give every cell its local count
//                 (the tile cells upstream of it along in-tile paths, itself
//                 included), its tail and its hops to the tail; exits become
//                 nodes; border cells record the node of their tail (link)
//   k_deep_links  every node's parent, and the number of children per node
//   k_deep_tiers  every node's tier: 1 + its parent's, 1 under a root tail.
//                 Cells whose tail is a tier-t exit are eroded by launch t;
//                 per tier, the list of tiles that hold such cells
//   k_deep_sum    the count through every exit: its local count plus its
//                 children's, leaf to root (the last child to arrive carries
//                 the sum on, so each node is summed once)
//   k_deep_final  per tile: the counts finished by the inflows of the ring
//                 cells that drain into the tile, each walked down its
//                 in-tile chain; then the erosion of the cells with a root
//                 tail, level by level from the root
//   k_deep_erode  launch t >= 1, per listed tile: the erosion of the cells
//                 whose tail is a tier-t exit, from the exit's receiver
//                 (eroded by launch t-1) up
//
// Counts are integers, so the order of the additions is free: they equal
// the reference's ComputeFlowAcc (RB:262-279) counts.  Every cell is eroded
// after its receiver with the reference's Newton iteration (RB:306-328) from
// the same inputs (its elevation plus uplift, the receiver's new elevation,
// the factor of its count), so the new elevations are the reference's bits.
#ifndef FS_DEEP_H
#define FS_DEEP_H

namespace deep {

constexpr int kSide = 64;
constexpr int kCells = kSide * kSide;
constexpr int kBlock = 1024;
constexpr int kPer = kCells / kBlock;
constexpr int kNodeCap = 256;             // nodes per tile: a border has at most 252 cells
constexpr uint16_t kTermNone = 0xFFFD;    // beyond the grid
constexpr uint16_t kTermExit = 0xFFFE;    // the receiver lies in another tile
constexpr uint16_t kTermRoot = 0xFFFF;    // NO_FLOW
constexpr uint32_t kNoNode = 0xFFFFFFFFu;
constexpr int kSumShift = 40;             // node word: arrivals << 40 | count
constexpr uint64_t kSumMask = (uint64_t{1} << kSumShift) - 1;
constexpr uint64_t kOneArrival = uint64_t{1} << kSumShift;
constexpr int kMaxTier = 1023;            // erosion launches per step after the first
constexpr int kTierWords = kMaxTier + 2;  // [0] error bits, [1..kMaxTier] tiles, [kMaxTier+1] max
constexpr int32_t kErrTier = 1, kErrChain = 2;
static_assert(kCells <= 4096, "tail and hops fit 12 bits each");

// dinfo word of a cell: tail (12 bits) | hops to the tail << 12 | exit tail
constexpr uint32_t kInfoExit = 1u << 24;
__device__ __forceinline__ int info_tail(uint32_t w) { return static_cast<int>(w & 0xFFFu); }
__device__ __forceinline__ int info_hops(uint32_t w) { return static_cast<int>((w >> 12) & 0xFFFu); }

// Receiver direction r = 0..7 (W, NW, N, NE, E, SE, S, SW) as a column and
// row step.
__device__ __forceinline__ int dir_dx(int r) {
  constexpr unsigned kDx = (0u << 0) | (0u << 2) | (1u << 4) | (2u << 6) |
                           (2u << 8) | (2u << 10) | (1u << 12) | (0u << 14);
  return static_cast<int>((kDx >> (2 * r)) & 3u) - 1;
}
__device__ __forceinline__ int dir_dy(int r) {
  constexpr unsigned kDy = (1u << 0) | (0u << 2) | (0u << 4) | (0u << 6) |
                           (1u << 8) | (2u << 10) | (2u << 12) | (2u << 14);
  return static_cast<int>((kDy >> (2 * r)) & 3u) - 1;
}

struct Box {
  int x0, y0, tw, th;   // origin and extent of the tile (partial at the grid's edge)
};
__device__ __forceinline__ Box tile_box(int t, int tiles_x, int W, int H) {
  Box b;
  b.x0 = (t % tiles_x) * kSide;
  b.y0 = (t / tiles_x) * kSide;
  b.tw = min(kSide, W - b.x0);
  b.th = min(kSide, H - b.y0);
  return b;
}

// In-tile successor of tile cell i (column lx, row ly) with receiver byte r.
__device__ __forceinline__ uint16_t tile_next(uint8_t r, int lx, int ly, const Box &b) {
  if (r >= 8) return kTermRoot;
  const int nx = lx + dir_dx(r), ny = ly + dir_dy(r);
  if (static_cast<unsigned>(nx) >= static_cast<unsigned>(b.tw) ||
      static_cast<unsigned>(ny) >= static_cast<unsigned>(b.th))
    return kTermExit;
  return static_cast<uint16_t>(ny * kSide + nx);
}

__device__ __forceinline__ int block_max(int v, int *slot) {
#pragma unroll
  for (int d = 1; d < kWave; d <<= 1) v = max(v, __shfl_xor(v, d));
  if ((threadIdx.x & (kWave - 1)) == 0) atomicMax(slot, v);
  __syncthreads();
  return *slot;
}

// ---------------------------------------------------------------------------
// k_deep_local: one tile per workgroup
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(kBlock) void k_deep_local(
    const uint8_t *__restrict__ rcv, uint32_t *__restrict__ cnt, uint32_t *__restrict__ dinfo,
    uint32_t *__restrict__ link, uint64_t *__restrict__ node_word,
    uint32_t *__restrict__ node_kids, int32_t *__restrict__ node_cell,
    int32_t *__restrict__ nexit, int32_t *__restrict__ tstat, int W, int H, int tiles_x) {
  __shared__ uint16_t nx[kCells];
  __shared__ uint32_t lc[kCells];
  __shared__ uint16_t nid[kCells];   // node number of each exit within the tile
  __shared__ int nnode;
  const int t = blockIdx.x;
  const Box b = tile_box(t, tiles_x, W, H);
  if (threadIdx.x == 0) nnode = 0;
#pragma unroll
  for (int j = 0; j < kPer; j++) {
    const int i = j * kBlock + threadIdx.x;
    const int lx = i % kSide, ly = i / kSide;
    uint16_t n = kTermNone;
    if (lx < b.tw && ly < b.th)
      n = tile_next(rcv[static_cast<size_t>(b.y0 + ly) * W + b.x0 + lx], lx, ly, b);
    nx[i] = n;
    lc[i] = 0;
  }
  __syncthreads();
  // every cell walks its in-tile chain and adds 1 to each cell it passes,
  // itself and its tail included; the thread's walkers advance together
  int cur[kPer], hops[kPer];
  uint32_t live = 0;
#pragma unroll
  for (int j = 0; j < kPer; j++) {
    const int i = j * kBlock + threadIdx.x;
    cur[j] = i;
    hops[j] = 0;
    if (nx[i] != kTermNone) live |= 1u << j;
  }
  bool bad = false;
  while (live) {
#pragma unroll
    for (int j = 0; j < kPer; j++) {
      if (!((live >> j) & 1u)) continue;
      atomicAdd(&lc[cur[j]], 1u);
      const uint16_t n = nx[cur[j]];
      if (n >= kTermNone) {
        live &= ~(1u << j);
      } else if (hops[j] + 1 >= kCells) {   // a receiver cycle: not a forest
        bad = true;
        live &= ~(1u << j);
      } else {
        cur[j] = n;
        hops[j]++;
      }
    }
  }
  if (bad) atomicOr(&tstat[0], kErrChain);
  __syncthreads();
  // exits become the tile's nodes
#pragma unroll
  for (int j = 0; j < kPer; j++) {
    const int i = j * kBlock + threadIdx.x;
    const bool ex = nx[i] == kTermExit;
    const uint64_t m = __ballot(ex);
    int base = 0;
    if ((threadIdx.x & (kWave - 1)) == 0 && m) base = atomicAdd(&nnode, __popcll(m));
    base = __shfl(base, 0);
    if (ex) {
      const int lane = threadIdx.x & (kWave - 1);
      const int k = base + __popcll(m & ((uint64_t{1} << lane) - 1));
      nid[i] = static_cast<uint16_t>(k);
      const int id = t * kNodeCap + k;
      node_cell[id] = (b.y0 + i / kSide) * W + b.x0 + i % kSide;
      node_word[id] = lc[i];
      node_kids[id] = 0;
    }
  }
  __syncthreads();
#pragma unroll
  for (int j = 0; j < kPer; j++) {
    const int i = j * kBlock + threadIdx.x;
    if (nx[i] == kTermNone) continue;
    const int lx = i % kSide, ly = i / kSide;
    const size_t g = static_cast<size_t>(b.y0 + ly) * W + b.x0 + lx;
    const bool exit_tail = nx[cur[j]] == kTermExit;
    cnt[g] = lc[i];
    dinfo[g] = static_cast<uint32_t>(cur[j]) | (static_cast<uint32_t>(hops[j]) << 12) |
               (exit_tail ? kInfoExit : 0u);
    if (lx == 0 || ly == 0 || lx == b.tw - 1 || ly == b.th - 1)
      link[g] = exit_tail ? static_cast<uint32_t>(t * kNodeCap + nid[cur[j]]) : kNoNode;
  }
  if (threadIdx.x == 0) nexit[t] = nnode;
}

// ---------------------------------------------------------------------------
// the node forest: one tile's nodes per workgroup of kNodeCap threads
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(kNodeCap) void k_deep_links(
    const uint8_t *__restrict__ rcv, const uint32_t *__restrict__ link,
    const int32_t *__restrict__ node_cell, const int32_t *__restrict__ nexit,
    uint32_t *__restrict__ node_parent, uint32_t *__restrict__ node_kids, int W) {
  const int t = blockIdx.x, k = threadIdx.x;
  if (k >= nexit[t]) return;
  const int id = t * kNodeCap + k;
  const int32_t g = node_cell[id];
  // the receiver lies on the border of the neighbouring tile
  const uint32_t p = link[g + dir_offset(rcv[g], W)];
  node_parent[id] = p;
  if (p != kNoNode) atomicAdd(&node_kids[p], 1u);
}

// Tier of every node, by walking its parents (a chain of tiles crossed by
// one drainage path); tiles listed once per tier they hold.
__global__ __launch_bounds__(kNodeCap) void k_deep_tiers(
    const uint32_t *__restrict__ node_parent, const int32_t *__restrict__ nexit,
    uint16_t *__restrict__ node_tier, int32_t *__restrict__ tstat, int32_t *__restrict__ tlist,
    int ntiles) {
  __shared__ int tier_s[kNodeCap];
  const int t = blockIdx.x, k = threadIdx.x;
  const int n = nexit[t];
  int tier = 0;
  if (k < n) {
    const int id = t * kNodeCap + k;
    tier = 1;
    uint32_t p = node_parent[id];
    while (p != kNoNode && tier <= kMaxTier) {
      tier++;
      p = node_parent[p];
    }
    node_tier[id] = static_cast<uint16_t>(min(tier, kMaxTier + 1));
    tier_s[k] = tier;
  }
  __syncthreads();
  if (k >= n) return;
  if (tier > kMaxTier) {
    atomicOr(&tstat[0], kErrTier);
    return;
  }
  for (int q = 0; q < k; q++)
    if (tier_s[q] == tier) return;   // an earlier node of the tile lists it
  const int pos = atomicAdd(&tstat[tier], 1);
  tlist[static_cast<size_t>(tier) * ntiles + pos] = t;
  atomicMax(&tstat[kMaxTier + 1], tier);
}

// Count through every exit, leaf to root.
__global__ __launch_bounds__(kNodeCap) void k_deep_sum(
    uint64_t *__restrict__ node_word, const uint32_t *__restrict__ node_kids,
    const uint32_t *__restrict__ node_parent, const int32_t *__restrict__ nexit,
    uint32_t *__restrict__ node_out, int32_t *__restrict__ tstat, int nnodes) {
  const int t = blockIdx.x, k = threadIdx.x;
  if (k >= nexit[t]) return;
  uint32_t id = static_cast<uint32_t>(t * kNodeCap + k);
  if (node_kids[id] != 0) return;          // summed by its last child's thread
  uint64_t val = node_word[id] & kSumMask;
  for (int guard = 0; guard <= nnodes; guard++) {
    node_out[id] = static_cast<uint32_t>(val);
    const uint32_t p = node_parent[id];
    if (p == kNoNode) return;
    const uint64_t old = atomicAdd(&node_word[p], val + kOneArrival);
    if ((old >> kSumShift) + 1 != node_kids[p]) return;   // not the last child
    val += old & kSumMask;
    id = p;
  }
  atomicOr(&tstat[0], kErrChain);
}

// ---------------------------------------------------------------------------
// erosion
// ---------------------------------------------------------------------------
struct FinalShared {
  double hs[kCells];
  uint32_t cs[kCells];
  uint16_t nx[kCells];
  int maxhops;
};

template <int NMODE>
__global__ __launch_bounds__(kBlock) void k_deep_final(
    const uint8_t *__restrict__ rcv, uint32_t *__restrict__ cnt,
    const uint32_t *__restrict__ dinfo, const uint32_t *__restrict__ link,
    const uint32_t *__restrict__ node_out, const double *__restrict__ ftab,
    const double *__restrict__ h_old, double *__restrict__ h_new, int W, int H, int tiles_x,
    ErodeConst k, int32_t *hdr) {
  __shared__ FinalShared sh;
  const int t = blockIdx.x;
  const Box b = tile_box(t, tiles_x, W, H);
  uint32_t info[kPer];
  uint8_t rr[kPer];
  if (threadIdx.x == 0) sh.maxhops = 0;
#pragma unroll
  for (int j = 0; j < kPer; j++) {
    const int i = j * kBlock + threadIdx.x;
    const int lx = i % kSide, ly = i / kSide;
    uint16_t n = kTermNone;
    rr[j] = kNoFlow;
    info[j] = 0;
    if (lx < b.tw && ly < b.th) {
      const size_t g = static_cast<size_t>(b.y0 + ly) * W + b.x0 + lx;
      rr[j] = rcv[g];
      n = tile_next(rr[j], lx, ly, b);
      info[j] = dinfo[g];
      sh.cs[i] = cnt[g];
      sh.hs[i] = h_old[g];
    }
    sh.nx[i] = n;
  }
  __syncthreads();
  // inflows: a ring cell that drains into the tile is an exit of its own
  // tile; its count is added along the in-tile chain from the cell it
  // drains into
  {
    const int q = threadIdx.x;
    const int top = b.tw + 2, nring = 2 * top + 2 * b.th;
    if (q < nring) {
      int gx, gy;
      if (q < top) {
        gx = b.x0 - 1 + q;
        gy = b.y0 - 1;
      } else if (q < 2 * top) {
        gx = b.x0 - 1 + (q - top);
        gy = b.y0 + b.th;
      } else if (q < 2 * top + b.th) {
        gx = b.x0 - 1;
        gy = b.y0 + (q - 2 * top);
      } else {
        gx = b.x0 + b.tw;
        gy = b.y0 + (q - 2 * top - b.th);
      }
      if (gx >= 0 && gy >= 0 && gx < W && gy < H) {
        const size_t g = static_cast<size_t>(gy) * W + gx;
        const uint8_t r = rcv[g];
        if (r < 8) {
          const int tx = gx + dir_dx(r) - b.x0, ty = gy + dir_dy(r) - b.y0;
          if (static_cast<unsigned>(tx) < static_cast<unsigned>(b.tw) &&
              static_cast<unsigned>(ty) < static_cast<unsigned>(b.th)) {
            const uint32_t v = node_out[link[g]];
            int cur = ty * kSide + tx;
            for (int hop = 0; hop < kCells; hop++) {
              atomicAdd(&sh.cs[cur], v);
              const uint16_t n = sh.nx[cur];
              if (n >= kTermNone) break;
              cur = n;
            }
          }
        }
      }
    }
  }
  int mh = 0;
#pragma unroll
  for (int j = 0; j < kPer; j++)
    if (sh.nx[j * kBlock + threadIdx.x] != kTermNone && !(info[j] & kInfoExit))
      mh = max(mh, info_hops(info[j]));
  const int maxhops = block_max(mh, &sh.maxhops);   // ends with a barrier
#pragma unroll
  for (int j = 0; j < kPer; j++) {
    const int i = j * kBlock + threadIdx.x;
    if (sh.nx[i] == kTermNone) continue;
    cnt[static_cast<size_t>(b.y0 + i / kSide) * W + b.x0 + i % kSide] = sh.cs[i];
  }
  // cells with a root tail, level by level from the root (hops 0): the root
  // keeps its elevation, plus uplift off rings 0 and 1
  for (int L = 0; L <= maxhops; L++) {
#pragma unroll
    for (int j = 0; j < kPer; j++) {
      const int i = j * kBlock + threadIdx.x;
      if (sh.nx[i] == kTermNone || (info[j] & kInfoExit) || info_hops(info[j]) != L) continue;
      const int lx = i % kSide, ly = i / kSide;
      if (L == 0) {
        const int gx = b.x0 + lx, gy = b.y0 + ly;
        if (gx >= 2 && gx < W - 2 && gy >= 2 && gy < H - 2) sh.hs[i] += k.ud;
      } else {
        const int r = rr[j];
        const int nb = i + dir_dy(r) * kSide + dir_dx(r);
        const double fact = ftab[2 * static_cast<size_t>(sh.cs[i]) + (r & 1)];
        sh.hs[i] = newton_update<NMODE>(sh.hs[i] + k.ud, sh.hs[nb], fact, k, hdr);
      }
    }
    __syncthreads();
  }
#pragma unroll
  for (int j = 0; j < kPer; j++) {
    const int i = j * kBlock + threadIdx.x;
    if (sh.nx[i] == kTermNone || (info[j] & kInfoExit)) continue;
    h_new[static_cast<size_t>(b.y0 + i / kSide) * W + b.x0 + i % kSide] = sh.hs[i];
  }
}

struct ErodeShared {
  double hs[kCells];
  uint16_t tier_at[kCells];   // tier of each exit cell
  int maxhops;
};

// Launch `tier` >= 1: the listed tiles' cells whose tail is a tier-`tier`
// exit.  h_new holds every cell of lower tiers (earlier launches).
template <int NMODE>
__global__ __launch_bounds__(kBlock) void k_deep_erode(
    const uint8_t *__restrict__ rcv, const uint32_t *__restrict__ cnt,
    const uint32_t *__restrict__ dinfo, const int32_t *__restrict__ node_cell,
    const uint16_t *__restrict__ node_tier, const int32_t *__restrict__ nexit,
    const int32_t *__restrict__ list, const double *__restrict__ ftab,
    const double *__restrict__ h_old, double *h_new, int W, int H, int tiles_x, int tier,
    ErodeConst k, int32_t *hdr) {
  __shared__ ErodeShared sh;
  const int t = list[blockIdx.x];
  const Box b = tile_box(t, tiles_x, W, H);
  uint32_t info[kPer];
  uint8_t rr[kPer];
  bool valid[kPer];
  if (threadIdx.x == 0) sh.maxhops = 0;
#pragma unroll
  for (int j = 0; j < kPer; j++) {
    const int i = j * kBlock + threadIdx.x;
    const int lx = i % kSide, ly = i / kSide;
    valid[j] = lx < b.tw && ly < b.th;
    info[j] = 0;
    rr[j] = kNoFlow;
    if (valid[j]) {
      const size_t g = static_cast<size_t>(b.y0 + ly) * W + b.x0 + lx;
      info[j] = dinfo[g];
      rr[j] = rcv[g];
      sh.hs[i] = h_old[g];
    }
  }
  const int n = nexit[t];
  if (static_cast<int>(threadIdx.x) < n) {
    const int id = t * kNodeCap + threadIdx.x;
    const int32_t g = node_cell[id];
    const int i = (g / W - b.y0) * kSide + (g % W - b.x0);
    sh.tier_at[i] = node_tier[id];
  }
  __syncthreads();
  bool mine[kPer];
  int mh = 0;
#pragma unroll
  for (int j = 0; j < kPer; j++) {
    mine[j] = valid[j] && (info[j] & kInfoExit) && sh.tier_at[info_tail(info[j])] == tier;
    if (mine[j]) mh = max(mh, info_hops(info[j]));
  }
  const int maxhops = block_max(mh, &sh.maxhops);
  for (int L = 0; L <= maxhops; L++) {
#pragma unroll
    for (int j = 0; j < kPer; j++) {
      if (!mine[j] || info_hops(info[j]) != L) continue;
      const int i = j * kBlock + threadIdx.x;
      const int r = rr[j];
      const size_t g = static_cast<size_t>(b.y0 + i / kSide) * W + b.x0 + i % kSide;
      const double hr = L == 0 ? h_new[static_cast<long long>(g) + dir_offset(r, W)]
                               : sh.hs[i + dir_dy(r) * kSide + dir_dx(r)];
      const double fact = ftab[2 * static_cast<size_t>(cnt[g]) + (r & 1)];
      sh.hs[i] = newton_update<NMODE>(sh.hs[i] + k.ud, hr, fact, k, hdr);
    }
    __syncthreads();
  }
#pragma unroll
  for (int j = 0; j < kPer; j++) {
    if (!mine[j]) continue;
    const int i = j * kBlock + threadIdx.x;
    h_new[static_cast<size_t>(b.y0 + i / kSide) * W + b.x0 + i % kSide] = sh.hs[i];
  }
}

__global__ void k_deep_clear(int32_t *tstat) {
  for (int i = threadIdx.x; i < kTierWords; i += blockDim.x) tstat[i] = 0;
}

}  // namespace deep

#endif  // FS_DEEP_H
