// fs_deep.h -- the tile-graph path: a time-step for drainage of any depth,
// tile by tile with no halo, for one grid or a list of grids.  Included by
// fastscape_hip.hip and fs_ensemble.hip after fs_device.h; everything here
// has internal linkage.
//
// Within a 64x64 tile, each cell's receiver chain is followed only while it
// stays in the tile.  The chain ends at the cell's tail: a NO_FLOW cell of
// the tile (a root) or the last tile cell before the chain leaves the tile
// (an exit).  Exits are the nodes of a forest that spans the tiles: the
// parent of exit x is the tail of x's receiver when that tail is an exit.
// A node's tier is 1 + its parent's (1 when x's receiver has a root tail);
// a cell's tier is its tail's (0 for a root tail).  A tier-t cell can be
// eroded once the tier t-1 cells are: its chain ends, inside its tile, at
// an exit whose receiver is a tier t-1 cell of another tile.
//
//   k_deep_local  per tile: the in-tile walks give every cell its local count
//                 (the tile cells upstream of it along in-tile paths, itself
//                 included), its tail and its hops to the tail; exits become
//                 nodes; border cells record the node of their tail (link);
//                 each exit's receiver gets the "not yet eroded" mark in the
//                 new elevation field; the cells whose tail is a root start
//                 the tile's cell list, sorted by hops
//   k_deep_links  every node's parent, and the number of children per node
//   k_deep_tiers  every node's tier (walking its parents), and the number of
//                 (tile, tier) work items per tier
//   k_deep_scan   the work items' offsets: up to 64 queues, each holding its
//                 tiles' items in tier order (one workgroup)
//   k_deep_sum    the count through every exit: its local count plus its
//                 children's, leaf to root (the last child to arrive carries
//                 the sum on, so each node is summed once)
//   k_deep_counts per tile: the counts finished by the inflows of the ring
//                 cells that drain into the tile (each walked down its
//                 in-tile chain); the erosion of the cells with a root tail
//                 (tier 0) in LDS, from the roots up; the tile's other cells
//                 grouped by tier into the cell list, one work item per tier
//                 placed in tier order
//   k_deep_erode  a persistent grid takes the work items (tiers 1 and up)
//                 from the queues, each in tier order, and erodes each item's
//                 cells from its exits up:
//                 an exit waits until its receiver's new elevation replaces
//                 the mark; a wave's cells wait on their receivers' done
//                 flags in LDS
//
// Counts are integers, so the order of the additions is free: they equal
// the reference's ComputeFlowAcc (RB:262-279) counts.  Every cell is eroded
// after its receiver with the reference's Newton iteration (RB:306-328) from
// the same inputs (its elevation plus uplift, the receiver's new elevation,
// the factor of its count), so the new elevations are the reference's bits.
// Work items wait only on items of lower tiers, which their queues hand out
// before any item of a higher tier to workgroups that are running, so the
// persistent grid always drains; every wait is bounded all the same and
// reported in hdr[2].
#ifndef FS_DEEP_H
#define FS_DEEP_H

namespace {
namespace deep {

constexpr int kSide = 64;
constexpr int kCells = kSide * kSide;
constexpr int kBlock = 1024;
constexpr int kPer = kCells / kBlock;
constexpr int kNodeCap = 256;             // nodes per tile: a border has at most 252 cells
constexpr int kTiersPerTile = kNodeCap;   // distinct tiers in one tile: its nodes' and 0
constexpr uint16_t kTermNone = 0xFFFD;    // beyond the grid
constexpr uint16_t kTermExit = 0xFFFE;    // the receiver lies in another tile
constexpr uint16_t kTermRoot = 0xFFFF;    // NO_FLOW
constexpr uint32_t kNoNode = 0xFFFFFFFFu;
constexpr int kSumShift = 40;             // node word: arrivals << 40 | count
constexpr uint64_t kSumMask = (uint64_t{1} << kSumShift) - 1;
constexpr uint64_t kOneArrival = uint64_t{1} << kSumShift;
constexpr int kTierCap = FS_MAX_LEVELS;   // a tier is at most the depth of the drainage
// per-tier item counters are split over kShards words (by tile) to keep
// tens of thousands of atomics per step off a few addresses; the work items
// are handed out by up to kShards queues for the same reason
constexpr int kShards = 64;
// tstat words: error bits, deepest tier, items; the first item of each
// queue (and the end of the last); each queue's cursor on a cache line of
// its own (kQStride words apart).  Steps clear the first kStClear words.
constexpr int kStErr = 0, kStMaxTier = 1, kStItems = 2, kStClear = 3;
constexpr int kStQStart = 4;
constexpr int kQStride = 32;
constexpr int kStQCur = 128;
constexpr int kStWords = kStQCur + kShards * kQStride;
static_assert(kStQStart + kShards + 1 <= kStQCur, "queue starts fit before the cursors");
constexpr int32_t kErrTier = 1, kErrChain = 2;
// the "not yet eroded" mark: a signalling NaN that no arithmetic produces
constexpr unsigned long long kUnsolved = 0x7FF4A5A5A5A5A5A5ull;
constexpr int kSpinCap = 1 << 24;
static_assert(kCells <= 4096, "tail and hops fit 12 bits each");
static_assert(kTierCap < 65536, "tiers fit 16 bits");

// cell list entry: the tile cell (12 bits) | the position of its receiver in
// the cell's work item << 12 (13 bits; kNoPos: the receiver lies in another
// tile, or the cell has a root tail)
constexpr uint32_t kEntCell = 0xFFFu;
constexpr int kEntShift = 12;
constexpr uint32_t kNoPos = 0x1FFFu;
static_assert(kCells - 1 <= static_cast<int>(kEntCell) && kCells < static_cast<int>(kNoPos),
              "list entries");

// dinfo word of a cell: tail (12 bits) | hops to the tail << 12 | exit tail
constexpr uint32_t kInfoExit = 1u << 24;
__device__ __forceinline__ int info_tail(uint32_t w) { return static_cast<int>(w & 0xFFFu); }
__device__ __forceinline__ int info_hops(uint32_t w) { return static_cast<int>((w >> 12) & 0xFFFu); }

// work item: tile (27 bits) | first cell in the tile's sorted list << 27 |
// number of cells << 40
__device__ __forceinline__ uint64_t item_pack(int t, int first, int n) {
  return static_cast<uint64_t>(t) | (static_cast<uint64_t>(first) << 27) |
         (static_cast<uint64_t>(n) << 40);
}

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

// A grid the kernels step: its cells from `off` in the per-cell arrays, its
// tiles from `tile0` in the per-tile arrays (node ids are global: tile *
// kNodeCap + k), its erosion factors from `ftab_off`.
struct GraphGrid {
  uint64_t off, ftab_off;
  int32_t *hdr;       // hdr[2]: error bits; hdr[0]: deepest level (when levels are kept)
  int32_t *levels;    // cells per level (nlev entries), or nullptr
  int W, H, tiles_x, tile0, nlev, nmode;
  ErodeConst ek;
};
// The grids of one step: the tiles of the listed grids, numbered from 0 in
// list order (the work index).
struct GraphWork {
  const GraphGrid *grid;
  const int32_t *first;    // per entry: its first work index; entry size[0] holds the total
  const int32_t *which;    // per entry: its grid
  const int32_t *size;     // [0] entries, [1] tiles
};
// The values a workgroup needs of its tile's grid, held in scalar registers
// (the descriptor is read once per tile; readfirstlane tells the compiler
// the values are the same in every lane).
__device__ __forceinline__ int uni(int v) { return __builtin_amdgcn_readfirstlane(v); }
__device__ __forceinline__ uint64_t uni64(uint64_t v) {
  return (static_cast<uint64_t>(static_cast<uint32_t>(uni(static_cast<int>(v >> 32)))) << 32) |
         static_cast<uint32_t>(uni(static_cast<int>(v)));
}
template <class T>
__device__ __forceinline__ T *uniptr(T *p) {
  return reinterpret_cast<T *>(uni64(reinterpret_cast<uint64_t>(p)));
}
__device__ __forceinline__ double unid(double v) {
  return __longlong_as_double(static_cast<long long>(uni64(static_cast<uint64_t>(__double_as_longlong(v)))));
}
struct WorkTile {
  uint64_t off, ftab_off;
  int32_t *hdr, *levels;
  int W, H, tiles_x, nlev, nmode;
  int t;    // tile within its grid
  int tg;   // tile in the per-tile arrays
  ErodeConst ek;
};
__device__ __forceinline__ WorkTile work_tile(const GraphWork &w, int wi) {
  int lo = 0, hi = uni(w.size[0]) - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (uni(w.first[mid]) <= wi) lo = mid;
    else hi = mid - 1;
  }
  const GraphGrid *const g = w.grid + uni(w.which[lo]);
  WorkTile r;
  r.t = wi - uni(w.first[lo]);
  r.tg = uni(g->tile0) + r.t;
  r.off = uni64(g->off);
  r.ftab_off = uni64(g->ftab_off);
  r.hdr = uniptr(g->hdr);
  r.levels = uniptr(g->levels);
  r.W = uni(g->W);
  r.H = uni(g->H);
  r.tiles_x = uni(g->tiles_x);
  r.nlev = uni(g->nlev);
  r.nmode = uni(g->nmode);
  r.ek.kdt = unid(g->ek.kdt);
  r.ek.lenpow0 = unid(g->ek.lenpow0);
  r.ek.lenpow1 = unid(g->ek.lenpow1);
  r.ek.ud = unid(g->ek.ud);
  r.ek.neq = unid(g->ek.neq);
  r.ek.tol = unid(g->ek.tol);
  return r;
}

// The Newton solve in the grid's pow mode (NMODE < 0: chosen per grid).
template <int NMODE>
__device__ __forceinline__ double solve(int nm, double h0, double hn, double fact,
                                        const ErodeConst &k, int32_t *hdr) {
  if constexpr (NMODE >= 0) {
    return newton_update<NMODE>(h0, hn, fact, k, hdr);
  } else {
    if (nm == 1) return newton_update<1>(h0, hn, fact, k, hdr);
    if (nm == 2) return newton_update<2>(h0, hn, fact, k, hdr);
    return newton_update<0>(h0, hn, fact, k, hdr);
  }
}

__device__ __forceinline__ int block_max(int v, int *slot) {
#pragma unroll
  for (int d = 1; d < kWave; d <<= 1) v = max(v, __shfl_xor(v, d));
  if ((threadIdx.x & (kWave - 1)) == 0) atomicMax(slot, v);
  __syncthreads();
  return *slot;
}

// Exclusive scan of v over the workgroup (BLOCK threads); *total gets the
// sum.  sh holds BLOCK / kWave + 1 ints.  Ends with a barrier.
template <int BLOCK = kBlock>
__device__ __forceinline__ int block_scan(int v, int *sh, int *total) {
  const int lane = threadIdx.x & (kWave - 1), wave = threadIdx.x / kWave;
  int x = v;
#pragma unroll
  for (int d = 1; d < kWave; d <<= 1) {
    const int y = __shfl_up(x, d);
    if (lane >= d) x += y;
  }
  if (lane == kWave - 1) sh[wave] = x;
  __syncthreads();
  if (threadIdx.x == 0) {
    int run = 0;
    for (int w = 0; w < BLOCK / kWave; w++) {
      const int c = sh[w];
      sh[w] = run;
      run += c;
    }
    sh[BLOCK / kWave] = run;
  }
  __syncthreads();
  const int r = sh[wave] + x - v;
  *total = sh[BLOCK / kWave];
  __syncthreads();
  return r;
}

// Border position of a border cell of the tile: rows 0 and th-1, then
// columns 0 and tw-1 (exits lie on the border; a corner takes its row's).
__device__ __forceinline__ int border_pos(int lx, int ly, const Box &b) {
  if (ly == 0) return lx;
  if (ly == b.th - 1) return kSide + lx;
  if (lx == 0) return 2 * kSide + ly;
  return 3 * kSide + ly;
}

// The cell at border position k, or -1 (a position beyond a partial tile,
// or a corner counted with its row).
__device__ __forceinline__ int border_cell(int k, const Box &b) {
  int lx, ly;
  bool ok;
  if (k < kSide) {
    lx = k;
    ly = 0;
    ok = lx < b.tw;
  } else if (k < 2 * kSide) {
    lx = k - kSide;
    ly = b.th - 1;
    ok = lx < b.tw && b.th > 1;
  } else if (k < 3 * kSide) {
    lx = 0;
    ly = k - 2 * kSide;
    ok = ly >= 1 && ly < b.th - 1;
  } else {
    lx = b.tw - 1;
    ly = k - 3 * kSide;
    ok = ly >= 1 && ly < b.th - 1 && b.tw > 1;
  }
  return ok ? ly * kSide + lx : -1;
}

// ---------------------------------------------------------------------------
// k_deep_local: one tile per workgroup of kLocalBlock threads (half the
// other tile kernels': its LDS fits four tiles per CU, which 1024-thread
// workgroups cannot use)
// ---------------------------------------------------------------------------
constexpr int kLocalBlock = 512;
constexpr int kLocalPer = kCells / kLocalBlock;
__global__ __launch_bounds__(kLocalBlock) void k_deep_local(
    GraphWork w, const uint8_t *__restrict__ rcv0, uint32_t *__restrict__ cnt0,
    uint32_t *__restrict__ dinfo0, uint32_t *__restrict__ link0, uint64_t *__restrict__ node_word,
    uint32_t *__restrict__ node_kids, int32_t *__restrict__ node_cell,
    int32_t *__restrict__ nexit, int32_t *__restrict__ tile_n0, uint32_t *__restrict__ cells,
    double *__restrict__ h_new0, int32_t *__restrict__ tstat) {
  __shared__ uint16_t nx[kCells];
  // walk words: (next cell + 1) << 16 | local count; 0 in the top half at a
  // tail (later the root list's counters)
  __shared__ uint32_t jw[kCells];
  __shared__ uint16_t slot_node[4 * kSide];   // node number of the exit at each border position
  __shared__ int wave_nodes[4 * kSide / kWave];
  __shared__ int maxhops, scan[kLocalBlock / kWave + 1];
  const int wi = blockIdx.x;   // one tile per workgroup; the grid covers the list exactly
  if (wi >= w.size[1]) return;
  const WorkTile wt = work_tile(w, wi);
  const WorkTile &G = wt;
  const int W = G.W, H = G.H;
  const uint8_t *const rcv = rcv0 + G.off;
  uint32_t *const cnt = cnt0 + G.off, *const dinfo = dinfo0 + G.off, *const link = link0 + G.off;
  double *const h_new = h_new0 + G.off;
  const int t = wt.tg;   // per-tile arrays and node ids
  const Box b = tile_box(wt.t, G.tiles_x, W, H);
  if (threadIdx.x == 0) maxhops = 0;
  if (wt.t == 0 && G.levels) {   // the grid's level counts start from zero
    if (threadIdx.x == 0) G.hdr[0] = 0;
    for (int q = threadIdx.x; q < G.nlev; q += kLocalBlock) G.levels[q] = 0;
  }
  // receiver bytes, kept for the exits' marks; all requested before any is
  // used (a cell past a partial tile's edge reads one inside it)
  uint8_t rr[kLocalPer];
#pragma unroll
  for (int j = 0; j < kLocalPer; j++) {
    const int i = j * kLocalBlock + threadIdx.x;
    const int lx = min(i % kSide, b.tw - 1), ly = min(i / kSide, b.th - 1);
    rr[j] = rcv[static_cast<size_t>(b.y0 + ly) * W + b.x0 + lx];
  }
#pragma unroll
  for (int j = 0; j < kLocalPer; j++) {
    const int i = j * kLocalBlock + threadIdx.x;
    const int lx = i % kSide, ly = i / kSide;
    uint16_t n = kTermNone;
    if (lx < b.tw && ly < b.th) n = tile_next(rr[j], lx, ly, b);
    else rr[j] = 0xFF;
    nx[i] = n;
    jw[i] = n < kTermNone ? (static_cast<uint32_t>(n) + 1u) << 16 : 0u;
  }
  __syncthreads();
  // every tile cell walks its in-tile chain to its tail, adding 1 to each
  // cell it passes, itself included: one LDS atomic per step returns the
  // cell's word, whose top half names the next cell.  The walk gives the
  // cell's tail and hops, the adds every cell's local count (the tile cells
  // upstream of it, itself included).  A thread's walks advance together,
  // one step per round, with no barrier.
  int cur[kLocalPer], hops[kLocalPer];
  uint32_t alive = 0;
#pragma unroll
  for (int j = 0; j < kLocalPer; j++) {
    const int i = j * kLocalBlock + threadIdx.x;
    cur[j] = i;
    hops[j] = 0;
    if (i % kSide < b.tw && i / kSide < b.th) alive |= 1u << j;
  }
  for (int step = 0; alive != 0 && step < kCells; step++) {
#pragma unroll
    for (int j = 0; j < kLocalPer; j++) {
      if (!(alive & (1u << j))) continue;
      const uint32_t nxt = atomicAdd(&jw[cur[j]], 1u) >> 16;
      if (nxt == 0) {
        alive &= ~(1u << j);
      } else {
        cur[j] = static_cast<int>(nxt) - 1;
        hops[j]++;
      }
    }
  }
  if (alive != 0) atomicOr(&tstat[kStErr], kErrChain);   // a receiver cycle
  int mh = 0;
#pragma unroll
  for (int j = 0; j < kLocalPer; j++) mh = max(mh, hops[j]);
  const int top = block_max(mh, &maxhops);   // its barrier ends every walk
  // exits become the tile's nodes, numbered in border-position order: the
  // first four waves take one border position per lane
  constexpr int kSlotWaves = 4 * kSide / kWave;
  int slot_cell = -1, slot_rank = 0;
  uint64_t slot_mask = 0;
  if (threadIdx.x < 4 * kSide) {
    slot_cell = border_cell(threadIdx.x, b);
    const bool ex = slot_cell >= 0 && nx[slot_cell] == kTermExit;
    slot_mask = __ballot(ex);
    const int lane = threadIdx.x & (kWave - 1);
    slot_rank = ex ? __popcll(slot_mask & ((uint64_t{1} << lane) - 1)) : -1;
    if (lane == 0) wave_nodes[threadIdx.x / kWave] = __popcll(slot_mask);
  }
#pragma unroll
  for (int j = 0; j < kLocalPer; j++) {   // the marks on the exits' receivers
    const int i = j * kLocalBlock + threadIdx.x;
    if (nx[i] != kTermExit) continue;
    const size_t g = static_cast<size_t>(b.y0 + i / kSide) * W + b.x0 + i % kSide;
    h_new[g + dir_offset(rr[j], W)] = __longlong_as_double(static_cast<long long>(kUnsolved));
  }
  __syncthreads();
  int nnode = 0;
  if (threadIdx.x < 4 * kSide) {
    int base = 0;
#pragma unroll
    for (int q = 0; q < kSlotWaves; q++) {
      base += q < static_cast<int>(threadIdx.x) / kWave ? wave_nodes[q] : 0;
      nnode += wave_nodes[q];
    }
    if (slot_rank >= 0) {
      const int k = base + slot_rank;
      slot_node[threadIdx.x] = static_cast<uint16_t>(k);
      const int id = t * kNodeCap + k;
      node_cell[id] = (b.y0 + slot_cell / kSide) * W + b.x0 + slot_cell % kSide;
      node_word[id] = jw[slot_cell] & 0xFFFFu;
      node_kids[id] = 0;
    }
  }
  __syncthreads();
#pragma unroll
  for (int j = 0; j < kLocalPer; j++) {
    const int i = j * kLocalBlock + threadIdx.x;
    if (nx[i] == kTermNone) continue;
    const int lx = i % kSide, ly = i / kSide;
    const size_t g = static_cast<size_t>(b.y0 + ly) * W + b.x0 + lx;
    const bool exit_tail = nx[cur[j]] == kTermExit;
    cnt[g] = jw[i] & 0xFFFFu;
    dinfo[g] = static_cast<uint32_t>(cur[j]) | (static_cast<uint32_t>(hops[j]) << 12) |
               (exit_tail ? kInfoExit : 0u);
    if (lx == 0 || ly == 0 || lx == b.tw - 1 || ly == b.th - 1)
      link[g] = exit_tail ? static_cast<uint32_t>(t * kNodeCap +
                                                  slot_node[border_pos(cur[j] % kSide,
                                                                       cur[j] / kSide, b)])
                          : kNoNode;
  }
  if (threadIdx.x == 0) nexit[t] = nnode;   // every slot thread summed the same count
  // the cells with a root tail, sorted by hops, start the tile's cell list
  // (jw is free once every count is stored: its words count the cells per
  // hops)
  __syncthreads();
  for (int q = threadIdx.x; q <= top; q += kLocalBlock) jw[q] = 0;
  __syncthreads();
  bool rooted[kLocalPer];
#pragma unroll
  for (int j = 0; j < kLocalPer; j++) {
    const int i = j * kLocalBlock + threadIdx.x;
    rooted[j] = nx[i] != kTermNone && nx[cur[j]] == kTermRoot;
    if (rooted[j]) atomicAdd(&jw[hops[j]], 1u);
  }
  __syncthreads();
  int carry = 0;
  for (int base = 0; base <= top; base += kLocalBlock) {
    const int q = base + threadIdx.x;
    const int v = q <= top ? static_cast<int>(jw[q]) : 0;
    int total;
    const int ex = block_scan<kLocalBlock>(v, scan, &total);
    if (q <= top) jw[q] = static_cast<uint32_t>(carry + ex);
    carry += total;
  }
  __syncthreads();   // every cursor is set before any thread takes a slot
  uint32_t *const list = cells + static_cast<size_t>(t) * kCells;
#pragma unroll
  for (int j = 0; j < kLocalPer; j++)
    if (rooted[j])
      list[atomicAdd(&jw[hops[j]], 1u)] =
          static_cast<uint32_t>(j * kLocalBlock + threadIdx.x) | kNoPos << kEntShift;
  // n0, and the tile's largest hops (which bounds the sort keys of k_deep_counts)
  if (threadIdx.x == 0) tile_n0[t] = carry | top << 16;
}

// ---------------------------------------------------------------------------
// the node forest: one tile's nodes per workgroup of kNodeCap threads
// ---------------------------------------------------------------------------
// With node_up: 1 + the hops of each exit's receiver to its tail (the
// levels a node adds to its parent's).
__global__ __launch_bounds__(kNodeCap) void k_deep_links(
    GraphWork w, const uint8_t *__restrict__ rcv0, const uint32_t *__restrict__ link0,
    const uint32_t *__restrict__ dinfo0, const int32_t *__restrict__ node_cell,
    const int32_t *__restrict__ nexit, uint32_t *__restrict__ node_parent,
    uint32_t *__restrict__ node_kids, uint32_t *__restrict__ node_up) {
  const int k = threadIdx.x, wi = blockIdx.x;
  if (wi >= w.size[1]) return;
  const WorkTile wt = work_tile(w, wi);
  const int t = wt.tg;
  if (k >= nexit[t]) return;
  const int id = t * kNodeCap + k;
  const int32_t g = node_cell[id];
  // the receiver lies on the border of the neighbouring tile
  const int32_t y = g + dir_offset(rcv0[wt.off + g], wt.W);
  const uint32_t p = link0[wt.off + y];
  node_parent[id] = p;
  if (p != kNoNode) atomicAdd(&node_kids[p], 1u);
  if (node_up) node_up[id] = 1u + static_cast<uint32_t>(info_hops(dinfo0[wt.off + y]));
}

// Tier of every node, by walking its parents (a chain of tiles crossed by
// one drainage path); per tier, the number of tiles holding cells of it.
__global__ __launch_bounds__(kNodeCap) void k_deep_tiers(
    GraphWork w, const uint32_t *__restrict__ node_parent, const int32_t *__restrict__ nexit,
    uint16_t *__restrict__ node_tier, const uint32_t *__restrict__ node_up,
    uint32_t *__restrict__ node_level, int32_t *__restrict__ thist,
    int32_t *__restrict__ tstat) {
  __shared__ uint32_t seen[(kTierCap + 32) / 32];   // tiers already counted for the tile
  const int k = threadIdx.x, wi = blockIdx.x;
  if (wi >= w.size[1]) return;
  {
    const int t = work_tile(w, wi).tg;
    const int n = nexit[t];
    int tier = 0;
    if (k < n) {
      const int id = t * kNodeCap + k;
      tier = 1;
      uint32_t lev = node_up ? node_up[id] : 0u;
      uint32_t p = node_parent[id];
      while (p != kNoNode && tier <= kTierCap) {
        tier++;
        if (node_up) lev += node_up[p];
        p = node_parent[p];
      }
      node_tier[id] = static_cast<uint16_t>(min(tier, kTierCap));
      if (node_level) node_level[id] = lev;
    }
    // clear the words the tile's tiers can touch (the block's own entries)
    if (k < n && tier <= kTierCap) seen[tier >> 5] = 0;
    __syncthreads();
    if (k < n) {
      if (tier > kTierCap) {
        atomicOr(&tstat[kStErr], kErrTier);
      } else {
        const uint32_t bit = 1u << (tier & 31);
        if (!(atomicOr(&seen[tier >> 5], bit) & bit)) {   // the tile's first node of the tier
          atomicAdd(&thist[tier * kShards + (t & (kShards - 1))], 1);
          if (tier > __hip_atomic_load(&tstat[kStMaxTier], __ATOMIC_RELAXED,
                                       __HIP_MEMORY_SCOPE_AGENT))
            atomicMax(&tstat[kStMaxTier], tier);
        }
      }
    }
    __syncthreads();
  }
}

// Offsets of the work items: thist (counts per tier and shard) -> tcur
// (the first slot of each (tier, shard), then the cursor k_deep_counts
// appends at).  The items are laid out queue by queue -- queue q holds the
// shards s with s % nq == q -- and each queue's tier by tier, so a queue
// hands out its items in tier order.  Writes each queue's first item and
// cursor, clears thist for the next step.  One workgroup of kBlock threads;
// nq divides kShards.
__global__ __launch_bounds__(kBlock) void k_deep_scan(int32_t *__restrict__ thist,
                                                      int32_t *__restrict__ tcur,
                                                      int32_t *__restrict__ tstat, int nq) {
  __shared__ int sh[kBlock / kWave + 1];
  const int spq = kShards / nq;                       // shards per queue
  const int per_q = (tstat[kStMaxTier] + 1) * spq;    // (tier, shard) words per queue
  const int n = nq * per_q;
  int carry = 0;
  for (int base = 0; base < n; base += kBlock) {
    const int k = base + threadIdx.x;
    int idx = 0, v = 0;
    if (k < n) {
      const int q = k / per_q, rem = k - q * per_q;
      const int tier = rem / spq, s = q + (rem - tier * spq) * nq;
      idx = tier * kShards + s;
      v = thist[idx];
    }
    int total;
    const int ex = block_scan(v, sh, &total);
    if (k < n) {
      tcur[idx] = carry + ex;
      thist[idx] = 0;
      if (k % per_q == 0) {   // the queue's first word
        tstat[kStQStart + k / per_q] = carry + ex;
        tstat[kStQCur + (k / per_q) * kQStride] = carry + ex;
      }
    }
    carry += total;
  }
  if (threadIdx.x == 0) {
    tstat[kStItems] = carry;
    tstat[kStQStart + nq] = carry;
  }
}

// Count through every exit, leaf to root.
__global__ __launch_bounds__(kNodeCap) void k_deep_sum(
    GraphWork w, uint64_t *__restrict__ node_word, const uint32_t *__restrict__ node_kids,
    const uint32_t *__restrict__ node_parent, const int32_t *__restrict__ nexit,
    uint32_t *__restrict__ node_out, int32_t *__restrict__ tstat) {
  const int k = threadIdx.x, wi = blockIdx.x;
  if (wi >= w.size[1]) return;
  {
    const int t = work_tile(w, wi).tg;
    if (k >= nexit[t]) return;
    uint32_t id = static_cast<uint32_t>(t * kNodeCap + k);
    if (node_kids[id] != 0) return;          // summed by its last child's thread
    uint64_t val = node_word[id] & kSumMask;
    bool ended = false;
    for (int guard = 0; guard <= kTierCap; guard++) {
      node_out[id] = static_cast<uint32_t>(val);
      const uint32_t p = node_parent[id];
      if (p == kNoNode) {
        ended = true;
        break;
      }
      const uint64_t old = atomicAdd(&node_word[p], val + kOneArrival);
      if ((old >> kSumShift) + 1 != node_kids[p]) {   // not the last child
        ended = true;
        break;
      }
      val += old & kSumMask;
      id = p;
    }
    if (!ended) atomicOr(&tstat[kStErr], kErrChain);
  }
}

// ---------------------------------------------------------------------------
// k_deep_levels: per tile, each cell's level (hops to its root across tiles:
// its hops to its tail, plus its tail node's levels) counted into its grid's
// level histogram, and the grid's deepest level; for grids that keep them
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(kBlock) void k_deep_levels(GraphWork w,
                                                        const uint32_t *__restrict__ dinfo0,
                                                        const uint32_t *__restrict__ link0,
                                                        const uint32_t *__restrict__ node_level) {
  __shared__ int hist[kCells];
  __shared__ int lo_s, hi_s;
  const int wi = blockIdx.x;
  if (wi >= w.size[1]) return;
  {
    const WorkTile wt = work_tile(w, wi);
    const WorkTile &G = wt;
    if (!G.levels) return;   // block-uniform
    const Box b = tile_box(wt.t, G.tiles_x, G.W, G.H);
    if (threadIdx.x == 0) {
      lo_s = 0x7FFFFFFF;
      hi_s = 0;
    }
    __syncthreads();
    int lev[kPer];
    int lo = 0x7FFFFFFF, hi = 0;
#pragma unroll
    for (int j = 0; j < kPer; j++) {
      const int i = j * kBlock + threadIdx.x;
      const int lx = i % kSide, ly = i / kSide;
      lev[j] = -1;
      if (lx >= b.tw || ly >= b.th) continue;
      const uint32_t inf = dinfo0[G.off + static_cast<size_t>(b.y0 + ly) * G.W + b.x0 + lx];
      int L = info_hops(inf);
      if (inf & kInfoExit) {
        const int tl = info_tail(inf);
        const size_t tg = static_cast<size_t>(b.y0 + tl / kSide) * G.W + b.x0 + tl % kSide;
        L += static_cast<int>(node_level[link0[G.off + tg]]);   // an exit's link is its node
      }
      lev[j] = L;
      lo = min(lo, L);
      hi = max(hi, L);
    }
#pragma unroll
    for (int d = 1; d < kWave; d <<= 1) {
      lo = min(lo, __shfl_xor(lo, d));
      hi = max(hi, __shfl_xor(hi, d));
    }
    if ((threadIdx.x & (kWave - 1)) == 0) {
      atomicMin(&lo_s, lo);
      atomicMax(&hi_s, hi);
    }
    __syncthreads();
    const int base = lo_s, span = min(hi_s - lo_s + 1, kCells);
    for (int q = threadIdx.x; q < span; q += kBlock) hist[q] = 0;
    __syncthreads();
#pragma unroll
    for (int j = 0; j < kPer; j++) {
      if (lev[j] < 0 || lev[j] >= G.nlev) continue;
      if (lev[j] - base < span) atomicAdd(&hist[lev[j] - base], 1);
      else atomicAdd(&G.levels[lev[j]], 1);
    }
    __syncthreads();
    for (int q = threadIdx.x; q < span; q += kBlock)
      if (hist[q] && base + q < G.nlev) atomicAdd(&G.levels[base + q], hist[q]);
    if (threadIdx.x == 0) atomicMax(&G.hdr[0], hi_s);
    __syncthreads();
  }
}

// ---------------------------------------------------------------------------
// k_deep_counts: one tile per workgroup of kCntBlock threads
// ---------------------------------------------------------------------------
// Half the threads of the other tile kernels: with the time-shared region
// below the tile's LDS is a third of a CU's, so three tiles run per CU and
// one tile's loads overlap another's wavefront.
constexpr int kCntBlock = 512;
constexpr int kCntPer = kCells / kCntBlock;
constexpr int kCntWaves = kCntBlock / kWave;
constexpr int kHashCap = 512;    // open-addressing set of the tile's exit tiers
// Each cell's step inside the tile: its receiver direction 0..7 when the
// receiver is a tile cell, else one of these.
constexpr uint8_t kCodeExit = 8, kCodeRoot = 9, kCodeNone = 10;
constexpr uint8_t kCodeStep = 15, kCodeExitTail = 16;   // the step; the cell's tail is an exit
struct CountShared {
  union {
    // last: the root-tailed cells' unsolved marks, then their new elevations (bits)
    uint64_t hs[kCells];
    struct {
      int keys[kCells];             // sort cursors, then each cell's list position (16 bits)
      int tset[kHashCap];
      int tiers[kTiersPerTile];     // the tile's distinct tiers, then sorted
      int sorted[kTiersPerTile];
      int hist[kTiersPerTile];      // cells per tier rank, then their first slot
      uint16_t tier_b[4 * kSide];   // tier of the exit at each border position
    } s;
  } u;
  uint32_t cs[kCells];
  uint8_t code[kCells];
  int scan[kCntWaves + 1];
  int ntier, maxhops;
};
static_assert(sizeof(CountShared) * 3 <= 160 * 1024, "three tiles per CU");

__device__ __forceinline__ uint8_t tile_code(uint8_t r, int lx, int ly, const Box &b) {
  if (r >= 8) return kCodeRoot;
  const int nx = lx + dir_dx(r), ny = ly + dir_dy(r);
  if (static_cast<unsigned>(nx) >= static_cast<unsigned>(b.tw) ||
      static_cast<unsigned>(ny) >= static_cast<unsigned>(b.th))
    return kCodeExit;
  return r;
}
// the tile cell one step from cell i in direction r (a tile code below 8)
__device__ __forceinline__ int tile_step(int i, int r) {
  return i + dir_dy(r) * kSide + dir_dx(r);
}

__device__ __forceinline__ int tier_rank(const int *sorted, int n, int tier) {
  int lo = 0, hi = n - 1;
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if (sorted[mid] < tier) lo = mid + 1;
    else hi = mid;
  }
  return lo;
}

template <int NMODE>
__global__ __launch_bounds__(kCntBlock) __attribute__((amdgpu_waves_per_eu(6))) void k_deep_counts(
    GraphWork w, const uint8_t *__restrict__ rcv0, uint32_t *__restrict__ cnt0,
    const uint32_t *__restrict__ dinfo0, const uint32_t *__restrict__ link0,
    const uint32_t *__restrict__ node_out, const int32_t *__restrict__ node_cell,
    const uint16_t *__restrict__ node_tier, const int32_t *__restrict__ nexit,
    const int32_t *__restrict__ tile_n0, uint32_t *__restrict__ cells,
    int32_t *__restrict__ tcur, uint64_t *__restrict__ items, const double *__restrict__ ftab0,
    const double *__restrict__ h_old0, double *__restrict__ h_new0) {
  __shared__ CountShared sh;
  auto &S = sh.u.s;
  const int wave = threadIdx.x / kWave, lane = threadIdx.x % kWave;
  const int wi = blockIdx.x;   // one tile per workgroup; the grid covers the list exactly
  if (wi >= w.size[1]) return;
  const WorkTile wt = work_tile(w, wi);
  const WorkTile &G = wt;
  const int W = G.W, H = G.H;
  const uint8_t *const rcv = rcv0 + G.off;
  uint32_t *const cnt = cnt0 + G.off;
  const uint32_t *const dinfo = dinfo0 + G.off, *const link = link0 + G.off;
  const double *const ftab = ftab0 + G.ftab_off, *const h_old = h_old0 + G.off;
  double *const h_new = h_new0 + G.off;
  const ErodeConst k = G.ek;
  int32_t *const hdr = G.hdr;
  const int t = wt.tg;   // per-tile arrays and node ids
  const Box b = tile_box(wt.t, G.tiles_x, W, H);
  const int n0 = uni(tile_n0[t]) & 0xFFFF;
  // the ring cell this thread looks after for the inflows: its receiver
  // byte is requested before the tile's loads, its link and count after
  // them, so the three dependent reads overlap the tile's
  int ring_gx = 0, ring_gy = 0;
  uint8_t ring_r = 0xFF;
  {
    const int q = threadIdx.x;
    const int ring_top = b.tw + 2, nring = 2 * ring_top + 2 * b.th;
    if (q < nring) {
      if (q < ring_top) {
        ring_gx = b.x0 - 1 + q;
        ring_gy = b.y0 - 1;
      } else if (q < 2 * ring_top) {
        ring_gx = b.x0 - 1 + (q - ring_top);
        ring_gy = b.y0 + b.th;
      } else if (q < 2 * ring_top + b.th) {
        ring_gx = b.x0 - 1;
        ring_gy = b.y0 + (q - 2 * ring_top);
      } else {
        ring_gx = b.x0 + b.tw;
        ring_gy = b.y0 + (q - 2 * ring_top - b.th);
      }
      if (ring_gx >= 0 && ring_gy >= 0 && ring_gx < W && ring_gy < H)
        ring_r = rcv[static_cast<size_t>(ring_gy) * W + ring_gx];
    }
  }
  uint32_t info[kCntPer];
  if (threadIdx.x == 0) sh.ntier = sh.maxhops = 0;
  if (threadIdx.x < kTiersPerTile) S.hist[threadIdx.x] = 0;
  for (int q = threadIdx.x; q < kHashCap; q += kCntBlock) S.tset[q] = -1;
  // every load is issued before any is used (a cell past a partial tile's
  // edge reads a cell inside it and ignores the value): a branch per cell
  // would make each cell's loads wait on the previous cell's
  uint8_t rv[kCntPer];
  uint32_t cv[kCntPer];
#pragma unroll
  for (int j = 0; j < kCntPer; j++) {
    const int i = j * kCntBlock + threadIdx.x;
    const int lx = min(i % kSide, b.tw - 1), ly = min(i / kSide, b.th - 1);
    const size_t g = static_cast<size_t>(b.y0 + ly) * W + b.x0 + lx;
    rv[j] = rcv[g];
    info[j] = dinfo[g];
    cv[j] = cnt[g];
  }
#pragma unroll
  for (int j = 0; j < kCntPer; j++) {
    const int i = j * kCntBlock + threadIdx.x;
    const int lx = i % kSide, ly = i / kSide;
    const bool in = lx < b.tw && ly < b.th;
    if (!in) info[j] = 0;
    sh.code[i] = in ? static_cast<uint8_t>(tile_code(rv[j], lx, ly, b) |
                                            ((info[j] & kInfoExit) ? kCodeExitTail : 0))
                    : kCodeNone;
    sh.cs[i] = in ? cv[j] : 0u;
  }
  int ring_cell = -1;   // the tile cell the ring cell drains into
  uint32_t ring_v = 0;  // and the count through the ring cell
  if (ring_r < 8) {
    const int tx = ring_gx + dir_dx(ring_r) - b.x0, ty = ring_gy + dir_dy(ring_r) - b.y0;
    if (static_cast<unsigned>(tx) < static_cast<unsigned>(b.tw) &&
        static_cast<unsigned>(ty) < static_cast<unsigned>(b.th)) {
      ring_cell = ty * kSide + tx;
      ring_v = node_out[link[static_cast<size_t>(ring_gy) * W + ring_gx]];
    }
  }
  const int n = nexit[t];
  int my_tier = -1;
  if (static_cast<int>(threadIdx.x) < n) {
    const int id = t * kNodeCap + threadIdx.x;
    const int32_t g = node_cell[id];
    my_tier = node_tier[id];
    S.tier_b[border_pos(g % W - b.x0, g / W - b.y0, b)] = static_cast<uint16_t>(my_tier);
  }
  __syncthreads();
  // the tile's distinct exit tiers: the node that inserts a tier into the
  // set lists it
  if (my_tier >= 0) {
    int slot = static_cast<int>((static_cast<uint32_t>(my_tier) * 2654435761u) >> 23);
    for (;;) {   // at most kNodeCap tiers in kHashCap slots: a free slot is always found
      const int old = atomicCAS(&S.tset[slot], -1, my_tier);
      if (old == -1) {
        S.sorted[atomicAdd(&sh.ntier, 1)] = my_tier;
        break;
      }
      if (old == my_tier) break;
      slot = (slot + 1) & (kHashCap - 1);
    }
  }
  // inflows: a ring cell that drains into the tile is an exit of its own
  // tile; its count is added along the in-tile chain from the cell it
  // drains into
  if (ring_cell >= 0) {
    int cur = ring_cell;
    for (int hop = 0; hop < kCells; hop++) {
      atomicAdd(&sh.cs[cur], ring_v);
      const int c = sh.code[cur] & kCodeStep;
      if (c >= 8) break;
      cur = tile_step(cur, c);
    }
  }
  __syncthreads();
  const int ntier = sh.ntier;
  if (static_cast<int>(threadIdx.x) < ntier) {
    const int v = S.sorted[threadIdx.x];
    int rank = 0;
    for (int q = 0; q < ntier; q++) rank += S.sorted[q] < v;
    S.tiers[rank] = v;
  }
  __syncthreads();
  // final counts; each exit-tailed cell's tier rank
  int rank[kCntPer];
#pragma unroll
  for (int j = 0; j < kCntPer; j++) {
    const int i = j * kCntBlock + threadIdx.x;
    rank[j] = -1;
    if ((sh.code[i] & kCodeStep) == kCodeNone) continue;
    cnt[static_cast<size_t>(b.y0 + i / kSide) * W + b.x0 + i % kSide] = sh.cs[i];
    if (!(info[j] & kInfoExit)) continue;
    const int tl = info_tail(info[j]);
    rank[j] = tier_rank(S.tiers, ntier, S.tier_b[border_pos(tl % kSide, tl / kSide, b)]);
    atomicAdd(&S.hist[rank[j]], 1);
  }
  __syncthreads();
  int mh = 0;
#pragma unroll
  for (int j = 0; j < kCntPer; j++)
    if (rank[j] >= 0) mh = max(mh, info_hops(info[j]));
  const int span = block_max(mh, &sh.maxhops) + 1;   // hops values per tier
  // the work items: tier ranks in order, each its cells' count; the slot
  // is taken now and written after the sort, which hides the atomic's
  // round trip
  int item_pos = -1;
  uint64_t item_word = 0;
  {
    const int v = static_cast<int>(threadIdx.x) < ntier ? S.hist[threadIdx.x] : 0;
    int total;
    const int ex = block_scan<kCntBlock>(v, sh.scan, &total);
    if (static_cast<int>(threadIdx.x) < ntier) {
      const int tier = S.tiers[threadIdx.x];
      item_pos = atomicAdd(&tcur[tier * kShards + (t & (kShards - 1))], 1);
      item_word = item_pack(wi, n0 + ex, v);
      S.hist[threadIdx.x] = n0 + ex;   // first slot of the rank
    }
  }
  // each rank's cells sorted by hops (a work item solves its cells in list
  // order): one counting sort over (rank, hops) keys when they fit the key
  // array, else one per rank (hops are below kCells, so one rank's always fit)
  uint32_t *const list = cells + static_cast<size_t>(t) * kCells;
  const int passes = ntier * span <= kCells ? 1 : ntier;
  int lp[kCntPer];   // each exit-tailed cell's position in the list
#pragma unroll
  for (int j = 0; j < kCntPer; j++) lp[j] = -1;
  for (int pass = 0; pass < passes; pass++) {
    const int nkeys = passes == 1 ? ntier * span : span;
    for (int q = threadIdx.x; q < nkeys; q += kCntBlock) S.keys[q] = 0;
    __syncthreads();
    int key[kCntPer];
#pragma unroll
    for (int j = 0; j < kCntPer; j++) {
      key[j] = -1;
      if (rank[j] < 0 || (passes > 1 && rank[j] != pass)) continue;
      key[j] = (passes == 1 ? rank[j] * span : 0) + info_hops(info[j]);
      atomicAdd(&S.keys[key[j]], 1);
    }
    __syncthreads();
    int carry = passes == 1 ? n0 : S.hist[pass];
    for (int base = 0; base < nkeys; base += kCntBlock) {
      const int q = base + threadIdx.x;
      const int v = q < nkeys ? S.keys[q] : 0;
      int total;
      const int ex = block_scan<kCntBlock>(v, sh.scan, &total);
      if (q < nkeys) S.keys[q] = carry + ex;
      carry += total;
    }
    __syncthreads();   // every cursor is set before any thread takes a slot
#pragma unroll
    for (int j = 0; j < kCntPer; j++)
      if (key[j] >= 0) lp[j] = atomicAdd(&S.keys[key[j]], 1);
    __syncthreads();
  }
  // the list entries: each cell with the position of its in-tile receiver
  // (same tail, so the same work item, one hop earlier) relative to the item
  {
    uint16_t *const lpos = reinterpret_cast<uint16_t *>(S.keys);   // free after the sort
#pragma unroll
    for (int j = 0; j < kCntPer; j++)
      if (lp[j] >= 0) lpos[j * kCntBlock + threadIdx.x] = static_cast<uint16_t>(lp[j]);
    __syncthreads();
#pragma unroll
    for (int j = 0; j < kCntPer; j++) {
      if (lp[j] < 0) continue;
      const int i = j * kCntBlock + threadIdx.x;
      const int c = sh.code[i] & kCodeStep;
      const uint32_t rp = c < 8 ? static_cast<uint32_t>(lpos[tile_step(i, c)] - S.hist[rank[j]])
                                : kNoPos;
      list[lp[j]] = static_cast<uint32_t>(i) | rp << kEntShift;
    }
  }
  if (item_pos >= 0) items[item_pos] = item_word;
  // the cells with a root tail, from the list k_deep_local sorted by hops;
  // each wave takes 64-cell chunks in order.  A lane takes its cells' old
  // elevations from memory and solves a cell once its receiver (an earlier
  // cell of the list) holds its new elevation in place of the unsolved mark,
  // which one 64-bit LDS read both signals and delivers
  int ci[kCntPer];
  double h0s[kCntPer];
#pragma unroll
  for (int c = 0; c < kCntPer; c++) {   // kCntWaves waves cover kCells in kCntPer rounds
    const int p = (wave + c * kCntWaves) * kWave + lane;
    ci[c] = static_cast<int>(list[p] & kEntCell);   // read whatever p: one wait for all
  }
#pragma unroll
  for (int c = 0; c < kCntPer; c++)
    if ((wave + c * kCntWaves) * kWave + lane >= n0) ci[c] = -1;
#pragma unroll
  for (int c = 0; c < kCntPer; c++)
    if (ci[c] >= 0)
      h0s[c] = h_old[static_cast<size_t>(b.y0 + ci[c] / kSide) * W + b.x0 + ci[c] % kSide];
  __syncthreads();   // the sort's and the entries' arrays are dead: the marks go down
  uint64_t *const hbits = sh.u.hs;
#pragma unroll
  for (int c = 0; c < kCntPer; c++)
    if (ci[c] >= 0) hbits[ci[c]] = kUnsolved;
  __syncthreads();
  auto fact_of = [&](int i) {
    const int r = sh.code[i];
    return r == kCodeRoot ? 0.0 : ftab[2 * static_cast<size_t>(sh.cs[i]) + (r & 1)];
  };
  double fnext = ci[0] >= 0 ? fact_of(ci[0]) : 0.0;
#pragma unroll
  for (int c = 0; c < kCntPer; c++) {
    const int i = ci[c];
    if (i < 0) break;   // later rounds lie further down the list
    const int r = sh.code[i];
    const bool root = r == kCodeRoot;
    const int ri = root ? i : tile_step(i, r);
    const double fact = fnext;
    if (c + 1 < kCntPer && ci[c + 1] >= 0) fnext = fact_of(ci[c + 1]);
    const double h0 = h0s[c];
    int spins = 0;
    for (;;) {
      const uint64_t rb = root ? 0
                               : __hip_atomic_load(&hbits[ri], __ATOMIC_RELAXED,
                                                   __HIP_MEMORY_SCOPE_WORKGROUP);
      const bool ready = root || rb != kUnsolved;
      const bool give_up = !ready && ++spins > kSpinCap;
      if (ready || give_up) {
        double hn = h0;
        if (give_up) {
          atomicOr(&hdr[2], kErrLookback);
        } else if (root) {
          const int gx = b.x0 + i % kSide, gy = b.y0 + i / kSide;
          if (gx >= 2 && gx < W - 2 && gy >= 2 && gy < H - 2) hn += k.ud;
        } else {
          hn = solve<NMODE>(G.nmode, h0 + k.ud, __longlong_as_double(static_cast<long long>(rb)),
                            fact, k, hdr);
        }
        __hip_atomic_store(&hbits[i], static_cast<uint64_t>(__double_as_longlong(hn)),
                           __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
        break;
      }
      __builtin_amdgcn_s_sleep(1);
    }
  }
  __syncthreads();
#pragma unroll
  for (int j = 0; j < kCntPer; j++) {
    const int i = j * kCntBlock + threadIdx.x;
    const int c = sh.code[i];
    if ((c & kCodeStep) == kCodeNone || (c & kCodeExitTail)) continue;
    h_new[static_cast<size_t>(b.y0 + i / kSide) * W + b.x0 + i % kSide] =
        __longlong_as_double(static_cast<long long>(hbits[i]));
  }
}

// ---------------------------------------------------------------------------
// k_deep_erode: persistent; one work item at a time per workgroup
// ---------------------------------------------------------------------------
constexpr int kEroBlock = 256;   // items are small and latency-bound: many workgroups per CU
constexpr int kWin = 1024;       // cells of an item held in LDS at once
struct ErodeShared {
  double hs[kWin];          // new elevation of each cell of the window, by position
  uint8_t done[kWin];
  uint64_t item;
};

// Work items (the exit-tailed cells); an item's cells come sorted by hops, so
// a cell's receiver in the tile (one hop nearer the exit) is always earlier.
// The item is solved in windows of kWin cells: a receiver in the window is
// read from LDS, one in an earlier window of the item from the new elevation
// field (final: stored before the window's barrier), one in another tile from
// the field once its mark is replaced.
//
// The items come from nq queues (k_deep_scan's layout); a workgroup takes
// one item at a time from queue blockIdx.x % nq and moves to the next queue
// only when its own is empty.  An item waits only on items of lower tiers;
// a queue hands out its items in tier order, and every queue keeps its own
// workgroups (nq <= the grid, all resident) until it is empty, so the
// lowest-tier item being waited on is always held or about to be taken by
// a running workgroup, and the grid drains.
template <int NMODE>
__global__ __launch_bounds__(kEroBlock) void k_deep_erode(
    GraphWork w, const uint8_t *__restrict__ rcv0, const uint32_t *__restrict__ cnt0,
    const uint32_t *__restrict__ cells, const uint64_t *__restrict__ items,
    const double *__restrict__ ftab0, const double *__restrict__ h_old0, double *h_new0,
    int32_t *tstat, int nq) {
  __shared__ ErodeShared sh;
  const int wave = threadIdx.x / kWave, lane = threadIdx.x % kWave;
  int q = static_cast<int>(blockIdx.x) % nq, empty = 0;   // thread 0's queue
  for (;;) {
    if (threadIdx.x == 0) {
      uint64_t it = ~uint64_t{0};
      while (empty < nq) {
        int32_t *const cur = &tstat[kStQCur + q * kQStride];
        const int end = tstat[kStQStart + q + 1];
        // another queue is looked at before it is claimed from: near the
        // end most of them are empty
        if (empty == 0 ||
            __hip_atomic_load(cur, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < end) {
          const int p = atomicAdd(cur, 1);
          if (p < end) {
            it = items[p];
            break;
          }
        }
        q = q + 1 == nq ? 0 : q + 1;
        empty++;
      }
      sh.item = it;
    }
    __syncthreads();
    const uint64_t item = sh.item;
    if (item == ~uint64_t{0}) return;
    const WorkTile wt = work_tile(w, static_cast<int>(item & ((1u << 27) - 1)));
    const WorkTile &G = wt;
    const int first = static_cast<int>((item >> 27) & 0x1FFF);
    const int n = static_cast<int>(item >> 40);
    const int W = G.W;
    const uint8_t *const rcv = rcv0 + G.off;
    const uint32_t *const cnt = cnt0 + G.off;
    const double *const ftab = ftab0 + G.ftab_off, *const h_old = h_old0 + G.off;
    double *const h_new = h_new0 + G.off;
    const ErodeConst k = G.ek;
    int32_t *const hdr = G.hdr;
    const Box b = tile_box(wt.t, G.tiles_x, W, G.H);
    const uint32_t *const tl = cells + static_cast<size_t>(wt.tg) * kCells + first;
    for (int w0 = 0; w0 < n; w0 += kWin) {
      const int wn = min(kWin, n - w0);
      for (int q = threadIdx.x; q < wn; q += kEroBlock) sh.done[q] = 0;
      __syncthreads();
      // wavefront: each wave takes 64-cell chunks in order; a lane solves its
      // cell as soon as its receiver is done.  The wait, the solve and the
      // publication share one loop, so a lane whose receiver is on its own
      // wave gets it on a later pass.
      for (int ch = wave; ch * kWave < wn; ch += kEroBlock / kWave) {
        const int q = ch * kWave + lane;
        if (q >= wn) continue;
        const uint32_t e = tl[w0 + q];
        const int i = static_cast<int>(e & kEntCell);
        const int rp = static_cast<int>(e >> kEntShift);
        const bool local = rp != static_cast<int>(kNoPos) && rp >= w0;
        const int lx = i % kSide, ly = i / kSide;
        const size_t g = static_cast<size_t>(b.y0 + ly) * W + b.x0 + lx;
        const uint8_t r = rcv[g];
        const double h0 = h_old[g];
        const double fact = ftab[2 * static_cast<size_t>(cnt[g]) + (r & 1)];
        const unsigned long long *rp_g = reinterpret_cast<const unsigned long long *>(
            h_new + static_cast<long long>(g) + dir_offset(r, W));
        int spins = 0;
        for (;;) {
          bool ready;
          double hr = 0.0;
          if (local) {
            ready = __hip_atomic_load(&sh.done[rp - w0], __ATOMIC_ACQUIRE,
                                      __HIP_MEMORY_SCOPE_WORKGROUP) != 0;
            if (ready) hr = sh.hs[rp - w0];
          } else {
            const unsigned long long v =
                __hip_atomic_load(rp_g, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            ready = v != kUnsolved;
            hr = __longlong_as_double(static_cast<long long>(v));
          }
          const bool give_up = !ready && ++spins > kSpinCap;
          if (ready || give_up) {
            double hn = h0;
            if (give_up) atomicOr(&hdr[2], kErrLookback);
            else hn = solve<NMODE>(G.nmode, h0 + k.ud, hr, fact, k, hdr);
            sh.hs[q] = hn;
            __hip_atomic_store(&sh.done[q], static_cast<uint8_t>(1), __ATOMIC_RELEASE,
                               __HIP_MEMORY_SCOPE_WORKGROUP);
            // every new elevation goes past this XCD's cache: border cells
            // may be the receivers of other tiles' exits, and a cache line
            // holds cells of two tiles wherever a row does not start on a
            // line, so a plain store beside such a receiver would leave this
            // XCD's L2 with a copy of the line that its polls could read
            __hip_atomic_store(reinterpret_cast<unsigned long long *>(h_new + g),
                               static_cast<unsigned long long>(__double_as_longlong(hn)),
                               __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            break;
          }
          __builtin_amdgcn_s_sleep(1);
        }
      }
      __syncthreads();
    }
  }
}

// The persistent erosion grid: as many workgroups as fit on the device at
// once (host side).
template <int NMODE>
int erode_grid(int ncu) {
  int per_cu = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_deep_erode<NMODE>, kEroBlock, 0) !=
          hipSuccess ||
      per_cu < 1)
    per_cu = 1;
  return per_cu * (ncu < 1 ? 1 : ncu);
}

__global__ void k_deep_clear(int32_t *tstat) {
  if (threadIdx.x < kStClear) tstat[threadIdx.x] = 0;
}

// Queues for a persistent erosion grid of `blocks` workgroups: the largest
// power of two up to kShards that leaves every queue a workgroup.
inline int erode_queues(int blocks) {
  int q = kShards;
  while (q > 1 && q > blocks) q >>= 1;
  return q;
}

}  // namespace deep
}  // namespace

#endif  // FS_DEEP_H
