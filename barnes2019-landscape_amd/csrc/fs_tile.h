// fs_tile.h -- the tile path of the time-step: region kernels that keep a
// 64x64 tile plus a halo of FS_TILE_HALO cells in LDS.  Included by
// fs_device.h with a 16-cell halo and, inside its own namespace, with a
// 32-cell halo for drainage up to 32 levels deep; it therefore has no
// include guard and relies on fs_device.h's declarations above it.
#ifndef FS_TILE_HALO
#error "define FS_TILE_HALO before including fs_tile.h"
#endif

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
constexpr int kHalo = FS_TILE_HALO;
constexpr int kReg = kTile + 2 * kHalo;       // region side
constexpr int kRegCells = kReg * kReg;
// tile accumulation: 512-thread workgroups, four tiles per CU (39 KB of LDS
// each) with the 16-cell halo; 1024 threads where a grid has too few tiles
// to give every CU four
constexpr int kAccBlock = kHalo <= 16 ? 512 : 1024;
constexpr int kAccBlockWide = 1024;
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

// Region-local index offset of direction r (0..7): one packed byte constant
// while the offsets fit a byte (kReg <= 126), else from the step tables.
__device__ __forceinline__ int region_offset(int r) {
  if constexpr (kReg <= 126) {
    constexpr int kOff[8] = {-1, -kReg - 1, -kReg, -kReg + 1, 1, kReg + 1, kReg, kReg - 1};
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
  } else {
    int lx = 0, ly = 0;
    (void)region_step(r, lx, ly);
    return lx + ly * kReg;
  }
}

// Pass 1: upstream cell counts (ComputeFlowAcc, RB:262-279) and levels
// (receiver hops to an outlet, the level GenerateOrder assigns, RB:213-254)
// of the tile's cells.  Every region cell walks its receiver chain once and
// adds 1 to each cell it passes, itself included, so a cell's count is the
// number of region cells draining through it -- its whole upstream area when
// every level is at most kHalo.  A tile cell deeper than that (or whose chain
// leaves the region) raises hdr[3]; the host then runs the level path.
// Halo cells the erosion pass solves per tile at most: those on the receiver
// chains of the tile's cells (with a 16-cell halo: mean 60-66, at most 92 per
// tile of 5120 over 120 steps at the benchmark preset).
constexpr int kHaloCap = kHalo <= 16 ? 256 : 512;

__device__ __forceinline__ bool in_tile(int lx, int ly) {
  return lx >= kHalo && lx < kHalo + kTile && ly >= kHalo && ly < kHalo + kTile;
}

// Accumulation word of a region cell in LDS: bits 0-14 its count, bit 15
// set for a tile cell, bits 16-31 four times the signed region-index offset
// of its receiver (the byte offset in the word array) or of a stop code.  One
// atomic add both counts a walker's visit and tells it where to go next: the
// returned word shifted right by 16 is the step of the walker's byte address.
// The stop codes lie below every offset (-128..-126 for a 16-cell halo, so
// that they also fit a byte).
constexpr uint32_t kAccCount = 0x7FFFu;
constexpr uint32_t kAccTile = 1u << 15;
constexpr int kStopRoot = -(kReg + 32);        // an outlet (NO_FLOW)
constexpr int kStopLeaves = kStopRoot + 1;     // the receiver lies outside the region
constexpr int kStopOutside = kStopRoot + 2;    // beyond the grid
constexpr int kStopMax = kStopOutside;         // codes <= kStopMax end a walk
static_assert(kRegCells <= static_cast<int>(kAccCount) && kReg + 1 < -kStopMax,
              "accumulation word layout");
// offsets and stop codes fit one signed byte: the dword loader's lookups apply
constexpr bool kByteCodes = kStopRoot >= -128;

static_assert(4 * kStopRoot >= -32768, "byte offsets fit the word's top half");
// words above this (as signed integers) carry an offset, not a stop code
constexpr int32_t kStopWordMax = static_cast<int32_t>(
    static_cast<uint32_t>(4 * kStopMax) << 16 | 0xFFFFu);
__device__ __forceinline__ int acc_code(uint32_t w) { return static_cast<int32_t>(w) >> 18; }
__device__ __forceinline__ uint32_t acc_word(int code, bool tile) {
  return static_cast<uint32_t>((4 * code) & 0xFFFF) << 16 | (tile ? kAccTile : 0u);
}

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
    w_s[i] = acc_word(code, in_tile(lx, ly));
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
    // byte b of code, sign-extended and times four, into bits 16-31 of word b
    uint4 q;
    q.x = (static_cast<uint32_t>(static_cast<int32_t>(code << 24) >> 6) & 0xFFFC0000u) | t;
    q.y = (static_cast<uint32_t>(static_cast<int32_t>(code << 16) >> 6) & 0xFFFC0000u) | t;
    q.z = (static_cast<uint32_t>(static_cast<int32_t>(code << 8) >> 6) & 0xFFFC0000u) | t;
    q.w = (static_cast<uint32_t>(static_cast<int32_t>(code) >> 6) & 0xFFFC0000u) | t;
    *reinterpret_cast<uint4 *>(&w_s[4 * d]) = q;
  }
}

// Mark (bitmask over region cells) the halo cells on the receiver chains of
// the tile's cells: walks start at the tile's border cells and stop at an
// outlet, the region edge, a tile cell, an already marked cell or after
// kHalo hops.  Each newly marked cell is also appended to list (the
// first kHaloCap of them; *nlist counts them all).  mark and *nlist must be
// zero and w_s loaded before the call (counts may still be changing: only the
// offset and the tile bit are read).
template <int BLOCK>
__device__ __forceinline__ void mark_halo_chains(const uint32_t *w_s, uint32_t *mark,
                                                 uint16_t *list, uint32_t *nlist) {
  // A chain leaves the tile from a cell of its border and re-enters (if it
  // does) at another border cell, so walks from the 4*kTile-4 border cells,
  // one per thread, mark every halo cell on the chains of all tile cells.
  constexpr int kBorder = 4 * kTile - 4;
  static_assert(kBorder <= BLOCK, "one border cell per thread");
  const int b = threadIdx.x;
  if (b >= kBorder) return;
  int tx, ty;
  if (b < kTile) {
    tx = b;
    ty = 0;
  } else if (b < 2 * kTile) {
    tx = b - kTile;
    ty = kTile - 1;
  } else if (b < 3 * kTile - 2) {
    tx = 0;
    ty = b - 2 * kTile + 1;
  } else {
    tx = kTile - 1;
    ty = b - (3 * kTile - 2) + 1;
  }
  int c = (kHalo + ty) * kReg + kHalo + tx;
  int code = acc_code(w_s[c]);
  for (int hop = 0; hop <= kHalo && code > kStopMax; hop++) {
    c += code;
    const uint32_t wd = w_s[c];
    if (wd & kAccTile) break;
    const uint32_t bit = 1u << (c & 31);
    if (atomicOr(&mark[c >> 5], bit) & bit) break;   // marked by another walk
    const uint32_t k = atomicAdd(nlist, 1u);
    if (k < static_cast<uint32_t>(kHaloCap)) list[k] = static_cast<uint16_t>(c);
    code = acc_code(wd);
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
  uint32_t nlist;                 // marked cells appended to hlist
  uint16_t hlist[kHaloCap];       // marked cells in marking order
  int over, maxlev;
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
// tile cells' level histogram to it; with max_level, raises *max_level to the
// deepest level among the tile's cells that reach an outlet.
template <int BLOCK>
__device__ __forceinline__ int tile_acc_body(AccShared &sh, const uint8_t *__restrict__ rcv,
                                             uint32_t *__restrict__ cnt,
                                             uint8_t *__restrict__ lvl,
                                             TilePlan *__restrict__ plan, int W, int H, int t,
                                             int tiles_x, int32_t *level_hist,
                                             int32_t *max_level,
                                             unsigned long long *phase = nullptr) {
  static_assert(kRegCells / 32 <= BLOCK && kHalo + 1 <= BLOCK, "one mark word per thread");
  uint32_t *w_s = sh.w_s;
  unsigned long long *const ph = phase_row(phase);
  const uint64_t t0 = ph ? wall_clock64() : 0;
  const int x0 = (t % tiles_x) * kTile, y0 = (t / tiles_x) * kTile;
  const int rx0 = x0 - kHalo, ry0 = y0 - kHalo;
  if (threadIdx.x == 0) {
    sh.over = sh.maxlev = 0;
    sh.nlist = 0;
  }
  if (threadIdx.x < kHalo + 1) sh.hist[threadIdx.x] = 0;
  if (threadIdx.x < kRegCells / 32) sh.mark[threadIdx.x] = 0;
  bool done = false;
  if constexpr (kByteCodes) {
    if (rx0 >= 0 && ry0 >= 0 && rx0 + kReg <= W && ry0 + kReg <= H && W % 4 == 0) {
      load_region_words_x4<BLOCK>(rcv, w_s, rx0, ry0, W);
      done = true;
    }
  }
  if (!done) load_region_words<BLOCK>(rcv, w_s, rx0, ry0, W, H);
  __syncthreads();
  const uint64_t t1 = ph ? wall_clock64() : 0;
  // Every region cell walks its receiver chain and adds 1 to each cell it
  // passes, itself included; the thread's chains advance together, one cell
  // per round.  A walker stops at an outlet (its level is the round it got
  // there), where its chain leaves the region, or after kHalo hops (kHalo + 1
  // rounds).  Walkers hold byte addresses into w_s; a stopped walker holds
  // the complement of the address of the cell that stopped it, so whether it
  // reached an outlet is read back from that cell's code afterwards.
  constexpr int kPer = kRegCells / BLOCK;
  static_assert(kPer <= 32, "walkers per thread (bits of live)");
  int xa[kPer], hops[kPer];
  uint32_t live = 0;
#pragma unroll
  for (int j = 0; j < kPer; j++) {
    const int i = j * BLOCK + threadIdx.x;
    hops[j] = 0;
    if (acc_code(w_s[i]) != kStopOutside) {
      xa[j] = 4 * i;
      live |= 1u << j;
    } else {
      xa[j] = ~(4 * i);
    }
  }
  for (int round = 0; round <= kHalo && live; round++) {
#pragma unroll
    for (int j = 0; j < kPer; j++) {
      if (xa[j] < 0) continue;
      uint32_t *w = reinterpret_cast<uint32_t *>(reinterpret_cast<char *>(w_s) + xa[j]);
      const int32_t wd = static_cast<int32_t>(atomicAdd(w, 1u));
      if (wd > kStopWordMax) {
        xa[j] += wd >> 16;
      } else {
        xa[j] = ~xa[j];
        hops[j] = round;
        live &= ~(1u << j);
      }
    }
  }
  // the level of a cell is the length of its walk; only slots whose region
  // rows can meet the tile's rows hold tile cells
  int over = 0, deepest = 0;
  constexpr int kJ0 = (kHalo * kReg) / BLOCK;
  constexpr int kJ1 = ((kHalo + kTile) * kReg - 1) / BLOCK;
  // region column and row of slot j's cell, stepped from slot kJ0's
  constexpr int kStepY = BLOCK / kReg, kStepX = BLOCK % kReg;
  int lx = (kJ0 * BLOCK + static_cast<int>(threadIdx.x)) % kReg;
  int ly = (kJ0 * BLOCK + static_cast<int>(threadIdx.x)) / kReg;
#pragma unroll
  for (int j = kJ0; j <= kJ1; j++) {
    const int tx = lx - kHalo, ty = ly - kHalo;
    lx += kStepX;
    ly += kStepY;
    if (lx >= kReg) {
      lx -= kReg;
      ly++;
    }
    if (static_cast<unsigned>(tx) >= static_cast<unsigned>(kTile) ||
        static_cast<unsigned>(ty) >= static_cast<unsigned>(kTile))
      continue;
    if (x0 + tx >= W || y0 + ty >= H) continue;     // beyond the grid
    const bool rooted = xa[j] < 0 && acc_code(w_s[~xa[j] >> 2]) == kStopRoot;
    // a level when rooted; stored straight away (a wave's cells are a run of
    // region cells, so the bytes mostly share lines)
    lvl[static_cast<size_t>(y0 + ty) * W + x0 + tx] = static_cast<uint8_t>(hops[j]);
    if (rooted) {
      if (level_hist) atomicAdd(&sh.hist[hops[j]], 1);
      deepest = max(deepest, hops[j]);
    } else {
      over = 1;
    }
  }
  if (over) sh.over = 1;
  if (max_level) {
#pragma unroll
    for (int d = 1; d < kWave; d <<= 1) deepest = max(deepest, __shfl_xor(deepest, d));
    if ((threadIdx.x & (kWave - 1)) == 0 && deepest > 0) atomicMax(&sh.maxlev, deepest);
  }
  uint64_t t2 = 0;
  if (ph) {
    __syncthreads();
    t2 = wall_clock64();
  }
  mark_halo_chains<BLOCK>(w_s, sh.mark, sh.hlist, &sh.nlist);
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
  if (max_level && threadIdx.x == 0 && sh.maxlev > 0) atomicMax(max_level, sh.maxlev);
  if (level_hist && threadIdx.x < kHalo + 1 && sh.hist[threadIdx.x])
    atomicAdd(&level_hist[threadIdx.x], sh.hist[threadIdx.x]);
  if (!tile_over) {
    // the plan: every needed cell's receiver slot; halo cells' region index
    // and level (the hops from the cell to its outlet, all inside the region
    // since the tile is not over)
    static_assert(kHaloCap <= BLOCK, "one pass over the marked cells");
    if (threadIdx.x < nhalo) {
      const int c = sh.hlist[threadIdx.x];
      const int s = kTileCells + sh.pre[c >> 5] + __popc(sh.mark[c >> 5] & ((1u << (c & 31)) - 1u));
      plan->rslot[s] = plan_receiver(w_s, sh.mark, sh.pre, c);
      plan->hreg[s - kTileCells] = static_cast<uint16_t>(c);
      int lev = 0, cc = c, code = acc_code(w_s[c]);
      while (code > kStopMax && lev <= kHalo) {
        cc += code;
        code = acc_code(w_s[cc]);
        lev++;
      }
      plan->hlvl[s - kTileCells] = static_cast<uint8_t>(lev);
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
  uint64_t t3 = 0;
  if (ph) {
    __syncthreads();
    t3 = wall_clock64();
  }
  for (int it = 0; it < kTileCells / BLOCK; it++) {
    const int i = it * BLOCK + threadIdx.x;
    const int tx = i % kTile, ty = i / kTile;
    const int gx = x0 + tx, gy = y0 + ty;
    if (gx >= W || gy >= H) continue;
    const size_t gi = static_cast<size_t>(gy) * W + gx;
    cnt[gi] = w_s[(ty + kHalo) * kReg + tx + kHalo] & kAccCount;
  }
  if (ph) {
    __syncthreads();
    if (threadIdx.x == 0) {
      const uint64_t t4 = wall_clock64();
      atomicAdd(&ph[kPhAccLoad], t1 - t0);
      atomicAdd(&ph[kPhAccWalk], t2 - t1);
      atomicAdd(&ph[kPhAccPlan], t3 - t2);
      atomicAdd(&ph[kPhAccStore], t4 - t3);
    }
  }
  return tile_over;
}

// kTimed: this launch records phase ticks (the host runs it on a sample of
// steps; the untimed instantiation carries no timing code at all).
template <bool kTimed, int BLOCK = kAccBlock>
__global__ __launch_bounds__(BLOCK) void k_tile_acc(const uint8_t *__restrict__ rcv,
                                                        uint32_t *__restrict__ cnt,
                                                        uint8_t *__restrict__ lvl,
                                                        TilePlan *__restrict__ plan, int32_t *hdr,
                                                        int W, int H, int tiles_x, int ntiles,
                                                        unsigned long long *phase) {
  __shared__ AccShared sh;
  const int t = xcd_tile(blockIdx.x, ntiles);
  // with the deep halo, hdr[0] (cleared by k_fill_counts) receives the
  // deepest tile-cell level, so the host can step back to the 16-cell halo
  const int over = tile_acc_body<BLOCK>(sh, rcv, cnt, lvl, plan + t, W, H, t, tiles_x, nullptr,
                                            kHalo > 16 ? hdr : nullptr,
                                            kTimed ? phase : nullptr);
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
  uint16_t c_s[kSlots];          // counts (at most kRegCells when the tile path holds)
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
                                                int32_t *hdr,
                                                unsigned long long *phase = nullptr) {
  constexpr int kWaves = kEroBlock / kWave;
  unsigned long long *const ph = phase_row(phase);
  const uint64_t t0 = ph ? wall_clock64() : 0;
  constexpr int kKeys = kHalo + 1;
  constexpr int kTilePer = kTileCells / kEroBlock;
  constexpr int kSlotPer = (kSlots + kEroBlock - 1) / kEroBlock;
  constexpr int kKeyBits = kKeys <= 32 ? 5 : 6;
  static_assert(kKeys <= 64 && kRegCells <= 65535, "level keys and 16-bit counts");
  static_assert(kHaloCap <= kEroBlock && 2 * kEroFactCache <= kEroBlock, "block shape");
  static_assert(kEroFactCache <= kFactCache, "factor table holds the cached counts");
  double *hs = sh.hs, *ft_s = sh.ft_s;
  uint16_t *c_s = sh.c_s;
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
    const bool halo = static_cast<int>(threadIdx.x) < nh;
    const int hr = halo ? plan->hreg[threadIdx.x] : 0;
    const uint8_t hl = halo ? plan->hlvl[threadIdx.x] : 0;
    const size_t hg = static_cast<size_t>(ry0 + hr / kReg) * W + (rx0 + hr % kReg);
    const double hh = halo ? h_old[hg] : 0.0;
    const uint32_t hc = halo ? cnt[hg] : 0u;
#pragma unroll
    for (int j = 0; j < kTilePer; j++) {
      hs[j * kEroBlock + threadIdx.x] = th[j];
      c_s[j * kEroBlock + threadIdx.x] = static_cast<uint16_t>(tc[j]);
      d_s[j * kEroBlock + threadIdx.x] = tl[j];
    }
#pragma unroll
    for (int q = 0; q < kSlotPer; q++) {
      const int sl = q * kEroBlock + threadIdx.x;
      if (sl < kSlots) rs_s[sl] = rv[q];
    }
    if (halo) {
      hs[kTileCells + threadIdx.x] = hh;
      c_s[kTileCells + threadIdx.x] = static_cast<uint16_t>(hc);
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
  // with kKeyBits ballots, no atomics; then level bases
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
    for (int b = 0; b < kKeyBits; b++) {
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
  const uint64_t t1 = ph ? wall_clock64() : 0;
  // level 0 first: uplift of the interior outlets; rings 0 and 1 stay fixed
  int ch = wave;
  for (; ch < root_chunks; ch += kWaves) {
    const int sl = list[ch * kWave + lane];
    if (sl == 0xFFFF) continue;
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
    __hip_atomic_store(&d_s[sl], kDone, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_WORKGROUP);
  }
  uint64_t tm = 0;
  if (ph) {
    __syncthreads();
    tm = wall_clock64();
  }
  // levels 1 and up, each cell as soon as its receiver is done
  for (; ch < nchunks; ch += kWaves) {
    const int sl = list[ch * kWave + lane];
    if (sl == 0xFFFF) continue;
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
    __hip_atomic_store(&d_s[sl], kDone, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_WORKGROUP);
  }
  __syncthreads();
  const uint64_t t2 = ph ? wall_clock64() : 0;
#pragma unroll
  for (int j = 0; j < kTilePer; j++) {
    const int tj = j * kEroBlock + threadIdx.x;
    const int gx = x0 + tj % kTile, gy = y0 + tj / kTile;
    if (gx >= W || gy >= H) continue;
    h_new[static_cast<size_t>(gy) * W + gx] = hs[tj];
  }
  if (ph) {
    __syncthreads();
    if (threadIdx.x == 0) {
      const uint64_t t3 = wall_clock64();
      atomicAdd(&ph[kPhEroBucket], t1 - t0);
      atomicAdd(&ph[kPhEroRoots], tm - t1);
      atomicAdd(&ph[kPhEroLoop], t2 - tm);
      atomicAdd(&ph[kPhEroStore], t3 - t2);
    }
  }
}

template <int NMODE, bool kTimed>
__global__ __launch_bounds__(kEroBlock) void k_tile_erode(
    const double *__restrict__ h_old, double *__restrict__ h_new,
    const TilePlan *__restrict__ plan, const uint8_t *__restrict__ lvl,
    const uint32_t *__restrict__ cnt, const double *__restrict__ ftab, int W, int H,
    int tiles_x, int ntiles, ErodeConst k, int32_t *hdr, unsigned long long *phase) {
  __shared__ EroShared sh;
  // launched right behind the accumulation, before the host has read its
  // outcome: when a tile was too deep (hdr[3], the plans are incomplete)
  // the whole grid does nothing
  if (hdr[3] != 0) return;
  const int t = xcd_tile(blockIdx.x, ntiles);
  tile_erode_body<NMODE>(sh, h_old, h_new, plan + t, lvl, cnt, ftab, W, H, t, tiles_x, k, hdr,
                         kTimed ? phase : nullptr);
}

