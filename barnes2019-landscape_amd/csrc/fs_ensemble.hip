// fs_ensemble.hip -- many FastScape grids ("realizations") of different
// sizes and parameters advanced by ONE persistent cooperative launch per
// time-step.
//
// Every realization takes the reference's time-step (fastscape_RB.cpp:366-376)
// and gets the same bits as a single-grid model with its size, parameters and
// initial terrain.  A step is one launch of k_ens_step, a grid of one
// 1024-thread workgroup per CU, in phases separated by grid barriers:
//
//   R  receivers of every realization (RB:160-182), 256-column x recv_rows
//      work items dealt round-robin; block 0 lists the tile candidates
//      (realizations whose drainage was at most kHalo levels deep last step)
//   A  tile accumulation (upstream counts and levels, RB:213-279) of every
//      candidate's 64x64 tiles; a tile deeper than kHalo marks its
//      realization deep
//   -  block 0 lists the erosion work (the shallow realizations' tiles) and
//      the deep realizations (a tile deeper than kHalo, or deeper than kHalo
//      last step) for the tile-graph kernels
//   E  work items claimed in list order from one counter: tile uplift +
//      erosion in LDS (RB:285-328)
//
// The deep realizations then take one tile-graph step together (fs_deep.h):
// per 64x64 tile with no halo, every deep realization's tiles in the same
// launches, the tiles joined by the graph of the cells where drainage leaves
// a tile; the step also counts their levels, so the next step knows which
// are shallow again.
//
// Results do not depend on the schedule (work item sizes): each cell's
// elevation depends only on its own inputs and its receiver's final value,
// and upstream counts are integers.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <map>
#include <new>
#include <numeric>
#include <string>
#include <thread>
#include <type_traits>
#include <vector>

#include "fs_internal.h"
#include "fs_hostutil.h"

#include "fs_device.h"
#include "fs_deep.h"

namespace {

constexpr int kEnsBlock = 1024;
constexpr int kRecvLanes = 256;                     // columns per receiver work item
constexpr int kRecvQuarters = kEnsBlock / kRecvLanes;
constexpr int kHistLen = kHalo + 1;
constexpr int kSpinCap = 1 << 22;                   // bounded waits (seconds)
constexpr int kMaxRealizations = 1 << 20;

// per-realization state words (kRH ints each)
constexpr int kRH = 16;
constexpr int kRhDepth = 0;       // deepest level of the last step
constexpr int kRhOver = 1;        // tiles deeper than kHalo this step
constexpr int kRhErr = 2;         // error bits (same slot as a model's hdr[2])
constexpr int kRhPath = 7;        // path of the last step: 1 tiles, 3 tile graph
constexpr int kRhCand = 8;        // tile candidate this step
constexpr int kRhTileSteps = 9;
constexpr int kRhLevelSteps = 10;   // steps off the tiles (the tile graph)

// global control words
constexpr int kCtlBarrier = 0, kCtlClaim = 1, kCtlAbort = 2;
constexpr int kCtlEntriesA = 3, kCtlItemsA = 4, kCtlEntriesE = 5, kCtlItemsE = 6;
constexpr int kCtlGraph = 8;      // the tile-graph work list's size: [8] entries, [9] tiles
constexpr int kCtlInts = 10;

struct EnsReal {
  uint64_t off;        // first cell of the realization in the cell arrays
  uint64_t ftab_off;   // first entry of its erosion factor table
  uint64_t lev_off;    // first entry of its level counts
  uint64_t plan_off;   // first tile plan of the realization
  int W, H, tiles_x, ntiles;
  int lev_cap;         // level entries available
  int nmode;
  int recv_xb;         // receiver column blocks
  int pad;
  double d_card, d_diag;
  ErodeConst ek;
};

struct EnsArgs {
  const EnsReal *real;
  int R;
  int recv_rows, tiles_per_claim;
  double *h_old, *h_new;
  uint8_t *rcv, *lvl;
  uint32_t *cnt;
  TilePlan *plan;
  const double *ftab;
  int32_t *rh, *hist, *ctl;
  const int32_t *recv_first;   // R+1 prefix of receiver work items
  int32_t *tabA_off, *tabA_real, *tabE_off, *tabE_real;
  int32_t *tabG_off, *tabG_real;   // the deep realizations (tile-graph work list)
  uint32_t bar_base;
};

__device__ __forceinline__ int32_t ld_agent(const int32_t *p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Barrier of `members` workgroups on a monotonically increasing counter;
// `target` is this workgroup's running arrival target.  The wait is bounded:
// on expiry the launch is marked aborted and every later wait returns at
// once, so the grid always drains.
__device__ __forceinline__ void counter_barrier(int32_t *bar, uint32_t &target, int members,
                                                int32_t *ctl) {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
  __syncthreads();
  target += static_cast<uint32_t>(members);
  if (threadIdx.x == 0) {
    __hip_atomic_fetch_add(bar, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    int spins = 0;
    while (static_cast<int32_t>(static_cast<uint32_t>(ld_agent(bar)) - target) < 0) {
      if (ld_agent(&ctl[kCtlAbort])) break;
      if (++spins > kSpinCap) {
        atomicOr(&ctl[kCtlAbort], 1);
        break;
      }
      __builtin_amdgcn_s_sleep(2);
    }
  }
  __syncthreads();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
}

// Last index i in [0, n) with off[i] <= v (off ascending, off[0] <= v).
__device__ __forceinline__ int find_entry(const int32_t *off, int n, int v) {
  int lo = 0, hi = n - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (off[mid] <= v) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

struct TableShared {
  uint32_t scan[kEnsBlock / kWave + 1];
};

union EnsShared {
  AccShared acc;
  EroShared ero;
  TableShared tab;
};

// Block 0: list the realizations with count(r) > 0, in increasing r among
// those with pick(r), appending after `entries`/`items`; off[e] is the first
// work item of entry e.
template <class Pick, class Count>
__device__ void append_entries(int R, Pick pick, Count count, int32_t *off, int32_t *real,
                               int &entries, int &items, TableShared &ts) {
  for (int base = 0; base < R; base += kEnsBlock) {
    const int r = base + threadIdx.x;
    int v = 0;
    if (r < R && pick(r)) v = count(r);
    uint32_t te, ti;
    const uint32_t ee = block_excl_scan<kEnsBlock>(v > 0 ? 1u : 0u, &te, ts.scan);
    const uint32_t ei = block_excl_scan<kEnsBlock>(static_cast<uint32_t>(v), &ti, ts.scan);
    if (v > 0) {
      off[entries + ee] = items + static_cast<int>(ei);
      real[entries + ee] = r;
    }
    entries += static_cast<int>(te);
    items += static_cast<int>(ti);
  }
}

template <int NM, class F>
__device__ __forceinline__ void with_nmode(int nm, F f) {
  if constexpr (NM >= 0) {
    f(std::integral_constant<int, NM>{});
  } else {
    if (nm == 1) f(std::integral_constant<int, 1>{});
    else if (nm == 2) f(std::integral_constant<int, 2>{});
    else f(std::integral_constant<int, 0>{});
  }
}

// One time-step of every realization.  NM is the pow mode shared by all
// realizations, or -1 for a mix.
template <int NM>
__global__ __launch_bounds__(kEnsBlock) void k_ens_step(EnsArgs a) {
  __shared__ EnsShared sh;
  __shared__ int s_item;
  const int R = a.R;
  uint32_t bar = a.bar_base;
  // ---- phase R: per-step state, the tile candidates, receivers -----------
  for (int r = blockIdx.x * kEnsBlock + threadIdx.x; r < R; r += gridDim.x * kEnsBlock) {
    int32_t *q = a.rh + static_cast<size_t>(r) * kRH;
    q[kRhOver] = 0;
    for (int l = 0; l < kHistLen; l++) a.hist[static_cast<size_t>(r) * kHistLen + l] = 0;
  }
  if (blockIdx.x == 0) {
    if (threadIdx.x == 0) a.ctl[kCtlClaim] = 0;
    int entries = 0, items = 0;
    append_entries(
        R, [&](int r) { return a.rh[static_cast<size_t>(r) * kRH + kRhDepth] <= kHalo; },
        [&](int r) { return a.real[r].ntiles; }, a.tabA_off, a.tabA_real, entries, items,
        sh.tab);
    if (threadIdx.x == 0) {
      a.tabA_off[entries] = items;
      a.ctl[kCtlEntriesA] = entries;
      a.ctl[kCtlItemsA] = items;
    }
    for (int r = threadIdx.x; r < R; r += kEnsBlock)
      a.rh[static_cast<size_t>(r) * kRH + kRhCand] =
          a.rh[static_cast<size_t>(r) * kRH + kRhDepth] <= kHalo;
  }
  {
    const int quarter = threadIdx.x / kRecvLanes, lane = threadIdx.x % kRecvLanes;
    const int nitems = a.recv_first[R];
    const int spi = a.recv_rows / kRecvRows;
    for (int item = blockIdx.x * kRecvQuarters + quarter; item < nitems;
         item += gridDim.x * kRecvQuarters) {
      const int r = find_entry(a.recv_first, R + 1, item);
      const EnsReal &g = a.real[r];
      const int local = item - a.recv_first[r];
      const int x = (local % g.recv_xb) * kRecvLanes + lane + 2;
      if (x >= g.W - 2) continue;
      const int nstrips = (g.H - 4 + kRecvRows - 1) / kRecvRows;
      const int s0 = (local / g.recv_xb) * spi, s1 = min(nstrips, s0 + spi);
      receivers_strips(a.h_old + g.off, a.rcv + g.off, g.W, g.H, x, s0, s1, g.d_card, g.d_diag);
    }
  }
  counter_barrier(&a.ctl[kCtlBarrier], bar, gridDim.x, a.ctl);
  // ---- phase A: tile accumulation of the candidates -----------------------
  {
    const int entries = ld_agent(&a.ctl[kCtlEntriesA]), nitems = ld_agent(&a.ctl[kCtlItemsA]);
    for (int b = blockIdx.x; b < nitems; b += gridDim.x) {
      const int item = xcd_tile(b, nitems);
      const int e = find_entry(a.tabA_off, entries + 1, item);
      const int r = a.tabA_real[e];
      const EnsReal &g = a.real[r];
      const int t = item - a.tabA_off[e];
      const int over = tile_acc_body<kEnsBlock>(
          sh.acc, a.rcv + g.off, a.cnt + g.off, a.lvl + g.off, a.plan + g.plan_off + t, g.W, g.H,
          t, g.tiles_x, a.hist + static_cast<size_t>(r) * kHistLen, nullptr);
      if (threadIdx.x == 0 && over) atomicAdd(&a.rh[static_cast<size_t>(r) * kRH + kRhOver], 1);
      __syncthreads();
    }
  }
  counter_barrier(&a.ctl[kCtlBarrier], bar, gridDim.x, a.ctl);
  // ---- erosion work list ----------------------------------------------------
  if (blockIdx.x == 0) {
    auto deep = [&](int r) {
      const int32_t *q = a.rh + static_cast<size_t>(r) * kRH;
      return !q[kRhCand] || q[kRhOver] > 0;
    };
    for (int r = threadIdx.x; r < R; r += kEnsBlock) {
      int32_t *q = a.rh + static_cast<size_t>(r) * kRH;
      if (deep(r)) {
        q[kRhPath] = 3;
        q[kRhLevelSteps]++;
      } else {
        int D = 0;
        for (int l = 0; l < kHistLen; l++)
          if (a.hist[static_cast<size_t>(r) * kHistLen + l]) D = l;
        q[kRhDepth] = D;
        q[kRhPath] = 1;
        q[kRhTileSteps]++;
      }
    }
    __syncthreads();
    int entries = 0, items = 0;
    // the deep realizations' tiles: the tile-graph kernels' work list
    append_entries(R, deep, [&](int r) { return a.real[r].ntiles; }, a.tabG_off, a.tabG_real,
                   entries, items, sh.tab);
    if (threadIdx.x == 0) {
      a.tabG_off[entries] = items;
      a.ctl[kCtlGraph] = entries;
      a.ctl[kCtlGraph + 1] = items;
    }
    entries = items = 0;
    const int tpc = a.tiles_per_claim;
    append_entries(
        R, [&](int r) { return !deep(r); },
        [&](int r) { return (a.real[r].ntiles + tpc - 1) / tpc; }, a.tabE_off, a.tabE_real,
        entries, items, sh.tab);
    if (threadIdx.x == 0) {
      a.tabE_off[entries] = items;
      a.ctl[kCtlEntriesE] = entries;
      a.ctl[kCtlItemsE] = items;
    }
  }
  counter_barrier(&a.ctl[kCtlBarrier], bar, gridDim.x, a.ctl);
  // ---- phase E: erosion -------------------------------------------------------
  const int entries = ld_agent(&a.ctl[kCtlEntriesE]), nitems = ld_agent(&a.ctl[kCtlItemsE]);
  for (;;) {
    if (threadIdx.x == 0) s_item = atomicAdd(&a.ctl[kCtlClaim], 1);
    __syncthreads();
    const int item = s_item;
    __syncthreads();
    if (item >= nitems) break;
    const int e = find_entry(a.tabE_off, entries + 1, item);
    const int r = a.tabE_real[e];
    const EnsReal &g = a.real[r];
    int32_t *q = a.rh + static_cast<size_t>(r) * kRH;
    const int j = item - a.tabE_off[e];
    const int t1 = min(g.ntiles, (j + 1) * a.tiles_per_claim);
    const ErodeConst ek = g.ek;
    for (int tile = j * a.tiles_per_claim; tile < t1; tile++) {
      with_nmode<NM>(g.nmode, [&](auto c) {
        tile_erode_body<decltype(c)::value>(sh.ero, a.h_old + g.off, a.h_new + g.off,
                                            a.plan + g.plan_off + tile, a.lvl + g.off,
                                            a.cnt + g.off,
                                            a.ftab + g.ftab_off, g.W, g.H, tile, g.tiles_x,
                                            ek, q);
      });
      __syncthreads();
    }
    __syncthreads();
  }
}

constexpr int kBarriersPerStep = 3;

}  // namespace

// ===========================================================================
// host side
// ===========================================================================
struct fs_ensemble {
  int device = 0;
  int R = 0;
  size_t cells = 0;
  hipStream_t stream = nullptr;
  std::vector<fs_realization> spec;   // internal order
  std::vector<int> to_internal;       // user index -> internal index
  std::vector<EnsReal> real;
  int nm_all = -1;                    // shared pow mode, -1 when mixed
  int grid = 0;

  EnsReal *real_d = nullptr;
  double *h[2] = {nullptr, nullptr};
  uint8_t *rcv = nullptr, *lvl = nullptr;
  uint32_t *cnt = nullptr;
  TilePlan *plan = nullptr;
  int32_t *lev = nullptr;                     // per-level counts of tile-graph steps
  double *ftab = nullptr;
  int32_t *rh = nullptr, *hist = nullptr, *ctl = nullptr;
  int32_t *tabA_off = nullptr, *tabA_real = nullptr, *tabE_off = nullptr, *tabE_real = nullptr;
  int32_t *tabG_off = nullptr, *tabG_real = nullptr;
  int32_t *graph_n_h = nullptr;               // pinned: the tile-graph work list's size
  // the tile-graph step of the deep realizations (fs_deep.h)
  deep::GraphGrid *graph_grid = nullptr;
  uint32_t *dinfo = nullptr, *link = nullptr;
  uint64_t *node_word = nullptr, *items = nullptr;
  uint32_t *node_kids = nullptr, *node_parent = nullptr, *node_out = nullptr;
  uint32_t *node_up = nullptr, *node_level = nullptr;
  int32_t *node_cell = nullptr, *nexit = nullptr, *tile_n0 = nullptr;
  uint16_t *node_tier = nullptr;
  uint32_t *cells_list = nullptr;
  int32_t *thist = nullptr, *tcur = nullptr, *tstat = nullptr;
  int graph_blocks = 0, graph_queues = 1;
  uint64_t graph_tiles = 0;                   // tile-graph tiles stepped so far
  std::map<int, int32_t *> recv_first;   // per recv_rows
  int cur = 0;
  uint32_t bar_base = 0;
  fs_schedule sched{64, 1, 0};
  bool broken = false, have_step = false;
  uint64_t steps = 0;
  double step_ms = 0;
  hipEvent_t ev[2] = {nullptr, nullptr};
  // autotuning
  int tune_trial = 0, tune_interval = 0;
  uint64_t tuned_at = ~0ull;
  std::vector<fs_schedule> tune_cand;
  std::vector<double> tune_ms;
};

namespace {

void free_ensemble(fs_ensemble *e) {
  if (!e) return;
  (void)hipSetDevice(e->device);
  if (e->stream) (void)hipStreamSynchronize(e->stream);
  for (void *p : {static_cast<void *>(e->real_d), static_cast<void *>(e->h[0]),
                  static_cast<void *>(e->h[1]), static_cast<void *>(e->rcv),
                  static_cast<void *>(e->lvl), static_cast<void *>(e->cnt),
                  static_cast<void *>(e->plan), static_cast<void *>(e->lev),
                  static_cast<void *>(e->ftab), static_cast<void *>(e->rh),
                  static_cast<void *>(e->hist), static_cast<void *>(e->ctl),
                  static_cast<void *>(e->tabA_off), static_cast<void *>(e->tabA_real),
                  static_cast<void *>(e->tabE_off), static_cast<void *>(e->tabE_real),
                  static_cast<void *>(e->tabG_off), static_cast<void *>(e->tabG_real),
                  static_cast<void *>(e->graph_grid), static_cast<void *>(e->dinfo),
                  static_cast<void *>(e->link), static_cast<void *>(e->node_word),
                  static_cast<void *>(e->items), static_cast<void *>(e->node_kids),
                  static_cast<void *>(e->node_parent), static_cast<void *>(e->node_out),
                  static_cast<void *>(e->node_up), static_cast<void *>(e->node_level),
                  static_cast<void *>(e->node_cell), static_cast<void *>(e->nexit),
                  static_cast<void *>(e->tile_n0), static_cast<void *>(e->node_tier),
                  static_cast<void *>(e->cells_list), static_cast<void *>(e->thist),
                  static_cast<void *>(e->tcur), static_cast<void *>(e->tstat)})
    if (p) (void)hipFree(p);
  if (e->graph_n_h) (void)hipHostFree(e->graph_n_h);
  for (auto &kv : e->recv_first)
    if (kv.second) (void)hipFree(kv.second);
  for (auto &ev : e->ev)
    if (ev) (void)hipEventDestroy(ev);
  if (e->stream) (void)hipStreamDestroy(e->stream);
  delete e;
}

int check_ensemble(fs_ensemble *e) {
  if (!e) return FS_ERR_INVALID_ARGUMENT;
  if (e->broken) {
    fs::set_last_error("ensemble is unusable after an earlier device error");
    return FS_ERR_INTERNAL;
  }
  FS_HIP_TRY(hipSetDevice(e->device));
  return FS_OK;
}

int check_schedule(const fs_schedule &s) {
  if (s.recv_rows < kRecvRows || s.recv_rows > 256 || s.recv_rows % kRecvRows != 0 ||
      s.tiles_per_claim < 1 || s.tiles_per_claim > 16 || s.group_cells < 0)
    return FS_ERR_INVALID_ARGUMENT;
  return FS_OK;   // group_cells is accepted and has no effect (no level path any more)
}

// Receiver work items of every realization for `rows` rows per item.
int recv_table(fs_ensemble *e, int rows, const int32_t **out) {
  auto it = e->recv_first.find(rows);
  if (it == e->recv_first.end()) {
    std::vector<int32_t> first(e->R + 1);
    long long run = 0;
    for (int r = 0; r < e->R; r++) {
      first[r] = static_cast<int32_t>(run);
      const EnsReal &g = e->real[r];
      const long long strips = (g.H - 4 + kRecvRows - 1) / kRecvRows;
      const long long per = rows / kRecvRows;
      run += static_cast<long long>(g.recv_xb) * ((strips + per - 1) / per);
    }
    if (run >= (1ll << 31)) return FS_ERR_DIMENSIONS;
    first[e->R] = static_cast<int32_t>(run);
    int32_t *d = nullptr;
    FS_TRY(fs::dev_alloc(&d, first.size()));
    e->recv_first[rows] = d;
    FS_HIP_TRY(hipMemcpy(d, first.data(), first.size() * sizeof(int32_t), hipMemcpyHostToDevice));
    it = e->recv_first.find(rows);
  }
  *out = it->second;
  return FS_OK;
}

template <int NM>
int launch_step_nm(fs_ensemble *e, const EnsArgs &a) {
  EnsArgs args = a;
  void *params[] = {&args};
  FS_HIP_TRY(hipLaunchCooperativeKernel(reinterpret_cast<const void *>(&k_ens_step<NM>),
                                        dim3(e->grid), dim3(kEnsBlock), params, 0, e->stream));
  return FS_OK;
}

// The tile-graph step of the realizations the step kernel listed as deep
// (fs_deep.h), on the same stream.  The host reads the list's tile count
// (one wait per step) so every per-tile kernel is launched with exactly one
// workgroup per tile.
template <int NM>
int graph_launches(fs_ensemble *e, const deep::GraphWork &w, int T) {
  hipStream_t s = e->stream;
  const double *h_old = e->h[e->cur];
  double *h_new = e->h[e->cur ^ 1];
  hipLaunchKernelGGL(deep::k_deep_clear, dim3(1), dim3(64), 0, s, e->tstat);
  hipLaunchKernelGGL(deep::k_deep_local, dim3(T), dim3(deep::kLocalBlock), 0, s, w, e->rcv, e->cnt,
                     e->dinfo, e->link, e->node_word, e->node_kids, e->node_cell, e->nexit,
                     e->tile_n0, e->cells_list, h_new, e->tstat);
  hipLaunchKernelGGL(deep::k_deep_links, dim3(T), dim3(deep::kNodeCap), 0, s, w, e->rcv, e->link,
                     e->dinfo, e->node_cell, e->nexit, e->node_parent, e->node_kids, e->node_up);
  hipLaunchKernelGGL(deep::k_deep_tiers, dim3(T), dim3(deep::kNodeCap), 0, s, w, e->node_parent,
                     e->nexit, e->node_tier, e->node_up, e->node_level, e->thist, e->tstat);
  hipLaunchKernelGGL(deep::k_deep_scan, dim3(1), dim3(deep::kBlock), 0, s, e->thist, e->tcur,
                     e->tstat, e->graph_queues);
  hipLaunchKernelGGL(deep::k_deep_sum, dim3(T), dim3(deep::kNodeCap), 0, s, w, e->node_word,
                     e->node_kids, e->node_parent, e->nexit, e->node_out, e->tstat);
  hipLaunchKernelGGL(deep::k_deep_levels, dim3(T), dim3(deep::kBlock), 0, s, w, e->dinfo, e->link,
                     e->node_level);
  hipLaunchKernelGGL(deep::k_deep_counts<NM>, dim3(T), dim3(deep::kCntBlock), 0, s, w, e->rcv,
                     e->cnt, e->dinfo, e->link, e->node_out, e->node_cell, e->node_tier, e->nexit,
                     e->tile_n0, e->cells_list, e->tcur, e->items, e->ftab, h_old, h_new);
  hipLaunchKernelGGL(deep::k_deep_erode<NM>, dim3(static_cast<unsigned>(e->graph_blocks)),
                     dim3(deep::kEroBlock), 0, s, w, e->rcv, e->cnt, e->cells_list, e->items,
                     e->ftab, h_old, h_new, e->tstat, e->graph_queues);
  const hipError_t err = hipGetLastError();
  if (err != hipSuccess) {
    fs::set_last_error(std::string("tile-graph launch: ") + hipGetErrorString(err));
    return FS_ERR_HIP;
  }
  return FS_OK;
}

int graph_step(fs_ensemble *e) {
  FS_HIP_TRY(hipMemcpyAsync(e->graph_n_h, e->ctl + kCtlGraph, 2 * sizeof(int32_t),
                            hipMemcpyDeviceToHost, e->stream));
  FS_HIP_TRY(hipStreamSynchronize(e->stream));
  const int T = e->graph_n_h[1];
  if (T <= 0) return FS_OK;
  deep::GraphWork w;
  w.grid = e->graph_grid;
  w.first = e->tabG_off;
  w.which = e->tabG_real;
  w.size = e->ctl + kCtlGraph;
  e->graph_tiles += static_cast<uint64_t>(T);
  if (e->nm_all == 1) return graph_launches<1>(e, w, T);
  if (e->nm_all == 2) return graph_launches<2>(e, w, T);
  return graph_launches<-1>(e, w, T);
}

int launch_step(fs_ensemble *e) {
  const int32_t *rf = nullptr;
  FS_TRY(recv_table(e, e->sched.recv_rows, &rf));
  EnsArgs a{};
  a.real = e->real_d;
  a.R = e->R;
  a.recv_rows = e->sched.recv_rows;
  a.tiles_per_claim = e->sched.tiles_per_claim;
  a.h_old = e->h[e->cur];
  a.h_new = e->h[e->cur ^ 1];
  a.rcv = e->rcv;
  a.lvl = e->lvl;
  a.cnt = e->cnt;
  a.plan = e->plan;
  a.ftab = e->ftab;
  a.rh = e->rh;
  a.hist = e->hist;
  a.ctl = e->ctl;
  a.recv_first = rf;
  a.tabA_off = e->tabA_off;
  a.tabA_real = e->tabA_real;
  a.tabE_off = e->tabE_off;
  a.tabE_real = e->tabE_real;
  a.tabG_off = e->tabG_off;
  a.tabG_real = e->tabG_real;
  a.bar_base = e->bar_base;
  int st;
  if (e->nm_all == 1) st = launch_step_nm<1>(e, a);
  else if (e->nm_all == 2) st = launch_step_nm<2>(e, a);
  else st = launch_step_nm<-1>(e, a);
  if (st == FS_OK) st = graph_step(e);
  if (st != FS_OK) {
    e->broken = true;
    return st;
  }
  e->bar_base += static_cast<uint32_t>(kBarriersPerStep * e->grid);
  e->cur ^= 1;
  e->steps++;
  e->have_step = true;
  return FS_OK;
}

// Device error bits of every realization and the abort flag, after a sync.
int ensemble_errors(fs_ensemble *e) {
  FS_HIP_TRY(hipStreamSynchronize(e->stream));
  int32_t ctl[kCtlInts];
  FS_HIP_TRY(hipMemcpy(ctl, e->ctl, sizeof(ctl), hipMemcpyDeviceToHost));
  if (ctl[kCtlAbort]) {
    e->broken = true;
    fs::set_last_error("a grid barrier timed out");
    return FS_ERR_INTERNAL;
  }
  int32_t ts[deep::kStWords];
  FS_HIP_TRY(hipMemcpy(ts, e->tstat, sizeof(ts), hipMemcpyDeviceToHost));
  if (ts[deep::kStErr]) {
    e->broken = true;
    fs::set_last_error(ts[deep::kStErr] & deep::kErrTier
                           ? "tile-graph step: a drainage path crosses more tile borders than levels allow"
                           : "tile-graph step: receivers do not form a forest");
    return ts[deep::kStErr] & deep::kErrTier ? FS_ERR_LEVEL_OVERFLOW : FS_ERR_INTERNAL;
  }
  std::vector<int32_t> rh(static_cast<size_t>(e->R) * kRH);
  FS_HIP_TRY(hipMemcpy(rh.data(), e->rh, rh.size() * sizeof(int32_t), hipMemcpyDeviceToHost));
  for (int r = 0; r < e->R; r++) {
    const int32_t err = rh[static_cast<size_t>(r) * kRH + kRhErr];
    if (!err) continue;
    e->broken = true;
    fs::set_last_error("realization error flags " + std::to_string(err));
    if (err & kErrNewton) return FS_ERR_NEWTON;
    if (err & kErrDepth) return FS_ERR_LEVEL_OVERFLOW;
    return FS_ERR_INTERNAL;
  }
  return FS_OK;
}

double elapsed_ms(hipEvent_t a, hipEvent_t b) {
  float ms = 0.f;
  if (hipEventElapsedTime(&ms, a, b) != hipSuccess) return -1.0;
  return ms;
}

// Time `steps` real steps with schedule s; returns ms per step (< 0 on error).
int timed_steps(fs_ensemble *e, const fs_schedule &s, int steps, double *ms_per_step) {
  e->sched = s;
  FS_HIP_TRY(hipEventRecord(e->ev[0], e->stream));
  for (int i = 0; i < steps; i++) FS_TRY(launch_step(e));
  FS_HIP_TRY(hipEventRecord(e->ev[1], e->stream));
  FS_HIP_TRY(hipEventSynchronize(e->ev[1]));
  const double ms = elapsed_ms(e->ev[0], e->ev[1]);
  e->step_ms += ms;
  *ms_per_step = ms / steps;
  return FS_OK;
}

// Coordinate search over the schedule knobs, each candidate timed on
// tune_trial real steps; stops early when `budget` steps are used.
// Returns the number of steps taken in *used.
int autotune(fs_ensemble *e, int budget, int *used) {
  *used = 0;
  const std::vector<int> rows_vals = {16, 64, 256};
  const std::vector<int> tpc_vals = {1, 2, 4};
  fs_schedule best = e->sched;
  double best_ms = -1;
  for (int knob = 1; knob < 3; knob++) {
    const std::vector<int> &vals = knob == 1 ? rows_vals : tpc_vals;
    for (int v : vals) {
      if (*used + e->tune_trial > budget) {
        e->sched = best;
        return FS_OK;
      }
      fs_schedule c = best;
      (knob == 0 ? c.group_cells : knob == 1 ? c.recv_rows : c.tiles_per_claim) = v;
      if (best_ms >= 0 && c.group_cells == best.group_cells && c.recv_rows == best.recv_rows &&
          c.tiles_per_claim == best.tiles_per_claim)
        continue;
      double ms = 0;
      FS_TRY(timed_steps(e, c, e->tune_trial, &ms));
      *used += e->tune_trial;
      e->tune_cand.push_back(c);
      e->tune_ms.push_back(ms);
      if (best_ms < 0 || ms < best_ms) {
        best_ms = ms;
        best = c;
      }
    }
  }
  e->sched = best;
  return FS_OK;
}

int run_steps(fs_ensemble *e, long long total) {
  long long done = 0;
  FS_HIP_TRY(hipEventRecord(e->ev[0], e->stream));
  bool open = true;   // ev[0] marks the start of untimed-by-tuner steps
  double plain_ms = 0;
  while (done < total) {
    const bool due = e->tune_trial > 0 &&
                     (e->tuned_at == ~0ull ||
                      (e->tune_interval > 0 && e->steps >= e->tuned_at + e->tune_interval));
    if (due) {
      if (open) {
        FS_HIP_TRY(hipEventRecord(e->ev[1], e->stream));
        FS_HIP_TRY(hipEventSynchronize(e->ev[1]));
        plain_ms += elapsed_ms(e->ev[0], e->ev[1]);
        open = false;
      }
      e->tuned_at = e->steps;
      int used = 0;
      FS_TRY(autotune(e, static_cast<int>(std::min<long long>(total - done, 1 << 30)), &used));
      done += used;
      continue;
    }
    if (!open) {
      FS_HIP_TRY(hipEventRecord(e->ev[0], e->stream));
      open = true;
    }
    FS_TRY(launch_step(e));
    done++;
  }
  if (open) {
    FS_HIP_TRY(hipEventRecord(e->ev[1], e->stream));
    FS_HIP_TRY(hipEventSynchronize(e->ev[1]));
    plain_ms += elapsed_ms(e->ev[0], e->ev[1]);
  }
  e->step_ms += plain_ms;
  return ensemble_errors(e);
}

int check_index(fs_ensemble *e, int r, size_t count, int *ri) {
  if (r < 0 || r >= e->R) return FS_ERR_INVALID_ARGUMENT;
  *ri = e->to_internal[r];
  const EnsReal &g = e->real[*ri];
  if (count != static_cast<size_t>(g.W) * g.H) return FS_ERR_BUFFER_SIZE;
  return FS_OK;
}

template <class T>
int download(fs_ensemble *e, T *dst, const T *src, size_t count) {
  FS_HIP_TRY(hipStreamSynchronize(e->stream));
  FS_HIP_TRY(hipMemcpy(dst, src, count * sizeof(T), hipMemcpyDeviceToHost));
  return FS_OK;
}

size_t align_up(size_t v, size_t a) { return (v + a - 1) / a * a; }

}  // namespace

extern "C" {

int fs_ensemble_create(const fs_realization *reals, int count, int device, fs_ensemble **out) {
  if (!out) return FS_ERR_INVALID_ARGUMENT;
  *out = nullptr;
  if (!reals || count <= 0 || count > kMaxRealizations) return FS_ERR_INVALID_ARGUMENT;
  for (int r = 0; r < count; r++) FS_TRY(fs::validate_params(reals[r].width, reals[r].height, &reals[r].params));
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0 || device < 0 || device >= ndev) {
    fs::set_last_error("no HIP device " + std::to_string(device));
    return FS_ERR_NO_DEVICE;
  }
  fs_ensemble *e = new (std::nothrow) fs_ensemble();
  if (!e) return FS_ERR_OUT_OF_MEMORY;
  struct Guard {
    fs_ensemble *e;
    ~Guard() { if (e) free_ensemble(e); }
  } guard{e};
  e->device = device;
  e->R = count;
  FS_HIP_TRY(hipSetDevice(device));
  int coop = 0;
  FS_HIP_TRY(hipDeviceGetAttribute(&coop, hipDeviceAttributeCooperativeLaunch, device));
  if (!coop) {
    fs::set_last_error("device does not support cooperative launches");
    return FS_ERR_NO_DEVICE;
  }
  // largest realizations first: the erosion list starts with them
  std::vector<int> ord(count);
  std::iota(ord.begin(), ord.end(), 0);
  std::stable_sort(ord.begin(), ord.end(), [&](int a, int b) {
    return static_cast<long long>(reals[a].width) * reals[a].height >
           static_cast<long long>(reals[b].width) * reals[b].height;
  });
  e->to_internal.resize(count);
  e->spec.resize(count);
  e->real.resize(count);
  size_t cells = 0, ftab_n = 0, lev_n = 0;
  int nm0 = -2;
  for (int i = 0; i < count; i++) {
    const fs_realization &s = reals[ord[i]];
    e->to_internal[ord[i]] = i;
    e->spec[i] = s;
    EnsReal &g = e->real[i];
    const size_t n = static_cast<size_t>(s.width) * s.height;
    g.off = cells;
    g.W = s.width;
    g.H = s.height;
    g.tiles_x = (s.width + kTile - 1) / kTile;
    g.ntiles = g.tiles_x * ((s.height + kTile - 1) / kTile);
    g.recv_xb = (s.width - 4 + kRecvLanes - 1) / kRecvLanes;
    g.d_card = s.params.dx;
    g.d_diag = s.params.dx * fs::kSqrt2;
    g.ek = make_erode_const(s.params, &g.nmode);
    g.ftab_off = ftab_n;
    g.lev_off = lev_n;
    g.lev_cap = static_cast<int>(std::min<size_t>(n, FS_MAX_LEVELS) + 3);
    nm0 = nm0 == -2 ? g.nmode : (nm0 == g.nmode ? nm0 : -1);
    cells += align_up(n, 64);
    ftab_n += align_up(2 * (std::max<size_t>(n, kFactCache - 1) + 1), 64);
    lev_n += align_up(static_cast<size_t>(g.lev_cap), 64);
  }
  e->nm_all = nm0 == 0 ? -1 : nm0;
  e->cells = cells;
  long long tiles_total = 0;
  for (EnsReal &g : e->real) {
    g.plan_off = static_cast<uint64_t>(tiles_total);
    tiles_total += g.ntiles;
  }
  if (tiles_total >= (1ll << 31)) return FS_ERR_DIMENSIONS;

  FS_HIP_TRY(hipStreamCreateWithFlags(&e->stream, hipStreamNonBlocking));
  for (auto &ev : e->ev) FS_HIP_TRY(hipEventCreate(&ev));
  FS_TRY(fs::dev_alloc(&e->real_d, count));
  FS_TRY(fs::dev_alloc(&e->h[0], cells));
  FS_TRY(fs::dev_alloc(&e->h[1], cells));
  FS_TRY(fs::dev_alloc(&e->rcv, cells + 16));
  FS_TRY(fs::dev_alloc(&e->lvl, cells));
  FS_TRY(fs::dev_alloc(&e->cnt, cells));
  FS_TRY(fs::dev_alloc(&e->plan, static_cast<size_t>(tiles_total)));
  FS_TRY(fs::dev_alloc(&e->lev, lev_n));
  FS_TRY(fs::dev_alloc(&e->ftab, ftab_n));
  FS_TRY(fs::dev_alloc(&e->rh, static_cast<size_t>(count) * kRH));
  FS_TRY(fs::dev_alloc(&e->hist, static_cast<size_t>(count) * kHistLen));
  FS_TRY(fs::dev_alloc(&e->ctl, kCtlInts));
  FS_TRY(fs::dev_alloc(&e->tabA_off, count + 1));
  FS_TRY(fs::dev_alloc(&e->tabA_real, count + 1));
  FS_TRY(fs::dev_alloc(&e->tabE_off, count + 1));
  FS_TRY(fs::dev_alloc(&e->tabE_real, count + 1));
  FS_TRY(fs::dev_alloc(&e->tabG_off, count + 1));
  FS_TRY(fs::dev_alloc(&e->tabG_real, count + 1));
  // the tile-graph step (deep realizations): the realizations' tiles are
  // numbered by plan_off, so the per-tile arrays cover every tile
  {
    const size_t T = static_cast<size_t>(tiles_total);
    const size_t nodes = T * deep::kNodeCap;
    const size_t tier_words = static_cast<size_t>(deep::kTierCap + 1) * deep::kShards;
    FS_TRY(fs::dev_alloc(&e->graph_grid, count));
    FS_TRY(fs::dev_alloc(&e->dinfo, cells));
    FS_TRY(fs::dev_alloc(&e->link, cells));
    FS_TRY(fs::dev_alloc(&e->node_word, nodes));
    FS_TRY(fs::dev_alloc(&e->node_kids, nodes));
    FS_TRY(fs::dev_alloc(&e->node_parent, nodes));
    FS_TRY(fs::dev_alloc(&e->node_out, nodes));
    FS_TRY(fs::dev_alloc(&e->node_up, nodes));
    FS_TRY(fs::dev_alloc(&e->node_level, nodes));
    FS_TRY(fs::dev_alloc(&e->node_cell, nodes));
    FS_TRY(fs::dev_alloc(&e->node_tier, nodes));
    FS_TRY(fs::dev_alloc(&e->nexit, T));
    FS_TRY(fs::dev_alloc(&e->tile_n0, T));
    FS_TRY(fs::dev_alloc(&e->cells_list, T * deep::kCells));
    FS_TRY(fs::dev_alloc(&e->items, T * deep::kTiersPerTile));
    FS_TRY(fs::dev_alloc(&e->thist, tier_words));
    FS_TRY(fs::dev_alloc(&e->tcur, tier_words));
    FS_TRY(fs::dev_alloc(&e->tstat, deep::kStWords));
    FS_HIP_TRY(hipMemset(e->thist, 0, tier_words * sizeof(int32_t)));
    FS_HIP_TRY(hipMemset(e->tstat, 0, deep::kStWords * sizeof(int32_t)));
    if (hipHostMalloc(reinterpret_cast<void **>(&e->graph_n_h), 2 * sizeof(int32_t)) !=
        hipSuccess) {
      e->graph_n_h = nullptr;
      fs::set_last_error("hipHostMalloc failed");
      return FS_ERR_OUT_OF_MEMORY;
    }
  }
  FS_HIP_TRY(hipMemcpy(e->real_d, e->real.data(), count * sizeof(EnsReal), hipMemcpyHostToDevice));
  {
    std::vector<deep::GraphGrid> gg(count);
    for (int r = 0; r < count; r++) {
      const EnsReal &g = e->real[r];
      deep::GraphGrid &d = gg[r];
      d.off = g.off;
      d.ftab_off = g.ftab_off;
      d.hdr = e->rh + static_cast<size_t>(r) * kRH;   // [0] depth, [2] error bits
      d.levels = e->lev + g.lev_off;
      d.W = g.W;
      d.H = g.H;
      d.tiles_x = g.tiles_x;
      d.tile0 = static_cast<int>(g.plan_off);
      d.nlev = g.lev_cap;
      d.nmode = g.nmode;
      d.ek = g.ek;
    }
    FS_HIP_TRY(hipMemcpy(e->graph_grid, gg.data(), gg.size() * sizeof(deep::GraphGrid),
                         hipMemcpyHostToDevice));
  }
  FS_HIP_TRY(hipMemset(e->rcv, 0xFF, cells + 16));   // rings never get a receiver, RB:361-362
  FS_HIP_TRY(hipMemset(e->cnt, 0, cells * sizeof(uint32_t)));
  FS_HIP_TRY(hipMemset(e->rh, 0, static_cast<size_t>(count) * kRH * sizeof(int32_t)));
  FS_HIP_TRY(hipMemset(e->hist, 0, static_cast<size_t>(count) * kHistLen * sizeof(int32_t)));
  FS_HIP_TRY(hipMemset(e->ctl, 0, kCtlInts * sizeof(int32_t)));

  // Initial terrains and erosion factor tables.  Host threads fill host
  // buffers, one realization each per round; this thread alone makes every
  // HIP call (no HIP state is created on threads that exit).
  unsigned nt = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
  nt = std::min<unsigned>(nt, static_cast<unsigned>(count));
  std::vector<std::vector<double>> terr(nt), tabs(nt);
  for (int r0 = 0; r0 < count; r0 += static_cast<int>(nt)) {
    const int nr = std::min(static_cast<int>(nt), count - r0);
    std::vector<int> status(static_cast<size_t>(nr), FS_OK);
    auto fill = [&](int w) {
      const int i = r0 + w;
      const EnsReal &g = e->real[i];
      const fs_realization &s = e->spec[i];
      const size_t n = static_cast<size_t>(g.W) * g.H;
      const size_t kmax = std::max<size_t>(n, kFactCache - 1);
      try {
        terr[w].resize(n);
        tabs[w].resize(2 * (kmax + 1));
      } catch (const std::bad_alloc &) {
        status[w] = FS_ERR_OUT_OF_MEMORY;
        return;
      }
      fs::generate_terrain(terr[w].data(), g.W, g.H, s.engine_seed);
      fs::fill_fact_table(tabs[w].data(), kmax, s.params.cell_area, s.params.meq, g.ek.kdt,
                          g.ek.lenpow0, g.ek.lenpow1);
    };
    {
      std::vector<std::thread> pool;
      for (int w = 1; w < nr; w++) pool.emplace_back(fill, w);
      fill(0);
      for (auto &th : pool) th.join();
    }
    for (int w = 0; w < nr; w++) {
      if (status[w] != FS_OK) {
        fs::set_last_error("initialising realizations failed");
        return status[w];
      }
      const EnsReal &g = e->real[r0 + w];
      FS_HIP_TRY(hipMemcpy(e->h[0] + g.off, terr[w].data(), terr[w].size() * sizeof(double),
                           hipMemcpyHostToDevice));
      FS_HIP_TRY(hipMemcpy(e->ftab + g.ftab_off, tabs[w].data(), tabs[w].size() * sizeof(double),
                           hipMemcpyHostToDevice));
    }
  }

  int per_cu = 0;
  hipError_t oe;
  if (e->nm_all == 1)
    oe = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_ens_step<1>, kEnsBlock, 0);
  else if (e->nm_all == 2)
    oe = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_ens_step<2>, kEnsBlock, 0);
  else
    oe = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_ens_step<-1>, kEnsBlock, 0);
  FS_HIP_TRY(oe);
  int ncu = 0;
  FS_HIP_TRY(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, device));
  if (per_cu < 1 || ncu < 1) {
    fs::set_last_error("the step kernel does not fit on a CU");
    return FS_ERR_INTERNAL;
  }
  e->grid = per_cu * ncu;
  e->graph_blocks = e->nm_all == 1   ? deep::erode_grid<1>(ncu)
                    : e->nm_all == 2 ? deep::erode_grid<2>(ncu)
                                     : deep::erode_grid<-1>(ncu);
  e->graph_queues = deep::erode_queues(e->graph_blocks);
  const int32_t *rf = nullptr;
  FS_TRY(recv_table(e, e->sched.recv_rows, &rf));
  FS_HIP_TRY(hipDeviceSynchronize());
  guard.e = nullptr;
  *out = e;
  return FS_OK;
}

int fs_ensemble_destroy(fs_ensemble *e) {
  if (!e) return FS_ERR_INVALID_ARGUMENT;
  free_ensemble(e);
  return FS_OK;
}

int fs_ensemble_size(fs_ensemble *e, int *count, size_t *cells) {
  if (!e) return FS_ERR_INVALID_ARGUMENT;
  if (count) *count = e->R;
  if (cells) {
    size_t c = 0;
    for (const EnsReal &g : e->real) c += static_cast<size_t>(g.W) * g.H;
    *cells = c;
  }
  return FS_OK;
}

int fs_ensemble_run(fs_ensemble *e, int nstep) {
  FS_TRY(check_ensemble(e));
  if (nstep < 0) return FS_ERR_INVALID_ARGUMENT;
  return run_steps(e, static_cast<long long>(nstep) + 1);
}

int fs_ensemble_step(fs_ensemble *e) {
  FS_TRY(check_ensemble(e));
  return run_steps(e, 1);
}

int fs_ensemble_get_h(fs_ensemble *e, int r, double *h, size_t count) {
  FS_TRY(check_ensemble(e));
  if (!h) return FS_ERR_INVALID_ARGUMENT;
  int ri;
  FS_TRY(check_index(e, r, count, &ri));
  return download(e, h, e->h[e->cur] + e->real[ri].off, count);
}

int fs_ensemble_set_h(fs_ensemble *e, int r, const double *h, size_t count) {
  FS_TRY(check_ensemble(e));
  if (!h) return FS_ERR_INVALID_ARGUMENT;
  int ri;
  FS_TRY(check_index(e, r, count, &ri));
  FS_HIP_TRY(hipStreamSynchronize(e->stream));
  FS_HIP_TRY(hipMemcpy(e->h[e->cur] + e->real[ri].off, h, count * sizeof(double),
                       hipMemcpyHostToDevice));
  return FS_OK;
}

int fs_ensemble_get_receivers(fs_ensemble *e, int r, int32_t *rec, size_t count) {
  FS_TRY(check_ensemble(e));
  if (!rec) return FS_ERR_INVALID_ARGUMENT;
  int ri;
  FS_TRY(check_index(e, r, count, &ri));
  std::vector<uint8_t> v(count);
  FS_TRY(download(e, v.data(), e->rcv + e->real[ri].off, count));
  for (size_t i = 0; i < count; i++) rec[i] = v[i] == kNoFlow ? FS_NO_FLOW : v[i];
  return FS_OK;
}

int fs_ensemble_get_accum(fs_ensemble *e, int r, double *accum, size_t count) {
  FS_TRY(check_ensemble(e));
  if (!accum) return FS_ERR_INVALID_ARGUMENT;
  int ri;
  FS_TRY(check_index(e, r, count, &ri));
  const EnsReal &g = e->real[ri];
  int32_t q[kRH];
  FS_TRY(download(e, q, e->rh + static_cast<size_t>(ri) * kRH, kRH));
  std::vector<uint32_t> v(count);
  FS_TRY(download(e, v.data(), e->cnt + g.off, count));
  const double area = e->spec[ri].params.cell_area;
  for (size_t i = 0; i < count; i++) accum[i] = static_cast<double>(v[i]) * area;
  return FS_OK;
}

int fs_ensemble_get_levels(fs_ensemble *e, int r, int32_t *levels, size_t capacity,
                           int32_t *nlevel) {
  FS_TRY(check_ensemble(e));
  if (!nlevel || r < 0 || r >= e->R) return FS_ERR_INVALID_ARGUMENT;
  *nlevel = 0;
  if (!e->have_step) return FS_OK;
  const int ri = e->to_internal[r];
  const EnsReal &g = e->real[ri];
  const int32_t n = g.W * g.H;
  int32_t q[kRH];
  FS_TRY(download(e, q, e->rh + static_cast<size_t>(ri) * kRH, kRH));
  const int D = q[kRhDepth];
  if (D < 0 || D + 3 > g.lev_cap) return FS_ERR_INTERNAL;
  std::vector<int32_t> lv(D + 3);
  if (q[kRhPath] == 3) {
    // the tile-graph step counted cells per level
    std::vector<int32_t> c(D + 1);
    FS_TRY(download(e, c.data(), e->lev + g.lev_off, c.size()));
    lv[0] = 0;
    for (int l = 0; l <= D; l++) lv[l + 1] = lv[l] + c[l];
    lv[D + 2] = n;
  } else {
    int32_t hist[kHistLen];
    FS_TRY(download(e, hist, e->hist + static_cast<size_t>(ri) * kHistLen, kHistLen));
    if (D >= kHistLen) return FS_ERR_INTERNAL;
    lv[0] = 0;
    for (int l = 0; l <= D; l++) lv[l + 1] = lv[l] + hist[l];
    lv[D + 2] = n;
  }
  if (lv[D + 1] != n) {
    fs::set_last_error("level sizes cover " + std::to_string(lv[D + 1]) + " of " +
                       std::to_string(n) + " cells");
    return FS_ERR_INTERNAL;
  }
  *nlevel = D + 3;
  if (!levels) return FS_OK;
  if (capacity < static_cast<size_t>(*nlevel)) return FS_ERR_BUFFER_SIZE;
  std::copy(lv.begin(), lv.end(), levels);
  return FS_OK;
}

int fs_ensemble_path_stats(fs_ensemble *e, int r, uint64_t *tile_steps, uint64_t *level_steps) {
  FS_TRY(check_ensemble(e));
  if (r < 0 || r >= e->R) return FS_ERR_INVALID_ARGUMENT;
  int32_t q[kRH];
  FS_TRY(download(e, q, e->rh + static_cast<size_t>(e->to_internal[r]) * kRH, kRH));
  if (tile_steps) *tile_steps = static_cast<uint32_t>(q[kRhTileSteps]);
  if (level_steps) *level_steps = static_cast<uint32_t>(q[kRhLevelSteps]);
  return FS_OK;
}

int fs_ensemble_set_schedule(fs_ensemble *e, const fs_schedule *s) {
  FS_TRY(check_ensemble(e));
  if (!s) return FS_ERR_INVALID_ARGUMENT;
  FS_TRY(check_schedule(*s));
  const int32_t *rf = nullptr;
  FS_TRY(recv_table(e, s->recv_rows, &rf));
  e->sched = *s;
  return FS_OK;
}

int fs_ensemble_get_schedule(fs_ensemble *e, fs_schedule *s) {
  if (!e || !s) return FS_ERR_INVALID_ARGUMENT;
  *s = e->sched;
  return FS_OK;
}

int fs_ensemble_autotune(fs_ensemble *e, int trial_steps, int interval) {
  FS_TRY(check_ensemble(e));
  if (trial_steps < 0 || interval < 0) return FS_ERR_INVALID_ARGUMENT;
  e->tune_trial = trial_steps;
  e->tune_interval = interval;
  e->tuned_at = ~0ull;
  if (trial_steps > 0) {
    const int32_t *rf = nullptr;
    for (int rows : {16, 64, 256}) FS_TRY(recv_table(e, rows, &rf));
  }
  return FS_OK;
}

int fs_ensemble_tune_log(fs_ensemble *e, fs_schedule *cand, double *ms_per_step, int capacity,
                         int *count) {
  if (!e || !count) return FS_ERR_INVALID_ARGUMENT;
  *count = static_cast<int>(e->tune_cand.size());
  if (!cand && !ms_per_step) return FS_OK;
  if (capacity < *count) return FS_ERR_BUFFER_SIZE;
  for (int i = 0; i < *count; i++) {
    if (cand) cand[i] = e->tune_cand[i];
    if (ms_per_step) ms_per_step[i] = e->tune_ms[i];
  }
  return FS_OK;
}

int fs_ensemble_stats(fs_ensemble *e, uint64_t *steps, double *step_ms) {
  if (!e) return FS_ERR_INVALID_ARGUMENT;
  if (steps) *steps = e->steps;
  if (step_ms) *step_ms = e->step_ms;
  return FS_OK;
}

int fs_ensemble_synchronize(fs_ensemble *e) {
  FS_TRY(check_ensemble(e));
  FS_HIP_TRY(hipStreamSynchronize(e->stream));
  return FS_OK;
}

}  // extern "C"
