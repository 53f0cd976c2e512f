// fastscape_hip.hip -- FastScape time-step on MI355X (gfx950).
//
// One time-step (the body of the reference's loop, fastscape_RB.cpp:366-376)
// is a short chain of kernels on the model's stream:
//
//   k_receivers  D8 steepest-descent receiver of every interior cell
//                (RB:160-182) -> rcv, one byte per cell
//   k_topology   donor bitmask of every cell (the inverse of rcv, RB:188-203)
//                and the level of every cell, i.e. its number of receiver
//                hops to an outlet, with a per-level histogram; cells deeper
//                than kWalkCap hops are finished by k_jump (pointer jumping)
//   -- host reads the level histogram once per step, so every later launch
//      is sized exactly --
//   k_roots      level 0: every NO_FLOW cell in ascending index (RB:223-226),
//                a single-pass stream compaction (decoupled look-back)
//   k_bfs        level L+1 = donors of level L's cells in level-L order, each
//                donor list in ascending cell index (RB:234-247), one
//                single-pass scan per level; runs of small levels go to one
//                workgroup (k_bfs_small), the analogue of RB+P's
//                `if(lvlsize>500)` serial fallback (fastscape_RB+P.cpp:328)
//   k_flowacc    upstream cell counts, pulled level by level from the top
//                (pull form of RB:262-279, as fastscape_RB+PI.cpp:306-323)
//   k_uplift     uplift of the interior outlets (pits) (RB:285-296); every
//                other interior cell receives its uplift inside k_erode
//   k_erode      implicit stream-power update of each level from level 1 up,
//                the reference's Newton iteration replayed operation by
//                operation (RB:306-328)
//
// Arithmetic is IEEE double with contraction off (-ffp-contract=off): the
// slope is a true division, the erosion factor is (K*dt)*A^m / dr^n with
// A^m read from a table filled by the host libm (the same pow the reference
// calls), so receivers, levels, the stack and, for n = 1, elevations are
// bit-identical to the reference's.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <memory>
#include <mutex>
#include <new>
#include <string>
#include <vector>

#include "fs_internal.h"
#include "fs_hostutil.h"

#include "fs_device.h"

// The same tile kernels with a 32-cell halo, for drainage 17..32 levels deep.
namespace {
namespace tile32 {
#define FS_TILE_HALO 32
#include "fs_tile.h"
#undef FS_TILE_HALO
}  // namespace tile32
}  // namespace

#include "fs_deep.h"

namespace {
constexpr int kGraphRetry = 64;   // graph steps between attempts of the 32-cell halo
}  // namespace

// ===========================================================================
// host side
// ===========================================================================
struct KTimed {
  int kid;
  hipEvent_t a, b;
};

// A run of levels swept by one launch.
struct Span {
  int lo, hi;   // inclusive level range
  bool small;
};

struct fs_model {
  int W = 0, H = 0;
  int32_t N = 0;
  int device = 0;
  int ncu = 0;                     // the device's CUs (read on first use)
  fs_params p{};
  double d_card = 0, d_diag = 0;
  ErodeConst ek{};
  int nmode = 1;
  hipStream_t stream = nullptr;
  hipStream_t side = nullptr;     // header read-back of the tile path
  hipEvent_t acc_done = nullptr;  // tile accumulation finished on `stream`

  // device state; h and h_alt swap after each tile-path step
  double *h = nullptr, *h_alt = nullptr;
  uint8_t *rcv = nullptr, *dmask = nullptr, *lvl = nullptr;
  uint16_t *depth = nullptr;
  unsigned long long *jump = nullptr;
  int32_t *list = nullptr, *stack = nullptr, *hdr = nullptr, *levels_d = nullptr;
  uint32_t *cnt = nullptr;
  // the level path's records in stack order (index = position in `stack`)
  uint8_t *rdir = nullptr;   // receiver direction
  int32_t *fdon = nullptr;   // position of the first donor
  uint8_t *ndon = nullptr;   // number of donors
  uint32_t *cnt_s = nullptr; // upstream cell count
  double *ftab = nullptr;   // erosion factor per (upstream count, diagonal)
  TilePlan *plan = nullptr; // per-tile erosion plan of the tile path
  tile32::TilePlan *plan32 = nullptr;  // the same for the 32-cell halo (on first use)
  uint64_t *status = nullptr;
  unsigned long long *ticket = nullptr;
  unsigned long long *phase = nullptr;   // tile kernels' phase ticks (stage timers)
  int32_t *fill_flag = nullptr;          // a fill launch changed a cell
  long long steps_in_call = 0;           // steps taken by the current fs_run / fs_step
  bool fill_mode = false;                // fill depressions before every step

  // host mirrors
  int32_t *hdr_h = nullptr;      // pinned: header + level counts
  int32_t *levels_h = nullptr;   // pinned: levels uploaded each step
  std::vector<int32_t> counts, levels;
  int32_t nlevel = 0, maxdepth = 0, prev_maxdepth = FS_MAX_LEVELS;
  bool have_step = false;
  bool order_current = false;      // stack and donor masks match rcv
  bool cnt_stack = false;          // the last step's counts are in cnt_s (level path)
  int path = FS_PATH_AUTO;
  uint64_t tile_steps = 0, level_steps = 0, tile32_steps = 0;
  // halo the next step tries first: 16 or 32; 0 = level path (the last
  // step's drainage was deeper than 32 levels)
  int tile_halo = kHalo;
  unsigned long long ticket_base = 0;
  uint32_t epoch = 0;

  // timers
  uint64_t timer_us[FS_NUM_TIMERS] = {};
  double stage_ms[FS_NUM_TIMERS] = {};
  hipEvent_t ev[2][7] = {};
  bool ev_pending[2] = {false, false};
  bool ev_tile[2] = {false, false};   // that step took the tile path
  int ev_cur = 0;
  double tile_acc_ms = 0, tile_ero_ms = 0;   // tile kernels' time not yet split

  // false for a model whose device buffers are lent to it (the entry point
  // that allocates nothing): no lazy allocations, no events, no side stream
  bool owns_memory = true;
  std::vector<Span> spans;  // level plan of the current sweep

  // the tile-graph path (allocated on first use)
  uint32_t *dinfo = nullptr;       // per cell: tail, hops to it, exit tail
  uint32_t *link = nullptr;        // per tile-border cell: node of its tail
  uint64_t *node_word = nullptr;   // per node: arrivals and count
  uint32_t *node_kids = nullptr, *node_parent = nullptr, *node_out = nullptr;
  int32_t *node_cell = nullptr, *nexit = nullptr;
  uint16_t *node_tier = nullptr;
  int32_t *tile_n0 = nullptr;      // per tile: cells with a root tail | largest hops << 16
  int32_t *thist = nullptr, *tcur = nullptr;   // per tier: work items, their cursor
  uint32_t *cells = nullptr;       // per tile: its cells sorted by tier (list entries)
  uint64_t *items = nullptr;       // work items in tier order
  int32_t *tstat = nullptr;        // error bits, deepest tier, items, queue head
  int graph_blocks = 0;            // persistent erosion grid
  int graph_queues = 1;            // its work-item queues
  deep::GraphGrid *graph_grid = nullptr;   // this model as the tile-graph kernels' one grid
  int32_t *graph_list = nullptr;           // its work list: first[2], which[1], size[2]
  uint64_t graph_steps = 0;
  int graph_run = 0;               // graph steps since the last tile attempt
  bool graph_in_call = false;      // the current fs_run / fs_step took a graph step

  bool ktiming = false;
  std::vector<KTimed> kpending;
  std::vector<hipEvent_t> kfree;
  double kms[FS_NUM_KERNELS] = {};
  uint64_t klaunches[FS_NUM_KERNELS] = {};
};

namespace {

uint32_t next_epoch(fs_model *m) {
  m->epoch = (m->epoch + 1) & 0x3fffffffu;
  if (m->epoch == 0) m->epoch = 1;
  return m->epoch;
}

int kt_begin(fs_model *m, int kid) {
  if (!m->ktiming) return FS_OK;
  KTimed t{kid, nullptr, nullptr};
  for (hipEvent_t *e : {&t.a, &t.b}) {
    if (!m->kfree.empty()) {
      *e = m->kfree.back();
      m->kfree.pop_back();
    } else {
      FS_HIP_TRY(hipEventCreate(e));
    }
  }
  FS_HIP_TRY(hipEventRecord(t.a, m->stream));
  m->kpending.push_back(t);
  return FS_OK;
}

int kt_end(fs_model *m) {
  if (!m->ktiming) return FS_OK;
  FS_HIP_TRY(hipEventRecord(m->kpending.back().b, m->stream));
  return FS_OK;
}

// Fold the timed launches that have finished (in launch order; the tile
// path's erosion may still be running when this is called).
int kt_collect(fs_model *m) {
  size_t done = 0;
  for (; done < m->kpending.size(); done++) {
    const KTimed &t = m->kpending[done];
    const hipError_t q = hipEventQuery(t.b);
    if (q == hipErrorNotReady) break;
    FS_HIP_TRY(q);
    float ms = 0.f;
    FS_HIP_TRY(hipEventElapsedTime(&ms, t.a, t.b));
    m->kms[t.kid] += ms;
    m->klaunches[t.kid]++;
    m->kfree.push_back(t.a);
    m->kfree.push_back(t.b);
  }
  m->kpending.erase(m->kpending.begin(), m->kpending.begin() + static_cast<std::ptrdiff_t>(done));
  return FS_OK;
}

int stage_collect(fs_model *m, int set) {
  if (!m->ev_pending[set]) return FS_OK;
  float ms[6] = {};
  for (int i = 0; i < 6; i++)
    FS_HIP_TRY(hipEventElapsedTime(&ms[i], m->ev[set][i], m->ev[set][i + 1]));
  if (m->ev_tile[set]) {
    // donors..erosion ran fused in the two tile kernels; split_tile_time
    // assigns their time to the stages
    m->stage_ms[FS_T_RECEIVERS] += ms[0];
    m->tile_acc_ms += ms[1] + ms[2] + ms[3] + ms[4];
    m->tile_ero_ms += ms[5];
  } else {
    for (int i = 0; i < 6; i++) m->stage_ms[FS_T_RECEIVERS + i] += ms[i];
  }
  m->ev_pending[set] = false;
  return FS_OK;
}

// Split the tile kernels' event time over the reference's stages in
// proportion to the phase ticks their sampled workgroups recorded on the
// timed steps:
// k_tile_acc's region load -> DetermineDonors, its chain walks and
// count/level stores -> FlowAcc, its halo-chain marking and erosion plan ->
// GenerateOrder; k_tile_erode's level bucketing -> GenerateOrder, its outlet
// uplift -> Uplift, its wavefront and stores -> Erosion.
int split_tile_time(fs_model *m) {
  if (m->tile_acc_ms <= 0 && m->tile_ero_ms <= 0) return FS_OK;
  double ph[kPhaseStride] = {};
  if (m->phase) {
    std::vector<unsigned long long> raw(static_cast<size_t>(kPhaseSlots) * kPhaseStride);
    FS_HIP_TRY(hipMemcpy(raw.data(), m->phase, raw.size() * sizeof(raw[0]),
                         hipMemcpyDeviceToHost));
    for (int r = 0; r < kPhaseSlots; r++)
      for (int i = 0; i < kPhaseStride; i++)
        ph[i] += static_cast<double>(raw[static_cast<size_t>(r) * kPhaseStride + i]);
    FS_HIP_TRY(hipMemset(m->phase, 0, raw.size() * sizeof(raw[0])));
  }
  const double acc = ph[kPhAccLoad] + ph[kPhAccWalk] + ph[kPhAccPlan] + ph[kPhAccStore];
  if (acc > 0) {
    m->stage_ms[FS_T_DONORS] += m->tile_acc_ms * ph[kPhAccLoad] / acc;
    m->stage_ms[FS_T_FLOWACC] += m->tile_acc_ms * (ph[kPhAccWalk] + ph[kPhAccStore]) / acc;
    m->stage_ms[FS_T_ORDER] += m->tile_acc_ms * ph[kPhAccPlan] / acc;
  } else {
    m->stage_ms[FS_T_FLOWACC] += m->tile_acc_ms;
  }
  const double ero = ph[kPhEroBucket] + ph[kPhEroRoots] + ph[kPhEroLoop] + ph[kPhEroStore];
  if (ero > 0) {
    m->stage_ms[FS_T_ORDER] += m->tile_ero_ms * ph[kPhEroBucket] / ero;
    m->stage_ms[FS_T_UPLIFT] += m->tile_ero_ms * ph[kPhEroRoots] / ero;
    m->stage_ms[FS_T_EROSION] += m->tile_ero_ms * (ph[kPhEroLoop] + ph[kPhEroStore]) / ero;
  } else {
    m->stage_ms[FS_T_EROSION] += m->tile_ero_ms;
  }
  m->tile_acc_ms = m->tile_ero_ms = 0;
  return FS_OK;
}

int check_launch() {
  FS_HIP_TRY(hipGetLastError());
  return FS_OK;
}

// Record a stage event; a model without events (lent buffers) skips it.
int mark(hipEvent_t e, hipStream_t s) {
  if (e) FS_HIP_TRY(hipEventRecord(e, s));
  return FS_OK;
}

int device_error_status(int32_t err) {
  if (err & kErrNewton) return FS_ERR_NEWTON;
  if (err & kErrDepth) return FS_ERR_LEVEL_OVERFLOW;
  if (err) {
    fs::set_last_error("device consistency flag " + std::to_string(err));
    return FS_ERR_INTERNAL;
  }
  return FS_OK;
}

void free_model(fs_model *m) {
  if (!m) return;
  (void)hipSetDevice(m->device);
  if (m->stream) (void)hipStreamSynchronize(m->stream);
  for (void *ptr : {static_cast<void *>(m->h), static_cast<void *>(m->h_alt),
                    static_cast<void *>(m->rcv),
                    static_cast<void *>(m->dmask), static_cast<void *>(m->lvl),
                    static_cast<void *>(m->depth),
                    static_cast<void *>(m->jump), static_cast<void *>(m->list),
                    static_cast<void *>(m->stack), static_cast<void *>(m->hdr),
                    static_cast<void *>(m->rdir), static_cast<void *>(m->fdon),
                    static_cast<void *>(m->ndon), static_cast<void *>(m->cnt_s),
                    static_cast<void *>(m->levels_d), static_cast<void *>(m->cnt),
                    static_cast<void *>(m->ftab), static_cast<void *>(m->status),
                    static_cast<void *>(m->ticket), static_cast<void *>(m->plan),
                    static_cast<void *>(m->plan32), static_cast<void *>(m->phase),
                    static_cast<void *>(m->fill_flag), static_cast<void *>(m->dinfo),
                    static_cast<void *>(m->link), static_cast<void *>(m->node_word),
                    static_cast<void *>(m->node_kids), static_cast<void *>(m->node_parent),
                    static_cast<void *>(m->node_out), static_cast<void *>(m->node_cell),
                    static_cast<void *>(m->nexit), static_cast<void *>(m->node_tier),
                    static_cast<void *>(m->tstat), static_cast<void *>(m->tile_n0),
                    static_cast<void *>(m->thist), static_cast<void *>(m->tcur),
                    static_cast<void *>(m->cells), static_cast<void *>(m->items),
                    static_cast<void *>(m->graph_grid), static_cast<void *>(m->graph_list)})
    if (ptr) (void)hipFree(ptr);
  if (m->hdr_h) (void)hipHostFree(m->hdr_h);
  if (m->levels_h) (void)hipHostFree(m->levels_h);
  for (auto &set : m->ev)
    for (auto &e : set)
      if (e) (void)hipEventDestroy(e);
  for (auto &t : m->kpending) {
    (void)hipEventDestroy(t.a);
    (void)hipEventDestroy(t.b);
  }
  for (auto e : m->kfree) (void)hipEventDestroy(e);
  if (m->acc_done) (void)hipEventDestroy(m->acc_done);
  if (m->side) (void)hipStreamDestroy(m->side);
  if (m->stream) (void)hipStreamDestroy(m->stream);
  delete m;
}

int create_model(int W, int H, const fs_params *p, const double *h_init, uint32_t seed,
                 int device, fs_model **out) {
  if (!out) return FS_ERR_INVALID_ARGUMENT;
  *out = nullptr;
  FS_TRY(fs::validate_params(W, H, p));
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0 || device < 0 || device >= ndev) {
    fs::set_last_error("no HIP device " + std::to_string(device));
    return FS_ERR_NO_DEVICE;
  }
  const auto t0 = std::chrono::steady_clock::now();
  fs_model *m = new (std::nothrow) fs_model();
  if (!m) return FS_ERR_OUT_OF_MEMORY;
  struct Guard {
    fs_model *m;
    ~Guard() { if (m) free_model(m); }
  } guard{m};
  m->W = W;
  m->H = H;
  m->N = W * H;
  m->device = device;
  m->p = *p;
  m->d_card = p->dx;
  m->d_diag = p->dx * fs::kSqrt2;
  m->ek = make_erode_const(*p, &m->nmode);
  FS_HIP_TRY(hipSetDevice(device));
  FS_HIP_TRY(hipStreamCreateWithFlags(&m->stream, hipStreamNonBlocking));
  FS_HIP_TRY(hipStreamCreateWithFlags(&m->side, hipStreamNonBlocking));
  FS_HIP_TRY(hipEventCreateWithFlags(&m->acc_done, hipEventDisableTiming));
  for (auto &set : m->ev)
    for (auto &e : set) FS_HIP_TRY(hipEventCreate(&e));

  const size_t N = static_cast<size_t>(m->N);
  const size_t ntiles = (N + kBfsTile - 1) / kBfsTile + 1;
  FS_TRY(fs::dev_alloc(&m->h, N));
  FS_TRY(fs::dev_alloc(&m->h_alt, N));
  FS_TRY(fs::dev_alloc(&m->rcv, N + 16));
  FS_TRY(fs::dev_alloc(&m->dmask, N));
  FS_TRY(fs::dev_alloc(&m->lvl, N));
  FS_TRY(fs::dev_alloc(&m->depth, N));
  FS_TRY(fs::dev_alloc(&m->jump, N));
  FS_TRY(fs::dev_alloc(&m->list, N));
  FS_TRY(fs::dev_alloc(&m->stack, N));
  FS_TRY(fs::dev_alloc(&m->cnt, N));
  FS_TRY(fs::dev_alloc(&m->rdir, N));
  FS_TRY(fs::dev_alloc(&m->fdon, N));
  FS_TRY(fs::dev_alloc(&m->ndon, N));
  FS_TRY(fs::dev_alloc(&m->cnt_s, N));
  // at least kFactCache counts, so the tile kernels' LDS cache stays in bounds
  const size_t kmax = std::max<size_t>(N, kFactCache - 1);
  FS_TRY(fs::dev_alloc(&m->ftab, 2 * (kmax + 1)));
  FS_TRY(fs::dev_alloc(&m->hdr, kHdrAlloc));
  FS_TRY(fs::dev_alloc(&m->levels_d, FS_MAX_LEVELS + 4));
  FS_TRY(fs::dev_alloc(&m->status, ntiles));
  FS_TRY(fs::dev_alloc(&m->ticket, 1));
  const size_t plan_tiles = static_cast<size_t>((m->W + kTile - 1) / kTile) *
                            static_cast<size_t>((m->H + kTile - 1) / kTile);
  FS_TRY(fs::dev_alloc(&m->plan, plan_tiles));
  FS_TRY(fs::dev_alloc(&m->phase, static_cast<size_t>(kPhaseSlots) * kPhaseStride));
  FS_TRY(fs::dev_alloc(&m->fill_flag, 1));
  if (hipHostMalloc(reinterpret_cast<void **>(&m->hdr_h),
                    (kHdrInts + FS_MAX_LEVELS + 3) * sizeof(int32_t)) != hipSuccess ||
      hipHostMalloc(reinterpret_cast<void **>(&m->levels_h),
                    (FS_MAX_LEVELS + 4) * sizeof(int32_t)) != hipSuccess) {
    fs::set_last_error("hipHostMalloc failed");
    return FS_ERR_OUT_OF_MEMORY;
  }
  FS_HIP_TRY(hipMemsetAsync(m->rcv, 0xFF, N + 16, m->stream));   // RB:361-362
  FS_HIP_TRY(hipMemsetAsync(m->hdr, 0, kHdrAlloc * sizeof(int32_t), m->stream));
  FS_HIP_TRY(hipMemsetAsync(m->status, 0, ntiles * sizeof(uint64_t), m->stream));
  FS_HIP_TRY(hipMemsetAsync(m->ticket, 0, sizeof(unsigned long long), m->stream));
  FS_HIP_TRY(hipMemsetAsync(m->plan, 0, plan_tiles * sizeof(TilePlan), m->stream));
  FS_HIP_TRY(hipMemsetAsync(m->phase, 0,
                            static_cast<size_t>(kPhaseSlots) * kPhaseStride * sizeof(uint64_t),
                            m->stream));
  FS_HIP_TRY(hipMemsetAsync(m->cnt, 0, N * sizeof(uint32_t), m->stream));
  {
    std::vector<double> tab;
    try {
      tab.resize(2 * (kmax + 1));
    } catch (const std::bad_alloc &) {
      return FS_ERR_OUT_OF_MEMORY;
    }
    fs::fill_fact_table(tab.data(), kmax, p->cell_area, p->meq, m->ek.kdt, m->ek.lenpow0,
                        m->ek.lenpow1);
    FS_HIP_TRY(hipMemcpy(m->ftab, tab.data(), tab.size() * sizeof(double), hipMemcpyHostToDevice));
  }
  if (h_init) {
    FS_HIP_TRY(hipMemcpy(m->h, h_init, N * sizeof(double), hipMemcpyHostToDevice));
  } else {
    std::vector<double> h0;
    try {
      h0.resize(N);
    } catch (const std::bad_alloc &) {
      return FS_ERR_OUT_OF_MEMORY;
    }
    fs::generate_terrain(h0.data(), W, H, seed);
    FS_HIP_TRY(hipMemcpy(m->h, h0.data(), N * sizeof(double), hipMemcpyHostToDevice));
  }
  FS_HIP_TRY(hipStreamSynchronize(m->stream));
  const auto us = std::chrono::duration_cast<std::chrono::microseconds>(
                      std::chrono::steady_clock::now() - t0).count();
  m->timer_us[FS_T_INITIALIZE] += static_cast<uint64_t>(us);
  m->timer_us[FS_T_OVERALL] += static_cast<uint64_t>(us);
  guard.m = nullptr;
  *out = m;
  return FS_OK;
}

// Read the header (max level, unfinished count, error bits) and the level
// histogram; synchronises the stream.
int fetch_header(fs_model *m, int ncounts) {
  ncounts = std::min(ncounts, FS_MAX_LEVELS + 3);
  FS_HIP_TRY(hipMemcpyAsync(m->hdr_h, m->hdr, (kHdrInts + ncounts) * sizeof(int32_t),
                            hipMemcpyDeviceToHost, m->stream));
  FS_HIP_TRY(hipStreamSynchronize(m->stream));
  return FS_OK;
}

// Plan a level sweep into out: maximal runs of levels whose size is at most
// kSmallLevel go to one single-workgroup launch; every larger level gets its
// own grid-wide launch.
void plan_levels(const std::vector<int32_t> &counts, int lo, int hi, std::vector<Span> &out) {
  out.clear();
  int L = lo;
  while (L <= hi) {
    if (counts[L] > kSmallLevel) {
      out.push_back({L, L, false});
      L++;
    } else {
      int e = L;
      while (e + 1 <= hi && counts[e + 1] <= kSmallLevel) e++;
      out.push_back({L, e, true});
      L = e + 1;
    }
  }
}

template <int NMODE>
int launch_erosion(fs_model *m, const std::vector<Span> &spans) {
  int32_t *hdr = m->hdr;
  for (const Span &s : spans) {
    if (s.small) {
      FS_TRY(kt_begin(m, FS_K_EROSION));
      hipLaunchKernelGGL(k_erode_small<NMODE>, dim3(1), dim3(kSmallBlock), 0, m->stream, m->stack,
                         m->rdir, m->cnt_s, m->ftab, m->h, m->levels_d, s.lo, s.hi, m->W, m->N,
                         m->ek, hdr);
      FS_TRY(check_launch());
      FS_TRY(kt_end(m));
    } else {
      const int32_t ls = m->levels[s.lo], le = m->levels[s.lo + 1];
      const unsigned blocks = static_cast<unsigned>((le - ls + 255) / 256);
      FS_TRY(kt_begin(m, FS_K_EROSION));
      hipLaunchKernelGGL(k_erode<NMODE>, dim3(blocks), dim3(256), 0, m->stream, m->stack,
                         m->rdir, m->cnt_s, m->ftab, m->h, ls, le, m->W, m->N, m->ek, hdr);
      FS_TRY(check_launch());
      FS_TRY(kt_end(m));
    }
  }
  return FS_OK;
}

int launch_receivers(fs_model *m) {
  const unsigned gx = static_cast<unsigned>((m->W - 4 + 255) / 256);
  const int strips = (m->H - 4 + kRecvRows - 1) / kRecvRows;
  const unsigned gy = static_cast<unsigned>(std::min((strips + kRecvStrips - 1) / kRecvStrips, 65535));
  FS_TRY(kt_begin(m, FS_K_RECEIVERS));
  hipLaunchKernelGGL(k_receivers, dim3(gx, gy), dim3(256), 0, m->stream, m->h, m->rcv, m->W, m->H,
                     m->d_card, m->d_diag, m->hdr);
  FS_TRY(check_launch());
  return kt_end(m);
}

int launch_fill_counts(fs_model *m) {
  const int nclear = m->prev_maxdepth + 1;
  hipLaunchKernelGGL(k_fill_counts, dim3((nclear + 255) / 256), dim3(256), 0, m->stream, m->hdr,
                     nclear);
  return check_launch();
}

// levels[] exactly as GenerateOrder leaves it (RB:218-251): offsets of
// levels 0..D, the stack size, and the trailing duplicate; nlevel = D+3.
// Reads the histogram from the header already fetched into hdr_h.
int take_levels(fs_model *m, int D) {
  if (D < 0 || D > FS_MAX_LEVELS) return FS_ERR_LEVEL_OVERFLOW;
  if (D + 1 > kHdrCounts) FS_TRY(fetch_header(m, D + 1));
  m->prev_maxdepth = D;
  m->counts.assign(m->hdr_h + kHdrInts, m->hdr_h + kHdrInts + D + 1);
  long long total = 0;
  for (int32_t c : m->counts) total += c;
  if (total != m->N) {
    fs::set_last_error("level histogram covers " + std::to_string(total) + " of " +
                       std::to_string(m->N) + " cells");
    return m->hdr_h[1] > 0 ? FS_ERR_LEVEL_OVERFLOW : FS_ERR_INTERNAL;
  }
  m->levels.resize(D + 3);
  m->levels[0] = 0;
  for (int L = 0; L <= D; L++) m->levels[L + 1] = m->levels[L] + m->counts[L];
  m->levels[D + 2] = m->N;
  m->nlevel = D + 3;
  std::memcpy(m->levels_h, m->levels.data(), m->levels.size() * sizeof(int32_t));
  FS_HIP_TRY(hipMemcpyAsync(m->levels_d, m->levels_h, m->levels.size() * sizeof(int32_t),
                            hipMemcpyHostToDevice, m->stream));
  return FS_OK;
}

// ComputeDonors + GenerateOrder (RB:188-254) from the receivers in rcv:
// donor masks, levels (with pointer jumping for deep cells), the level
// histogram, and the exact stack.  Synchronises once (twice with deep cells).
int build_order(fs_model *m, hipEvent_t ev_topology_done) {
  hipStream_t s = m->stream;
  const int W = m->W, H = m->H;
  const int32_t N = m->N;
  FS_TRY(launch_fill_counts(m));
  FS_TRY(kt_begin(m, FS_K_TOPOLOGY));
  hipLaunchKernelGGL(k_topology, dim3(std::min(H, kTopoBlocks)), dim3(256), 0, s, m->rcv,
                     m->dmask, m->depth, m->jump, m->list, m->hdr, W, H);
  FS_TRY(check_launch());
  FS_TRY(kt_end(m));
  if (ev_topology_done) FS_HIP_TRY(hipEventRecord(ev_topology_done, s));
  FS_TRY(fetch_header(m, kHdrCounts));
  FS_TRY(stage_collect(m, m->ev_cur ^ 1));
  FS_TRY(kt_collect(m));
  FS_TRY(device_error_status(m->hdr_h[2]));
  if (m->hdr_h[1] > 0) {
    for (int pass = 0; pass < kJumpPasses; pass++) {
      FS_TRY(kt_begin(m, FS_K_JUMP));
      hipLaunchKernelGGL(k_jump, dim3(kJumpBlocks), dim3(256), 0, s, m->list, m->depth, m->jump,
                         m->hdr, pass);
      FS_TRY(check_launch());
      FS_TRY(kt_end(m));
    }
    FS_TRY(fetch_header(m, kHdrCounts));
    FS_TRY(device_error_status(m->hdr_h[2]));
  }
  const int D = m->hdr_h[0];
  FS_TRY(take_levels(m, D));
  {
    const unsigned tiles = static_cast<unsigned>((N + kRootTile - 1) / kRootTile);
    FS_TRY(kt_begin(m, FS_K_ROOTS));
    hipLaunchKernelGGL(k_roots, dim3(tiles), dim3(kScanBlock), 0, s, m->rcv, m->stack,
                       m->status, m->ticket, m->ticket_base, next_epoch(m), N, m->hdr);
    FS_TRY(check_launch());
    FS_TRY(kt_end(m));
    m->ticket_base += tiles;
  }
  if (D >= 1) {
    plan_levels(m->counts, 0, D - 1, m->spans);
    for (const Span &sp : m->spans) {
      if (sp.small) {
        FS_TRY(kt_begin(m, FS_K_BFS));
        hipLaunchKernelGGL(k_bfs_small, dim3(1), dim3(kSmallBlock), 0, s, m->stack, m->dmask,
                           m->rdir, m->fdon, m->ndon, m->levels_d, sp.lo, sp.hi + 1, W, N,
                           m->hdr);
        FS_TRY(check_launch());
        FS_TRY(kt_end(m));
      } else {
        const int32_t ls = m->levels[sp.lo], le = m->levels[sp.lo + 1];
        const unsigned tiles = static_cast<unsigned>((le - ls + kBfsTile - 1) / kBfsTile);
        FS_TRY(kt_begin(m, FS_K_BFS));
        hipLaunchKernelGGL(k_bfs, dim3(tiles), dim3(kScanBlock), 0, s, m->stack, m->dmask,
                           m->rdir, m->fdon, m->ndon, ls, le, W, N, m->status, m->ticket,
                           m->ticket_base, next_epoch(m), m->hdr);
        FS_TRY(check_launch());
        FS_TRY(kt_end(m));
        m->ticket_base += tiles;
      }
    }
  }
  m->order_current = true;
  return FS_OK;
}

// The level path: the reference's stage sequence with each level a launch
// (or a run of small levels one single-workgroup launch).  Receivers are
// already in rcv.
int level_path(fs_model *m, hipEvent_t *ev) {
  hipStream_t s = m->stream;
  const int W = m->W, H = m->H;
  const int32_t N = m->N;
  FS_TRY(build_order(m, ev[2]));
  const int D = static_cast<int>(m->counts.size()) - 1;
  FS_TRY(mark(ev[3], s));
  // ---- FlowAcc: levels D down to 0 ----
  {
    plan_levels(m->counts, 0, D, m->spans);
    for (auto it = m->spans.rbegin(); it != m->spans.rend(); ++it) {
      if (it->small) {
        FS_TRY(kt_begin(m, FS_K_FLOWACC));
        hipLaunchKernelGGL(k_flowacc_small, dim3(1), dim3(kSmallBlock), 0, s, m->fdon, m->ndon,
                           m->cnt_s, m->levels_d, it->hi, it->lo, D, N, m->hdr);
        FS_TRY(check_launch());
        FS_TRY(kt_end(m));
      } else {
        const int32_t ls = m->levels[it->lo], le = m->levels[it->lo + 1];
        FS_TRY(kt_begin(m, FS_K_FLOWACC));
        hipLaunchKernelGGL(k_flowacc, dim3(static_cast<unsigned>((le - ls + 255) / 256)), dim3(256),
                           0, s, m->fdon, m->ndon, m->cnt_s, ls, le, it->lo == D, N, m->hdr);
        FS_TRY(check_launch());
        FS_TRY(kt_end(m));
      }
    }
  }
  FS_TRY(mark(ev[4], s));
  // ---- Uplift of the interior outlets ----
  {
    const int32_t n0 = m->counts[0];
    FS_TRY(kt_begin(m, FS_K_UPLIFT));
    hipLaunchKernelGGL(k_uplift_roots, dim3(static_cast<unsigned>((n0 + 255) / 256)), dim3(256), 0,
                       s, m->stack, n0, m->h, W, H, m->ek.ud, N, m->hdr);
    FS_TRY(check_launch());
    FS_TRY(kt_end(m));
  }
  FS_TRY(mark(ev[5], s));
  // ---- Erode: levels 1..D ----
  if (D >= 1) {
    plan_levels(m->counts, 1, D, m->spans);
    if (m->nmode == 1) FS_TRY(launch_erosion<1>(m, m->spans));
    else if (m->nmode == 2) FS_TRY(launch_erosion<2>(m, m->spans));
    else FS_TRY(launch_erosion<0>(m, m->spans));
  }
  m->level_steps++;
  m->cnt_stack = true;
  return FS_OK;
}

template <int NMODE, bool kTimed>
int launch_tile_erode(fs_model *m, int tiles, int tiles_x, int halo) {
  if (halo == kHalo)
    hipLaunchKernelGGL((k_tile_erode<NMODE, kTimed>), dim3(tiles), dim3(kEroBlock), 0, m->stream,
                       m->h, m->h_alt, m->plan, m->lvl, m->cnt, m->ftab, m->W, m->H, tiles_x,
                       tiles, m->ek, m->hdr, m->phase);
  else
    hipLaunchKernelGGL((tile32::k_tile_erode<NMODE, kTimed>), dim3(tiles), dim3(tile32::kEroBlock),
                       0, m->stream, m->h, m->h_alt, m->plan32, m->lvl, m->cnt, m->ftab, m->W,
                       m->H, tiles_x, tiles, m->ek, m->hdr, m->phase);
  return check_launch();
}

template <int NMODE>
int launch_tile_erode(fs_model *m, int tiles, int tiles_x, int halo, bool timed) {
  return timed ? launch_tile_erode<NMODE, true>(m, tiles, tiles_x, halo)
               : launch_tile_erode<NMODE, false>(m, tiles, tiles_x, halo);
}

// Accumulation with the given halo (16 or 32), then the erosion with the
// same halo launched behind it at once: the erosion grid does nothing when a
// tile was too deep, so the host reads the accumulation's header on a side
// stream while the erosion runs.  *fits = false when some level exceeds the
// halo (or a tile's chains need more halo cells than its plan holds); the
// erosion was then empty and hdr_h[0] is the deepest level among the cells
// that reach an outlet within the region.  clear: reset the header first
// (k_receivers already did for the first accumulation of a step).
int tile_attempt(fs_model *m, int halo, int tiles, int tiles_x, bool clear, hipEvent_t *ev,
                 bool *fits) {
  hipStream_t s = m->stream;
  *fits = false;
  if (clear) {
    hipLaunchKernelGGL(k_fill_counts, dim3(1), dim3(64), 0, s, m->hdr, 0);
    FS_TRY(check_launch());
  }
  // one step in kPhaseEvery (and the first of each call) runs the timed
  // instantiations that sample the phase ticks of the stage timers
  const bool timed = m->phase && (m->steps_in_call == 0 || m->tile_steps % kPhaseEvery == 0);
  if (!m->ncu)
    FS_HIP_TRY(hipDeviceGetAttribute(&m->ncu, hipDeviceAttributeMultiprocessorCount, m->device));
  FS_TRY(kt_begin(m, FS_K_TILE_ACC));
  if (halo == kHalo) {
    // fewer tiles than four per CU: one 1024-thread tile per CU finishes
    // sooner than a 512-thread one
    const bool wide = tiles < 4 * m->ncu;
    if (wide && timed)
      hipLaunchKernelGGL((k_tile_acc<true, kAccBlockWide>), dim3(tiles), dim3(kAccBlockWide), 0, s,
                         m->rcv, m->cnt, m->lvl, m->plan, m->hdr, m->W, m->H, tiles_x, tiles,
                         m->phase);
    else if (wide)
      hipLaunchKernelGGL((k_tile_acc<false, kAccBlockWide>), dim3(tiles), dim3(kAccBlockWide), 0,
                         s, m->rcv, m->cnt, m->lvl, m->plan, m->hdr, m->W, m->H, tiles_x, tiles,
                         m->phase);
    else if (timed)
      hipLaunchKernelGGL(k_tile_acc<true>, dim3(tiles), dim3(kAccBlock), 0, s, m->rcv, m->cnt,
                         m->lvl, m->plan, m->hdr, m->W, m->H, tiles_x, tiles, m->phase);
    else
      hipLaunchKernelGGL(k_tile_acc<false>, dim3(tiles), dim3(kAccBlock), 0, s, m->rcv, m->cnt,
                         m->lvl, m->plan, m->hdr, m->W, m->H, tiles_x, tiles, m->phase);
  } else {
    if (!m->plan32) {
      if (!m->owns_memory) return FS_ERR_INTERNAL;   // lent buffers include plan32
      FS_TRY(fs::dev_alloc(&m->plan32, static_cast<size_t>(tiles)));
    }
    if (timed)
      hipLaunchKernelGGL(tile32::k_tile_acc<true>, dim3(tiles), dim3(tile32::kAccBlock), 0, s,
                         m->rcv, m->cnt, m->lvl, m->plan32, m->hdr, m->W, m->H, tiles_x, tiles,
                         m->phase);
    else
      hipLaunchKernelGGL(tile32::k_tile_acc<false>, dim3(tiles), dim3(tile32::kAccBlock), 0, s,
                         m->rcv, m->cnt, m->lvl, m->plan32, m->hdr, m->W, m->H, tiles_x, tiles,
                         m->phase);
  }
  FS_TRY(check_launch());
  FS_TRY(kt_end(m));
  FS_TRY(mark(ev[2], s));
  if (m->side) {
    FS_HIP_TRY(hipEventRecord(m->acc_done, s));
    FS_HIP_TRY(hipStreamWaitEvent(m->side, m->acc_done, 0));
    FS_HIP_TRY(hipMemcpyAsync(m->hdr_h, m->hdr, kHdrInts * sizeof(int32_t),
                              hipMemcpyDeviceToHost, m->side));
  }
  FS_TRY(mark(ev[3], s));
  FS_TRY(mark(ev[4], s));
  FS_TRY(mark(ev[5], s));
  FS_TRY(kt_begin(m, FS_K_TILE_ERODE));
  if (m->nmode == 1) FS_TRY(launch_tile_erode<1>(m, tiles, tiles_x, halo, timed));
  else if (m->nmode == 2) FS_TRY(launch_tile_erode<2>(m, tiles, tiles_x, halo, timed));
  else FS_TRY(launch_tile_erode<0>(m, tiles, tiles_x, halo, timed));
  FS_TRY(kt_end(m));
  if (m->side) {
    FS_HIP_TRY(hipStreamSynchronize(m->side));
  } else {
    // no side stream: read the header behind the erosion on the model's stream
    FS_HIP_TRY(hipMemcpyAsync(m->hdr_h, m->hdr, kHdrInts * sizeof(int32_t),
                              hipMemcpyDeviceToHost, s));
    FS_HIP_TRY(hipStreamSynchronize(s));
  }
  FS_TRY(device_error_status(m->hdr_h[2]));
  *fits = m->hdr_h[3] == 0;
  return FS_OK;
}

// The tile path (every level <= 32): accumulation with a 16-cell halo, or a
// 32-cell one when the drainage is deeper, and erosion with the same halo.
// Returns FS_OK with *done = false when the levels are too deep for either
// (both erosions were empty); the caller then runs the level path on the
// same receivers.  The halo that fitted is tried first at the next step.
int tile_path(fs_model *m, hipEvent_t *ev, bool *done) {
  *done = false;
  const int tiles_x = (m->W + kTile - 1) / kTile, tiles_y = (m->H + kTile - 1) / kTile;
  const int tiles = tiles_x * tiles_y;
  int halo = m->tile_halo == tile32::kHalo ? tile32::kHalo : kHalo;
  bool fits = false;
  FS_TRY(tile_attempt(m, halo, tiles, tiles_x, false, ev, &fits));
  FS_TRY(stage_collect(m, m->ev_cur ^ 1));
  FS_TRY(kt_collect(m));
  if (!fits && halo == kHalo) {
    halo = tile32::kHalo;
    FS_TRY(tile_attempt(m, halo, tiles, tiles_x, true, ev, &fits));
  }
  if (!fits) {
    m->tile_halo = kHalo;                   // the level path sets it from the depth it finds
    return FS_OK;
  }
  // a 32-halo step whose levels all fit 16 tries the 16-cell halo next
  m->tile_halo = (halo == tile32::kHalo && m->hdr_h[0] <= kHalo) ? kHalo : halo;
  m->order_current = false;                 // levels, stack, donor masks on request
  m->cnt_stack = false;
  std::swap(m->h, m->h_alt);
  m->tile_steps++;
  if (halo != kHalo) m->tile32_steps++;
  *done = true;
  return FS_OK;
}

// Buffers of the tile-graph path, on its first step.
int ensure_graph(fs_model *m) {
  if (m->dinfo) return FS_OK;
  if (!m->owns_memory) return FS_ERR_INTERNAL;
  const size_t N = static_cast<size_t>(m->N);
  const size_t tiles = static_cast<size_t>((m->W + deep::kSide - 1) / deep::kSide) *
                       static_cast<size_t>((m->H + deep::kSide - 1) / deep::kSide);
  const size_t nodes = tiles * deep::kNodeCap;
  int ncu = 0;
  FS_HIP_TRY(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, m->device));
  m->graph_blocks = m->nmode == 1   ? deep::erode_grid<1>(ncu)
                    : m->nmode == 2 ? deep::erode_grid<2>(ncu)
                                    : deep::erode_grid<0>(ncu);
  m->graph_queues = deep::erode_queues(m->graph_blocks);
  FS_TRY(fs::dev_alloc(&m->dinfo, N));
  FS_TRY(fs::dev_alloc(&m->link, N));
  FS_TRY(fs::dev_alloc(&m->node_word, nodes));
  FS_TRY(fs::dev_alloc(&m->node_kids, nodes));
  FS_TRY(fs::dev_alloc(&m->node_parent, nodes));
  FS_TRY(fs::dev_alloc(&m->node_out, nodes));
  FS_TRY(fs::dev_alloc(&m->node_cell, nodes));
  FS_TRY(fs::dev_alloc(&m->node_tier, nodes));
  FS_TRY(fs::dev_alloc(&m->nexit, tiles));
  FS_TRY(fs::dev_alloc(&m->tile_n0, tiles));
  const size_t tier_words = static_cast<size_t>(deep::kTierCap + 1) * deep::kShards;
  FS_TRY(fs::dev_alloc(&m->thist, tier_words));
  FS_TRY(fs::dev_alloc(&m->tcur, tier_words));
  FS_TRY(fs::dev_alloc(&m->cells, tiles * deep::kCells));
  FS_TRY(fs::dev_alloc(&m->items, tiles * deep::kTiersPerTile));
  FS_TRY(fs::dev_alloc(&m->tstat, deep::kStWords));
  FS_HIP_TRY(hipMemsetAsync(m->thist, 0, tier_words * sizeof(int32_t), m->stream));
  deep::GraphGrid g{};
  g.off = 0;
  g.ftab_off = 0;
  g.hdr = m->hdr;
  g.levels = nullptr;
  g.W = m->W;
  g.H = m->H;
  g.tiles_x = (m->W + deep::kSide - 1) / deep::kSide;
  g.tile0 = 0;
  g.nlev = 0;
  g.nmode = m->nmode;
  g.ek = m->ek;
  const int32_t list[5] = {0, static_cast<int32_t>(tiles), 0, 1, static_cast<int32_t>(tiles)};
  FS_TRY(fs::dev_alloc(&m->graph_grid, 1));
  FS_TRY(fs::dev_alloc(&m->graph_list, 5));
  FS_HIP_TRY(hipMemcpy(m->graph_grid, &g, sizeof(g), hipMemcpyHostToDevice));
  FS_HIP_TRY(hipMemcpy(m->graph_list, list, sizeof(list), hipMemcpyHostToDevice));
  return FS_OK;
}

deep::GraphWork graph_work(fs_model *m) {
  deep::GraphWork w;
  w.grid = m->graph_grid;
  w.first = m->graph_list;        // {0, tiles}
  w.which = m->graph_list + 2;    // {0}
  w.size = m->graph_list + 3;     // {1 entry, tiles}
  return w;
}

// The counts' completion with the erosion of the root-tailed cells, then the
// persistent erosion of the other tiers.
template <int NMODE>
int launch_graph_erosion(fs_model *m, int tiles) {
  hipStream_t s = m->stream;
  const deep::GraphWork w = graph_work(m);
  FS_TRY(kt_begin(m, FS_K_GRAPH_COUNTS));
  hipLaunchKernelGGL(deep::k_deep_counts<NMODE>, dim3(tiles), dim3(deep::kCntBlock), 0, s, w, m->rcv,
                     m->cnt, m->dinfo, m->link, m->node_out, m->node_cell, m->node_tier,
                     m->nexit, m->tile_n0, m->cells, m->tcur, m->items, m->ftab, m->h, m->h_alt);
  FS_TRY(check_launch());
  FS_TRY(kt_end(m));
  FS_TRY(kt_begin(m, FS_K_GRAPH_ERODE));
  hipLaunchKernelGGL(deep::k_deep_erode<NMODE>, dim3(static_cast<unsigned>(m->graph_blocks)),
                     dim3(deep::kEroBlock), 0, s, w, m->rcv, m->cnt, m->cells, m->items, m->ftab,
                     m->h, m->h_alt, m->tstat, m->graph_queues);
  FS_TRY(check_launch());
  return kt_end(m);
}

// The tile-graph path (fs_deep.h): any depth, no halo, no host round trip.
// Receivers are already in rcv.  Errors (a tier beyond kTierCap, a receiver
// cycle, a wait that timed out) surface when the call finishes.
int graph_path(fs_model *m, hipEvent_t *ev) {
  FS_TRY(ensure_graph(m));
  hipStream_t s = m->stream;
  const int tiles = ((m->W + deep::kSide - 1) / deep::kSide) *
                    ((m->H + deep::kSide - 1) / deep::kSide);
  const deep::GraphWork w = graph_work(m);
  hipLaunchKernelGGL(deep::k_deep_clear, dim3(1), dim3(64), 0, s, m->tstat);
  FS_TRY(check_launch());
  FS_TRY(kt_begin(m, FS_K_GRAPH_LOCAL));
  hipLaunchKernelGGL(deep::k_deep_local, dim3(tiles), dim3(deep::kLocalBlock), 0, s, w, m->rcv, m->cnt,
                     m->dinfo, m->link, m->node_word, m->node_kids, m->node_cell, m->nexit,
                     m->tile_n0, m->cells, m->h_alt, m->tstat);
  FS_TRY(check_launch());
  FS_TRY(kt_end(m));
  FS_TRY(mark(ev[2], s));
  FS_TRY(kt_begin(m, FS_K_GRAPH_NODES));
  hipLaunchKernelGGL(deep::k_deep_links, dim3(tiles), dim3(deep::kNodeCap), 0, s, w, m->rcv,
                     m->link, m->dinfo, m->node_cell, m->nexit, m->node_parent, m->node_kids,
                     static_cast<uint32_t *>(nullptr));
  FS_TRY(check_launch());
  hipLaunchKernelGGL(deep::k_deep_tiers, dim3(tiles), dim3(deep::kNodeCap), 0, s, w,
                     m->node_parent, m->nexit, m->node_tier, static_cast<const uint32_t *>(nullptr),
                     static_cast<uint32_t *>(nullptr), m->thist, m->tstat);
  FS_TRY(check_launch());
  hipLaunchKernelGGL(deep::k_deep_scan, dim3(1), dim3(deep::kBlock), 0, s, m->thist, m->tcur,
                     m->tstat, m->graph_queues);
  FS_TRY(check_launch());
  FS_TRY(mark(ev[3], s));
  hipLaunchKernelGGL(deep::k_deep_sum, dim3(tiles), dim3(deep::kNodeCap), 0, s, w, m->node_word,
                     m->node_kids, m->node_parent, m->nexit, m->node_out, m->tstat);
  FS_TRY(check_launch());
  FS_TRY(kt_end(m));
  FS_TRY(mark(ev[4], s));
  FS_TRY(mark(ev[5], s));
  if (m->nmode == 1) FS_TRY(launch_graph_erosion<1>(m, tiles));
  else if (m->nmode == 2) FS_TRY(launch_graph_erosion<2>(m, tiles));
  else FS_TRY(launch_graph_erosion<0>(m, tiles));
  std::swap(m->h, m->h_alt);
  m->order_current = false;                 // levels, stack, donor masks on request
  m->cnt_stack = false;
  m->graph_steps++;
  m->graph_in_call = true;
  return FS_OK;
}

// Error bits of the graph steps of this call (read once the stream is idle).
int graph_status(fs_model *m) {
  if (!m->graph_in_call) return FS_OK;
  m->graph_in_call = false;
  int32_t err = 0;
  FS_HIP_TRY(hipMemcpy(&err, m->tstat + deep::kStErr, sizeof(err), hipMemcpyDeviceToHost));
  if (err & deep::kErrTier) {
    fs::set_last_error("tile-graph path: a drainage path crosses more tile borders than levels allow");
    return FS_ERR_LEVEL_OVERFLOW;
  }
  if (err) {
    fs::set_last_error("tile-graph path: receivers do not form a forest");
    return FS_ERR_INTERNAL;
  }
  return FS_OK;
}

// Depression filling of the current field (an extension; the reference has
// none): the Priority-Flood+epsilon surface, relaxed tile by tile into the
// second buffer until a launch changes no cell, then made current.  Each
// launch finalises every cell whose flood path crosses one more tile
// border, so the launches are bounded.
int fill_depressions(fs_model *m) {
  if (!m->fill_flag) return FS_ERR_INVALID_ARGUMENT;   // models over lent buffers
  hipStream_t s = m->stream;
  const size_t N = static_cast<size_t>(m->N);
  const unsigned ib = static_cast<unsigned>(std::min<size_t>((N + 255) / 256, 1u << 16));
  hipLaunchKernelGGL(k_fill_init, dim3(ib), dim3(256), 0, s, m->h, m->h_alt, m->W, m->H);
  FS_TRY(check_launch());
  const int tiles_x = (m->W + kFillTile - 1) / kFillTile;
  const int tiles = tiles_x * ((m->H + kFillTile - 1) / kFillTile);
  const long long cap = 8ll * tiles + 16;
  for (long long launch = 0;; launch++) {
    if (launch > cap) {
      fs::set_last_error("depression filling did not settle");
      return FS_ERR_INTERNAL;
    }
    FS_HIP_TRY(hipMemsetAsync(m->fill_flag, 0, sizeof(int32_t), s));
    hipLaunchKernelGGL(k_fill, dim3(tiles), dim3(kFillBlock), 0, s, m->h, m->h_alt, m->W, m->H,
                       tiles_x, m->fill_flag);
    FS_TRY(check_launch());
    int32_t changed = 0;
    FS_HIP_TRY(hipMemcpyAsync(&changed, m->fill_flag, sizeof(changed), hipMemcpyDeviceToHost, s));
    FS_HIP_TRY(hipStreamSynchronize(s));
    if (!changed) break;
  }
  std::swap(m->h, m->h_alt);
  return FS_OK;
}

// One time-step.
int step_once(fs_model *m) {
  hipEvent_t *ev = m->ev[m->ev_cur];
  if (m->fill_mode) FS_TRY(fill_depressions(m));
  FS_TRY(mark(ev[0], m->stream));
  FS_TRY(launch_receivers(m));
  FS_TRY(mark(ev[1], m->stream));
  bool done = false;
  if (m->path == FS_PATH_AUTO && m->tile_halo != 0) FS_TRY(tile_path(m, ev, &done));
  const bool tiled = done;
  if (!done && (m->path == FS_PATH_TILE_GRAPH || (m->path == FS_PATH_AUTO && m->owns_memory))) {
    FS_TRY(graph_path(m, ev));
    done = true;
    if (m->path == FS_PATH_AUTO) {
      // drainage this deep rarely shallows again: try the 32-cell halo
      // every kGraphRetry graph steps
      m->tile_halo = ++m->graph_run % kGraphRetry == 0 ? tile32::kHalo : 0;
    }
  }
  if (!done) {
    FS_TRY(level_path(m, ev));
    // the next step tries the tile halo this step's depth would have fitted
    const int D = static_cast<int>(m->counts.size()) - 1;
    m->tile_halo = D <= kHalo ? kHalo : D <= tile32::kHalo ? tile32::kHalo : 0;
  }
  FS_TRY(mark(ev[6], m->stream));
  m->ev_pending[m->ev_cur] = ev[0] != nullptr;
  m->ev_tile[m->ev_cur] = tiled;
  m->ev_cur ^= 1;
  m->have_step = true;
  return FS_OK;
}

int finish(fs_model *m) {
  FS_HIP_TRY(hipStreamSynchronize(m->stream));
  FS_TRY(stage_collect(m, 0));
  FS_TRY(stage_collect(m, 1));
  FS_TRY(kt_collect(m));
  FS_TRY(graph_status(m));
  FS_TRY(fetch_header(m, 0));
  return device_error_status(m->hdr_h[2]);
}

int run_steps(fs_model *m, long long iterations) {
  if (!m) return FS_ERR_INVALID_ARGUMENT;
  if (iterations < 0) return FS_ERR_INVALID_ARGUMENT;
  FS_HIP_TRY(hipSetDevice(m->device));
  const auto t0 = std::chrono::steady_clock::now();
  // clear the device error word left by a previous call
  FS_HIP_TRY(hipMemcpyAsync(m->hdr_h, m->hdr, kHdrInts * sizeof(int32_t), hipMemcpyDeviceToHost,
                            m->stream));
  FS_HIP_TRY(hipStreamSynchronize(m->stream));
  if (m->hdr_h[2] != 0) {
    const int32_t zero = 0;
    FS_HIP_TRY(hipMemcpy(m->hdr + 2, &zero, sizeof(zero), hipMemcpyHostToDevice));
  }
  int st = FS_OK;
  for (long long i = 0; i < iterations && st == FS_OK; i++) {
    m->steps_in_call = i;
    st = step_once(m);
  }
  const int fst = finish(m);
  if (st == FS_OK) st = fst;
  const int sst = split_tile_time(m);
  if (st == FS_OK) st = sst;
  for (int i = 0; i < 6; i++) {
    const double ms = m->stage_ms[FS_T_RECEIVERS + i];
    const uint64_t whole = static_cast<uint64_t>(ms * 1000.0);
    m->timer_us[FS_T_RECEIVERS + i] += whole;
    m->stage_ms[FS_T_RECEIVERS + i] = ms - static_cast<double>(whole) / 1000.0;
  }
  m->timer_us[FS_T_OVERALL] += static_cast<uint64_t>(
      std::chrono::duration_cast<std::chrono::microseconds>(std::chrono::steady_clock::now() - t0)
          .count());
  return st;
}

int check_model(fs_model *m) {
  if (!m) return FS_ERR_INVALID_ARGUMENT;
  const hipError_t e = hipSetDevice(m->device);
  if (e != hipSuccess) {
    fs::set_last_error(std::string("hipSetDevice: ") + hipGetErrorString(e));
    return FS_ERR_HIP;
  }
  return FS_OK;
}

// Rebuild the stack and donor masks of the last step if it took the tile path.
int ensure_order(fs_model *m) {
  if (m->order_current || !m->have_step) return FS_OK;
  FS_HIP_TRY(hipStreamSynchronize(m->stream));
  FS_TRY(build_order(m, nullptr));
  FS_HIP_TRY(hipStreamSynchronize(m->stream));
  FS_TRY(fetch_header(m, 0));
  return device_error_status(m->hdr_h[2]);
}

template <class T>
int download(fs_model *m, T *dst, const T *src, size_t n) {
  FS_HIP_TRY(hipStreamSynchronize(m->stream));
  FS_HIP_TRY(hipMemcpy(dst, src, n * sizeof(T), hipMemcpyDeviceToHost));
  return FS_OK;
}

}  // namespace

// ===========================================================================
// C ABI
// ===========================================================================
extern "C" {

int fs_params_preset(int preset, fs_params *out) {
  if (!out) return FS_ERR_INVALID_ARGUMENT;
  if (preset == FS_PRESET_REFERENCE) {       // RB:58-65
    *out = fs_params{2e-6, 2, 0.8, 2e-3, 1000., 1., 40000, 1e-3};
    return FS_OK;
  }
  if (preset == FS_PRESET_BENCHMARK) {
    *out = fs_params{1e-5, 1, 0.4, 2e-3, 1000., 100., 1e4, 1e-3};
    return FS_OK;
  }
  return FS_ERR_INVALID_ARGUMENT;
}

int fs_create(int width, int height, const fs_params *params, uint32_t engine_seed, int device,
              fs_model **out) {
  return create_model(width, height, params, nullptr, engine_seed, device, out);
}

int fs_create_from_h(int width, int height, const fs_params *params, const double *h, int device,
                     fs_model **out) {
  if (!h) return FS_ERR_INVALID_ARGUMENT;
  return create_model(width, height, params, h, 0, device, out);
}

int fs_destroy(fs_model *m) {
  if (!m) return FS_ERR_INVALID_ARGUMENT;
  free_model(m);
  return FS_OK;
}

int fs_run(fs_model *m, int nstep) {
  if (nstep < 0) return FS_ERR_INVALID_ARGUMENT;
  return run_steps(m, static_cast<long long>(nstep) + 1);   // RB:366
}

int fs_step(fs_model *m, int iterations) {
  if (iterations < 0) return FS_ERR_INVALID_ARGUMENT;
  return run_steps(m, iterations);
}

int fs_get_h(fs_model *m, double *out, size_t count) {
  FS_TRY(check_model(m));
  if (!out) return FS_ERR_INVALID_ARGUMENT;
  if (count != static_cast<size_t>(m->N)) return FS_ERR_BUFFER_SIZE;
  return download(m, out, m->h, count);
}

int fs_set_h(fs_model *m, const double *in, size_t count) {
  FS_TRY(check_model(m));
  if (!in) return FS_ERR_INVALID_ARGUMENT;
  if (count != static_cast<size_t>(m->N)) return FS_ERR_BUFFER_SIZE;
  FS_HIP_TRY(hipStreamSynchronize(m->stream));
  FS_HIP_TRY(hipMemcpy(m->h, in, count * sizeof(double), hipMemcpyHostToDevice));
  return FS_OK;
}

int fs_get_timers(fs_model *m, uint64_t us[FS_NUM_TIMERS]) {
  if (!m || !us) return FS_ERR_INVALID_ARGUMENT;
  for (int i = 0; i < FS_NUM_TIMERS; i++) us[i] = m->timer_us[i];
  return FS_OK;
}

int fs_reset_timers(fs_model *m) {
  if (!m) return FS_ERR_INVALID_ARGUMENT;
  for (int i = 0; i < FS_NUM_TIMERS; i++) {
    m->timer_us[i] = 0;
    m->stage_ms[i] = 0;
  }
  return FS_OK;
}

int fs_get_receivers(fs_model *m, int32_t *rec, size_t count) {
  FS_TRY(check_model(m));
  if (!rec) return FS_ERR_INVALID_ARGUMENT;
  if (count != static_cast<size_t>(m->N)) return FS_ERR_BUFFER_SIZE;
  std::vector<uint8_t> r(count);
  FS_TRY(download(m, r.data(), m->rcv, count));
  for (size_t i = 0; i < count; i++) rec[i] = r[i] == kNoFlow ? FS_NO_FLOW : r[i];
  return FS_OK;
}

int fs_get_donors(fs_model *m, int32_t *ndon, int32_t *donor, size_t count) {
  FS_TRY(check_model(m));
  if (!ndon || !donor) return FS_ERR_INVALID_ARGUMENT;
  if (count != static_cast<size_t>(m->N)) return FS_ERR_BUFFER_SIZE;
  FS_TRY(ensure_order(m));
  std::vector<uint8_t> mk(count);
  FS_TRY(download(m, mk.data(), m->dmask, count));
  const int W = m->W;
  const int off[8] = {-W - 1, -W, -W + 1, -1, 1, W - 1, W, W + 1};
  for (size_t c = 0; c < count; c++) {
    int k = 0;
    for (int j = 0; j < 8; j++)
      if (mk[c] & (1u << j)) donor[8 * c + k++] = static_cast<int32_t>(c) + off[j];
    ndon[c] = k;
    for (; k < 8; k++) donor[8 * c + k] = 0;
  }
  return FS_OK;
}

int fs_get_stack(fs_model *m, int32_t *stack, size_t count) {
  FS_TRY(check_model(m));
  if (!stack) return FS_ERR_INVALID_ARGUMENT;
  if (count != static_cast<size_t>(m->N)) return FS_ERR_BUFFER_SIZE;
  FS_TRY(ensure_order(m));
  return download(m, stack, m->stack, count);
}

int fs_get_levels(fs_model *m, int32_t *levels, size_t capacity, int32_t *nlevel) {
  if (!m || !nlevel) return FS_ERR_INVALID_ARGUMENT;
  FS_TRY(check_model(m));
  FS_TRY(ensure_order(m));
  *nlevel = m->have_step ? m->nlevel : 0;
  if (!levels) return FS_OK;
  if (capacity < static_cast<size_t>(*nlevel)) return FS_ERR_BUFFER_SIZE;
  std::copy(m->levels.begin(), m->levels.begin() + *nlevel, levels);
  return FS_OK;
}

int fs_get_accum(fs_model *m, double *accum, size_t count) {
  FS_TRY(check_model(m));
  if (!accum) return FS_ERR_INVALID_ARGUMENT;
  if (count != static_cast<size_t>(m->N)) return FS_ERR_BUFFER_SIZE;
  std::vector<uint32_t> c(count);
  if (m->cnt_stack) {
    // the level path keeps counts in stack order: scatter them by cell
    std::vector<int32_t> st(count);
    std::vector<uint32_t> cs(count);
    FS_TRY(download(m, st.data(), m->stack, count));
    FS_TRY(download(m, cs.data(), m->cnt_s, count));
    for (size_t p = 0; p < count; p++) {
      if (static_cast<uint32_t>(st[p]) >= count) return FS_ERR_INTERNAL;
      c[static_cast<size_t>(st[p])] = cs[p];
    }
  } else {
    FS_TRY(download(m, c.data(), m->cnt, count));
  }
  for (size_t i = 0; i < count; i++) accum[i] = static_cast<double>(c[i]) * m->p.cell_area;
  return FS_OK;
}

int fs_device_h(fs_model *m, double **dptr) {
  if (!m || !dptr) return FS_ERR_INVALID_ARGUMENT;
  *dptr = m->h;
  return FS_OK;
}

int fs_stream(fs_model *m, void **stream) {
  if (!m || !stream) return FS_ERR_INVALID_ARGUMENT;
  *stream = static_cast<void *>(m->stream);
  return FS_OK;
}

int fs_synchronize(fs_model *m) {
  FS_TRY(check_model(m));
  FS_HIP_TRY(hipStreamSynchronize(m->stream));
  return FS_OK;
}

int fs_set_kernel_timing(fs_model *m, int enable) {
  FS_TRY(check_model(m));
  FS_HIP_TRY(hipStreamSynchronize(m->stream));
  FS_TRY(kt_collect(m));
  m->ktiming = enable != 0;
  for (int i = 0; i < FS_NUM_KERNELS; i++) {
    m->kms[i] = 0;
    m->klaunches[i] = 0;
  }
  return FS_OK;
}

int fs_kernel_stats(fs_model *m, int kernel, double *total_ms, uint64_t *launches) {
  if (!m || kernel < 0 || kernel >= FS_NUM_KERNELS) return FS_ERR_INVALID_ARGUMENT;
  if (total_ms) *total_ms = m->kms[kernel];
  if (launches) *launches = m->klaunches[kernel];
  return FS_OK;
}

int fs_calibrate_copy(size_t nbytes, int device, double *ms) {
  if (!ms || nbytes < 16) return FS_ERR_INVALID_ARGUMENT;
  FS_HIP_TRY(hipSetDevice(device));
  const size_t n = nbytes / 16;
  uint4 *a = nullptr, *b = nullptr;
  FS_TRY(fs::dev_alloc(&a, n));
  if (fs::dev_alloc(&b, n) != FS_OK) {
    (void)hipFree(a);
    return FS_ERR_OUT_OF_MEMORY;
  }
  hipEvent_t e0, e1;
  int st = FS_OK;
  if (hipMemset(a, 1, n * 16) != hipSuccess || hipEventCreate(&e0) != hipSuccess ||
      hipEventCreate(&e1) != hipSuccess) {
    st = FS_ERR_HIP;
  } else {
    const unsigned blocks = static_cast<unsigned>(std::min<size_t>((n + 255) / 256, 1u << 20));
    hipLaunchKernelGGL(k_copy_calibration, dim3(blocks), dim3(256), 0, nullptr, a, b, n);
    (void)hipEventRecord(e0, nullptr);
    hipLaunchKernelGGL(k_copy_calibration, dim3(blocks), dim3(256), 0, nullptr, a, b, n);
    (void)hipEventRecord(e1, nullptr);
    float t = 0.f;
    if (hipEventSynchronize(e1) != hipSuccess || hipEventElapsedTime(&t, e0, e1) != hipSuccess)
      st = FS_ERR_HIP;
    *ms = t;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
  }
  (void)hipFree(a);
  (void)hipFree(b);
  return st;
}

int fs_fill_depressions(fs_model *m) {
  FS_TRY(check_model(m));
  if (!m->owns_memory) return FS_ERR_INVALID_ARGUMENT;
  FS_HIP_TRY(hipStreamSynchronize(m->stream));
  FS_TRY(fill_depressions(m));
  FS_HIP_TRY(hipStreamSynchronize(m->stream));
  return FS_OK;
}

int fs_set_depression_fill(fs_model *m, int enable) {
  if (!m || !m->owns_memory) return FS_ERR_INVALID_ARGUMENT;
  m->fill_mode = enable != 0;
  return FS_OK;
}

int fs_set_step_path(fs_model *m, int path) {
  if (!m || (path != FS_PATH_AUTO && path != FS_PATH_LEVELS && path != FS_PATH_TILE_GRAPH) ||
      (path == FS_PATH_TILE_GRAPH && !m->owns_memory))
    return FS_ERR_INVALID_ARGUMENT;
  m->path = path;
  return FS_OK;
}

int fs_path_stats(fs_model *m, uint64_t *tile_steps, uint64_t *level_steps) {
  if (!m) return FS_ERR_INVALID_ARGUMENT;
  if (tile_steps) *tile_steps = m->tile_steps;
  if (level_steps) *level_steps = m->level_steps;
  return FS_OK;
}

int fs_graph_stats(fs_model *m, uint64_t *graph_steps, int32_t *tiers_last_step) {
  if (!m) return FS_ERR_INVALID_ARGUMENT;
  if (graph_steps) *graph_steps = m->graph_steps;
  if (tiers_last_step) {
    *tiers_last_step = 0;
    if (m->tstat) {
      FS_HIP_TRY(hipSetDevice(m->device));
      FS_HIP_TRY(hipStreamSynchronize(m->stream));
      FS_HIP_TRY(hipMemcpy(tiers_last_step, m->tstat + deep::kStMaxTier, sizeof(int32_t),
                           hipMemcpyDeviceToHost));
    }
  }
  return FS_OK;
}

int fs_tile_stats(fs_model *m, uint64_t *halo16_steps, uint64_t *halo32_steps) {
  if (!m) return FS_ERR_INVALID_ARGUMENT;
  if (halo16_steps) *halo16_steps = m->tile_steps - m->tile32_steps;
  if (halo32_steps) *halo32_steps = m->tile32_steps;
  return FS_OK;
}

const char *fs_kernel_name(int kernel) {
  static const char *names[FS_NUM_KERNELS] = {"k_receivers", "k_topology",   "k_roots",
                                              "k_bfs",       "k_flowacc",    "k_uplift_roots",
                                              "k_erode",     "k_jump",       "k_tile_acc",
                                              "k_tile_erode", "k_deep_local", "k_deep_nodes",
                                              "k_deep_counts", "k_deep_erode"};
  if (kernel < 0 || kernel >= FS_NUM_KERNELS) return "unknown";
  return names[kernel];
}

}  // extern "C"

// ===========================================================================
// the entry point that allocates nothing (fs_step_noalloc)
// ===========================================================================
// Every work array of a time-step lives in one file-scope device array,
// `scratch`, sized for grids of at most kNaMaxCells cells; each GPU holds its
// own copy for the life of the process.  Host code never passes `scratch`
// itself (in host code the name is a stand-in, not a device address): at
// set-up, once per device with that device current, hipGetSymbolAddress
// gives that device's address, and the work arrays are carved from it.
namespace {

constexpr size_t kNaMaxCells = FS_NOALLOC_MAX_CELLS;
constexpr int kNaMaxTiles = 512;        // tile plans held; wider grids take the level path
constexpr int kNaMaxDevices = 64;

constexpr size_t na_align(size_t v) { return (v + 255) & ~size_t{255}; }

// Byte offsets of the work arrays in `scratch` (each 256-byte aligned).
struct NaLayout {
  size_t h_alt, rcv, dmask, lvl, depth, jump, list, stack, cnt, rdir, fdon, ndon, cnt_s, ftab,
      hdr, levels, status, ticket, plan, plan32, total;
};

constexpr NaLayout na_layout() {
  NaLayout L{};
  size_t o = 0;
  const size_t n = kNaMaxCells;
  L.h_alt = o;  o = na_align(o + 8 * n);
  L.rcv = o;    o = na_align(o + n + 16);
  L.dmask = o;  o = na_align(o + n);
  L.lvl = o;    o = na_align(o + n);
  L.depth = o;  o = na_align(o + 2 * n);
  L.jump = o;   o = na_align(o + 8 * n);
  L.list = o;   o = na_align(o + 4 * n);
  L.stack = o;  o = na_align(o + 4 * n);
  L.cnt = o;    o = na_align(o + 4 * n);
  L.rdir = o;   o = na_align(o + n);
  L.fdon = o;   o = na_align(o + 4 * n);
  L.ndon = o;   o = na_align(o + n);
  L.cnt_s = o;  o = na_align(o + 4 * n);
  L.ftab = o;   o = na_align(o + 16 * ((n > kFactCache - 1 ? n : kFactCache - 1) + 1));
  L.hdr = o;    o = na_align(o + 4 * static_cast<size_t>(kHdrAlloc));
  L.levels = o; o = na_align(o + 4 * static_cast<size_t>(FS_MAX_LEVELS + 4));
  L.status = o; o = na_align(o + 8 * (n / kBfsTile + 2));
  L.ticket = o; o = na_align(o + 8);
  L.plan = o;   o = na_align(o + sizeof(TilePlan) * kNaMaxTiles);
  L.plan32 = o; o = na_align(o + sizeof(tile32::TilePlan) * kNaMaxTiles);
  L.total = o;
  return L;
}
constexpr NaLayout kNa = na_layout();
static_assert(kNa.total <= (size_t{76} << 20), "scratch stays within 76 MiB");

__device__ __align__(256) unsigned char scratch[kNa.total];

// Vector-store fill of n bytes at a 16-byte aligned address with byte v
// (memory a caller's graph may hold is cleared by a kernel, not a memset).
__global__ __launch_bounds__(256) void k_fill_bytes(uint8_t *__restrict__ p, uint32_t v,
                                                    size_t n) {
  const uint32_t v4 = (v & 0xFFu) * 0x01010101u;
  const size_t nv = n / 16;
  for (size_t i = blockIdx.x * 256ull + threadIdx.x; i < nv; i += gridDim.x * 256ull)
    reinterpret_cast<uint4 *>(p)[i] = make_uint4(v4, v4, v4, v4);
  if (blockIdx.x == 0 && threadIdx.x < n % 16) p[nv * 16 + threadIdx.x] = static_cast<uint8_t>(v);
}

struct NaDevice {
  std::mutex lock;               // calls on one device run one at a time
  bool ready = false;
  unsigned char *base = nullptr; // this device's address of `scratch`
  fs_model *m = nullptr;         // model shell over the carved arrays
  std::vector<int32_t> hdr_host, levels_host;
  std::unique_ptr<double[]> tab_host;
  // parameters the factor table in `scratch` was filled for, and its size
  bool tab_valid = false;
  double tab_key[5] = {};
  size_t tab_kmax = 0;
};
NaDevice g_na[kNaMaxDevices];

// Once per device, with that device current.
int na_setup(int dev, NaDevice &d) {
  void *ptr = nullptr;
  FS_HIP_TRY(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
  if (!ptr) return FS_ERR_INTERNAL;
  unsigned char *b = static_cast<unsigned char *>(ptr);
  fs_model *m = new (std::nothrow) fs_model();
  if (!m) return FS_ERR_OUT_OF_MEMORY;
  try {
    m->counts.reserve(FS_MAX_LEVELS + 4);
    m->levels.reserve(FS_MAX_LEVELS + 4);
    m->spans.reserve(FS_MAX_LEVELS + 4);
    d.hdr_host.assign(kHdrInts + FS_MAX_LEVELS + 3, 0);
    d.levels_host.assign(FS_MAX_LEVELS + 4, 0);
    d.tab_host.reset(new double[2 * (kNaMaxCells + 1)]);
  } catch (const std::bad_alloc &) {
    delete m;
    return FS_ERR_OUT_OF_MEMORY;
  }
  m->owns_memory = false;
  m->device = dev;
  m->h_alt = reinterpret_cast<double *>(b + kNa.h_alt);
  m->rcv = b + kNa.rcv;
  m->dmask = b + kNa.dmask;
  m->lvl = b + kNa.lvl;
  m->depth = reinterpret_cast<uint16_t *>(b + kNa.depth);
  m->jump = reinterpret_cast<unsigned long long *>(b + kNa.jump);
  m->list = reinterpret_cast<int32_t *>(b + kNa.list);
  m->stack = reinterpret_cast<int32_t *>(b + kNa.stack);
  m->cnt = reinterpret_cast<uint32_t *>(b + kNa.cnt);
  m->rdir = b + kNa.rdir;
  m->fdon = reinterpret_cast<int32_t *>(b + kNa.fdon);
  m->ndon = b + kNa.ndon;
  m->cnt_s = reinterpret_cast<uint32_t *>(b + kNa.cnt_s);
  m->ftab = reinterpret_cast<double *>(b + kNa.ftab);
  m->hdr = reinterpret_cast<int32_t *>(b + kNa.hdr);
  m->levels_d = reinterpret_cast<int32_t *>(b + kNa.levels);
  m->status = reinterpret_cast<uint64_t *>(b + kNa.status);
  m->ticket = reinterpret_cast<unsigned long long *>(b + kNa.ticket);
  m->plan = reinterpret_cast<TilePlan *>(b + kNa.plan);
  m->plan32 = reinterpret_cast<tile32::TilePlan *>(b + kNa.plan32);
  m->hdr_h = d.hdr_host.data();
  m->levels_h = d.levels_host.data();
  d.base = b;
  d.m = m;
  d.ready = true;
  return FS_OK;
}

int na_fill(hipStream_t s, void *p, uint32_t v, size_t n) {
  const unsigned blocks = static_cast<unsigned>(std::min<size_t>((n / 16 + 255) / 256 + 1, 1024));
  hipLaunchKernelGGL(k_fill_bytes, dim3(blocks), dim3(256), 0, s, static_cast<uint8_t *>(p), v, n);
  return check_launch();
}

}  // namespace

extern "C" int fs_step_noalloc(double *h, int width, int height, const fs_params *params,
                               int iterations, void *stream) {
  if (!h || !params || iterations < 0) return FS_ERR_INVALID_ARGUMENT;
  FS_TRY(fs::validate_params(width, height, params));
  const size_t N = static_cast<size_t>(width) * static_cast<size_t>(height);
  if (N > kNaMaxCells) {
    fs::set_last_error("fs_step_noalloc holds grids of at most 1048576 cells");
    return FS_ERR_DIMENSIONS;
  }
  int dev = 0;
  FS_HIP_TRY(hipGetDevice(&dev));
  if (dev < 0 || dev >= kNaMaxDevices) return FS_ERR_NO_DEVICE;
  NaDevice &d = g_na[dev];
  std::lock_guard<std::mutex> guard(d.lock);
  if (!d.ready) FS_TRY(na_setup(dev, d));
  fs_model *m = d.m;
  hipStream_t s = static_cast<hipStream_t>(stream);
  const int tiles = ((width + kTile - 1) / kTile) * ((height + kTile - 1) / kTile);
  // the call's grid and parameters; no step taken yet
  m->W = width;
  m->H = height;
  m->N = static_cast<int32_t>(N);
  m->p = *params;
  m->d_card = params->dx;
  m->d_diag = params->dx * fs::kSqrt2;
  m->ek = make_erode_const(*params, &m->nmode);
  m->stream = s;
  m->h = h;
  m->h_alt = reinterpret_cast<double *>(d.base + kNa.h_alt);
  m->have_step = m->order_current = m->cnt_stack = false;
  m->fill_mode = false;
  m->path = tiles <= kNaMaxTiles ? FS_PATH_AUTO : FS_PATH_LEVELS;
  m->tile_halo = kHalo;
  m->prev_maxdepth = FS_MAX_LEVELS;
  m->ticket_base = 0;
  m->epoch = 0;
  m->nlevel = 0;
  m->tile_steps = m->level_steps = m->tile32_steps = 0;
  // state that must start defined: every receiver NO_FLOW (rings never get
  // one), counts, the header, the scan status words and the ticket
  const size_t ntiles = N / kBfsTile + 2;
  FS_TRY(na_fill(s, m->rcv, 0xFF, N + 16));
  FS_TRY(na_fill(s, m->cnt, 0, 4 * N));
  FS_TRY(na_fill(s, m->hdr, 0, 4 * static_cast<size_t>(kHdrAlloc)));
  FS_TRY(na_fill(s, m->status, 0, 8 * ntiles));
  FS_TRY(na_fill(s, m->ticket, 0, 8));
  // erosion factors for counts 0..max(N, kFactCache - 1), kept across calls
  // with the same parameters
  const size_t kmax = std::max<size_t>(N, kFactCache - 1);
  const double key[5] = {params->cell_area, params->meq, m->ek.kdt, m->ek.lenpow0,
                         m->ek.lenpow1};
  if (!d.tab_valid || d.tab_kmax < kmax || std::memcmp(key, d.tab_key, sizeof(key)) != 0) {
    d.tab_valid = false;
    fs::fill_fact_table(d.tab_host.get(), kmax, params->cell_area, params->meq, m->ek.kdt,
                        m->ek.lenpow0, m->ek.lenpow1, 1);
    FS_HIP_TRY(hipMemcpyAsync(m->ftab, d.tab_host.get(), 2 * (kmax + 1) * sizeof(double),
                              hipMemcpyHostToDevice, s));
    FS_HIP_TRY(hipStreamSynchronize(s));
    std::memcpy(d.tab_key, key, sizeof(key));
    d.tab_kmax = kmax;
    d.tab_valid = true;
  }
  int st = run_steps(m, iterations);
  if (st == FS_OK && m->h != h) {
    // an odd number of tile steps left the field in the second buffer
    const hipError_t e = hipMemcpyAsync(h, m->h, N * sizeof(double), hipMemcpyDeviceToDevice, s);
    if (e == hipSuccess) {
      const hipError_t e2 = hipStreamSynchronize(s);
      if (e2 != hipSuccess) {
        fs::set_last_error(std::string("hipStreamSynchronize: ") + hipGetErrorString(e2));
        st = FS_ERR_HIP;
      }
    } else {
      fs::set_last_error(std::string("hipMemcpyAsync: ") + hipGetErrorString(e));
      st = FS_ERR_HIP;
    }
  }
  m->h = nullptr;
  m->stream = nullptr;
  return st;
}
