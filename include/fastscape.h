/*
 * fastscape.h -- C ABI of the MI355X FastScape time-step library
 * (libfastscape_hip.so).
 *
 * The reference keeps the whole model in one C++ class per variant
 * (class FastScape_RB, fastscape_RB.cpp:46-406).  This header is the
 * language-neutral boundary that replaces it: plain pointers and sizes, no
 * C++ or torch types, status codes instead of asserts and exceptions.  The
 * C++ drop-in class with the reference's own surface is fastscape.hpp.
 *
 * Grid conventions are the reference's: row-major, cell c = y*W + x, ring 0
 * (the outer edge) is a halo, ring 1 is held at elevation 0, interior cells
 * are [2, W-2) x [2, H-2) (fastscape_RB.cpp:73-77, 118-128).
 *
 * Threading: one host thread per model at a time; models are independent.
 * Every model owns one HIP stream; all device work of a model is ordered
 * on it.
 */
#ifndef FASTSCAPE_H
#define FASTSCAPE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes -------------------------------------------------------
 * Replace assert(levels[nlevel-1]==nstack) (fastscape_RB.cpp:253), the
 * timer's runtime_error (CumulativeTimer.cpp:21) and the usage exit
 * (fastscape_RB.cpp:418-421). */
#define FS_OK                    0
#define FS_ERR_INVALID_ARGUMENT  1  /* null pointer, bad preset, nstep < 0  */
#define FS_ERR_DIMENSIONS        2  /* W or H < 5, or W*H >= 2^31           */
#define FS_ERR_OUT_OF_MEMORY     3  /* host or device allocation failed     */
#define FS_ERR_HIP               4  /* a HIP runtime call failed            */
#define FS_ERR_LEVEL_OVERFLOW    5  /* drainage deeper than FS_MAX_LEVELS   */
#define FS_ERR_NEWTON            6  /* a Newton solve did not converge      */
#define FS_ERR_NO_DEVICE         7  /* no HIP device / bad device ordinal   */
#define FS_ERR_BUFFER_SIZE       8  /* caller buffer has the wrong size     */
#define FS_ERR_INTERNAL          9  /* an internal consistency check failed */
#define FS_ERR_IO               10  /* file could not be written            */

/* Direction code of a cell without a receiver (fastscape_RB.cpp:50). */
#define FS_NO_FLOW (-1)
/* Deepest drainage supported: a cell may be at most this many receiver hops
 * from its outlet (the reference sizes its level buffer 2W+2H,
 * fastscape_RB.cpp:358, and overruns it silently beyond that). */
#define FS_MAX_LEVELS 65533

/* ---- parameters -----------------------------------------------------------
 * The reference's constants are const class members (fastscape_RB.cpp:58-65);
 * here they are a parameter block.  Receiver distances are
 * dr = {dx, dx*SQRT2} for cardinal / diagonal neighbours
 * (fastscape_RB.cpp:63, with dx = 1 there). */
typedef struct fs_params {
  double keq;        /* K   stream-power coefficient      RB:58 */
  double neq;        /* n   slope exponent                RB:59 */
  double meq;        /* m   drainage-area exponent        RB:60 */
  double ueq;        /* U   uplift rate                   RB:61 */
  double dt;         /*     time step                     RB:62 */
  double dx;         /*     cell spacing                  RB:63 */
  double cell_area;  /*     area of one cell              RB:65 */
  double tol;        /*     Newton tolerance              RB:64 */
} fs_params;

#define FS_PRESET_REFERENCE 0  /* the reference's constants (RB:58-65)      */
#define FS_PRESET_BENCHMARK 1  /* K=1e-5 n=1 m=0.4 U=2e-3 dt=1000 dx=100
                                  area=1e4                                  */

/* Timer slots, in the order of the reference's timers (RB:100-107). */
#define FS_T_INITIALIZE 0
#define FS_T_RECEIVERS  1
#define FS_T_DONORS     2
#define FS_T_ORDER      3
#define FS_T_FLOWACC    4
#define FS_T_UPLIFT     5
#define FS_T_EROSION    6
#define FS_T_OVERALL    7
#define FS_NUM_TIMERS   8

typedef struct fs_model fs_model;

/* Fill *out with a preset.  No counterpart: the reference hard-codes the
 * constants (RB:58-65). */
int fs_params_preset(int preset, fs_params *out);

/* Construct a model: allocate device state on `device` and fill the terrain
 * with GenerateRandomTerrain's recipe (RB:111-130): one uniform[0,1) double
 * per cell in row-major order from mt19937(engine_seed) via libstdc++'s
 * generate_canonical (random.cpp:45-49), rings 0 and 1 set to 0.
 * Replaces FastScape_RB::FastScape_RB(width, height) (RB:135-151).  The
 * reference seeds its engine with seed*thread (random.cpp:31), so a nonzero
 * seed_rand() argument means engine_seed 0; fastscape.hpp keeps that mapping. */
int fs_create(int width, int height, const fs_params *params,
              uint32_t engine_seed, int device, fs_model **out);

/* Same, with the caller's initial elevations (W*H doubles, row-major). */
int fs_create_from_h(int width, int height, const fs_params *params,
                     const double *h, int device, fs_model **out);

/* Free all host and device memory of the model.  Replaces ~FastScape_RBGPU
 * (fastscape_RB+GPU.cpp:123-125). */
int fs_destroy(fs_model *m);

/* run(nstep): nstep+1 time-steps, as the reference's
 * `for(step=0; step<=nstep; step++)` (RB:341-398).  Resumable: the state is
 * the elevation field, so consecutive calls continue the same run.
 * Timers accumulate.  Prints nothing (the C++ wrapper prints the reference's
 * progress and timer lines). */
int fs_run(fs_model *m, int nstep);

/* Exactly `iterations` time-steps (no +1).  No reference counterpart; used
 * by benchmarks and per-step parity checks. */
int fs_step(fs_model *m, int iterations);

/* The time-step of fs_step on a caller's elevation field, allocating
 * nothing.  No counterpart in the reference: it is for callers that cannot
 * allocate device or pinned memory (serving loops, harnesses that forbid
 * hipMalloc).  h: width*height doubles on the current HIP device, updated in
 * place by `iterations` time-steps on `stream` (a hipStream_t; NULL = the
 * null stream); same results, bit for bit, as fs_create_from_h + fs_step.
 * Every work array comes from one device array built into the library,
 * sized for at most FS_NOALLOC_MAX_CELLS cells (about 73 MB, held on each
 * GPU the library runs on until the process ends); larger grids return
 * FS_ERR_DIMENSIONS.  Grids of more than 512 64x64 tiles take the level
 * path.  All calls on one GPU share that array, which keeps the last call's
 * contents: calls on one device are serialised by a lock (overlapping calls
 * would race on the shared array), and each returns after its device work
 * has finished.  Stage timers and per-stage readouts are not kept. */
#define FS_NOALLOC_MAX_CELLS 1048576
int fs_step_noalloc(double *h, int width, int height, const fs_params *params, int iterations,
                    void *stream);

/* Copy the elevations (W*H doubles) device -> host.  Replaces getH()
 * (RB:403-405), which returns the reference's internal vector. */
int fs_get_h(fs_model *m, double *out, size_t count);

/* Overwrite the elevations (W*H doubles) host -> device (resume / edit, the
 * use RB:334-340 describes). */
int fs_set_h(fs_model *m, const double *in, size_t count);

/* Per-stage times in microseconds, in FS_T_* order (RB:380-387).  Stage
 * times are device time between HIP events bracketing each stage; Overall is
 * host wall time of fs_run / fs_step plus construction.  On the tile path
 * (below) donors, levels and accumulation are one kernel and uplift and
 * erosion another; their time is split over the stages by phase timestamps
 * that the kernels record on one step in 32.  On the tile-graph path,
 * DetermineDonors is the in-tile walks, GenerateOrder the tile graph,
 * FlowAcc the sums over it, Uplift is fused into Erosion (which also
 * finishes the counts).  FS_PATH_LEVELS gives the reference's stage split. */
int fs_get_timers(fs_model *m, uint64_t us[FS_NUM_TIMERS]);
int fs_reset_timers(fs_model *m);

/* ---- step path ------------------------------------------------------------
 * FS_PATH_AUTO (default): when every cell is at most FS_TILE_HALO receiver
 * hops from its outlet, accumulation and erosion run per 64x64 tile with a
 * 16-cell halo in LDS (two kernels per step); when every cell is at most
 * FS_TILE_HALO_DEEP hops from it, the same two kernels run with a 32-cell
 * halo; otherwise (or when a tile's chains need more halo cells than its
 * plan holds: 256 with the 16-cell halo, 512 with the 32-cell one) the step
 * takes the tile-graph path: per 64x64 tile with no halo, the tiles joined
 * by the graph of the cells where drainage leaves a tile, the erosion one
 * persistent launch that takes (tile, tier) work items in the order of the
 * tile borders crossed on the way to an outlet.  Deep drainage stays on the
 * tile-graph path and retries the 32-cell halo every 64 steps.
 * Models over lent buffers (fs_step_noalloc) take the level path instead.
 * FS_PATH_LEVELS: always the level path -- one launch per level for order,
 * accumulation and erosion, as the reference's stages (RB:366-372).
 * FS_PATH_TILE_GRAPH: always the tile-graph path.  All give bit-identical
 * results.  No reference counterpart. */
#define FS_PATH_AUTO   0
#define FS_PATH_LEVELS 1
#define FS_PATH_TILE_GRAPH 2
#define FS_TILE_HALO   16
#define FS_TILE_HALO_DEEP 32
int fs_set_step_path(fs_model *m, int path);
/* Number of time-steps taken by each path since creation; tile_steps counts
 * both halos. */
int fs_path_stats(fs_model *m, uint64_t *tile_steps, uint64_t *level_steps);
/* Tile-path steps split by halo (16-cell, 32-cell).  No reference
 * counterpart. */
int fs_tile_stats(fs_model *m, uint64_t *halo16_steps, uint64_t *halo32_steps);
/* Tile-graph path steps, and the most tile borders one drainage path
 * crossed in the last of them (synchronises the model's stream).  No
 * reference counterpart. */
int fs_graph_stats(fs_model *m, uint64_t *graph_steps, int32_t *tiers_last_step);

/* ---- depression filling ----------------------------------------------------
 * No counterpart in the reference's code: its README claims depression
 * handling (README.md:35-37), but pits stay NO_FLOW there and never erode.
 * The fill raises every interior cell to the Priority-Flood+epsilon surface
 * (Barnes, Lehman & Mulla 2014): the lowest surface at or above h on which
 * every interior cell has a strictly lower D8 neighbour (one ulp, by
 * nextafter), rings 0 and 1 unchanged -- the same bits as a priority-flood
 * on the CPU.  fs_fill_depressions fills the current field once;
 * fs_set_depression_fill(m, 1) fills before the receivers of every
 * following time-step, so every interior cell drains to the rings (off by
 * default: the reference's behaviour).  Not available to models stepped by
 * fs_step_noalloc. */
int fs_fill_depressions(fs_model *m);
int fs_set_depression_fill(fs_model *m, int enable);

/* ---- state of the last completed time-step (parity checks) -------------- */
/* After a tile-path step, levels[], the stack and the donor lists are
 * rebuilt from that step's receivers on the first request (the level path's
 * kernels); accumulation and receivers are always kept. */
/* rec[] in the reference encoding: 0..7 or FS_NO_FLOW (RB:86, 160-182). */
int fs_get_receivers(fs_model *m, int32_t *rec, size_t count);
/* ndon[count] and donor[8*count] as ComputeDonors leaves them
 * (RB:87-88, 188-203): donor lists in ascending donor index; slots beyond
 * ndon are written as 0. */
int fs_get_donors(fs_model *m, int32_t *ndon, int32_t *donor, size_t count);
/* stack[] (RB:89, 213-254): the full level order, `count` = W*H. */
int fs_get_stack(fs_model *m, int32_t *stack, size_t count);
/* levels[] (RB:96) and nlevel (RB:97).  `*nlevel` receives the count; pass
 * levels = NULL to query it. */
int fs_get_levels(fs_model *m, int32_t *levels, size_t capacity, int32_t *nlevel);
/* accum[] (RB:85, 262-279), cell_area times the upstream cell count. */
int fs_get_accum(fs_model *m, double *accum, size_t count);

/* ---- device access --------------------------------------------------------
 * Device pointer of the elevation field (W*H doubles on the model's device)
 * and the model's stream (a hipStream_t).  Work a caller enqueues on other
 * streams must be ordered against this stream by the caller.
 * Lifetime: the pointer names the current field only until the next
 * fs_run / fs_step.  The tile path reads one buffer and writes the other
 * (neighbouring tiles read each other's cells as halo), then exchanges
 * them, so after a step the field may live in the other buffer: fetch the
 * pointer again after every step.  Writes through a stale pointer are
 * lost; reads through it see an older field. */
int fs_device_h(fs_model *m, double **dptr);
int fs_stream(fs_model *m, void **stream);
int fs_synchronize(fs_model *m);

/* ---- kernel timing ---------------------------------------------------------
 * When enabled, every launch of each kernel is bracketed by HIP events on the
 * model's stream and its duration accumulated per kernel.  Kernel ids:
 * FS_K_* below.  fs_kernel_stats returns total milliseconds and launches. */
#define FS_K_RECEIVERS 0
#define FS_K_TOPOLOGY  1
#define FS_K_ROOTS     2
#define FS_K_BFS       3
#define FS_K_FLOWACC   4
#define FS_K_UPLIFT    5
#define FS_K_EROSION   6
#define FS_K_JUMP      7
#define FS_K_TILE_ACC  8
#define FS_K_TILE_ERODE 9
#define FS_K_GRAPH_LOCAL 10
#define FS_K_GRAPH_NODES 11
#define FS_K_GRAPH_COUNTS 12
#define FS_K_GRAPH_ERODE 13
#define FS_NUM_KERNELS 14
int fs_set_kernel_timing(fs_model *m, int enable);
int fs_kernel_stats(fs_model *m, int kernel, double *total_ms, uint64_t *launches);
const char *fs_kernel_name(int kernel);

/* Copy `nbytes` (rounded down to 16) between two device buffers with a plain
 * streaming kernel, twice; *ms receives the second launch's time.  No
 * reference counterpart: it calibrates hardware traffic counters (each
 * launch reads and writes exactly nbytes). */
int fs_calibrate_copy(size_t nbytes, int device, double *ms);

/* ---- ensembles --------------------------------------------------------------
 * Many independent grids ("realizations") of different sizes and parameters,
 * advanced together by one persistent cooperative launch per time-step.  No
 * reference counterpart: the reference runs one grid per process; its README
 * motivates ensembles for inverse problems (README.md:28-29).  Each
 * realization's time-step is the reference's (RB:366-376) and gives the same
 * bits as an fs_model of the same size, parameters and initial terrain.
 * Realization r's initial terrain is GenerateRandomTerrain (RB:111-130) from
 * an mt19937 seeded with its engine_seed (0 = the reference's terrain). */
typedef struct fs_ensemble fs_ensemble;
typedef struct fs_realization {
  int width, height;
  fs_params params;
  uint32_t engine_seed;
} fs_realization;

/* Work schedule of a step; every schedule gives the same results.
 *   recv_rows        rows of 256 columns per receiver work item (16..256,
 *                    a multiple of 16)
 *   tiles_per_claim  64x64 tiles per erosion work item (1..16)
 *   group_cells      accepted (>= 0) and without effect: deep realizations
 *                    take the tile-graph step, whose work spreads over every
 *                    CU */
typedef struct fs_schedule {
  int recv_rows;
  int tiles_per_claim;
  int group_cells;
} fs_schedule;

int fs_ensemble_create(const fs_realization *reals, int count, int device, fs_ensemble **out);
int fs_ensemble_destroy(fs_ensemble *e);
int fs_ensemble_size(fs_ensemble *e, int *count, size_t *cells);
/* nstep+1 time-steps of every realization (RB:366), resumable. */
int fs_ensemble_run(fs_ensemble *e, int nstep);
int fs_ensemble_step(fs_ensemble *e);
int fs_ensemble_get_h(fs_ensemble *e, int r, double *h, size_t count);
int fs_ensemble_set_h(fs_ensemble *e, int r, const double *h, size_t count);
/* State of realization r after the last step, in fs_get_* conventions. */
int fs_ensemble_get_receivers(fs_ensemble *e, int r, int32_t *rec, size_t count);
int fs_ensemble_get_accum(fs_ensemble *e, int r, double *accum, size_t count);
int fs_ensemble_get_levels(fs_ensemble *e, int r, int32_t *levels, size_t capacity,
                           int32_t *nlevel);
/* Steps realization r took on the halo tiles and off them (level_steps:
 * drainage deeper than 16 levels, the tile-graph step). */
int fs_ensemble_path_stats(fs_ensemble *e, int r, uint64_t *tile_steps, uint64_t *level_steps);
int fs_ensemble_set_schedule(fs_ensemble *e, const fs_schedule *s);
int fs_ensemble_get_schedule(fs_ensemble *e, fs_schedule *s);
/* Online autotuning inside fs_ensemble_run: at the run's first step and
 * then every `interval` steps (0 = only at the first), each candidate
 * schedule of a coordinate search runs `trial_steps` real time-steps and
 * the fastest is kept.  trial_steps = 0 turns tuning off (the default). */
int fs_ensemble_autotune(fs_ensemble *e, int trial_steps, int interval);
/* Tuning record: candidates timed so far and the ms/step of each; pass
 * NULL arrays to query the count. */
int fs_ensemble_tune_log(fs_ensemble *e, fs_schedule *cand, double *ms_per_step,
                         int capacity, int *count);
/* Steps taken and summed device milliseconds of their launches. */
int fs_ensemble_stats(fs_ensemble *e, uint64_t *steps, double *step_ms);
int fs_ensemble_synchronize(fs_ensemble *e);

/* ---- learned proposal network ----------------------------------------------
 * A 4-level convolutional U-Net (widths 32/64/128/160, ~1.4M parameters) in
 * BF16 on the MFMA units with FP32 accumulation, trained with Adam on FP32
 * master weights.  It maps (initial terrain, 3 normalised parameters) on an
 * H x W grid (multiples of 32) to one output channel; the inversion driver
 * uses it to rank candidate (K, U, m) sets so that only the best few are run
 * through the exact time-step.  No reference counterpart.
 * terrain: batch*H*W floats; params: batch*3 floats; outputs batch*H*W. */
#define FS_UNET_IN_CHANNELS 8
typedef struct fs_unet fs_unet;
int fs_unet_create(int height, int width, int max_batch, uint64_t seed, int device,
                   fs_unet **out);
int fs_unet_destroy(fs_unet *u);
int fs_unet_param_count(fs_unet *u, size_t *n);
/* Parameter vector: per conv layer W[tap][out][in] then bias[out], then the
 * head's 32 weights and its bias.  Layer i has cin[i] -> cout[i] channels at
 * resolution (H >> level[i]) x (W >> level[i]). */
int fs_unet_layer_shapes(fs_unet *u, int *cin, int *cout, int *level, int capacity, int *count);
int fs_unet_get_params(fs_unet *u, float *p, size_t n);
int fs_unet_set_params(fs_unet *u, const float *p, size_t n);
int fs_unet_forward(fs_unet *u, const float *terrain, const float *params, int batch, float *y);
/* loss = mean((y - target)^2) over batch*H*W, and its gradient w.r.t. every
 * parameter (grad may be NULL).  batch <= max_batch. */
int fs_unet_loss_grad(fs_unet *u, const float *terrain, const float *params, const float *target,
                      int batch, float *loss, float *grad, size_t n);
/* Training set kept in device memory; fs_unet_train runs `epochs` passes of
 * shuffled minibatches, one Adam step each, learning rate decayed by a
 * cosine from lr to lr/20; epoch_loss[epochs] receives mean losses. */
int fs_unet_set_data(fs_unet *u, const float *terrain, const float *params, const float *target,
                     int count);
int fs_unet_train(fs_unet *u, int epochs, int batch, float lr, uint64_t seed, float *epoch_loss);
/* Misfit of `count` parameter sets sharing one terrain against one target:
 * misfit[i] = RMS over the interior cells (rings 0 and 1 excluded) of
 * output_i - target.  Outputs stay on the device. */
int fs_unet_rank(fs_unet *u, const float *terrain, const float *params, int count,
                 const float *target, float *misfit);
int fs_unet_stats(fs_unet *u, double *train_ms, double *flops_per_sample, uint64_t *samples);

/* ---- host helpers ---------------------------------------------------------- */
/* PrintDEM (RB:20-40): ArcGIS ASCII grid of the inner (W-2)x(H-2) cells,
 * byte-identical to the reference writer. */
int fs_write_dem(const char *filename, const double *h, int width, int height);
/* GenerateRandomTerrain's recipe on the host (RB:111-130). */
int fs_generate_terrain(double *h, int width, int height, uint32_t engine_seed);
/* Message for a status code; never NULL. */
const char *fs_strerror(int status);
/* Detail of the last error raised on this thread (e.g. the failing HIP
 * call); empty string if none. */
const char *fs_last_error(void);

#ifdef __cplusplus
}
#endif
#endif /* FASTSCAPE_H */
