/*
 * fastscape_oracle.c -- CPU oracle (checker) for the FastScape time-step.
 *
 * TEST INFRASTRUCTURE ONLY: a plain-C restatement of the reference's
 * level-ordered FastScape step, used to check the HIP implementation.  It is
 * loaded only by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
 * leg, never by the product path.
 *
 * Each function cites the reference lines it restates (RB = fastscape_RB.cpp,
 * P = fastscape_RB+P.cpp, random.cpp, CumulativeTimer.cpp).  Build with
 * -ffp-contract=off (see oracle/Makefile).
 *
 * Pinning (what checks each function besides retyping it):
 *   fso_mt_next / fso_canonical : mt19937 known answer (10000th output of the
 *       default-seeded engine = 4123659995) and libstdc++'s own
 *       generate_canonical / uniform_real_distribution on the same engine
 *       (tests/test_oracle.py).
 *   fso_init_terrain .. fso_erode, fso_model_run : per-stage arrays and final
 *       elevations of the compiled reference (oracle/_ref, driven by
 *       oracle/ref_harness.cpp) committed as fixtures in tests/golden/, plus
 *       brute-force and closed-form checks on small grids.
 *   fso_erode_levels_omp : equality with fso_erode on the same input.
 *   fso_write_dem : byte equality with the reference PrintDEM output.
 */
#include "fastscape_oracle.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* RB:51 -- the literal rounds to the same double as sqrt(2). */
static const double FSO_SQRT2 = 1.414213562373095048801688724209698078569671875376948;

int fso_preset(int which, fso_params *out) {
  if (!out) return -1;
  if (which == 0) {            /* RB:58-65 */
    out->keq = 2e-6; out->neq = 2; out->meq = 0.8; out->ueq = 2e-3;
    out->dt = 1000.; out->dx = 1.; out->cell_area = 40000; out->tol = 1e-3;
    return 0;
  }
  if (which == 1) {            /* benchmark constants: dx=dy=100 m, area=dx^2 */
    out->keq = 1e-5; out->neq = 1; out->meq = 0.4; out->ueq = 2e-3;
    out->dt = 1000.; out->dx = 100.; out->cell_area = 1e4; out->tol = 1e-3;
    return 0;
  }
  return -1;
}

/* ------------------------------------------------------------------------ */
/* mt19937 (Matsumoto & Nishimura 1998), the engine of random.hpp:22.        */
/* ------------------------------------------------------------------------ */
void fso_mt_seed(fso_mt19937 *e, uint32_t seed) {
  e->mt[0] = seed;
  for (int i = 1; i < 624; i++)
    e->mt[i] = 1812433253u * (e->mt[i - 1] ^ (e->mt[i - 1] >> 30)) + (uint32_t)i;
  e->mti = 624;
}

uint32_t fso_mt_next(fso_mt19937 *e) {
  if (e->mti >= 624) {
    for (int k = 0; k < 624; k++) {
      uint32_t y = (e->mt[k] & 0x80000000u) | (e->mt[(k + 1) % 624] & 0x7fffffffu);
      e->mt[k] = e->mt[(k + 397) % 624] ^ (y >> 1) ^ ((y & 1u) ? 0x9908b0dfu : 0u);
    }
    e->mti = 0;
  }
  uint32_t y = e->mt[e->mti++];
  y ^= y >> 11;
  y ^= (y << 7) & 0x9d2c5680u;
  y ^= (y << 15) & 0xefc60000u;
  y ^= y >> 18;
  return y;
}

/* uniform_rand_real(0,1) (random.cpp:45-49) = libstdc++'s
 * generate_canonical<double,53> over a 32-bit engine: two draws x0, x1,
 * sum = x0 + x1*2^32 rounded to double, divided by 2^64, and 1.0 replaced by
 * the largest double below 1; then *(1-0)+0, which is exact. */
double fso_canonical(fso_mt19937 *e) {
  double sum = 0.0, tmp = 1.0;
  sum += (double)fso_mt_next(e) * tmp; tmp *= 4294967296.0;
  sum += (double)fso_mt_next(e) * tmp; tmp *= 4294967296.0;
  double r = sum / tmp;
  if (r >= 1.0) r = nextafter(1.0, 0.0);
  return r * (1.0 - 0.0) + 0.0;
}

/* GenerateRandomTerrain (RB:111-130): one uniform draw per cell in row-major
 * order, then rings 0 and 1 forced to 0.  seed_rand(seed) seeds thread 0's
 * engine with seed*0 (random.cpp:31), so any nonzero CLI seed gives
 * engine_seed 0 here. */
void fso_init_terrain(double *h, int W, int H, uint32_t engine_seed) {
  fso_mt19937 e;
  fso_mt_seed(&e, engine_seed);
  for (int y = 0; y < H; y++)
    for (int x = 0; x < W; x++) {
      const size_t c = (size_t)y * W + x;
      h[c] = fso_canonical(&e);
      if (x == 0 || y == 0 || x == W - 1 || y == H - 1) h[c] = 0;
      if (x == 1 || y == 1 || x == W - 2 || y == H - 2) h[c] = 0;
    }
}

static void fso_nshift(int W, int nshift[8]) {
  /* RB:137, directions 0..7 = W, NW, N, NE, E, SE, S, SW (RB:79-82) */
  nshift[0] = -1; nshift[1] = -W - 1; nshift[2] = -W; nshift[3] = -W + 1;
  nshift[4] = 1;  nshift[5] = W + 1;  nshift[6] = W;  nshift[7] = W - 1;
}

/* ComputeReceivers (RB:160-182): interior cells only; steepest positive
 * slope (h[c]-h[n])/dr[n], strict '>' from 0 so ties go to the lowest
 * direction and flat/pit cells stay NO_FLOW. */
void fso_receivers(const double *h, int *rec, int W, int H, const double dr[8]) {
  int nshift[8];
  fso_nshift(W, nshift);
  for (int y = 2; y < H - 2; y++)
    for (int x = 2; x < W - 2; x++) {
      const int c = y * W + x;
      double max_slope = 0;
      int max_n = FSO_NO_FLOW;
      for (int n = 0; n < 8; n++) {
        const double slope = (h[c] - h[c + nshift[n]]) / dr[n];
        if (slope > max_slope) { max_slope = slope; max_n = n; }
      }
      rec[c] = max_n;
    }
}

/* ComputeDonors (RB:188-203): invert rec by a push over cells in ascending
 * index, so each donor list is in ascending donor index. */
void fso_donors(const int *rec, int *ndon, int *donor, int W, int H) {
  int nshift[8];
  fso_nshift(W, nshift);
  const int size = W * H;
  for (int i = 0; i < size; i++) ndon[i] = 0;
  for (int c = 0; c < size; c++) {
    if (rec[c] == FSO_NO_FLOW) continue;
    const int n = c + nshift[rec[c]];
    donor[8 * n + ndon[n]] = c;
    ndon[n]++;
  }
}

/* GenerateOrder (RB:213-254): level 0 = every NO_FLOW cell in ascending
 * index (ring 0 included), level L+1 = the donor lists of level L's cells in
 * level-L order.  levels[] holds the level offsets and ends with a duplicate
 * of the stack size that RB's loop leaves behind (RB:249-251).  Returns 0, or
 * -1 if levels_cap is too small (the reference would overrun its buffer). */
int fso_order(const int *rec, const int *ndon, const int *donor, int *stack,
              int *levels, int levels_cap, int W, int H, int *nlevel_out) {
  const int size = W * H;
  int nstack = 0, nlevel;
  if (levels_cap < 3) return -1;
  levels[0] = 0;
  nlevel = 1;
  for (int c = 0; c < size; c++)
    if (rec[c] == FSO_NO_FLOW) stack[nstack++] = c;
  levels[nlevel++] = nstack;
  int level_bottom = -1, level_top = 0;
  while (level_bottom < level_top) {
    level_bottom = level_top;
    level_top = nstack;
    for (int si = level_bottom; si < level_top; si++) {
      const int c = stack[si];
      for (int k = 0; k < ndon[c]; k++) stack[nstack++] = donor[8 * c + k];
    }
    if (nlevel >= levels_cap) return -1;
    levels[nlevel++] = nstack;
  }
  nlevel--;
  if (levels[nlevel - 1] != nstack) return -2;   /* RB:253 */
  *nlevel_out = nlevel;
  return 0;
}

/* ComputeFlowAcc (RB:262-279): accum = cell_area, then push accum[c] into
 * its receiver in reverse stack order.  Every value is an integer multiple of
 * cell_area, so all partial sums are exact. */
void fso_flowacc(const int *stack, const int *rec, double *accum, int W, int H,
                 double cell_area) {
  int nshift[8];
  fso_nshift(W, nshift);
  const int size = W * H;
  for (int i = 0; i < size; i++) accum[i] = cell_area;
  for (int s = size - 1; s >= 0; s--) {
    const int c = stack[s];
    if (rec[c] != FSO_NO_FLOW) {
      const int n = c + nshift[rec[c]];
      accum[n] += accum[c];
    }
  }
}

/* AddUplift (RB:285-296): interior cells only. */
void fso_uplift(double *h, int W, int H, double ueq, double dt) {
  for (int y = 2; y < H - 2; y++)
    for (int x = 2; x < W - 2; x++) h[(size_t)y * W + x] += ueq * dt;
}

static inline double fso_powx(double x, double e, int integer_e, int powmode) {
  if (powmode == FSO_POW_EXACT) {
    if (integer_e == 0) return 1.0;
    if (integer_e == 1) return x;
    if (integer_e == 2) return x * x;
  }
  return pow(x, e);
}

/* The Newton solve of one cell (RB:313-326).  Returns the new elevation;
 * *iters receives the iteration count.  The reference has no iteration cap;
 * the oracle stops after 1000 iterations and reports it (never reached on the
 * reference's inputs). */
static inline double fso_newton(double h0, double hn, double accum, double length,
                                const fso_params *p, int powmode, int *iters) {
  const double neq = p->neq;
  const int in = (neq == 1.0) ? 1 : (neq == 2.0) ? 2 : -1;
  const int in1 = (neq == 1.0) ? 0 : (neq == 2.0) ? 1 : -1;
  const double fact = p->keq * p->dt * pow(accum, p->meq) / pow(length, neq);
  double hnew = h0, hp = h0, diff = 2 * p->tol;
  int it = 0;
  while (fabs(diff) > p->tol) {
    hnew -= (hnew - h0 + fact * fso_powx(hnew - hn, neq, in, powmode)) /
            (1. + fact * neq * fso_powx(hnew - hn, neq - 1, in1, powmode));
    diff = hnew - hp;
    hp = hnew;
    if (++it >= 1000) break;
  }
  *iters = it;
  return hnew;
}

/* Erode (RB:306-328): stack order, downstream first, cells without a
 * receiver skipped.  Returns 0, or -1 if a Newton solve hit the cap. */
int fso_erode(const int *stack, const int *rec, const double *accum, double *h,
              int W, int H, const fso_params *p, int powmode) {
  int nshift[8];
  fso_nshift(W, nshift);
  const double dr[8] = {p->dx, p->dx * FSO_SQRT2, p->dx, p->dx * FSO_SQRT2,
                        p->dx, p->dx * FSO_SQRT2, p->dx, p->dx * FSO_SQRT2};
  const int size = W * H;
  int bad = 0;
  for (int s = 0; s < size; s++) {
    const int c = stack[s];
    if (rec[c] == FSO_NO_FLOW) continue;
    const int n = c + nshift[rec[c]];
    int it;
    h[c] = fso_newton(h[c], h[n], accum[c], dr[rec[c]], p, powmode, &it);
    if (it >= 1000) bad = 1;
  }
  return bad ? -1 : 0;
}

/* Erode of fastscape_RB+P.cpp (P:313-351): level by level from level 1, each
 * level split over OpenMP threads when it holds more than 500 cells. */
int fso_erode_levels_omp(const int *stack, const int *levels, int nlevel,
                         const int *rec, const double *accum, double *h,
                         int W, int H, const fso_params *p, int powmode) {
  int nshift[8];
  (void)H;
  fso_nshift(W, nshift);
  const double dr[8] = {p->dx, p->dx * FSO_SQRT2, p->dx, p->dx * FSO_SQRT2,
                        p->dx, p->dx * FSO_SQRT2, p->dx, p->dx * FSO_SQRT2};
  int bad = 0;
  for (int li = 1; li < nlevel - 1; li++) {
    const int lvlstart = levels[li], lvlend = levels[li + 1];
    const int lvlsize = lvlend - lvlstart;
#pragma omp parallel for if (lvlsize > 500) reduction(| : bad)
    for (int si = lvlstart; si < lvlend; si++) {
      const int c = stack[si];
      if (rec[c] == FSO_NO_FLOW) continue;
      const int n = c + nshift[rec[c]];
      int it;
      h[c] = fso_newton(h[c], h[n], accum[c], dr[rec[c]], p, powmode, &it);
      if (it >= 1000) bad |= 1;
    }
  }
  return bad ? -1 : 0;
}

/* ------------------------------------------------------------------------ */
/* Whole model: constructor RB:135-151, run RB:341-398, timers              */
/* CumulativeTimer.cpp:9-37 (microseconds, truncated per interval).          */
/* ------------------------------------------------------------------------ */
static uint64_t fso_now_ns(void) {
  struct timespec ts;
  clock_gettime(CLOCK_MONOTONIC, &ts);
  return (uint64_t)ts.tv_sec * 1000000000ull + (uint64_t)ts.tv_nsec;
}

int fso_model_init(fso_model *m, int W, int H, const fso_params *p,
                   uint32_t engine_seed) {
  memset(m, 0, sizeof(*m));
  if (W < 5 || H < 5) return -1;
  m->W = W; m->H = H; m->size = W * H; m->p = *p;
  for (int i = 0; i < 8; i++) m->dr[i] = (i & 1) ? p->dx * FSO_SQRT2 : p->dx;
  fso_nshift(W, m->nshift);
  const size_t N = (size_t)m->size;
  m->h = (double *)malloc(N * sizeof(double));
  m->accum = (double *)malloc(N * sizeof(double));
  m->rec = (int *)malloc(N * sizeof(int));
  m->ndon = (int *)malloc(N * sizeof(int));
  m->donor = (int *)malloc(8 * N * sizeof(int));
  m->stack = (int *)malloc(N * sizeof(int));
  m->levels_cap = m->size + 3;   /* never overflows (one level per cell) */
  m->levels = (int *)malloc((size_t)m->levels_cap * sizeof(int));
  if (!m->h || !m->accum || !m->rec || !m->ndon || !m->donor || !m->stack ||
      !m->levels) {
    fso_model_free(m);
    return -2;
  }
  uint64_t t0 = fso_now_ns();
  fso_init_terrain(m->h, W, H, engine_seed);
  for (size_t i = 0; i < N; i++) m->rec[i] = FSO_NO_FLOW;   /* RB:361-362 */
  uint64_t us = (fso_now_ns() - t0) / 1000;
  m->timer_us[0] += us;
  m->timer_us[7] += us;
  return 0;
}

void fso_model_free(fso_model *m) {
  free(m->h); free(m->accum); free(m->rec); free(m->ndon); free(m->donor);
  free(m->stack); free(m->levels);
  memset(m, 0, sizeof(*m));
}

int fso_model_step(fso_model *m, int powmode, int parallel_erosion) {
  const int W = m->W, H = m->H;
  uint64_t t[7];
  int rc;
  t[0] = fso_now_ns();
  fso_receivers(m->h, m->rec, W, H, m->dr);
  t[1] = fso_now_ns();
  fso_donors(m->rec, m->ndon, m->donor, W, H);
  t[2] = fso_now_ns();
  rc = fso_order(m->rec, m->ndon, m->donor, m->stack, m->levels, m->levels_cap,
                 W, H, &m->nlevel);
  if (rc) return -3;
  t[3] = fso_now_ns();
  fso_flowacc(m->stack, m->rec, m->accum, W, H, m->p.cell_area);
  t[4] = fso_now_ns();
  fso_uplift(m->h, W, H, m->p.ueq, m->p.dt);
  t[5] = fso_now_ns();
  if (parallel_erosion)
    rc = fso_erode_levels_omp(m->stack, m->levels, m->nlevel, m->rec, m->accum,
                              m->h, W, H, &m->p, powmode);
  else
    rc = fso_erode(m->stack, m->rec, m->accum, m->h, W, H, &m->p, powmode);
  t[6] = fso_now_ns();
  for (int i = 0; i < 6; i++) m->timer_us[i + 1] += (t[i + 1] - t[i]) / 1000;
  return rc ? -4 : 0;
}

int fso_model_run(fso_model *m, int nstep, int powmode, int parallel_erosion) {
  uint64_t t0 = fso_now_ns();
  const size_t N = (size_t)m->size;
  for (size_t i = 0; i < N; i++) m->rec[i] = FSO_NO_FLOW;   /* RB:361-362 */
  m->timer_us[0] += (fso_now_ns() - t0) / 1000;
  for (int step = 0; step <= nstep; step++) {               /* RB:366 */
    int rc = fso_model_step(m, powmode, parallel_erosion);
    if (rc) return rc;
  }
  m->timer_us[7] += (fso_now_ns() - t0) / 1000;
  return 0;
}

/* PrintDEM (RB:20-40): the inner (W-2)x(H-2) cells, default ostream
 * formatting of a double (printf "%.6g"), each value followed by a space,
 * each row by a newline. */
/* ---- depression filling: Priority-Flood+epsilon (see the header) ---- */
typedef struct { double v; int c; } fso_heap_item;

/* Binary min-heap on (value, cell index): ties pop the lower index first, so
 * the order is deterministic (the filled values do not depend on it). */
static int fso_heap_less(fso_heap_item a, fso_heap_item b) {
  return a.v < b.v || (a.v == b.v && a.c < b.c);
}

static void fso_heap_push(fso_heap_item *hp, int *n, fso_heap_item x) {
  int i = (*n)++;
  while (i > 0) {
    const int pa = (i - 1) / 2;
    if (!fso_heap_less(x, hp[pa])) break;
    hp[i] = hp[pa];
    i = pa;
  }
  hp[i] = x;
}

static fso_heap_item fso_heap_pop(fso_heap_item *hp, int *n) {
  const fso_heap_item top = hp[0];
  const fso_heap_item last = hp[--(*n)];
  int i = 0;
  for (;;) {
    int ch = 2 * i + 1;
    if (ch >= *n) break;
    if (ch + 1 < *n && fso_heap_less(hp[ch + 1], hp[ch])) ch++;
    if (!fso_heap_less(hp[ch], last)) break;
    hp[i] = hp[ch];
    i = ch;
  }
  if (*n > 0) hp[i] = last;
  return top;
}

void fso_fill_depressions(double *h, int W, int H) {
  const int N = W * H;
  unsigned char *closed = (unsigned char *)calloc((size_t)N, 1);
  fso_heap_item *hp = (fso_heap_item *)malloc(sizeof(fso_heap_item) * (size_t)N);
  int n = 0;
  int nshift[8];
  fso_nshift(W, nshift);
  if (!closed || !hp) abort();
  /* seeds: rings 0 and 1 keep their elevations */
  for (int y = 0; y < H; y++)
    for (int x = 0; x < W; x++)
      if (x < 2 || y < 2 || x >= W - 2 || y >= H - 2) {
        const int c = y * W + x;
        closed[c] = 1;
        fso_heap_item it = {h[c], c};
        fso_heap_push(hp, &n, it);
      }
  while (n > 0) {
    const fso_heap_item it = fso_heap_pop(hp, &n);
    const int c = it.c, x = c % W, y = c / W;
    if (x < 1 || y < 1 || x >= W - 1 || y >= H - 1) continue;   /* ring 0: no interior neighbour */
    const double up = nextafter(h[c], INFINITY);
    for (int k = 0; k < 8; k++) {
      const int nb = c + nshift[k];
      if (closed[nb]) continue;
      closed[nb] = 1;
      h[nb] = fmax(h[nb], up);   /* fmax: a NaN elevation takes the spill value */
      fso_heap_item nx = {h[nb], nb};
      fso_heap_push(hp, &n, nx);
    }
  }
  free(hp);
  free(closed);
}

int fso_write_dem(const char *filename, const double *h, int W, int H) {
  FILE *f = fopen(filename, "w");
  if (!f) return -1;
  fprintf(f, "ncols %d\n", W - 2);
  fprintf(f, "nrows %d\n", H - 2);
  fputs("xllcorner 637500.000\n", f);
  fputs("yllcorner 206000.000\n", f);
  fputs("cellsize 500.000\n", f);
  fputs("NODATA_value -9999\n", f);
  for (int y = 1; y < H - 1; y++) {
    for (int x = 1; x < W - 1; x++) fprintf(f, "%.6g ", h[(size_t)y * W + x]);
    fputc('\n', f);
  }
  return fclose(f) == 0 ? 0 : -1;
}
