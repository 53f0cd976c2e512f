/*
 * fastscape_oracle.h -- CPU oracle for the FastScape time-step.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a plain-C restatement of the
 * reference's level-ordered FastScape step (fastscape_RB.cpp) used as the
 * checker for the HIP implementation.  Only tests/, __graft_entry__.smoke()
 * and bench.py's cpu_baseline leg may load it; the product path never does.
 *
 * Layout and encodings follow the reference exactly:
 *   - grid is row-major, cell c = y*W + x;
 *   - rec[c] is a direction 0..7 or FSO_NO_FLOW (-1)        (RB:50, RB:79-82)
 *   - direction n has offset nshift[n] = {-1,-W-1,-W,-W+1,1,W+1,W,W-1}[n]
 *                                                            (RB:137)
 *   - donor[8*c+k], k < ndon[c], lists the donors of c       (RB:87-88)
 *   - stack[] is the level order, levels[] its level offsets (RB:89, RB:96)
 *
 * Every function must be compiled with -ffp-contract=off: a fused
 * multiply-add changes the last bit of an elevation, and near-tied
 * receivers turn last-bit differences into different drainage.
 */
#ifndef FASTSCAPE_ORACLE_H
#define FASTSCAPE_ORACLE_H

#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define FSO_NO_FLOW (-1)

/* Model constants.  The reference holds these as const class members
 * (RB:58-65); here they are a parameter block. */
typedef struct fso_params {
  double keq;        /* K, stream-power coefficient            RB:58 */
  double neq;        /* n, slope exponent                      RB:59 */
  double meq;        /* m, area exponent                       RB:60 */
  double ueq;        /* U, uplift rate                         RB:61 */
  double dt;         /* time step                              RB:62 */
  double dx;         /* cell spacing; dr = {dx, dx*SQRT2}      RB:63 */
  double cell_area;  /* area of one cell                       RB:65 */
  double tol;        /* Newton tolerance                       RB:64 */
} fso_params;

/* How pow(x, n) and pow(x, n-1) are evaluated inside the Newton loop.
 *   FSO_POW_LIBM  : libm pow, exactly as the reference (RB:322).
 *   FSO_POW_EXACT : integer n evaluated by IEEE products (x^1 = x,
 *                   x^2 = x*x); this is what the GPU computes. */
enum { FSO_POW_LIBM = 0, FSO_POW_EXACT = 1 };

/* Presets: 0 = the reference's constants, 1 = the benchmark constants
 * (K=1e-5, n=1, m=0.4, U=2e-3, dt=1000, dx=100, area=1e4). */
int fso_preset(int which, fso_params *out);

/* ---- random initialisation (random.cpp:19-49, RB:111-130) ---- */
typedef struct fso_mt19937 { uint32_t mt[624]; int mti; } fso_mt19937;
void     fso_mt_seed(fso_mt19937 *e, uint32_t seed);
uint32_t fso_mt_next(fso_mt19937 *e);
double   fso_canonical(fso_mt19937 *e);
void     fso_init_terrain(double *h, int W, int H, uint32_t engine_seed);

/* ---- the six stages (RB:160-328) ---- */
void fso_receivers(const double *h, int *rec, int W, int H, const double dr[8]);
void fso_donors(const int *rec, int *ndon, int *donor, int W, int H);
int  fso_order(const int *rec, const int *ndon, const int *donor, int *stack,
               int *levels, int levels_cap, int W, int H, int *nlevel);
void fso_flowacc(const int *stack, const int *rec, double *accum, int W, int H,
                 double cell_area);
void fso_uplift(double *h, int W, int H, double ueq, double dt);
int  fso_erode(const int *stack, const int *rec, const double *accum, double *h,
               int W, int H, const fso_params *p, int powmode);
int  fso_erode_levels_omp(const int *stack, const int *levels, int nlevel,
                          const int *rec, const double *accum, double *h,
                          int W, int H, const fso_params *p, int powmode);

/* ---- whole model (RB:135-151, RB:341-398) ---- */
typedef struct fso_model {
  int W, H, size;
  fso_params p;
  double dr[8];
  int nshift[8];
  double *h, *accum;
  int *rec, *ndon, *donor, *stack, *levels;
  int levels_cap, nlevel;
  uint64_t timer_us[8];   /* Step1..Step7, Overall (RB:100-107) */
} fso_model;

int  fso_model_init(fso_model *m, int W, int H, const fso_params *p,
                    uint32_t engine_seed);
void fso_model_free(fso_model *m);
/* One time-step (RB:367-372).  parallel_erosion selects the level-parallel
 * OpenMP erosion of fastscape_RB+P.cpp:313-351 instead of RB's serial one. */
int  fso_model_step(fso_model *m, int powmode, int parallel_erosion);
/* run(nstep): nstep+1 time-steps (RB:366), timers accumulated. */
int  fso_model_run(fso_model *m, int nstep, int powmode, int parallel_erosion);

/* ---- depression filling (an extension: the reference has none) ----
 * Priority-Flood+epsilon (Barnes, Lehman & Mulla 2014, "Priority-Flood: An
 * optimal depression-filling and watershed-labeling algorithm for digital
 * elevation models", Computers & Geosciences 62:117-127, Algorithm 3 with a
 * single priority queue): rings 0 and 1 are the seeds and keep their
 * elevations; cells leave a min-heap in increasing order of their filled
 * value, and each unvisited D8 neighbour n of a popped cell c receives
 * max(h[n], nextafter(F[c], +inf)).  The reference's README claims
 * depression handling (README.md:35-37) but its code has none. */
void fso_fill_depressions(double *h, int W, int H);

/* ArcGIS ASCII grid writer, byte-compatible with PrintDEM (RB:20-40). */
int  fso_write_dem(const char *filename, const double *h, int W, int H);

#ifdef __cplusplus
}
#endif
#endif
