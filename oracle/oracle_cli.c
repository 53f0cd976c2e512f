/*
 * oracle_cli.c -- command-line front end of the CPU oracle.
 *
 * TEST INFRASTRUCTURE ONLY (the checker and the "port" CPU baseline).
 *
 *   oracle_cli <W> <H> <nstep> <engine_seed> <preset> <powmode> <parallel> <dem|->
 *
 * Runs construct + run(nstep) (nstep+1 time-steps, RB:366) and prints the
 * reference's "t" timer lines (RB:380-387) plus "m run_seconds".
 */
#include "fastscape_oracle.h"

#include <stdio.h>
#include <stdlib.h>
#include <time.h>
#ifdef _OPENMP
#include <omp.h>
#endif

int main(int argc, char **argv) {
  if (argc != 9) {
    fprintf(stderr,
            "usage: %s <W> <H> <nstep> <engine_seed> <preset> <powmode> "
            "<parallel> <dem|->\n", argv[0]);
    return 2;
  }
  const int W = atoi(argv[1]), H = atoi(argv[2]), nstep = atoi(argv[3]);
  const unsigned seed = (unsigned)strtoul(argv[4], NULL, 10);
  const int preset = atoi(argv[5]), powmode = atoi(argv[6]), par = atoi(argv[7]);
  fso_params p;
  if (fso_preset(preset, &p)) { fprintf(stderr, "bad preset\n"); return 2; }
  fso_model m;
  if (fso_model_init(&m, W, H, &p, seed)) { fprintf(stderr, "init failed\n"); return 1; }
  struct timespec a, b;
  clock_gettime(CLOCK_MONOTONIC, &a);
  int rc = fso_model_run(&m, nstep, powmode, par);
  clock_gettime(CLOCK_MONOTONIC, &b);
  if (rc) { fprintf(stderr, "run failed: %d\n", rc); fso_model_free(&m); return 1; }
  const char *names[8] = {"Step1: Initialize        ", "Step2: DetermineReceivers",
                          "Step3: DetermineDonors   ", "Step4: GenerateOrder     ",
                          "Step5: FlowAcc           ", "Step6: Uplift            ",
                          "Step7: Erosion           ", "Overall                  "};
  for (int i = 0; i < 8; i++)
    printf("t %s = %15llu microseconds\n", names[i], (unsigned long long)m.timer_us[i]);
  printf("m run_seconds = %.9f\n", (double)(b.tv_sec - a.tv_sec) + 1e-9 * (double)(b.tv_nsec - a.tv_nsec));
#ifdef _OPENMP
  printf("m omp_threads = %d\n", omp_get_max_threads());
#endif
  printf("m nlevel = %d\n", m.nlevel);
  if (argv[8][0] != '-' && fso_write_dem(argv[8], m.h, W, H)) {
    fprintf(stderr, "dem write failed\n");
    fso_model_free(&m);
    return 1;
  }
  fso_model_free(&m);
  return 0;
}
