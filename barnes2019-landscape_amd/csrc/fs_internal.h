// fs_internal.h -- declarations shared by the host and device parts of the
// FastScape library.  Not installed; the public boundary is fastscape.h.
#ifndef FS_INTERNAL_H
#define FS_INTERNAL_H

#include <cstddef>
#include <cstdint>
#include <string>

#include "fastscape.h"

namespace fs {

void set_last_error(const std::string &msg);
const char *last_error();
// Grid dimensions and parameter block accepted by the model (status code).
int validate_params(int W, int H, const fs_params *p);
void generate_terrain(double *h, int W, int H, uint32_t engine_seed);
void fill_fact_table(double *tab, size_t kmax, double cell_area, double meq, double kdt,
                     double lenpow_card, double lenpow_diag, unsigned max_threads = 16);
int write_dem(const char *filename, const double *h, int W, int H);
const char *status_string(int s);

// The reference's SQRT2 literal (RB:51) rounds to this double.
constexpr double kSqrt2 = 1.414213562373095048801688724209698078569671875376948;

}  // namespace fs

#endif
