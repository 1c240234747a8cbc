// pmc_stats.h -- in-situ size distributions of a box.
//
// bin_grid_histogram_1d (src/bin_grid.F90:225-248) over particle diameters
// with number-concentration weights, as extract_aero_size builds its number
// and mass distributions (src/extract_aero_size.F90:132-138): particle i
// adds w_i / width to the log-spaced diameter bin it falls in, where w_i is
// its number concentration (aero_state_num_concs, src/aero_state.F90:
// 1243-1261), times its mass or volume for the mass and volume
// distributions.  The sums are formed after a stable sort by bin, so a
// histogram does not depend on the order atomics would land in.
#pragma once

#include "pmc_exact.h"

namespace pmc {

constexpr int DIST_NUM = 0, DIST_MASS = 1, DIST_VOL = 2;

__global__ void k_dist_values(BoxDesc *boxes, Arrays A, int box, int quantity, int n_bin,
                              double log_lo, double log_hi, int *key, double *val) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const BoxDesc *B = &boxes[box];
  if (i >= B->n_part) return;
  const int64_t p = B->off + i;
  double v = 0.0, m = 0.0;
  for (int s = 0; s < A.n_spec; ++s) {
    const double x = A.vol[(int64_t)s * A.total_cap + p];
    v += x;
    m += x * A.density[s];
  }
  const double r = vol2rad(v);
  const int b = linspace_find(log_lo, log_hi, n_bin + 1, log(2.0 * r));
  double w = nc_at_radius(B->w, A.klass[p], r);
  if (quantity == DIST_MASS) w *= m;
  else if (quantity == DIST_VOL) w *= v;
  const double width = (log_hi - log_lo) / (double)n_bin;
  const bool in = (b >= 1 && b <= n_bin);
  key[i] = in ? b - 1 : n_bin;
  val[i] = in ? w / width : 0.0;
}

__global__ void k_dist_scatter(const int *ukey, const double *agg, const int *n_runs, int n_bin,
                               double *hist) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= *n_runs) return;
  if (ukey[i] < n_bin) hist[ukey[i]] = agg[i];
}

}  // namespace pmc
