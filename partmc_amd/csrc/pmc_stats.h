// pmc_stats.h -- in-situ size distributions of a box.
//
// bin_grid_histogram_1d (src/bin_grid.F90:225-248) over particle diameters
// with number-concentration weights, as extract_aero_size builds its number
// and mass distributions (src/extract_aero_size.F90:132-138): particle i
// adds w_i / width to the log-spaced diameter bin it falls in, where w_i is
// its number concentration (aero_state_num_concs, src/aero_state.F90:
// 1243-1261), times its mass or volume for the mass and volume
// distributions.
//
// Two passes, with no device-wide sort and no atomics, so a histogram does
// not depend on the order in which work lands:
//   k_dist_partial: DIST_BLOCKS workgroups each take a fixed slice of the
//     box's particles (slice b is [b n / NB, (b+1) n / NB)), in tiles of
//     DIST_TILE.  A tile is sorted by bin in LDS (block radix sort) and
//     summed per bin by a segmented scan; the last element of each run adds
//     the run's sum to the workgroup's LDS histogram, one writer per bin per
//     tile.  The workgroup writes its whole histogram window (zeros
//     included) to partials[b][.].
//   k_dist_reduce: one thread per bin adds partials[0..NB)[bin] in order.
// Bins are handled in windows of at most DIST_WIN, one pair of passes per
// window.  Both passes read the particle count from the box descriptor on
// the device, so the host never waits on the device to enqueue them.
#pragma once

#include <rocprim/block/block_radix_sort.hpp>
#include <rocprim/block/block_scan.hpp>

#include "pmc_exact.h"

namespace pmc {

constexpr int DIST_NUM = 0, DIST_MASS = 1, DIST_VOL = 2;
constexpr int DIST_THREADS = 256;
constexpr int DIST_ITEMS = 4;
constexpr int DIST_TILE = DIST_THREADS * DIST_ITEMS;
constexpr int DIST_BLOCKS = 512;  // workgroups of the partial pass: fixed, so sums do not depend on the device
constexpr int DIST_WIN = 2048;    // bins per window: the LDS histogram and a row of partials

struct DistSeg {  // segmented-sum element: starts a run, running sum
  int head;
  double v;
};
struct DistSegOp {
  __device__ DistSeg operator()(const DistSeg &a, const DistSeg &b) const {
    return b.head ? b : DistSeg{a.head, a.v + b.v};
  }
};

// bin (1-based, 0 / n_bin + 1 outside the grid) and weight / width of
// particle slot p of box B
__device__ inline int dist_item(const BoxDesc *B, const Arrays &A, int64_t p, int quantity,
                                int n_bin, double log_lo, double log_hi, double *val) {
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
  *val = w / width;
  return b;
}

// partials[blockIdx.x][0 .. win): this workgroup's sums over bins
// [win_lo, win_lo + win) (0-based) of its slice of the box's particles
__global__ void __launch_bounds__(DIST_THREADS)
k_dist_partial(const BoxDesc *boxes, Arrays A, int box, int quantity, int n_bin, double log_lo,
               double log_hi, int win_lo, int win, int key_bits, double *partials) {
  using Sort = rocprim::block_radix_sort<int, DIST_THREADS, DIST_ITEMS, double>;
  using Scan = rocprim::block_scan<DistSeg, DIST_THREADS>;
  __shared__ union {
    typename Sort::storage_type sort;
    typename Scan::storage_type scan;
  } st;
  __shared__ int s_key[DIST_TILE];
  __shared__ double s_hist[DIST_WIN];
  const BoxDesc *B = &boxes[box];
  const int64_t n = B->n_part;
  const int64_t lo = (int64_t)blockIdx.x * n / gridDim.x;
  const int64_t hi = (int64_t)(blockIdx.x + 1) * n / gridDim.x;
  const int t = threadIdx.x;
  for (int j = t; j < win; j += DIST_THREADS) s_hist[j] = 0.0;
  __syncthreads();
  for (int64_t base = lo; base < hi; base += DIST_TILE) {
    int key[DIST_ITEMS];
    double val[DIST_ITEMS];
    for (int k = 0; k < DIST_ITEMS; ++k) {
      const int64_t i = base + (int64_t)t * DIST_ITEMS + k;
      key[k] = win;  // padding and particles outside the window sort last
      val[k] = 0.0;
      if (i < hi) {
        double x;
        const int b = dist_item(B, A, B->off + i, quantity, n_bin, log_lo, log_hi, &x) - 1 - win_lo;
        if (b >= 0 && b < win && b + win_lo < n_bin) {
          key[k] = b;
          val[k] = x;
        }
      }
    }
    Sort().sort(key, val, st.sort, 0, key_bits);
    for (int k = 0; k < DIST_ITEMS; ++k) s_key[t * DIST_ITEMS + k] = key[k];
    __syncthreads();  // s_key complete; the sort's storage is free for the scan
    DistSeg seg[DIST_ITEMS];
    for (int k = 0; k < DIST_ITEMS; ++k) {
      const int i = t * DIST_ITEMS + k;
      seg[k] = DistSeg{i == 0 || s_key[i - 1] != key[k], val[k]};
    }
    Scan().inclusive_scan(seg, seg, st.scan, DistSegOp());
    for (int k = 0; k < DIST_ITEMS; ++k) {
      const int i = t * DIST_ITEMS + k;
      const bool tail = (i == DIST_TILE - 1) || s_key[i + 1] != key[k];
      if (tail && key[k] < win) s_hist[key[k]] += seg[k].v;
    }
    __syncthreads();  // before the next tile reuses s_key and the storage
  }
  double *out = partials + (int64_t)blockIdx.x * win;
  for (int j = t; j < win; j += DIST_THREADS) out[j] = s_hist[j];
}

// hist[win_lo + j] = sum over workgroups, in workgroup order, of partials[.][j]
__global__ void k_dist_reduce(const double *partials, int n_part_blocks, int win_lo, int win,
                              int n_bin, double *hist) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= win || win_lo + j >= n_bin) return;
  double s = 0.0;
  for (int b = 0; b < n_part_blocks; ++b) s += partials[(int64_t)b * win + j];
  hist[win_lo + j] = s;
}

}  // namespace pmc
