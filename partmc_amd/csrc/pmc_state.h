// pmc_state.h -- device-resident layout of the boxes and particles.
//
// All boxes share one structure-of-arrays particle store.  Box b owns the
// particle slots [off_b, off_b + cap_b); species volumes are species-major
// (vol[s * total_cap + p]) so per-species sweeps are coalesced; a
// particle's components are particle-major (csrc[p * MAX_COMP + c]), since
// they are only read and written a whole particle at a time.  Per-box
// tables (bin grid, k bounds, pair tables, (bin, class) lists) live in
// flat arrays with per-box strides fixed at handle creation.
#pragma once

#include "pmc_device.h"

namespace pmc {

struct BoxDesc {
  int64_t off, cap;          // particle slot range
  int64_t pool_off;          // list pool base (two halves of pool_half ints)
  int64_t pool_half;
  int64_t info_off, info_cap;
  int64_t next_id, id_stride, id_offset;
  int64_t n_info;
  // counters accumulated since the last reset
  int64_t n_samp, n_coag, n_warn, n_recheck, n_flip;
  double max_rel_err;
  double max_p_err;
  // counters of the last step (exact mode)
  int64_t step_samp, step_coag;
  int64_t part_steps;        // sum over steps of n_part at step start
  int n_part;
  int n_bin;
  int pool_cur, pool_top;
  int valid_sort, bounds_valid, tables_valid, regridded;
  int xlists_valid;          // exact-mode lists and forward maps are current
  int pkeys_valid;           // parallel-mode keys and live counts are current
  int rekey;                 // parallel mode: recompute every key this step
  int terms_kernel;          // kernel type the FP32 terms were made for (0 = none)
  int64_t n_accel_pairs;     // parallel mode: pairs the per-particle path would take
  unsigned long long rmin_bits, rmax_bits;  // radius extremes (ordered bits)
  int bounds_kernel;         // kernel type the cached bounds were made for
  int grid_req_n_bin;        // explicit grid request (-1 = none)
  double grid_req_min, grid_req_max;
  int error;
  int n_rounds, n_dropped;   // parallel mode: tail claim rounds, dropped samples
  int64_t n_tail_tests;      // parallel mode: tests finished in the tail
  int64_t n_arrivals;        // parallel mode: particles that joined a list within a step
  int64_t n_arr_over;        // ... of them beyond the entries their list reserved
  int64_t n_acc_samp, n_acc_coag;  // samples and coagulations of the per-particle path
  double log_lo, log_hi;     // log(edges[0]), log(edges[n_bin])
  Tape tape;                 // exact mode: uniforms recorded or replayed in call order
  Weights w;
  Env env;
};

// device pointers, passed by value to kernels
struct Arrays {
  int64_t total_cap;
  int n_spec, max_bin, max_keys;
  int tab_classes;           // class dimension of the pair tables
  double d_min, d_max;       // min/max species density (Brownian bounds)
  double *density;           // [n_spec]
  double *vol;               // [n_spec][total_cap]
  double *vtot;              // [total_cap] cached total volume
  int64_t *id;
  int *group, *klass, *nprim, *ncomp;
  int *csrc;                 // [total_cap][MAX_COMP]: a particle's components are contiguous
  double *ctime;             // [total_cap][MAX_COMP]
  double *least, *greatest;
  // exact-mode sorted structure (integer_rmap2 size_class)
  int *fwd_bin, *fwd_class, *fwd_idx;
  int *lstart, *llen, *lcap; // [n_box][max_keys]
  int *pool;
  // grids and tables
  double *edges;             // [n_box][max_bin + 1]
  double *centers, *widths;  // [n_box][max_bin]
  double *kmin, *kmax;       // [n_box][max_bin * max_bin]
  double *kw;                // [n_box][tab_classes^2 * max_bin^2] weighted k_max
  int8_t *cc;                // same layout, destination class
  int *sort_scratch;         // [n_box][SORT_THREADS][max_keys]
  // removal log
  int64_t *info_id, *info_other;
  int *info_action;
};

// component c (0-based) of slot p: a particle's MAX_COMP components sit
// together, so a merge reads and writes a few lines instead of one a
// component
__host__ __device__ inline int64_t comp_at(int64_t p, int c) { return p * MAX_COMP + c; }


constexpr int SORT_THREADS = 256;

// library error codes
constexpr int ERR_ARG = 1;
constexpr int ERR_HIP = 2;
constexpr int ERR_CAPACITY = 3;
constexpr int ERR_UNSUPPORTED = 4;
constexpr int ERR_STATE = 5;

}  // namespace pmc
