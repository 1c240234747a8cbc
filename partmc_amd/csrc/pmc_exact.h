// pmc_exact.h -- per-box preparation (sort / regrid), majorant and
// pair tables, and the exact sequential-order sweep.
//
// The exact sweep replays mc_coag's loop order (src/coagulation.F90:74-94)
// inside one wavefront per box.  Tests of a bin pair are evaluated 64 at a
// time speculatively -- the particle state cannot change before the first
// accepted test of a batch -- and the first accepted test is merged by
// lane 0 with the reference's swap-last bookkeeping, after which the
// speculation restarts.  With the same per-site Philox streams as the
// oracle, the resulting integer state is bit-identical.
#pragma once

#include "pmc_state.h"

namespace pmc {

// ------------------------------------------------------------ helpers
// Species are read SPEC_CHUNK at a time, the loads of a chunk issued
// before any is used (index clamped to the last species): a load in the
// body of a loop with a runtime count costs one memory round trip each.
// Sums run in species order, as one at a time would.
constexpr int SPEC_CHUNK = 4;

__device__ inline double particle_volume(const Arrays &A, int64_t p) {
  const int ns = A.n_spec;
  double v = 0.0;
  for (int s0 = 0; s0 < ns; s0 += SPEC_CHUNK) {
    double x[SPEC_CHUNK];
#pragma unroll
    for (int u = 0; u < SPEC_CHUNK; ++u)
      x[u] = A.vol[(int64_t)(s0 + u < ns ? s0 + u : ns - 1) * A.total_cap + p];
#pragma unroll
    for (int u = 0; u < SPEC_CHUNK; ++u)
      if (s0 + u < ns) v += x[u];
  }
  return v;
}

__device__ inline double particle_density(const Arrays &A, int64_t p) {
  const int ns = A.n_spec;
  double m = 0.0, v = 0.0;
  for (int s0 = 0; s0 < ns; s0 += SPEC_CHUNK) {
    double x[SPEC_CHUNK], rho[SPEC_CHUNK];
#pragma unroll
    for (int u = 0; u < SPEC_CHUNK; ++u) {
      const int s = s0 + u < ns ? s0 + u : ns - 1;
      x[u] = A.vol[(int64_t)s * A.total_cap + p];
      rho[u] = A.density[s];
    }
#pragma unroll
    for (int u = 0; u < SPEC_CHUNK; ++u)
      if (s0 + u < ns) m += x[u] * rho[u];
  }
  v = particle_volume(A, p);
  return m / v;
}

// kernel (src/coag_kernel.F90:83-125) times coag_num_conc_factor
// (num_conc_weighted_kernel, src/coag_kernel.F90:175-209), FP64 from the
// raw species volumes.
__device__ inline int weighted_kernel64(const Arrays &A, const Weights &w,
                                        const Env &env, int type, int64_t p1,
                                        int64_t p2, int c1, int c2, int cc,
                                        double *k_out) {
  double v1 = particle_volume(A, p1);
  double v2 = particle_volume(A, p2);
  double k = 0.0;
  int err = 0;
  switch (type) {
    case KERNEL_ADDITIVE: k = env.beta * (v1 + v2); break;
    case KERNEL_CONSTANT: k = C_BETA_0; break;
    case KERNEL_ZERO: k = 0.0; break;
    case KERNEL_BROWN: {
      double d1 = particle_density(A, p1);
      double d2 = particle_density(A, p2);
      k = kernel_brown_helper(v1, d1, v2, d2, env.temp, env.pressure);
      break;
    }
    case KERNEL_SEDI: err = kernel_sedi_helper(v1, v2, &k); break;
    default: err = 200724934;
  }
  double r1 = vol2rad(v1);
  double r2 = vol2rad(v2);
  *k_out = k * coag_num_conc_factor(w, r1, r2, c1, c2, cc);
  return err;
}

// bin_grid_make (log), src/bin_grid.F90:83-138 via logspace
// src/util.F90:497-520 (one thread)
__device__ inline void make_log_grid(int n_bin, double rmin, double rmax,
                                     double *edges, double *centers,
                                     double *widths) {
  if (n_bin <= 0) return;
  double lmin = log(rmin), lmax = log(rmax);
  int n = n_bin + 1;
  for (int i = 2; i <= n - 1; ++i) {
    double a = (double)(i - 1) / (double)(n - 1);
    edges[i - 1] = exp((1.0 - a) * lmin + a * lmax);
  }
  edges[0] = rmin;
  edges[n - 1] = rmax;
  double c1 = exp(interp_linear_disc(lmin, lmax, 2 * n_bin + 1, 2));
  double c2 = exp(interp_linear_disc(lmin, lmax, 2 * n_bin + 1, 2 * n_bin));
  double lc1 = log(c1), lc2 = log(c2);
  for (int i = 2; i <= n_bin - 1; ++i) {
    double a = (double)(i - 1) / (double)(n_bin - 1);
    centers[i - 1] = exp((1.0 - a) * lc1 + a * lc2);
  }
  centers[0] = c1;
  centers[n_bin - 1] = c2;
  for (int i = 0; i < n_bin; ++i) widths[i] = (lmax - lmin) / (double)n_bin;
}

// whole-record copy of particle slot `from` into slot `to`.  Fields are
// read in chunks before they are written: a store between two loads of
// the same arrays makes the second wait for the first, one memory round
// trip per field.
__device__ inline void copy_particle(const Arrays &A, int64_t from, int64_t to) {
  const int64_t tc = A.total_cap;
  for (int s0 = 0; s0 < A.n_spec; s0 += 8) {
    double x[8];
#pragma unroll
    for (int j = 0; j < 8; ++j)
      if (s0 + j < A.n_spec) x[j] = A.vol[(int64_t)(s0 + j) * tc + from];
#pragma unroll
    for (int j = 0; j < 8; ++j)
      if (s0 + j < A.n_spec) A.vol[(int64_t)(s0 + j) * tc + to] = x[j];
  }
  const double vtot = A.vtot[from];
  const int64_t id = A.id[from];
  const int group = A.group[from], klass = A.klass[from], nprim = A.nprim[from];
  const int nco = A.ncomp[from];
  const double least = A.least[from], greatest = A.greatest[from];
  const int fb = A.fwd_bin[from], fc = A.fwd_class[from], fi = A.fwd_idx[from];
  A.vtot[to] = vtot;
  A.id[to] = id;
  A.group[to] = group;
  A.klass[to] = klass;
  A.nprim[to] = nprim;
  A.ncomp[to] = nco;
  A.least[to] = least;
  A.greatest[to] = greatest;
  A.fwd_bin[to] = fb;
  A.fwd_class[to] = fc;
  A.fwd_idx[to] = fi;
  for (int c0 = 0; c0 < nco; c0 += 10) {  // entries past n_comp are never read
    int cs[10];
    double ct[10];
#pragma unroll
    for (int j = 0; j < 10; ++j) {
      if (c0 + j < nco) {
        cs[j] = A.csrc[comp_at(from, c0 + j)];
        ct[j] = A.ctime[comp_at(from, c0 + j)];
      }
    }
#pragma unroll
    for (int j = 0; j < 10; ++j) {
      if (c0 + j < nco) {
        A.csrc[comp_at(to, c0 + j)] = cs[j];
        A.ctime[comp_at(to, c0 + j)] = ct[j];
      }
    }
  }
}

// -------------------------------------------------------------- prepare
// One block of SORT_THREADS per box: aero_sorted_remake_if_needed
// (src/aero_sorted.F90:214-373), explicit-grid sorts (:254-262 with
// aero_sorted_discard_outside_grid :153-180) and aero_sorted_sort_particles
// (:185-209).  Lists are rebuilt into the other pool half every call so
// every list starts the sweep with slack.
__global__ void __launch_bounds__(SORT_THREADS)
k_prepare(BoxDesc *boxes, Arrays A, int only_box) {
  const int box = only_box >= 0 ? only_box : (int)blockIdx.x;
  const int t = threadIdx.x;
  BoxDesc *B = &boxes[box];
  __shared__ int s_need_sort, s_err, s_n;
  __shared__ double s_min[SORT_THREADS], s_max[SORT_THREADS];
  if (B->error) return;
  const int64_t base = B->off;
  double *edges = A.edges + (int64_t)box * (A.max_bin + 1);
  double *centers = A.centers + (int64_t)box * A.max_bin;
  double *widths = A.widths + (int64_t)box * A.max_bin;
  int *lstart = A.lstart + (int64_t)box * A.max_keys;
  int *llen = A.llen + (int64_t)box * A.max_keys;
  int *lcap = A.lcap + (int64_t)box * A.max_keys;
  const int n_class = B->w.n_class;
  if (t == 0) {
    s_need_sort = 0;
    s_err = 0;
    s_n = B->n_part;
  }
  __syncthreads();

  // refresh the cached total volumes (uploads and merges keep them too)
  for (int i = t; i < s_n; i += SORT_THREADS) A.vtot[base + i] = particle_volume(A, base + i);
  __syncthreads();

  if (B->grid_req_n_bin >= 0) {
    if (t == 0) {
      int nb = B->grid_req_n_bin;
      if (nb > A.max_bin) {
        s_err = ERR_CAPACITY;
      } else {
        make_log_grid(nb, B->grid_req_min, B->grid_req_max, edges, centers, widths);
        B->n_bin = nb;
        B->log_lo = log(edges[0]);
        B->log_hi = log(edges[nb]);
        B->bounds_valid = 0;
        B->tables_valid = 0;
        B->regridded = 1;
      }
    }
    __syncthreads();
    if (s_err) {
      if (t == 0) B->error = s_err;
      return;
    }
    const int nb = B->n_bin;
    int outside = 0;
    for (int i = t; i < s_n; i += SORT_THREADS) {
      const int b = bin_find_log(nb, B->log_lo, B->log_hi, vol2rad(A.vtot[base + i]));
      A.fwd_bin[base + i] = b;
      outside += (b < 1 || b > nb) ? 1 : 0;
    }
    __shared__ int s_outside;
    if (t == 0) s_outside = 0;
    __syncthreads();
    if (outside) atomicAdd(&s_outside, outside);
    __syncthreads();
    if (t == 0) {
      int n = s_n;
      // aero_sorted_discard_outside_grid's swap-last pass, needed only when
      // some particle lies outside the grid
      for (int i = s_outside ? n - 1 : -1; i >= 0; --i) {
        int b = A.fwd_bin[base + i];
        if (b < 1 || b > nb) {
          if (i < n - 1) copy_particle(A, base + n - 1, base + i);
          n--;
        }
      }
      s_n = n;
      B->n_part = n;
      B->grid_req_n_bin = -1;
      s_need_sort = 1;
    }
    __syncthreads();
  } else {
    // r_min / r_max; lists left stale by the parallel mode count as unsorted
    const bool use_lists = B->valid_sort && B->xlists_valid;
    double rmin = 0.0, rmax = 0.0;
    if (!use_lists) {
      double lo = 1e300, hi = -1e300;
      for (int i = t; i < s_n; i += SORT_THREADS) {
        double r = vol2rad(A.vtot[base + i]);
        lo = fmin(lo, r);
        hi = fmax(hi, r);
      }
      s_min[t] = lo;
      s_max[t] = hi;
      __syncthreads();
      for (int w = SORT_THREADS / 2; w > 0; w >>= 1) {
        if (t < w) {
          s_min[t] = fmin(s_min[t], s_min[t + w]);
          s_max[t] = fmax(s_max[t], s_max[t + w]);
        }
        __syncthreads();
      }
      if (s_n > 0) {
        rmin = s_min[0];
        rmax = s_max[0];
      }
    }
    if (t == 0) {
      B->regridded = 0;
      const int nb = B->n_bin;
      if (!use_lists && B->valid_sort && nb > 0) {
        // a valid sort without current lists: the occupied bins are those
        // of the radius extremes
        if (s_n == 0) {
          rmin = edges[nb];
          rmax = edges[0];
        } else {
          const int b_lo = bin_find_log(nb, B->log_lo, B->log_hi, rmin);
          const int b_hi = bin_find_log(nb, B->log_lo, B->log_hi, rmax);
          rmin = edges[max(b_lo, 1) - 1];
          rmax = edges[min(b_hi, nb)];
        }
      }
      if (use_lists) {
        int i_min = 0, i_max = 0;
        for (int b = 1; b <= nb; ++b) {
          int tot = 0;
          for (int c = 1; c <= n_class; ++c) tot += llen[(b - 1) + nb * (c - 1)];
          if (tot > 0) {
            if (i_min == 0) i_min = b;
            i_max = b;
          }
        }
        if (i_min == 0) {
          if (nb > 0) {
            rmin = edges[nb];
            rmax = edges[0];
          }
        } else {
          rmin = edges[i_min - 1];
          rmax = edges[i_max];
        }
      }
      bool done = false;
      if (rmax == 0.0) {
        if (!use_lists) {
          B->n_bin = 0;
          B->bounds_valid = 0;
          B->tables_valid = 0;
        }
        done = true;
      }
      if (!done) {
        bool need = false;
        if (nb < 1) {
          need = true;
        } else if ((rmin / edges[0] < 3.0) || (edges[nb] / rmax < 3.0)) {
          need = true;
        }
        if (need) {
          double gmin = rmin / 10.0, gmax = rmax * 10.0;
          double nbf = ceil((log10(gmax) - log10(gmin)) * 10.0);
          if (nbf < 0.0) {
            s_err = 538534122;
          } else if (nbf > (double)A.max_bin) {
            s_err = ERR_CAPACITY;
          } else {
            int nn = (int)nbf;
            make_log_grid(nn, gmin, gmax, edges, centers, widths);
            B->n_bin = nn;
            B->log_lo = log(edges[0]);
            B->log_hi = log(edges[nn]);
            B->bounds_valid = 0;
            B->tables_valid = 0;
            B->regridded = 1;
            s_need_sort = 1;
          }
        } else if (!use_lists) {
          s_need_sort = 1;
        }
      }
    }
    __syncthreads();
    if (s_err) {
      if (t == 0) B->error = s_err;
      return;
    }
  }

  const int nb = B->n_bin;
  const int K = nb * n_class;
  const int n = s_n;
  int *scr = A.sort_scratch + (int64_t)box * SORT_THREADS * A.max_keys + (int64_t)t * A.max_keys;
  if (s_need_sort) {
    // stable counting sort: thread t owns the contiguous segment
    // [t*S, (t+1)*S), so list order = particle order as appends produce.
    const int S = (n + SORT_THREADS - 1) / SORT_THREADS;
    const int i0 = t * S, i1 = min(n, i0 + S);
    for (int k = 0; k < K; ++k) scr[k] = 0;
    for (int i = i0; i < i1; ++i) {
      int b = bin_find_log(nb, B->log_lo, B->log_hi, vol2rad(A.vtot[base + i]));
      if (b < 1 || b > nb) {
        atomicExch(&s_err, 708651144);
        b = 1;
      }
      int c = A.klass[base + i];
      A.fwd_bin[base + i] = b;
      A.fwd_class[base + i] = c;
      scr[(b - 1) + nb * (c - 1)]++;
    }
    __syncthreads();
    int *scr0 = A.sort_scratch + (int64_t)box * SORT_THREADS * A.max_keys;
    for (int k = t; k < K; k += SORT_THREADS) {
      int run = 0;
      for (int tt = 0; tt < SORT_THREADS; ++tt) {
        int c = scr0[(int64_t)tt * A.max_keys + k];
        scr0[(int64_t)tt * A.max_keys + k] = run;
        run += c;
      }
      llen[k] = run;
    }
    __syncthreads();
  }
  if (t == 0) {
    int h = 1 - B->pool_cur;
    int64_t pbase = B->pool_off + (int64_t)h * B->pool_half;
    int64_t pos = 0;
    for (int k = 0; k < K; ++k) {
      lstart[k] = (int)(pbase + pos);
      lcap[k] = llen[k] + llen[k] / 4 + 8;
      pos += lcap[k];
    }
    if (pos > B->pool_half) s_err = ERR_CAPACITY;
    B->pool_top = (int)pos;
    B->pool_cur = h;
  }
  __syncthreads();
  if (s_err) {
    if (t == 0) B->error = s_err;
    return;
  }
  if (s_need_sort) {
    const int S = (n + SORT_THREADS - 1) / SORT_THREADS;
    const int i0 = t * S, i1 = min(n, i0 + S);
    for (int i = i0; i < i1; ++i) {
      int key = (A.fwd_bin[base + i] - 1) + nb * (A.fwd_class[base + i] - 1);
      int p = scr[key]++;
      A.pool[lstart[key] + p] = i + 1;
      A.fwd_idx[base + i] = p + 1;
    }
  } else {
    for (int i = t; i < n; i += SORT_THREADS) {
      int key = (A.fwd_bin[base + i] - 1) + nb * (A.fwd_class[base + i] - 1);
      A.pool[lstart[key] + A.fwd_idx[base + i] - 1] = i + 1;
    }
  }
  __syncthreads();
  if (t == 0) {
    B->valid_sort = 1;
    B->xlists_valid = 1;
    B->pkeys_valid = 0;
  }
}

// ------------------------------------------------------------- k bounds
// est_k_minmax_binned_unweighted / est_k_minmax_for_bin_unweighted
// (src/coag_kernel.F90:253-341): one thread per (box, b1, b2).  Bounds
// are recomputed only after a new grid (or kernel type), as in the
// reference where env changes do not invalidate them (:64-70).
__global__ void k_kbounds(BoxDesc *boxes, Arrays A, int type) {
  const int box = blockIdx.y;
  BoxDesc *B = &boxes[box];
  if (B->error || (B->bounds_valid && B->bounds_kernel == type)) return;
  const int nb = B->n_bin;
  const int idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= nb * nb) return;
  const int b1 = idx % nb + 1, b2 = idx / nb + 1;
  const double *edges = A.edges + (int64_t)box * (A.max_bin + 1);
  double v1_low = rad2vol(edges[b1 - 1]);
  double v1_high = rad2vol(edges[b1]);
  double v2_low = rad2vol(edges[b2 - 1]);
  double v2_high = rad2vol(edges[b2]);
  double kmin = 0.0, kmax = 0.0;
  Env env = B->env;
  for (int i = 1; i <= 3; ++i)
    for (int j = 1; j <= 3; ++j) {
      double v1 = interp_linear_disc(v1_low, v1_high, 3, i);
      double v2 = interp_linear_disc(v2_low, v2_high, 3, j);
      double a, b;
      int err = kernel_minmax(type, A.d_min, A.d_max, v1, v2, env, &a, &b);
      if (err) atomicExch(&B->error, err);
      if (i == 1 && j == 1) {
        kmin = a;
        kmax = b;
      } else {
        kmin = fmin(kmin, a);
        kmax = fmax(kmax, b);
      }
    }
  int64_t o = (int64_t)box * A.max_bin * A.max_bin + idx;
  A.kmin[o] = kmin;
  A.kmax[o] = kmax * 2.0;
}

__global__ void k_mark_bounds(BoxDesc *boxes, int n_box, int type) {
  int box = blockIdx.x * blockDim.x + threadIdx.x;
  if (box >= n_box) return;
  BoxDesc *B = &boxes[box];
  if (!B->bounds_valid || B->bounds_kernel != type) {
    B->bounds_valid = 1;
    B->bounds_kernel = type;
    B->tables_valid = 0;
  }
}

// coag_dest_class + max_coag_num_conc_factor per (c1, c2, b1, b2), times
// the unweighted k_max (mc_coag_for_bin, src/coagulation.F90:136-140).
// The reference evaluates these for every visited pair and step; they
// depend only on the grid and the weights, so they are tabled.
__global__ void k_pair_tables(BoxDesc *boxes, Arrays A) {
  const int box = blockIdx.y;
  BoxDesc *B = &boxes[box];
  if (B->error || B->tables_valid) return;
  const int nb = B->n_bin, nc = B->w.n_class;
  const int nn = nb * nb;
  const int idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= nc * nc * nn) return;
  const int pair = idx % nn, cls = idx / nn;
  const int b2 = pair % nb + 1, b1 = pair / nb + 1;
  const int c2 = cls % nc + 1, c1 = cls / nc + 1;
  const double *edges = A.edges + (int64_t)box * (A.max_bin + 1);
  const double *centers = A.centers + (int64_t)box * A.max_bin;
  Weights w = B->w;
  int cc = coag_dest_class(w, centers, b1, c1, c2);
  double f_max = max_coag_num_conc_factor(w, edges, b1, b2, c1, c2, cc);
  double k_max = A.kmax[(int64_t)box * A.max_bin * A.max_bin + (b1 - 1) + nb * (b2 - 1)];
  int64_t o = (int64_t)box * A.tab_classes * A.tab_classes * A.max_bin * A.max_bin + idx;
  A.kw[o] = k_max * f_max;
  A.cc[o] = (int8_t)cc;
}

__global__ void k_mark_tables(BoxDesc *boxes, int n_box) {
  int box = blockIdx.x * blockDim.x + threadIdx.x;
  if (box >= n_box) return;
  boxes[box].tables_valid = 1;
}

__device__ inline int64_t table_index(const Arrays &A, int box, int nb, int nc,
                                      int b1, int b2, int c1, int c2) {
  return (int64_t)box * A.tab_classes * A.tab_classes * A.max_bin * A.max_bin +
         (int64_t)((c1 - 1) * nc + (c2 - 1)) * nb * nb + (int64_t)(b1 - 1) * nb +
         (b2 - 1);
}

// --------------------------------------------------------- exact sweep
// A detached particle record, as aero_particle_t on the path.
struct Rec {
  double vol[MAX_SPEC];
  int64_t id;
  int group, klass, nprim, ncomp;
  int csrc[MAX_COMP];
  double ctime[MAX_COMP];
  double least, greatest;
};

struct ExactBox {
  const Arrays *A;
  BoxDesc *B;
  Tape *tape = nullptr;  // the box's tape when one records or replays, else null
  int box;
  int64_t base;
  int nb, nc;
  int *lstart, *llen, *lcap;
  uint64_t seed;
  uint32_t step;
  const double *edges;

  __device__ int key(int bin, int cls) const { return (bin - 1) + nb * (cls - 1); }
  __device__ int n_entry(int bin, int cls) const { return llen[key(bin, cls)]; }
  __device__ int entry(int bin, int cls, int i) const {  // 1-based domain
    return A->pool[lstart[key(bin, cls)] + i - 1];
  }

  __device__ void load(int dom0, Rec &r) const {
    int64_t p = base + dom0;
    for (int s = 0; s < A->n_spec; ++s) r.vol[s] = A->vol[(int64_t)s * A->total_cap + p];
    r.id = A->id[p];
    r.group = A->group[p];
    r.klass = A->klass[p];
    r.nprim = A->nprim[p];
    r.ncomp = A->ncomp[p];
    for (int c = 0; c < r.ncomp; ++c) {
      r.csrc[c] = A->csrc[comp_at(p, c)];
      r.ctime[c] = A->ctime[comp_at(p, c)];
    }
    r.least = A->least[p];
    r.greatest = A->greatest[p];
  }
  __device__ void store(int dom0, const Rec &r) const {
    int64_t p = base + dom0;
    double v = 0.0;
    for (int s = 0; s < A->n_spec; ++s) {
      A->vol[(int64_t)s * A->total_cap + p] = r.vol[s];
      v += r.vol[s];
    }
    A->vtot[p] = v;
    A->id[p] = r.id;
    A->group[p] = r.group;
    A->klass[p] = r.klass;
    A->nprim[p] = r.nprim;
    A->ncomp[p] = r.ncomp;
    for (int c = 0; c < r.ncomp; ++c) {
      A->csrc[comp_at(p, c)] = r.csrc[c];
      A->ctime[comp_at(p, c)] = r.ctime[c];
    }
    A->least[p] = r.least;
    A->greatest[p] = r.greatest;
  }
  __device__ double rec_volume(const Rec &r) const {
    double v = 0.0;
    for (int s = 0; s < A->n_spec; ++s) v += r.vol[s];
    return v;
  }
  __device__ double rec_radius(const Rec &r) const { return vol2rad(rec_volume(r)); }

  __device__ void add_info(int64_t id, int action, int64_t other) {
    int64_t k = B->n_info;
    if (k < B->info_cap) {
      A->info_id[B->info_off + k] = id;
      A->info_action[B->info_off + k] = action;
      A->info_other[B->info_off + k] = other;
    }
    B->n_info = k + 1;
  }

  // integer_varray_append with relocation when a list is full
  __device__ int list_append(int k, int val) {
    if (llen[k] == lcap[k]) {
      int ncap = lcap[k] * 2 + 8;
      if ((int64_t)B->pool_top + ncap > B->pool_half) return ERR_CAPACITY;
      int nstart = (int)(B->pool_off + (int64_t)B->pool_cur * B->pool_half + B->pool_top);
      for (int i = 0; i < llen[k]; ++i) A->pool[nstart + i] = A->pool[lstart[k] + i];
      lstart[k] = nstart;
      lcap[k] = ncap;
      B->pool_top += ncap;
    }
    A->pool[lstart[k] + llen[k]] = val;
    llen[k] += 1;
    return 0;
  }
  // integer_varray_remove_entry (1-based), src/integer_varray.F90:151-171
  __device__ void list_remove(int k, int index) {
    int n = llen[k];
    if (index < n) A->pool[lstart[k] + index - 1] = A->pool[lstart[k] + n - 1];
    A->pool[lstart[k] + n - 1] = 0;
    llen[k] = n - 1;
  }

  // aero_particle_array_remove_particle + integer_rmap2_remove
  // (src/aero_particle_array.F90:182-199, src/integer_rmap2.F90:197-246);
  // dom is 1-based.
  __device__ void remove(int dom) {
    int64_t pd = base + dom - 1;
    int r1_old = A->fwd_bin[pd], r2_old = A->fwd_class[pd], idx_old = A->fwd_idx[pd];
    int n_dom = B->n_part;
    int shifted = n_dom;
    int64_t ps = base + shifted - 1;
    if (shifted != dom) {
      int kf = key(A->fwd_bin[ps], A->fwd_class[ps]);
      A->pool[lstart[kf] + A->fwd_idx[ps] - 1] = dom;
      copy_particle(*A, ps, pd);  // particle record and forward maps
    }
    B->n_part = n_dom - 1;
    int kl = key(r1_old, r2_old);
    int idx_fix = llen[kl];
    int dom_fix = A->pool[lstart[kl] + idx_fix - 1];
    list_remove(kl, idx_old);
    if (idx_fix != idx_old) A->fwd_idx[base + dom_fix - 1] = idx_old;
  }

  // aero_sorted_add_particle (src/aero_sorted.F90:397-449) for a valid sort
  __device__ int add(const Rec &r) {
    if (B->n_part >= B->cap) return ERR_CAPACITY;
    int bin = bin_find_log(nb, B->log_lo, B->log_hi, rec_radius(r));
    int dom0 = B->n_part;
    store(dom0, r);
    B->n_part = dom0 + 1;
    if (bin < 1 || bin > nb) return 134572570;
    int k = key(bin, r.klass);
    int err = list_append(k, dom0 + 1);
    if (err) return err;
    A->fwd_bin[base + dom0] = bin;
    A->fwd_class[base + dom0] = r.klass;
    A->fwd_idx[base + dom0] = llen[k];
    return 0;
  }

  // integer_rmap2_change with the reference's range-1-only early return
  // (src/integer_rmap2.F90:136-191)
  __device__ int change(int dom, int bin, int cls) {
    int64_t pd = base + dom - 1;
    int r1_old = A->fwd_bin[pd], r2_old = A->fwd_class[pd];
    if (r1_old == bin) return 0;
    int idx_old = A->fwd_idx[pd];
    int kl = key(r1_old, r2_old);
    list_remove(kl, idx_old);
    if (idx_old <= llen[kl]) {
      int shifted = A->pool[lstart[kl] + idx_old - 1];
      A->fwd_idx[base + shifted - 1] = idx_old;
    }
    int kn = key(bin, cls);
    int err = list_append(kn, dom);
    if (err) return err;
    A->fwd_bin[pd] = bin;
    A->fwd_class[pd] = cls;
    A->fwd_idx[pd] = llen[kn];
    return 0;
  }

  __device__ int64_t new_id() {
    int64_t id = (B->next_id - 1) * B->id_stride + B->id_offset + 1;
    B->next_id += 1;
    return id;
  }

  __device__ int rand_group(Stream &s, const Weights &w, int cls, double r) const {
    double pdf[MAX_GROUP];
    for (int g = 0; g < w.n_group; ++g) {
      int idx = g + w.n_group * (cls - 1);
      pdf[g] = 1.0 / weight_nc(w.type[idx], w.mag[idx], w.expo[idx], r);
    }
    return sample_cts_pdf(s, w.n_group, pdf);
  }
};

// aero_particle_coagulate, src/aero_particle.F90:917-1033
__device__ inline int rec_coagulate(const ExactBox &E, Stream &s, const Rec &p1,
                                    const Rec &p2, Rec &out) {
  Rec o;
  for (int k = 0; k < E.A->n_spec; ++k) o.vol[k] = p1.vol[k] + p2.vol[k];
  o.group = 0;
  o.klass = 0;
  o.id = 0;
  int n1 = p1.ncomp, n2 = p2.ncomp;
  if (p1.nprim < n1) return 465791384;
  if (p2.nprim < n2) return 465791385;
  for (int c = 0; c < MAX_COMP; ++c) {
    o.csrc[c] = 0;
    o.ctime[c] = 0.0;
  }
  if (n1 + n2 > MAX_COMP) {
    int n1_new = prob_round(s, (double)p1.nprim / (double)(p1.nprim + p2.nprim) *
                                   (double)MAX_COMP);
    int n2_new = MAX_COMP - n1_new;
    int sample[MAX_COMP];
    sample_without_replacement(s, n1_new, n1, sample);
    for (int i = 0; i < n1_new; ++i) {
      o.csrc[i] = p1.csrc[sample[i] - 1];
      o.ctime[i] = p1.ctime[sample[i] - 1];
    }
    sample_without_replacement(s, n2_new, n2, sample);
    for (int i = 0; i < n2_new; ++i) {
      o.csrc[i + n1_new] = p2.csrc[sample[i] - 1];
      o.ctime[i + n1_new] = p2.ctime[sample[i] - 1];
    }
    o.ncomp = MAX_COMP;
  } else {
    for (int i = 0; i < n1; ++i) {
      o.csrc[i] = p1.csrc[i];
      o.ctime[i] = p1.ctime[i];
    }
    for (int i = 0; i < n2; ++i) {
      o.csrc[n1 + i] = p2.csrc[i];
      o.ctime[n1 + i] = p2.ctime[i];
    }
    o.ncomp = n1 + n2;
  }
  o.least = fmin(p1.least, p2.least);
  o.greatest = fmax(p1.greatest, p2.greatest);
  o.nprim = p1.nprim + p2.nprim;
  out = o;
  return 0;
}

// coagulate_weighting + coagulate, src/coagulation.F90:810-985
// (dom1, dom2 1-based)
__device__ inline int exact_coagulate(ExactBox &E, const Weights &w, Stream &s,
                                      int dom1, int dom2, int c1, int c2, int cc) {
  Rec pt1, pt2, ptc;
  E.load(dom1 - 1, pt1);
  E.load(dom2 - 1, pt2);
  if (pt1.id == pt2.id) return 371947172;
  double r1 = E.rec_radius(pt1);
  double r2 = E.rec_radius(pt2);
  double rc = vol2rad(rad2vol(r1) + rad2vol(r2));
  double nc1 = nc_at_radius(w, c1, r1);
  double nc2 = nc_at_radius(w, c2, r2);
  double ncc = nc_at_radius(w, cc, rc);
  int new_group = E.rand_group(s, w, cc, rc);
  double nc_min = fmin(nc1, fmin(nc2, ncc));
  bool remove_1 = s.uniform() < nc_min / nc1;
  bool remove_2 = s.uniform() < nc_min / nc2;
  bool create_new = s.uniform() < nc_min / ncc;
  bool id_1_lost, id_2_lost;
  if (create_new) {
    id_1_lost = false;
    id_2_lost = false;
    if (remove_1 && remove_2) {
      if (E.rec_volume(pt1) > E.rec_volume(pt2)) id_2_lost = true;
      else id_1_lost = true;
    }
  } else {
    id_1_lost = remove_1;
    id_2_lost = remove_2;
  }
  int64_t info_other = 0;
  if (create_new) {
    int err = rec_coagulate(E, s, pt1, pt2, ptc);
    if (err) return err;
    ptc.group = new_group;
    ptc.klass = cc;
    if (remove_1 && !id_1_lost) ptc.id = pt1.id;
    else if (remove_2 && !id_2_lost) ptc.id = pt2.id;
    else ptc.id = E.new_id();
    info_other = ptc.id;
  }
  if (dom2 > dom1) {
    if (remove_2) {
      E.remove(dom2);
      if (id_2_lost) E.add_info(pt2.id, 2, info_other);
    }
    if (remove_1) {
      E.remove(dom1);
      if (id_1_lost) E.add_info(pt1.id, 2, info_other);
    }
  } else {
    if (remove_1) {
      E.remove(dom1);
      if (id_1_lost) E.add_info(pt1.id, 2, info_other);
    }
    if (remove_2) {
      E.remove(dom2);
      if (id_2_lost) E.add_info(pt2.id, 2, info_other);
    }
  }
  if (create_new) return E.add(ptc);
  return 0;
}

// aero_state_dup_particle with random groups, src/aero_state.F90:488-539
__device__ inline int exact_dup(ExactBox &E, const Weights &w, Stream &s, int dom,
                                double n_part_mean) {
  int n_copies = prob_round(s, n_part_mean);
  if (n_copies == 0) {
    int64_t id = E.A->id[E.base + dom - 1];
    E.remove(dom);
    E.add_info(id, 4, 0);
  } else if (n_copies > 1) {
    for (int i_dup = 1; i_dup <= n_copies - 1; ++i_dup) {
      Rec np;
      E.load(dom - 1, np);
      np.id = E.new_id();
      np.group = E.rand_group(s, w, E.A->klass[E.base + dom - 1],
                              vol2rad(E.A->vtot[E.base + dom - 1]));
      int err = E.add(np);
      if (err) return err;
    }
  }
  return 0;
}

// sample_source_particle, src/coagulation.F90:333-517 (lane 0)
__device__ inline int exact_sample_source(ExactBox &E, const Weights &w,
                                          const Env &env, int type, uint32_t site,
                                          uint32_t ordinal, int bs, int cs,
                                          const Rec &target, double n_samp_mean,
                                          double accept_factor, int *n_samp_out,
                                          int *n_coag_out, Rec &source) {
  const int ns = E.A->n_spec;
  *n_samp_out = 0;
  *n_coag_out = 0;
  if (E.n_entry(bs, cs) == 0) return 0;
  Stream aps, smp, fin;
  aps.open(E.seed, E.box, E.step, site, ordinal, RNG_ACC_POISSON);
  smp.open(E.seed, E.box, E.step, site, ordinal, RNG_ACC_SAMPLE);
  fin.open(E.seed, E.box, E.step, site, ordinal, RNG_ACC_FINISH);
  aps.tape = smp.tape = fin.tape = E.tape;
  double nc_target = nc_at_radius(w, target.klass, E.rec_radius(target));
  int64_t target_id = target.id;
  int ct = target.klass;
  double nc_src_min = nc_at_radius(w, cs, E.edges[bs - 1]);
  double prm = nc_target / nc_src_min;
  if (!(prm <= 1.0)) return 653606684;
  int n_remove = rand_poisson(aps, prm * n_samp_mean);
  int n_extra = rand_poisson(aps, (1.0 - prm) * n_samp_mean);
  int n_total = n_remove + n_extra;
  int n_avg = 0, n_samp = 0;
  double prob_coag_tot = 0.0;
  double vol_sq[MAX_SPEC], vol_cv[MAX_SPEC];
  for (int k = 0; k < ns; ++k) {
    vol_sq[k] = 0.0;
    vol_cv[k] = 0.0;
  }
  for (int k = 0; k < ns; ++k) source.vol[k] = 0.0;
  source.id = 0;
  source.group = source.klass = source.nprim = source.ncomp = 0;
  for (int c = 0; c < MAX_COMP; ++c) {
    source.csrc[c] = 0;
    source.ctime[c] = 0.0;
  }
  source.least = source.greatest = 0.0;
  for (int i_samp = 1; i_samp <= n_total; ++i_samp) {
    if (E.n_entry(bs, cs) == 0) break;
    if (n_samp > n_remove && n_avg >= 2) {
      double cvmax = 0.0;
      for (int k = 0; k < ns; ++k) {
        double vm = source.vol[k] / (double)n_avg;
        if (vm > 0.0) vol_cv[k] = sqrt(vol_sq[k] / (double)n_avg - vm * vm) / vm;
        if (k == 0 || vol_cv[k] > cvmax) cvmax = vol_cv[k];
      }
      double m95 = cvmax / sqrt((double)n_avg) * student_t_95_coeff(n_avg);
      if (m95 < 0.1) break;
    }
    n_samp++;
    int i_unif = rand_int(smp, E.n_entry(bs, cs));
    int dom = E.entry(bs, cs, i_unif);
    Rec pi;
    E.load(dom - 1, pi);
    // the target is a detached copy (src/coagulation.F90:226-229): source
    // removals may move its slot, so the kernel is taken against the copy
    double k;
    int err = 0;
    {
      double v1 = E.rec_volume(pi), v2 = E.rec_volume(target);
      double kk = 0.0;
      err = 0;
      switch (type) {
        case KERNEL_ADDITIVE: kk = env.beta * (v1 + v2); break;
        case KERNEL_CONSTANT: kk = C_BETA_0; break;
        case KERNEL_ZERO: kk = 0.0; break;
        case KERNEL_BROWN: {
          double m1 = 0.0, m2 = 0.0;
          for (int q = 0; q < ns; ++q) {
            m1 += pi.vol[q] * E.A->density[q];
            m2 += target.vol[q] * E.A->density[q];
          }
          kk = kernel_brown_helper(v1, m1 / v1, v2, m2 / v2, env.temp, env.pressure);
          break;
        }
        case KERNEL_SEDI: err = kernel_sedi_helper(v1, v2, &kk); break;
        default: err = 200724934;
      }
      k = kk * coag_num_conc_factor(w, vol2rad(v1), vol2rad(v2), cs, ct, ct);
    }
    if (err) return err;
    double prob_coag = k * accept_factor;
    prob_coag_tot += prob_coag;
    if (smp.uniform() < prob_coag) {
      n_avg++;
      err = rec_coagulate(E, smp, source, pi, source);
      if (err) return err;
      for (int q = 0; q < ns; ++q) vol_sq[q] += pi.vol[q] * pi.vol[q];
      if (i_samp <= n_remove) {
        double nc_i = nc_at_radius(w, pi.klass, E.rec_radius(pi));
        double prob_remove_i = nc_target / nc_i;
        if (smp.uniform() < prob_remove_i / prm) {
          E.remove(dom);
          E.add_info(pi.id, 2, target_id);
        }
      }
    }
  }
  *n_samp_out = n_samp;
  if (n_avg == 0) return 0;
  double prob_coag_mean = prob_coag_tot / (double)n_samp;
  if (!(prob_coag_mean <= 1.0)) E.B->n_warn += 1;
  int n_coag = rand_binomial(fin, n_total, prob_coag_mean);
  for (int q = 0; q < ns; ++q) source.vol[q] = source.vol[q] * ((double)n_coag / (double)n_avg);
  source.nprim = prob_round(fin, (double)source.nprim * ((double)n_coag / (double)n_avg));
  int n_comp = source.ncomp;
  int n_comp_new = source.nprim < MAX_COMP ? source.nprim : MAX_COMP;
  int sample[MAX_COMP];
  int cs_src[MAX_COMP];
  double ct_src[MAX_COMP];
  for (int c = 0; c < MAX_COMP; ++c) {
    cs_src[c] = source.csrc[c];
    ct_src[c] = source.ctime[c];
  }
  if (n_comp_new < n_comp) {
    sample_without_replacement(fin, n_comp_new, n_comp, sample);
    for (int i = 0; i < n_comp_new; ++i) {
      source.csrc[i] = cs_src[sample[i] - 1];
      source.ctime[i] = ct_src[sample[i] - 1];
    }
    source.ncomp = n_comp_new;
  } else if (n_comp < MAX_COMP) {
    int i_comp = 1;
    while (i_comp <= n_comp_new) {
      if (i_comp + n_comp <= n_comp_new) {
        for (int i = 1; i <= n_comp; ++i) {
          source.csrc[i_comp + i - 2] = cs_src[i - 1];
          source.ctime[i_comp + i - 2] = ct_src[i - 1];
        }
        i_comp += n_comp;
      } else {
        int sample_size = n_comp_new - i_comp + 1;
        sample_without_replacement(fin, sample_size, n_comp, sample);
        for (int i = 1; i <= sample_size; ++i) {
          source.csrc[i_comp + i - 2] = cs_src[sample[i - 1] - 1];
          source.ctime[i_comp + i - 2] = ct_src[sample[i - 1] - 1];
        }
        i_comp = n_comp_new + 1;
      }
    }
    source.ncomp = n_comp_new;
  }
  if (source.nprim < source.ncomp) return 808144130;
  *n_coag_out = n_coag;
  return 0;
}

// coag_target_with_source, src/coagulation.F90:522-582 (lane 0)
__device__ inline int exact_target_with_source(ExactBox &E, const Weights &w,
                                               Stream &s, int bt, int ct,
                                               int target_entry, const Rec &source,
                                               int cc) {
  int dom = E.entry(bt, ct, target_entry);
  Rec tp, np;
  E.load(dom - 1, tp);
  int64_t target_id = tp.id;
  double old_nc = nc_at_radius(w, tp.klass, E.rec_radius(tp));
  int err = rec_coagulate(E, s, tp, source, np);
  if (err) return err;
  np.id = target_id;
  np.group = E.rand_group(s, w, cc, E.rec_radius(np));
  np.klass = cc;
  E.store(dom - 1, np);
  int new_bin = bin_find_log(E.nb, E.B->log_lo, E.B->log_hi, E.rec_radius(np));
  if (new_bin < 1 || new_bin > E.nb) return 765620746;
  err = E.change(dom, new_bin, cc);
  if (err) return err;
  double new_nc = nc_at_radius(w, np.klass, E.rec_radius(np));
  err = exact_dup(E, w, s, dom, old_nc / new_nc);
  if (err) return err;
  if (E.A->id[E.base + dom - 1] != target_id) return 654300924;
  return 0;
}

__device__ inline uint32_t pair_site(int nb, int nc, int b1, int b2, int c1, int c2) {
  return (uint32_t)((((c1 - 1) * nc + (c2 - 1)) * nb + (b1 - 1)) * nb + (b2 - 1));
}

// Lane 0's writes become visible to the other lanes of its wavefront: the
// box lives in one workgroup, so a workgroup-scope fence suffices (a
// device-scope fence would write back and invalidate the whole L2).
__device__ inline void wave_fence() { __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup"); }

// The sequential sweep: one 64-lane wavefront per box.
__global__ void __launch_bounds__(64)
k_exact_sweep(BoxDesc *boxes, Arrays A, int type, double del_t, uint64_t seed,
              uint32_t step) {
  const int box = blockIdx.x;
  const int lane = threadIdx.x;
  BoxDesc *B = &boxes[box];
  __shared__ Weights s_w;
  __shared__ Env s_env;
  if (lane == 0) {
    s_w = B->w;
    s_env = B->env;
    B->step_samp = 0;
    B->step_coag = 0;
    B->part_steps += B->n_part;
  }
  __syncthreads();
  if (B->error) return;
  const Weights &w = s_w;
  const Env env = s_env;
  ExactBox E;
  E.A = &A;
  E.B = B;
  E.tape = B->tape.mode != TAPE_OFF ? &B->tape : nullptr;
  E.box = box;
  E.base = B->off;
  E.nb = B->n_bin;
  E.nc = w.n_class;
  E.lstart = A.lstart + (int64_t)box * A.max_keys;
  E.llen = A.llen + (int64_t)box * A.max_keys;
  E.lcap = A.lcap + (int64_t)box * A.max_keys;
  E.seed = seed;
  E.step = step;
  E.edges = A.edges + (int64_t)box * (A.max_bin + 1);
  const int nb = E.nb, nc = E.nc;
  int64_t tot_samp = 0, tot_coag = 0, n_warn = 0;
  int err = 0;

  bool mono_dec = true;
  for (int i = 0; i < w.n_group * w.n_class; ++i) {
    if (w.type[i] == WEIGHT_POWER && w.expo[i] > 0.0) mono_dec = false;
  }

  for (int c1 = 1; c1 <= nc && !err; ++c1)
    for (int c2 = 1; c2 <= c1 && !err; ++c2)
      for (int b1 = 1; b1 <= nb && !err; ++b1) {
        if (E.n_entry(b1, c1) == 0) continue;
        const int b2_start = (c1 == c2) ? b1 : 1;
        for (int b2 = b2_start; b2 <= nb && !err; ++b2) {
          if (E.n_entry(b2, c2) == 0) continue;
          const int64_t ti = table_index(A, box, nb, nc, b1, b2, c1, c2);
          const int cc = A.cc[ti];
          const double k_max = A.kw[ti];
          const uint32_t site = pair_site(nb, nc, b1, b2, c1, c2);

          // try_per_particle_coag, src/coagulation.F90:150-241
          bool accel = false;
          if (mono_dec) {
            double a1 = nc_at_radius(w, c1, E.edges[b1 - 1]);
            double a2 = nc_at_radius(w, c1, E.edges[b1]);
            double d1 = nc_at_radius(w, c2, E.edges[b2 - 1]);
            double d2 = nc_at_radius(w, c2, E.edges[b2]);
            double nc1_min = fmin(a1, a2), nc1_max = fmax(a1, a2);
            double nc2_min = fmin(d1, d2), nc2_max = fmax(d1, d2);
            int bt = 0, bs = 0, ct = 0, cs = 0;
            bool ok = false;
            if (nc1_max < nc2_min && cc == c1) {
              bt = b1; bs = b2; ct = c1; cs = c2; ok = true;
            }
            if (nc2_max < nc1_min && cc == c2) {
              bt = b2; bs = b1; ct = c2; cs = c1; ok = true;
            }
            if (ok) {
              double n_spt = k_max * del_t * (double)E.n_entry(bs, cs);
              double accept_factor = 1.0 / k_max;
              if (n_spt >= 1.0) {
                double min_tv = rad2vol(E.edges[bt - 1]);
                double max_sv = rad2vol(E.edges[bs]);
                double max_new = min_tv + max_sv * n_spt;
                if (!(max_new / min_tv > 1.5)) {
                  accel = true;
                  if (lane == 0) {
                    B->n_accel_pairs += 1;
                    for (int e = E.n_entry(bt, ct); e >= 1 && !err; --e) {
                      int dom = E.entry(bt, ct, e);
                      Rec target, source;
                      E.load(dom - 1, target);
                      int ns_, nco;
                      err = exact_sample_source(E, w, env, type, site, (uint32_t)e, bs, cs,
                                                target, n_spt, accept_factor, &ns_, &nco,
                                                source);
                      if (!err && nco > 0) {
                        Stream ts;
                        ts.open(seed, box, step, site, (uint32_t)e, RNG_ACC_TARGET);
                        ts.tape = E.tape;
                        err = exact_target_with_source(E, w, ts, bt, ct, e, source, cc);
                      }
                      tot_samp += ns_;
                      tot_coag += nco;
                      B->n_acc_samp += ns_;
                      B->n_acc_coag += nco;
                    }
                  }
                  err = __shfl(err, 0);
                  wave_fence();
                }
              }
            }
          }
          if (accel) continue;

          // per_set_coag + compute_n_samp, src/coagulation.F90:587-702
          const bool same = (b1 == b2) && (c1 == c2);
          int n1 = E.n_entry(b1, c1), n2 = E.n_entry(b2, c2);
          double n_possible = same ? (double)n1 * ((double)n2 - 1.0) / 2.0
                                   : (double)n1 * (double)n2;
          double mean = k_max * del_t * n_possible;
          int n_samp = 0;
          if (lane == 0) {
            Stream ps;
            ps.open(seed, box, step, site, 0, RNG_POISSON);
            ps.tape = E.tape;
            n_samp = rand_poisson(ps, mean);
          }
          n_samp = __shfl(n_samp, 0);
          const double accept_factor = 1.0 / k_max;
          tot_samp += n_samp;
          int i = 1;
          while (i <= n_samp && !err) {
            n1 = E.n_entry(b1, c1);
            n2 = E.n_entry(b2, c2);
            if ((n1 < 2 && same) || n1 < 1 || n2 < 1) break;
            // speculative batch: lane j evaluates test i + j (one test at a
            // time on a tape, whose draws follow the call order)
            const int ti_s = i + lane;
            bool valid = ti_s <= n_samp && (!E.tape || lane == 0);
            bool accept = false, warn = false;
            int dom1 = 0, dom2 = 0;
            int lerr = 0;
            if (valid) {
              Stream ts;
              ts.open(seed, box, step, site, (uint32_t)(ti_s - 1), RNG_TEST);
              ts.tape = E.tape;
              int i1 = rand_int(ts, n1);
              int i2;
              if (same) {
                i2 = rand_int(ts, n2 - 1);
                if (i2 == i1) i2 = n2;
              } else {
                i2 = rand_int(ts, n2);
              }
              dom1 = E.entry(b1, c1, i1);
              dom2 = E.entry(b2, c2, i2);
              double k;
              lerr = weighted_kernel64(A, w, env, type, E.base + dom1 - 1, E.base + dom2 - 1, c1,
                                       c2, cc, &k);
              double p = k * accept_factor;
              warn = !(p <= 1.0);
              accept = ts.uniform() < p;
            }
            unsigned long long am = __ballot(valid && (accept || lerr != 0));
            int first = am ? (__ffsll((long long)am) - 1) : 64;
            // count the tests that happened: lanes 0..first (or the batch)
            int last = first < 64 ? first : 63;
            unsigned long long wm = __ballot(valid && warn && lane <= last);
            n_warn += __popcll(wm);
            if (first == 64) {
              i += E.tape ? 1 : 64;
              continue;
            }
            int fe = __shfl(lerr, first);
            if (fe) {
              err = fe;
              break;
            }
            int a1 = __shfl(dom1, first), a2 = __shfl(dom2, first);
            if (lane == 0) {
              Stream ms;
              ms.open(seed, box, step, site, (uint32_t)(i + first - 1), RNG_MERGE);
              ms.tape = E.tape;
              err = exact_coagulate(E, w, ms, a1, a2, c1, c2, cc);
            }
            err = __shfl(err, 0);
            wave_fence();
            tot_coag += 1;
            i = i + first + 1;
          }
        }
      }
  if (lane == 0) {
    if (err) B->error = err;
    B->step_samp = tot_samp;
    B->step_coag = tot_coag;
    B->n_samp += tot_samp;
    B->n_coag += tot_coag;
    B->n_warn += n_warn;
  }
}

}  // namespace pmc
