// pmc_rebalance.h -- population control and particle sinks of the boxes
// on the device: aero_state_double and aero_state_halve
// (src/aero_state.F90:1722-1786), the per-(group, class) counts
// aero_state_rebalance decides from (src/aero_state.F90:287-310,
// 1790-1852), dilution out of the box (aero_state_sample_particles,
// src/aero_state.F90:856-929) and the scenario's particle loss
// (src/scenario.F90:398-745).  The host sequences the passes in the
// reference's order; each pass runs every box that needs it at once, over
// tiles of REBAL_TILE particles (grid x = tile, y = box).
#pragma once

#include "pmc_scan.h"

#include "pmc_exact.h"

namespace pmc {

constexpr int GC_KEYS = MAX_GROUP * MAX_CLASS;  // counts per box, then n_part
constexpr int REBAL_THREADS = 256;
constexpr int REBAL_ITEMS = 16;
constexpr int REBAL_TILE = REBAL_THREADS * REBAL_ITEMS;
constexpr int INFO_DILUTION = 1;                // AERO_INFO_DILUTION, src/aero_info.F90:21
constexpr int INFO_HALVED = 3;                  // AERO_INFO_HALVED, src/aero_info.F90:25

// scenario loss function codes, src/scenario.F90:26-36
constexpr int LOSS_INVALID = 0, LOSS_NONE = 1, LOSS_CONSTANT = 2, LOSS_VOLUME = 3,
              LOSS_DRYDEP = 4, LOSS_CHAMBER = 5;

// what a removal pass removes
constexpr int RM_HALVE = 0, RM_DILUTE = 1, RM_LOSS = 2;
struct Removal {
  int kind;               // RM_*
  int g, c, k;            // halving: the (group, class) and its pass
  int loss;               // particle loss: LOSS_*
  double del_t;           // particle loss: time step (s)
  const double *prob;     // dilution: per-box probability
  const double *height;   // particle loss: per-box mixing height (m)
};

// Device scratch of the population-control passes.
struct RebalWork {
  int *gc;         // [n_box][GC_KEYS + 1] counts, then n_part
  int *plan;       // [n_box] passes to run (doubling) or 0/1 (halving)
  int *tile_a;     // [n_box][max_tiles] per-tile count
  int *tile_off;   // [n_box][max_tiles] exclusive scan of tile_a within the box
  int *tile_kmax;  // [n_box][max_tiles] largest kept index in the tile (-1: none)
  int *box_tot;    // [n_box] sum of tile_a
  int *box_imax;   // [n_box] largest kept index of the box (-1: none)
  uint8_t *rm;     // [total_cap] removal flags of a halving pass
  int *rp;         // [total_cap] inclusive count of removals up to the slot
  double *bval;    // [n_box] per-box parameter of a removal pass
  int max_tiles;
};

__device__ inline bool in_gc(const Arrays &A, int64_t p, int g, int c) {
  return A.group[p] == g && A.klass[p] == c;
}

__global__ void k_zero_int(int *a, int64_t n) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) a[i] = 0;
}

// counts[box][(g-1) + n_group*(c-1)] += particles of (g, c) in the tile;
// counts[box][GC_KEYS] = n_part (written by tile 0).
__global__ void __launch_bounds__(REBAL_THREADS)
k_count_gc(const BoxDesc *boxes, Arrays A, int *counts) {
  const int box = blockIdx.y, t = threadIdx.x;
  const BoxDesc *B = &boxes[box];
  const int n = B->n_part, ng = B->w.n_group, nc = B->w.n_class;
  const int start = blockIdx.x * REBAL_TILE;
  int *out = counts + (int64_t)box * (GC_KEYS + 1);
  if (blockIdx.x == 0 && t == 0) out[GC_KEYS] = n;
  if (start >= n) return;
  __shared__ int s_cnt[GC_KEYS];
  for (int k = t; k < GC_KEYS; k += REBAL_THREADS) s_cnt[k] = 0;
  __syncthreads();
  const int end = min(n, start + REBAL_TILE);
  for (int i = start + t; i < end; i += REBAL_THREADS) {
    const int64_t p = B->off + i;
    const int g = A.group[p], c = A.klass[p];
    if (g >= 1 && g <= ng && c >= 1 && c <= nc) atomicAdd(&s_cnt[(g - 1) + ng * (c - 1)], 1);
  }
  __syncthreads();
  for (int k = t; k < ng * nc; k += REBAL_THREADS)
    if (s_cnt[k]) atomicAdd(&out[k], s_cnt[k]);
}

// Per box: exclusive scan of the tile counts, their total, and the largest
// kept index over the tiles (halving passes).
__global__ void __launch_bounds__(REBAL_THREADS)
k_tile_scan(const BoxDesc *boxes, RebalWork R, int with_kmax) {
  const int box = blockIdx.x, t = threadIdx.x;
  if (!R.plan[box]) return;
  __shared__ scan_storage<REBAL_THREADS, int> ts;
  const int n_tiles = (boxes[box].n_part + REBAL_TILE - 1) / REBAL_TILE;
  const int *a = R.tile_a + (int64_t)box * R.max_tiles;
  int *off = R.tile_off + (int64_t)box * R.max_tiles;
  int done = 0, kmax = -1;
  for (int c0 = 0; c0 < n_tiles; c0 += REBAL_THREADS) {
    const int k = c0 + t;
    const int v = k < n_tiles ? a[k] : 0;
    int total = 0;
    const int ex = block_sum_excl<REBAL_THREADS>(v, ts, &total);
    if (k < n_tiles) off[k] = done + ex;
    done += total;
    if (with_kmax && k < n_tiles) kmax = max(kmax, R.tile_kmax[(int64_t)box * R.max_tiles + k]);
  }
  if (with_kmax) {
    const int m = block_reduce<REBAL_THREADS>(kmax, OpMax(), ts);
    if (t == 0) R.box_imax[box] = m;
  }
  if (t == 0) R.box_tot[box] = done;
}

// ---------------------------------------------------------------- double
// One doubling pass of (g, c) in every box with plan[box] > pass
// (aero_state_double, src/aero_state.F90:1722-1750): the box's (g, c)
// particles present at the start are appended in index order with
// consecutive new ids (aero_particle_new_id, src/aero_particle.F90:157-165),
// the sort is invalidated and the magnitude halved.
__global__ void __launch_bounds__(REBAL_THREADS)
k_double_count(const BoxDesc *boxes, Arrays A, RebalWork R, int g, int c, int pass) {
  const int box = blockIdx.y, t = threadIdx.x;
  const BoxDesc *B = &boxes[box];
  if (R.plan[box] <= pass || B->error) return;
  const int n = B->n_part, start = blockIdx.x * REBAL_TILE;
  if (start >= n) return;
  __shared__ scan_storage<REBAL_THREADS, int> tr;
  int cnt = 0;
  const int end = min(n, start + REBAL_TILE);
  for (int i = start + t; i < end; i += REBAL_THREADS) cnt += in_gc(A, B->off + i, g, c) ? 1 : 0;
  const int tot = block_reduce<REBAL_THREADS>(cnt, OpSum(), tr);
  if (t == 0) R.tile_a[(int64_t)box * R.max_tiles + blockIdx.x] = tot;
}

__global__ void __launch_bounds__(REBAL_THREADS)
k_double_copy(BoxDesc *boxes, Arrays A, RebalWork R, int g, int c, int pass) {
  const int box = blockIdx.y, t = threadIdx.x;
  BoxDesc *B = &boxes[box];
  if (R.plan[box] <= pass || B->error) return;
  const int n = B->n_part, start = blockIdx.x * REBAL_TILE;
  if (start >= n) return;
  __shared__ scan_storage<REBAL_THREADS, int> ts;
  const int64_t base = B->off, cap = B->cap;
  const int64_t next0 = B->next_id, stride = B->id_stride, offset = B->id_offset;
  int done = R.tile_off[(int64_t)box * R.max_tiles + blockIdx.x];
  for (int k = 0; k < REBAL_ITEMS; ++k) {
    const int i = start + k * REBAL_THREADS + t;
    const int flag = (i < n && in_gc(A, base + i, g, c)) ? 1 : 0;
    int total = 0;
    const int rank = block_sum_excl<REBAL_THREADS>(flag, ts, &total);
    if (flag) {
      const int64_t dst = (int64_t)n + done + rank;
      if (dst < cap) {
        copy_particle(A, base + i, base + dst);
        A.id[base + dst] = (next0 + done + rank - 1) * stride + offset + 1;
      }
    }
    done += total;
  }
}

__global__ void k_double_finish(BoxDesc *boxes, RebalWork R, int n_box, int g, int c, int pass) {
  const int box = blockIdx.x * blockDim.x + threadIdx.x;
  if (box >= n_box) return;
  BoxDesc *B = &boxes[box];
  if (R.plan[box] <= pass || B->error) return;
  const int added = R.box_tot[box];
  if ((int64_t)B->n_part + added > B->cap) {
    B->error = ERR_CAPACITY;
    return;
  }
  B->n_part += added;
  B->next_id += added;
  B->valid_sort = 0;
  B->xlists_valid = 0;
  B->pkeys_valid = 0;
  B->terms_kernel = 0;
  B->tables_valid = 0;
  B->w.mag[(g - 1) + B->w.n_group * (c - 1)] *= 0.5;
}

// --------------------------------------------------------------- removal
// One removal pass (a halving pass of (g, c), aero_state_halve,
// src/aero_state.F90:1755-1786; dilution; particle loss): from the last
// particle down, a particle that draws removal is logged (AERO_INFO_HALVED
// or AERO_INFO_DILUTION, other id 0) and the last particle moves into its
// slot.  Site of halving pass k: (g-1) + n_group*((c-1) + n_class*k);
// dilution and loss use site 0.
__device__ inline uint32_t halve_site(const BoxDesc *B, int g, int c, int k) {
  return (uint32_t)(g - 1) +
         (uint32_t)B->w.n_group * ((uint32_t)(c - 1) + (uint32_t)B->w.n_class * (uint32_t)k);
}

// scenario_loss_rate_dry_dep, src/scenario.F90:444-533 (spherical
// particles: the diameter is 2 vol2rad)
__device__ inline double loss_rate_dry_dep(double vol, double density, const Env &env,
                                           double height) {
  const double u_mean = 5.0, z_ref = 20.0, z_rough = 0.1;
  const double Acol = 2.0 / 1000.0, alpha = 1.2, beta = 2.0, gamma = 0.54;
  const double temp = env.temp, pressure = env.pressure;
  const double d_p = 2.0 * vol2rad(vol);
  const double density_air = (C_AIR_MW * pressure) / (C_UNIV_GAS * temp);
  const double visc_d = 1.8325e-5 * (416.16 / (temp + 120.0)) * pow(temp / 296.16, 1.5);
  const double visc_k = visc_d / density_air;
  const double gas_speed = sqrt((8.0 * C_BOLTZMANN * temp * C_AVOGADRO) / (C_PI * C_AIR_MW));
  const double gas_mean_free_path = (2.0 * visc_d) / (density_air * gas_speed);
  const double knud = (2.0 * gas_mean_free_path) / d_p;
  const double cunning = 1.0 + knud * (1.257 + 0.4 * exp(-1.1 / knud));
  const double grav = 9.81;
  const double V_s = (density * pow(d_p, 2.0) * grav * cunning) / (18.0 * visc_d);
  const double u_star = 0.4 * u_mean / log(z_ref / z_rough);
  const double R_a = (1.0 / (0.4 * u_star)) * log(z_ref / z_rough);
  const double diff_p = (C_BOLTZMANN * temp * cunning) / (3.0 * C_PI * visc_d * d_p);
  const double Sc = visc_k / diff_p;
  const double E_B = pow(Sc, -gamma);
  const double E_IN = 0.5 * pow(d_p / Acol, 2.0);
  const double St = (V_s * u_star) / (grav * Acol);
  const double E_IM = pow(St / (alpha + St), beta);
  const double R1 = exp(-pow(St, 0.5));
  const double eps_0 = 3.0;
  const double R_s = 1.0 / (eps_0 * u_star * (E_B + E_IN + E_IM) * R1);
  const double V_d = V_s + (1.0 / (R_a + R_s + R_a * R_s * V_s));
  return V_d / height;
}

// scenario_loss_rate, src/scenario.F90:398-437 (the chamber function is
// refused on the host)
__device__ inline double loss_rate(int type, double vol, double density, const Env &env,
                                   double height) {
  switch (type) {
    case LOSS_CONSTANT: return 1e-3;
    case LOSS_VOLUME: return 1e15 * vol;
    case LOSS_DRYDEP: return loss_rate_dry_dep(vol, density, env, height);
    default: return 0.0;
  }
}

// Whether slot i of the box goes in this pass.  The particle at 1-based
// index i draws the first uniform of stream (step, site, i, kind) and goes
// when it is below its probability: 1/2 for a halving pass of its class,
// the box's probability for dilution, 1 - exp(-del_t rate) for particle
// loss.  Independent per-particle draws have the distribution of the
// reference's binomial count with uniform picks (dilution) and of its
// geometric-skip accept-reject (particle loss).
__device__ inline bool removal_draw(const Arrays &A, const BoxDesc *B, int box, int i,
                                    const Removal &R, uint64_t seed, uint32_t step) {
  const int64_t p = B->off + i;
  double prob;
  uint32_t site = 0, kind;
  if (R.kind == RM_HALVE) {
    if (!in_gc(A, p, R.g, R.c)) return false;
    prob = 0.5;
    site = halve_site(B, R.g, R.c, R.k);
    kind = RNG_HALVE;
  } else if (R.kind == RM_DILUTE) {
    prob = R.prob[box];
    kind = RNG_DILUTE;
  } else {
    double v = 0.0, m = 0.0;
    for (int s = 0; s < A.n_spec; ++s) {
      const double x = A.vol[(int64_t)s * A.total_cap + p];
      v += x;
      m += x * A.density[s];
    }
    const double h = R.height ? R.height[box] : 1.0;
    prob = 1.0 - exp(-R.del_t * loss_rate(R.loss, v, m / v, B->env, h));
    kind = RNG_LOSS;
  }
  if (!(prob > 0.0)) return false;
  Stream s;
  s.open(seed, (uint32_t)box, step, site, (uint32_t)(i + 1), kind);
  return s.uniform() < prob;
}

__device__ inline int removal_action(const Removal &R) {
  return R.kind == RM_HALVE ? INFO_HALVED : INFO_DILUTION;
}

// With current bin lists (exact mode) the removals go through
// aero_sorted_remove_particle, whose list order the next steps depend on:
// one thread walks the box in the reference's order.
__global__ void __launch_bounds__(64)
k_remove_lists(BoxDesc *boxes, Arrays A, int n_box, const int *plan, Removal R, uint64_t seed,
               uint32_t step) {
  const int box = blockIdx.x * blockDim.x + threadIdx.x;
  if (box >= n_box) return;
  BoxDesc *B = &boxes[box];
  if (!plan[box] || B->error || !(B->valid_sort && B->xlists_valid)) return;
  ExactBox E;
  E.A = &A;
  E.B = B;
  E.box = box;
  E.base = B->off;
  E.nb = B->n_bin;
  E.nc = B->w.n_class;
  E.lstart = A.lstart + (int64_t)box * A.max_keys;
  E.llen = A.llen + (int64_t)box * A.max_keys;
  E.lcap = A.lcap + (int64_t)box * A.max_keys;
  E.seed = seed;
  E.step = step;
  E.edges = A.edges + (int64_t)box * (A.max_bin + 1);
  for (int i = B->n_part - 1; i >= 0; --i) {
    if (!removal_draw(A, B, box, i, R, seed, step)) continue;
    E.add_info(A.id[B->off + i], removal_action(R), 0);
    E.remove(i + 1);
  }
  B->pkeys_valid = 0;
  B->terms_kernel = 0;
  if (R.kind == RM_HALVE) {
    B->tables_valid = 0;
    B->w.mag[(R.g - 1) + B->w.n_group * (R.c - 1)] *= 2.0;
  }
}

// Without lists the swap-last order has a closed form, so every slot is
// settled at once.  Let imax be the largest kept index and r(j) the number
// of removals at indices in (j, imax).  Visiting from the last index down,
// the kept particles after the current index form a run that every removal
// below imax rotates by one (its last particle moves into the hole).  The
// result: slots below the kept count K keep their particle when it is kept;
// a removed slot f receives the particle found by j <- f, then
// j <- imax - r(j) while slot j is a removal (j grows strictly, so the walk
// ends at a kept particle, after about log2(n) steps when half are removed).
// Slots >= K are only read, so the moves run in place in one pass.
__global__ void __launch_bounds__(REBAL_THREADS)
k_remove_mark(const BoxDesc *boxes, Arrays A, RebalWork R, Removal X, uint64_t seed,
              uint32_t step) {
  const int box = blockIdx.y, t = threadIdx.x;
  const BoxDesc *B = &boxes[box];
  if (!R.plan[box] || B->error) return;
  const int n = B->n_part, start = blockIdx.x * REBAL_TILE;
  if (start >= n) return;
  __shared__ scan_storage<REBAL_THREADS, int> tr;
  int cnt = 0, kmax = -1;
  const int end = min(n, start + REBAL_TILE);
  for (int i = start + t; i < end; i += REBAL_THREADS) {
    const bool r = removal_draw(A, B, box, i, X, seed, step);
    R.rm[B->off + i] = r ? 1 : 0;
    cnt += r ? 1 : 0;
    if (!r) kmax = max(kmax, i);
  }
  const int tot = block_reduce<REBAL_THREADS>(cnt, OpSum(), tr);
  const int km = block_reduce<REBAL_THREADS>(kmax, OpMax(), tr);
  if (t == 0) {
    R.tile_a[(int64_t)box * R.max_tiles + blockIdx.x] = tot;
    R.tile_kmax[(int64_t)box * R.max_tiles + blockIdx.x] = km;
  }
}

// Inclusive removal counts per slot, and the removal log in the reference's
// order (from the last index down).
__global__ void __launch_bounds__(REBAL_THREADS)
k_remove_prefix(BoxDesc *boxes, Arrays A, RebalWork R, int action) {
  const int box = blockIdx.y, t = threadIdx.x;
  const BoxDesc *B = &boxes[box];
  if (!R.plan[box] || B->error) return;
  const int n = B->n_part, start = blockIdx.x * REBAL_TILE;
  if (start >= n) return;
  __shared__ scan_storage<REBAL_THREADS, int> ts;
  const int64_t base = B->off;
  const int n_rm = R.box_tot[box];
  int done = R.tile_off[(int64_t)box * R.max_tiles + blockIdx.x];
  for (int k = 0; k < REBAL_ITEMS; ++k) {
    const int i = start + k * REBAL_THREADS + t;
    const int flag = i < n ? R.rm[base + i] : 0;
    int total = 0;
    const int incl = block_scan_incl<REBAL_THREADS>(flag, OpSum(), ts, &total);
    if (i < n) {
      R.rp[base + i] = done + incl;
      if (flag) {
        const int64_t slot = B->n_info + (n_rm - (done + incl));
        if (slot < B->info_cap) {
          A.info_id[B->info_off + slot] = A.id[base + i];
          A.info_action[B->info_off + slot] = action;
          A.info_other[B->info_off + slot] = 0;
        }
      }
    }
    done += total;
  }
}

__global__ void __launch_bounds__(REBAL_THREADS)
k_remove_move(const BoxDesc *boxes, Arrays A, RebalWork R) {
  const int box = blockIdx.y;
  const BoxDesc *B = &boxes[box];
  if (!R.plan[box] || B->error) return;
  const int n = B->n_part, n_rm = R.box_tot[box], imax = R.box_imax[box];
  const int K = n - n_rm;
  const int64_t base = B->off;
  const int start = blockIdx.x * REBAL_TILE;
  const int end = min(K, start + REBAL_TILE);
  // removals above imax: n - 1 - imax of them
  const int above = n - 1 - imax;
  for (int f = start + threadIdx.x; f < end; f += REBAL_THREADS) {
    if (!R.rm[base + f]) continue;
    int j = f;
    while (R.rm[base + j]) j = imax - ((n_rm - R.rp[base + j]) - above);
    copy_particle(A, base + j, base + f);
  }
}

__global__ void k_remove_finish(BoxDesc *boxes, RebalWork R, int n_box, Removal X) {
  const int box = blockIdx.x * blockDim.x + threadIdx.x;
  if (box >= n_box) return;
  BoxDesc *B = &boxes[box];
  if (!R.plan[box] || B->error) return;
  const int n_rm = R.box_tot[box];
  B->n_part -= n_rm;
  B->n_info += n_rm;
  B->xlists_valid = 0;
  B->pkeys_valid = 0;
  B->terms_kernel = 0;
  if (X.kind == RM_HALVE) {
    B->tables_valid = 0;
    B->w.mag[(X.g - 1) + B->w.n_group * (X.c - 1)] *= 2.0;
  }
}

// ---------------------------------------------------------------- append
// New ids for appended particles (aero_particle_new_id,
// src/aero_particle.F90:157-165), in order.
__global__ void k_assign_ids(BoxDesc *boxes, Arrays A, int box, int first, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const BoxDesc *B = &boxes[box];
  if (i >= n) return;
  A.id[B->off + first + i] = (B->next_id + i - 1) * B->id_stride + B->id_offset + 1;
}

// flag[0] = 1 when an appended particle lies outside the box's grid
__global__ void k_append_check(const BoxDesc *boxes, Arrays A, int box, int first, int n,
                               int *flag) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const BoxDesc *B = &boxes[box];
  if (i >= n || !B->valid_sort) return;
  const int nb = B->n_bin;
  const int b = nb > 0 ? bin_find_log(nb, B->log_lo, B->log_hi, vol2rad(A.vtot[B->off + first + i]))
                       : 0;
  if (b < 1 || b > nb) atomicOr(flag, 1);
}

// aero_state_add_particle for appended slots [first, first + n)
// (src/aero_state.F90:365-383, aero_sorted_add_particle
// src/aero_sorted.F90:397-449): with current lists each particle joins the
// end of its (bin, class) list in order; the first one outside the grid
// makes the sort invalid and the remake waits for the next step, which
// sees every appended particle.
__global__ void k_append_finish(BoxDesc *boxes, Arrays A, int box, int first, int n,
                                const int *flag, int n_id) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  BoxDesc *B = &boxes[box];
  if (B->valid_sort && B->xlists_valid) {
    ExactBox E;
    E.A = &A;
    E.B = B;
    E.box = box;
    E.base = B->off;
    E.nb = B->n_bin;
    E.nc = B->w.n_class;
    E.lstart = A.lstart + (int64_t)box * A.max_keys;
    E.llen = A.llen + (int64_t)box * A.max_keys;
    E.lcap = A.lcap + (int64_t)box * A.max_keys;
    E.edges = A.edges + (int64_t)box * (A.max_bin + 1);
    for (int i = 0; i < n; ++i) {
      const int64_t p = B->off + first + i;
      const int bin = bin_find_log(E.nb, B->log_lo, B->log_hi, vol2rad(A.vtot[p]));
      if (bin < 1 || bin > E.nb) {
        B->valid_sort = 0;
        B->xlists_valid = 0;
        break;
      }
      const int k = E.key(bin, A.klass[p]);
      const int err = E.list_append(k, first + i + 1);
      if (err) {
        B->error = err;
        break;
      }
      A.fwd_bin[p] = bin;
      A.fwd_class[p] = A.klass[p];
      A.fwd_idx[p] = E.llen[k];
    }
  } else if (B->valid_sort && *flag) {
    B->valid_sort = 0;
    B->xlists_valid = 0;
  }
  B->n_part = first + n;
  B->next_id += n_id;
  B->pkeys_valid = 0;
  B->terms_kernel = 0;
}

}  // namespace pmc
