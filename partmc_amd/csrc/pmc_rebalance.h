// pmc_rebalance.h -- population control of the boxes on the device:
// aero_state_double and aero_state_halve (src/aero_state.F90:1722-1786) and
// the per-(group, class) counts aero_state_rebalance decides from
// (src/aero_state.F90:287-310, 1790-1852).  The host sequences the passes
// in the reference's (group, class) order; each pass runs every box that
// needs it at once, over tiles of REBAL_TILE particles (grid x = tile,
// y = box).
#pragma once

#include <hipcub/hipcub.hpp>

#include "pmc_exact.h"

namespace pmc {

constexpr int GC_KEYS = MAX_GROUP * MAX_CLASS;  // counts per box, then n_part
constexpr int REBAL_THREADS = 256;
constexpr int REBAL_ITEMS = 16;
constexpr int REBAL_TILE = REBAL_THREADS * REBAL_ITEMS;
constexpr int INFO_HALVED = 3;                  // AERO_INFO_HALVED, src/aero_info.F90:25

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
  using Scan = hipcub::BlockScan<int, REBAL_THREADS>;
  using Red = hipcub::BlockReduce<int, REBAL_THREADS>;
  __shared__ typename Scan::TempStorage ts;
  __shared__ typename Red::TempStorage tr;
  const int n_tiles = (boxes[box].n_part + REBAL_TILE - 1) / REBAL_TILE;
  const int *a = R.tile_a + (int64_t)box * R.max_tiles;
  int *off = R.tile_off + (int64_t)box * R.max_tiles;
  int done = 0, kmax = -1;
  for (int c0 = 0; c0 < n_tiles; c0 += REBAL_THREADS) {
    const int k = c0 + t;
    const int v = k < n_tiles ? a[k] : 0;
    int ex = 0, total = 0;
    Scan(ts).ExclusiveSum(v, ex, total);
    if (k < n_tiles) off[k] = done + ex;
    done += total;
    if (with_kmax && k < n_tiles) kmax = max(kmax, R.tile_kmax[(int64_t)box * R.max_tiles + k]);
    __syncthreads();
  }
  if (with_kmax) {
    const int m = Red(tr).Reduce(kmax, hipcub::Max());
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
  using Red = hipcub::BlockReduce<int, REBAL_THREADS>;
  __shared__ typename Red::TempStorage tr;
  int cnt = 0;
  const int end = min(n, start + REBAL_TILE);
  for (int i = start + t; i < end; i += REBAL_THREADS) cnt += in_gc(A, B->off + i, g, c) ? 1 : 0;
  const int tot = Red(tr).Sum(cnt);
  if (t == 0) R.tile_a[(int64_t)box * R.max_tiles + blockIdx.x] = tot;
}

__global__ void __launch_bounds__(REBAL_THREADS)
k_double_copy(BoxDesc *boxes, Arrays A, RebalWork R, int g, int c, int pass) {
  const int box = blockIdx.y, t = threadIdx.x;
  BoxDesc *B = &boxes[box];
  if (R.plan[box] <= pass || B->error) return;
  const int n = B->n_part, start = blockIdx.x * REBAL_TILE;
  if (start >= n) return;
  using Scan = hipcub::BlockScan<int, REBAL_THREADS>;
  __shared__ typename Scan::TempStorage ts;
  const int64_t base = B->off, cap = B->cap;
  const int64_t next0 = B->next_id, stride = B->id_stride, offset = B->id_offset;
  int done = R.tile_off[(int64_t)box * R.max_tiles + blockIdx.x];
  for (int k = 0; k < REBAL_ITEMS; ++k) {
    const int i = start + k * REBAL_THREADS + t;
    const int flag = (i < n && in_gc(A, base + i, g, c)) ? 1 : 0;
    int rank = 0, total = 0;
    Scan(ts).ExclusiveSum(flag, rank, total);
    if (flag) {
      const int64_t dst = (int64_t)n + done + rank;
      if (dst < cap) {
        copy_particle(A, base + i, base + dst);
        A.id[base + dst] = (next0 + done + rank - 1) * stride + offset + 1;
      }
    }
    done += total;
    __syncthreads();
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

// ----------------------------------------------------------------- halve
// One halving pass of (g, c) (aero_state_halve, src/aero_state.F90:
// 1755-1786): from the last particle down, a (g, c) particle at 1-based
// index i is removed when the first uniform of stream (step, site, i,
// RNG_HALVE) is below 1/2, logged as AERO_INFO_HALVED with other id 0, and
// the last particle moves into its slot.  Site of pass k:
// (g-1) + n_group*((c-1) + n_class*k).
__device__ inline uint32_t halve_site(const BoxDesc *B, int g, int c, int k) {
  return (uint32_t)(g - 1) +
         (uint32_t)B->w.n_group * ((uint32_t)(c - 1) + (uint32_t)B->w.n_class * (uint32_t)k);
}

__device__ inline bool halve_removes(const Arrays &A, const BoxDesc *B, int box, int i, int g,
                                     int c, uint32_t site, uint64_t seed, uint32_t step) {
  if (!in_gc(A, B->off + i, g, c)) return false;
  Stream s;
  s.open(seed, (uint32_t)box, step, site, (uint32_t)(i + 1), RNG_HALVE);
  return s.uniform() < 0.5;
}

// With current bin lists (exact mode) the removals go through
// aero_sorted_remove_particle, whose list order the next steps depend on:
// one thread walks the box in the reference's order.
__global__ void __launch_bounds__(64)
k_halve_lists(BoxDesc *boxes, Arrays A, int n_box, const int *plan, int g, int c, int k,
              uint64_t seed, uint32_t step) {
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
  const uint32_t site = halve_site(B, g, c, k);
  for (int i = B->n_part - 1; i >= 0; --i) {
    if (!halve_removes(A, B, box, i, g, c, site, seed, step)) continue;
    E.add_info(A.id[B->off + i], INFO_HALVED, 0);
    E.remove(i + 1);
  }
  B->pkeys_valid = 0;
  B->terms_kernel = 0;
  B->tables_valid = 0;
  B->w.mag[(g - 1) + B->w.n_group * (c - 1)] *= 2.0;
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
k_halve_mark(const BoxDesc *boxes, Arrays A, RebalWork R, int g, int c, int k, uint64_t seed,
             uint32_t step) {
  const int box = blockIdx.y, t = threadIdx.x;
  const BoxDesc *B = &boxes[box];
  if (!R.plan[box] || B->error) return;
  const int n = B->n_part, start = blockIdx.x * REBAL_TILE;
  if (start >= n) return;
  using Red = hipcub::BlockReduce<int, REBAL_THREADS>;
  __shared__ typename Red::TempStorage tr;
  const uint32_t site = halve_site(B, g, c, k);
  int cnt = 0, kmax = -1;
  const int end = min(n, start + REBAL_TILE);
  for (int i = start + t; i < end; i += REBAL_THREADS) {
    const bool r = halve_removes(A, B, box, i, g, c, site, seed, step);
    R.rm[B->off + i] = r ? 1 : 0;
    cnt += r ? 1 : 0;
    if (!r) kmax = max(kmax, i);
  }
  const int tot = Red(tr).Sum(cnt);
  __syncthreads();
  const int km = Red(tr).Reduce(kmax, hipcub::Max());
  if (t == 0) {
    R.tile_a[(int64_t)box * R.max_tiles + blockIdx.x] = tot;
    R.tile_kmax[(int64_t)box * R.max_tiles + blockIdx.x] = km;
  }
}

// Inclusive removal counts per slot, and the removal log in the reference's
// order (from the last index down).
__global__ void __launch_bounds__(REBAL_THREADS)
k_halve_prefix(BoxDesc *boxes, Arrays A, RebalWork R) {
  const int box = blockIdx.y, t = threadIdx.x;
  const BoxDesc *B = &boxes[box];
  if (!R.plan[box] || B->error) return;
  const int n = B->n_part, start = blockIdx.x * REBAL_TILE;
  if (start >= n) return;
  using Scan = hipcub::BlockScan<int, REBAL_THREADS>;
  __shared__ typename Scan::TempStorage ts;
  const int64_t base = B->off;
  const int n_rm = R.box_tot[box];
  int done = R.tile_off[(int64_t)box * R.max_tiles + blockIdx.x];
  for (int k = 0; k < REBAL_ITEMS; ++k) {
    const int i = start + k * REBAL_THREADS + t;
    const int flag = i < n ? R.rm[base + i] : 0;
    int incl = 0, total = 0;
    Scan(ts).InclusiveSum(flag, incl, total);
    if (i < n) {
      R.rp[base + i] = done + incl;
      if (flag) {
        const int64_t slot = B->n_info + (n_rm - (done + incl));
        if (slot < B->info_cap) {
          A.info_id[B->info_off + slot] = A.id[base + i];
          A.info_action[B->info_off + slot] = INFO_HALVED;
          A.info_other[B->info_off + slot] = 0;
        }
      }
    }
    done += total;
    __syncthreads();
  }
}

__global__ void __launch_bounds__(REBAL_THREADS)
k_halve_move(const BoxDesc *boxes, Arrays A, RebalWork R) {
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

__global__ void k_halve_finish(BoxDesc *boxes, RebalWork R, int n_box, int g, int c) {
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
  B->tables_valid = 0;
  B->w.mag[(g - 1) + B->w.n_group * (c - 1)] *= 2.0;
}

}  // namespace pmc
