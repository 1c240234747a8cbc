// pmc_rebalance.h -- population control of the boxes on the device:
// aero_state_double and aero_state_halve (src/aero_state.F90:1722-1786) and
// the per-(group, class) counts aero_state_rebalance decides from
// (src/aero_state.F90:287-310, 1790-1852).  The host sequences the passes
// in the reference's (group, class) order; each pass runs every box that
// needs it at once.
#pragma once

#include <hipcub/hipcub.hpp>

#include "pmc_exact.h"

namespace pmc {

constexpr int GC_KEYS = MAX_GROUP * MAX_CLASS;  // counts per box, then n_part
constexpr int REBAL_THREADS = 256;
constexpr int INFO_HALVED = 3;                  // AERO_INFO_HALVED, src/aero_info.F90:25

// counts[box][(g-1) + n_group*(c-1)] = particles of (g, c);
// counts[box][GC_KEYS] = n_part.
__global__ void __launch_bounds__(REBAL_THREADS)
k_count_gc(BoxDesc *boxes, Arrays A, int *counts) {
  const int box = blockIdx.x, t = threadIdx.x;
  const BoxDesc *B = &boxes[box];
  __shared__ int s_cnt[GC_KEYS];
  for (int k = t; k < GC_KEYS; k += REBAL_THREADS) s_cnt[k] = 0;
  __syncthreads();
  const int n = B->n_part, ng = B->w.n_group, nc = B->w.n_class;
  for (int i = t; i < n; i += REBAL_THREADS) {
    const int64_t p = B->off + i;
    const int g = A.group[p], c = A.klass[p];
    if (g >= 1 && g <= ng && c >= 1 && c <= nc) atomicAdd(&s_cnt[(g - 1) + ng * (c - 1)], 1);
  }
  __syncthreads();
  int *out = counts + (int64_t)box * (GC_KEYS + 1);
  for (int k = t; k < GC_KEYS; k += REBAL_THREADS) out[k] = s_cnt[k];
  if (t == 0) out[GC_KEYS] = n;
}

// One doubling pass of (g, c) in every box with plan[box] > pass
// (aero_state_double, src/aero_state.F90:1722-1750): the box's (g, c)
// particles present at the start are appended in index order with
// consecutive new ids (aero_particle_new_id, src/aero_particle.F90:157-165),
// the sort is invalidated and the magnitude halved.  One block per box; a
// block scan over chunks gives each copy its slot.
__global__ void __launch_bounds__(REBAL_THREADS)
k_double_pass(BoxDesc *boxes, Arrays A, const int *plan, int g, int c, int pass) {
  const int box = blockIdx.x, t = threadIdx.x;
  BoxDesc *B = &boxes[box];
  if (plan[box] <= pass || B->error) return;
  using Scan = hipcub::BlockScan<int, REBAL_THREADS>;
  __shared__ typename Scan::TempStorage ts;
  __shared__ int s_overflow;
  if (t == 0) s_overflow = 0;
  const int n = B->n_part;
  const int64_t base = B->off, cap = B->cap;
  const int64_t next0 = B->next_id, stride = B->id_stride, offset = B->id_offset;
  __syncthreads();
  int done = 0;
  for (int c0 = 0; c0 < n; c0 += REBAL_THREADS) {
    const int i = c0 + t;
    const int flag = (i < n && A.group[base + i] == g && A.klass[base + i] == c) ? 1 : 0;
    int rank = 0, total = 0;
    Scan(ts).ExclusiveSum(flag, rank, total);
    if (flag) {
      const int64_t dst = (int64_t)n + done + rank;
      if (dst < cap) {
        copy_particle(A, base + i, base + dst);
        A.id[base + dst] = (next0 + done + rank - 1) * stride + offset + 1;
      } else {
        s_overflow = 1;
      }
    }
    done += total;
    __syncthreads();
  }
  if (t == 0) {
    if (s_overflow) {
      B->error = ERR_CAPACITY;
      return;
    }
    B->n_part = n + done;
    B->next_id = next0 + done;
    B->valid_sort = 0;
    B->xlists_valid = 0;
    B->pkeys_valid = 0;
    B->terms_kernel = 0;
    B->tables_valid = 0;
    B->w.mag[(g - 1) + B->w.n_group * (c - 1)] *= 0.5;
  }
}

// One halving pass of (g, c) in every box with plan[box] != 0
// (aero_state_halve, src/aero_state.F90:1755-1786): from the last particle
// down, a (g, c) particle at 1-based index i is removed when the first
// uniform of stream (step, site, i, RNG_HALVE) is below 1/2, logged as
// AERO_INFO_HALVED with other id 0.  The swap-last removal goes through
// the bin lists when they are current (aero_sorted_remove_particle), so the
// particle order and the lists stay the reference's.  Site of pass k:
// (g-1) + n_group*((c-1) + n_class*k).  One thread per box: the removal
// order is sequential by definition.
__global__ void __launch_bounds__(64)
k_halve_pass(BoxDesc *boxes, Arrays A, int n_box, const int *plan, int g, int c, int k,
             uint64_t seed, uint32_t step) {
  const int box = blockIdx.x * blockDim.x + threadIdx.x;
  if (box >= n_box) return;
  BoxDesc *B = &boxes[box];
  if (!plan[box] || B->error) return;
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
  const int ng = B->w.n_group, nc = B->w.n_class;
  const uint32_t site = (uint32_t)(g - 1) + (uint32_t)ng * ((uint32_t)(c - 1) + (uint32_t)nc * (uint32_t)k);
  const bool lists = B->valid_sort && B->xlists_valid;
  const int64_t base = B->off;
  for (int i = B->n_part - 1; i >= 0; --i) {
    const int64_t p = base + i;
    if (A.group[p] != g || A.klass[p] != c) continue;
    Stream s;
    s.open(seed, (uint32_t)box, step, site, (uint32_t)(i + 1), RNG_HALVE);
    if (s.uniform() < 0.5) {
      E.add_info(A.id[p], INFO_HALVED, 0);
      if (lists) {
        E.remove(i + 1);
      } else {
        const int last = B->n_part - 1;
        if (i < last) copy_particle(A, base + last, p);
        B->n_part = last;
      }
    }
  }
  if (!lists) B->xlists_valid = 0;
  B->pkeys_valid = 0;
  B->terms_kernel = 0;
  B->tables_valid = 0;
  B->w.mag[(g - 1) + ng * (c - 1)] *= 2.0;
}

}  // namespace pmc
