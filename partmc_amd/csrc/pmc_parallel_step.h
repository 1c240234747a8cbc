// pmc_parallel_step.h -- the parallel-mode coagulation step.
//
// Per step and for every box at once:
//   1. keys: each particle caches its (bin, class) list key, its total
//      volume and (FP32 mode) its FP64-derived kernel terms; they change
//      only when the particle is created, the grid is remade, or the
//      environment or kernel type changes.  Live list counts are kept
//      exact by atomics, so aero_sorted_remake_if_needed's occupancy test
//      (src/aero_sorted.F90:264-292,341-351) needs no particle pass.
//   2. sort: a stable counting sort of the keys (tile histograms, one
//      device-wide scan, tile-local radix ranks) gives every (bin, class)
//      list in particle order.
//   3. pairs: one thread per bin pair draws n_samp ~ Poisson(k_max dt n1 n2)
//      (compute_n_samp, src/coagulation.F90:648-702) from the pair's own
//      stream; a device scan lays the tests of all pairs of all boxes out in
//      the reference's loop order (src/coagulation.F90:74-94).
//   4. tests: one thread per test pair draws its two entries
//      (find_rand_pair, :766-803), evaluates the weighted kernel (FP64, or
//      FP32 with an FP64 recheck inside the band) and accepts with
//      probability k/k_max (maybe_coag_pair, :712-759).
//   5. claims: an accepted test claims both particles with atomicMin of its
//      sequential index; a test wins when it holds both claims, i.e. no
//      earlier accepted test in the reference order touched either
//      particle.  Winners merge (coagulate_weighting + coagulate,
//      :810-985); losers are re-drawn in another round against the
//      particles still untouched this step.
//   6. compaction: slots freed by merges are filled from the tail.
// The step is deterministic: every draw is addressed by (box, step, site,
// ordinal) and every conflict is resolved by sequential index.
#pragma once

#include <rocprim/block/block_load.hpp>
#include <rocprim/block/block_radix_sort.hpp>
#include <rocprim/device/device_radix_sort.hpp>

#include "pmc_scan.h"

#include <string>
#include <vector>

namespace pmc {

constexpr int TILE_THREADS = 256;
constexpr int TILE_ITEMS = 16;
constexpr int TILE = TILE_THREADS * TILE_ITEMS;
constexpr uint32_t UNCLAIMED = 0xFFFFFFFFu;
constexpr uint8_t ST_LIVE = 0, ST_HOLE = 1, ST_NEW = 2;
constexpr int PAR_ROUNDS = 2;  // device-wide claim rounds before the per-box tail
static_assert(PAR_ROUNDS >= 1, "the tail runs on the pending list of a round");
constexpr int MAX_REDRAW = 64;

struct PW {  // device view of the parallel work arrays
  int K_stride;     // keys per box of the list tables (the allocation's grid)
  int K_hist;       // keys per box of this step's histogram (the step's grid)
  int64_t P_stride; // pair slots per box
  int tab_classes, max_bin;
  int *key;         // [total_cap] list key, -1 for no key
  uint8_t *status;  // [total_cap] (epoch << 2) | state; another step's epoch reads as live
  int epoch;        // this step's status epoch, 1..63
  uint32_t *claim;  // [total_cap]
  uint32_t *claim_wr;  // [total_cap] tail: earliest claim of a test that changes the particle
  uint32_t *dead_t; // [total_cap] test index of the winner that consumed it
  int *dead_r;      // [total_cap] claim round (step-wide count) of that winner
  float4 *terms;    // [total_cap]
  int *perm;        // [total_cap] global slots sorted by (box, key)
  int *tile_box, *tile_start, *tiles_in_box;
  int64_t *tile_first;  // [n_box] first tile of the box; its histogram starts at K_hist * this
  uint32_t *hist, *hist_scan;
  int *lstart, *lcnt, *live;  // [n_box * K_stride]
  uint32_t *pair_n, *pair_exc;  // [n_box * P_stride (+1)]
  int *acc, *acc_scan, *tp1, *tp2;
  uint32_t *tid, *pend;
  uint32_t *ev_t;
  int *ev_p1, *ev_p2, *ev_box, *ev_out;
  int *f_app, *f_id, *f_info, *f_hole, *f_lose, *f_pend, *s_pend;
  int *s_app, *s_id, *s_info, *s_hole, *s_lose;
  int *box_first;
  uint64_t *holes, *holes_sorted;  // (box << 32 | slot)
  int *cmp_lo, *cmp_hi, *cmp_head;
  unsigned long long *acc_samp;  // samples drawn by the device-wide target walks  // [n_box] compaction: hole range and holes below the final count
  // accelerated per-particle path (try_per_particle_coag): one test slot per
  // target particle, one event record per target slot of a batch
  uint8_t *pair_acc;           // [NP] 0 per-set; 1 targets in (b1, c1); 2 targets in (b2, c2)
  uint32_t *pair_accn;         // [NP + 1] target slots (first accelerated pair of a list)
  uint32_t *acc_exc;           // [NP + 1] their exclusive scan
  int *trec;                   // [T_cap] record of a test slot, -1 for per-set tests
  int *tns;                    // [T_cap] n_samp (+ ACC_WARN) of a target without effect
  int *ev_rec;                 // [T_cap] record of an event, -1 for per-set events
  int acc_cap, pool_cap;
  // event records [acc_cap]
  int *r_nsamp, *r_ncoag, *r_nrm, *r_nchain, *r_head, *r_ndup, *r_dhead, *r_flags,
      *r_err;
  int *r_rdhead, *r_nrd;       // distinct sources read from small source lists
  // particle pool [pool_cap]: entry r < acc_cap is the final target of
  // record r; entries from acc_cap on are the copies (aero_state_dup_particle)
  double *q_vol;               // [n_spec][pool_cap]
  int *q_group, *q_nprim, *q_ncomp, *q_next;
  int *q_csrc;                 // [MAX_COMP][pool_cap]
  double *q_ctime;             // [MAX_COMP][pool_cap]
  double *q_least, *q_greatest;
  int *dup_top;                // [1] copy allocator of the round
  int *rm_buf;                 // chunks: [0] next chunk or -1, [1..RM_CHUNK) removed slots
  int *rm_top;                 // [1] chunk allocator of the round
  int rm_cap;
  uint32_t *misc;              // [8] batch bounds from k_batch_end
  unsigned long long *hole_top;  // [8] hole counter; then tail_rounds and rstat
  int *pend_k;                 // [T_cap] pair of each pending test entering the tail
  uint32_t *tp_in, *tpair;     // [T_cap] pair of every test of a single-batch step
  int tpair_ok;                // tpair is current for this step
  unsigned long long *tail_rounds;  // [1] claim rounds run by the tail, summed over boxes
  // Particles that join a list within the step ("arrivals": a merged
  // particle that lands in another (bin, class) than its slot's list, an
  // appended particle).  The draws of every pair reserve acap entries after
  // each list's n step-start members; the j-th reserved entry is the list's
  // j-th arrival before the pair begins (in sequence order).
  int *acap;          // [n_box * K_stride] reserved arrival entries of the list this step
  int *arr_n;         // [n_box * K_stride] arrivals recorded this step
  int *arr_head;      // [n_box * K_stride] the list's arrivals beyond A_MAX: first entry, -1 none
  // arrival pool: list L's first A_MAX arrivals at L * A_MAX + i, the rest
  // from ap_ovf on, chained from arr_head[L]
  int *ap_p;          // the arrival's slot
  uint32_t *ap_t;     // sequence index at which it joined the list
  uint32_t *ap_dep;   // at which it left again (UNCLAIMED: still there)
  int *ap_next;       // next overflow arrival of the same list
  uint32_t *ap_top;   // [1] overflow entries used this step
  int64_t ap_ovf;     // n_box * K_stride * A_MAX
  uint32_t ap_cap;    // overflow entries
  int *arr_of;        // [total_cap] pool entry of the arrival a slot holds
  int *key0;          // [total_cap] a touched slot's list key at step start
  uint32_t *depart_t; // [total_cap] when a touched slot's step-start particle left that list
  uint8_t *thin;      // [T_cap] the test's last draw was thinned (per batch)
  // Per (box, bin): the lowest index of an event pending at the start of
  // this round whose results land in that bin or a lower one (tests on an
  // empty arrival entry excluded; particles only grow, so an event on bins
  // b1, b2 adds particles to bins from max(b1, b2) up), as a running
  // minimum over the bins; and the same, per bin, for the next round.
  uint32_t *minpend, *minpend_next;  // [n_box * K_stride] each
  uint32_t *scan_part;  // [SCAN_COLS * SCAN_MAX_BLOCKS] chunk totals of the device-wide scans
  // The counts that drive a batch's claim rounds, kept on the device so the
  // host launches every round without reading them back: RC_N tests of the
  // current device-wide round (0: the round does not run), RC_E its events,
  // RC_TAIL the tests left to the tail.
  int *rc;                      // [RC_SIZE]
  unsigned long long *pstat;    // [2] device-wide rounds run, tests left to the tail
  // the batch's round sizes, read back with the hole count after the tail to
  // size the next step's grids: round 0's events, round 1's tests and
  // events, the tail's tests (in the allocation of hole_top, after
  // tail_rounds)
  unsigned long long *rstat;    // [4]
};

constexpr int RC_N = 0, RC_E = 1, RC_TAIL = 2, RC_SIZE = 4;

constexpr int ACC_CAP = 1 << 18;
constexpr int DUP_CAP = 1 << 18;
constexpr int RM_CHUNK = 16;
constexpr int CHAIN_WAVE_MIN = 32;  // removal chains longer than this are applied by a wavefront
constexpr int RM_CAP = 1 << 24;
constexpr int ACC_WARN = 1 << 30;
constexpr int READ_TRACK = 128;  // source lists up to this size have their reads tracked
constexpr int RF_WARN = 1, RF_DIED = 2, RF_MERGED = 4;
constexpr int A_MAX = 128;      // arrival entries a list can reserve
constexpr int ARR_REACH = 3;    // an empty list reserves entries when a list at most this many bins below is occupied
constexpr int TNS_ATHIN = 1 << 29;  // tns bit of a test or target event on an empty arrival entry
constexpr int TNS_SIDE2 = 1 << 28;  // ... of the pair's second list

struct ParallelWork {
  int n_box = 0;
  int64_t T_cap = 1 << 22;
  int64_t hole_cap = 0;
  int64_t H = 0, NP = 0;
  int n_tiles = 0;
  PW d{};
  void *sort_tmp = nullptr;
  size_t sort_tmp_bytes = 0;
  std::vector<int> tiles_in_box;
  uint32_t *h_u32 = nullptr;  // pinned scalars
  int *h_i32 = nullptr;
  unsigned long long *h_u64 = nullptr;
  // optional per-kernel timing hooks of the owning handle
  void (*tbegin)(void *, const char *, hipEvent_t *) = nullptr;
  void (*tend)(void *, const char *, hipEvent_t) = nullptr;
  void *tctx = nullptr;
  hipStream_t tst = nullptr;  // stream of the launch being timed, when not the handle's
  int epoch = 0;              // status epoch of the last step (0: the array is clear)
  // a second stream (default priority) for the target walks, which run
  // beside the per-set tests of the same round
  hipStream_t st2 = nullptr;
  hipEvent_t ev_fork = nullptr, ev_join = nullptr;
  int64_t n_tail_rounds = 0;  // the tail's claim rounds, summed over boxes
  int64_t rid = 1;        // round identifiers: a fresh one for every claim round
  unsigned long long *pstat = nullptr;  // [2] device-wide rounds run, tests left to the tail
  // the last batch's round 0 events, round 1 tests and events, tail tests
  // (-1: none yet): the grids of the next step's rounds are sized from them
  int64_t est[4] = {-1, -1, -1, -1};
};

#define PLAUNCH(W, name, ...)                                   \
  do {                                                          \
    hipEvent_t ev_;                                             \
    if ((W).tbegin) (W).tbegin((W).tctx, name, &ev_);           \
    hipLaunchKernelGGL(__VA_ARGS__);                            \
    if ((W).tend) (W).tend((W).tctx, name, ev_);                \
  } while (0)

static std::string g_parallel_error;
inline const char *parallel_last_error() { return g_parallel_error.c_str(); }

#define PCHECK(expr)                                                       \
  do {                                                                     \
    hipError_t e_ = (expr);                                                \
    if (e_ != hipSuccess) {                                                \
      g_parallel_error = std::string(#expr) + ": " + hipGetErrorString(e_); \
      return ERR_HIP;                                                      \
    }                                                                      \
  } while (0)

template <class T>
static hipError_t pmalloc(T **p, size_t n) {
  return hipMalloc((void **)p, (n > 0 ? n : 1) * sizeof(T));
}

// ------------------------------------------------------------ kernels
__device__ inline int par_key(int bin, int cls, int nb) { return (bin - 1) + nb * (cls - 1); }

// keys (and optionally FP32 terms and cached volumes) for every particle of
// boxes whose keys are stale; terms for boxes whose terms are stale
__global__ void __launch_bounds__(TILE_THREADS)
k_par_rekey(BoxDesc *boxes, Arrays A, PW W, int type, int want_terms, int phase) {
  const int tile = blockIdx.x;
  const int box = W.tile_box[tile];
  BoxDesc *B = &boxes[box];
  if (B->error) return;
  const bool do_keys = phase == 0 ? (!B->pkeys_valid && B->valid_sort && B->n_bin > 0)
                                  : (B->rekey != 0);
  const bool do_terms = want_terms && phase == 0 && B->terms_kernel != type;
  if (!do_keys && !do_terms) return;
  const int n = B->n_part;
  const int nb = B->n_bin;
  const Env env = B->env;
  const int64_t base = B->off;
  int *live = W.live + (int64_t)box * W.K_stride;
  for (int it = 0; it < TILE_ITEMS; ++it) {
    int i = W.tile_start[tile] + it * TILE_THREADS + threadIdx.x;
    if (i >= n) continue;
    int64_t p = base + i;
    double v = particle_volume(A, p);
    A.vtot[p] = v;
    if (do_keys) {
      int b = bin_find_log(nb, B->log_lo, B->log_hi, vol2rad(v));
      if (b < 1 || b > nb) {
        atomicExch(&B->error, 708651144);
        continue;
      }
      int cls = A.klass[p];
      if (cls < 1 || cls > B->w.n_class) {
        atomicExch(&B->error, 233133947);  // weight class out of range
        continue;
      }
      int k = par_key(b, cls, nb);
      W.key[p] = k;
      atomicAdd(&live[k], 1);
    }
    if (do_terms) {
      double m = 0.0;
      for (int s = 0; s < A.n_spec; ++s) m += A.vol[(int64_t)s * A.total_cap + p] * A.density[s];
      W.terms[p] = particle_terms(type, v, m / v, env);
    }
  }
}

__global__ void k_par_zero_live(BoxDesc *boxes, PW W, int n_box, int phase) {
  int box = blockIdx.y;
  BoxDesc *B = &boxes[box];
  bool do_keys = phase == 0 ? (!B->pkeys_valid && B->valid_sort && B->n_bin > 0) : (B->rekey != 0);
  if (!do_keys || B->error) return;
  for (int k = threadIdx.x; k < W.K_stride; k += blockDim.x) W.live[(int64_t)box * W.K_stride + k] = 0;
}

__global__ void k_par_finish_rekey(BoxDesc *boxes, int n_box, int type, int want_terms, int phase) {
  int box = blockIdx.x * blockDim.x + threadIdx.x;
  if (box >= n_box) return;
  BoxDesc *B = &boxes[box];
  if (phase == 0) {
    if (!B->pkeys_valid && B->valid_sort && B->n_bin > 0) B->pkeys_valid = 1;
    // FP64 steps do not keep the terms of new particles current
    B->terms_kernel = want_terms ? type : 0;
  } else if (B->rekey) {
    B->rekey = 0;
    B->pkeys_valid = 1;
  }
}

// radius extremes of boxes without a valid sort (r_min/r_max loop of
// src/aero_sorted.F90:293-304); positive doubles order like their bits
__global__ void __launch_bounds__(TILE_THREADS)
k_par_minmax(BoxDesc *boxes, Arrays A, PW W) {
  const int tile = blockIdx.x;
  const int box = W.tile_box[tile];
  BoxDesc *B = &boxes[box];
  if (B->error || (B->valid_sort && B->pkeys_valid)) return;
  const int n = B->n_part;
  unsigned long long lo = ~0ull, hi = 0ull;
  for (int it = 0; it < TILE_ITEMS; ++it) {
    int i = W.tile_start[tile] + it * TILE_THREADS + threadIdx.x;
    if (i >= n) continue;
    double r = vol2rad(particle_volume(A, B->off + i));
    unsigned long long bits = (unsigned long long)__double_as_longlong(r);
    lo = min(lo, bits);
    hi = max(hi, bits);
  }
  if (hi) {
    atomicMin(&B->rmin_bits, lo);
    atomicMax(&B->rmax_bits, hi);
  }
}

__global__ void k_par_reset_minmax(BoxDesc *boxes, int n_box) {
  int box = blockIdx.x * blockDim.x + threadIdx.x;
  if (box >= n_box) return;
  boxes[box].rmin_bits = ~0ull;
  boxes[box].rmax_bits = 0ull;
}

// aero_sorted_remake_if_needed, src/aero_sorted.F90:214-373, one thread
// per box, occupancy from the live counts
__global__ void k_par_remake(BoxDesc *boxes, Arrays A, PW W, int n_box) {
  int box = blockIdx.x * blockDim.x + threadIdx.x;
  if (box >= n_box) return;
  BoxDesc *B = &boxes[box];
  B->regridded = 0;
  B->rekey = 0;
  if (B->error) return;
  double *edges = A.edges + (int64_t)box * (A.max_bin + 1);
  double *centers = A.centers + (int64_t)box * A.max_bin;
  double *widths = A.widths + (int64_t)box * A.max_bin;
  const int nb = B->n_bin, nc = B->w.n_class;
  const bool use_counts = B->valid_sort && B->pkeys_valid;
  double rmin = 0.0, rmax = 0.0;
  if (use_counts) {
    const int *live = W.live + (int64_t)box * W.K_stride;
    int i_min = 0, i_max = 0;
    for (int b = 1; b <= nb; ++b) {
      int tot = 0;
      for (int c = 1; c <= nc; ++c) tot += live[par_key(b, c, nb)];
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
  } else if (B->valid_sort && nb > 0) {
    // a valid sort whose keys were dropped (population control moved
    // particles): the occupied bins are those of the radius extremes
    rmin = edges[nb];
    rmax = edges[0];
    if (B->n_part > 0) {
      const int b_lo = bin_find_log(nb, B->log_lo, B->log_hi,
                                    __longlong_as_double((long long)B->rmin_bits));
      const int b_hi = bin_find_log(nb, B->log_lo, B->log_hi,
                                    __longlong_as_double((long long)B->rmax_bits));
      rmin = edges[max(b_lo, 1) - 1];
      rmax = edges[min(b_hi, nb)];
    }
  } else if (B->n_part > 0) {
    rmin = __longlong_as_double((long long)B->rmin_bits);
    rmax = __longlong_as_double((long long)B->rmax_bits);
  }
  if (rmax == 0.0) {
    if (!use_counts) {
      B->n_bin = 0;
      B->bounds_valid = 0;
      B->tables_valid = 0;
    }
    B->valid_sort = 1;
    B->pkeys_valid = 1;
    return;
  }
  bool need = (nb < 1) || (rmin / edges[0] < 3.0) || (edges[nb] / rmax < 3.0);
  if (need) {
    double gmin = rmin / 10.0, gmax = rmax * 10.0;
    double nbf = ceil((log10(gmax) - log10(gmin)) * 10.0);
    if (nbf < 0.0) {
      B->error = 538534122;
      return;
    }
    if (nbf > (double)A.max_bin) {
      B->error = ERR_CAPACITY;
      return;
    }
    int nn = (int)nbf;
    make_log_grid(nn, gmin, gmax, edges, centers, widths);
    B->n_bin = nn;
    B->log_lo = log(edges[0]);
    B->log_hi = log(edges[nn]);
    B->bounds_valid = 0;
    B->tables_valid = 0;
    B->regridded = 1;
    B->rekey = 1;
  } else if (!use_counts) {
    B->rekey = 1;
  }
  B->valid_sort = 1;
  B->xlists_valid = 0;
}

// tile histogram of keys into hist[box][key][tile]
__global__ void __launch_bounds__(TILE_THREADS)
k_par_hist(BoxDesc *boxes, PW W) {
  const int tile = blockIdx.x;
  const int box = W.tile_box[tile];
  BoxDesc *B = &boxes[box];
  extern __shared__ int s_h[];
  const int K = W.K_hist;
  for (int k = threadIdx.x; k < K; k += TILE_THREADS) s_h[k] = 0;
  __syncthreads();
  const int n = B->error ? 0 : B->n_part;
  const int64_t base = B->off;
  for (int it = 0; it < TILE_ITEMS; ++it) {
    int i = W.tile_start[tile] + it * TILE_THREADS + threadIdx.x;
    if (i < n) atomicAdd(&s_h[W.key[base + i]], 1);
  }
  __syncthreads();
  const int t_in_box = W.tile_start[tile] / TILE;
  const int nt = W.tiles_in_box[box];
  uint32_t *h = W.hist + (int64_t)K * W.tile_first[box];
  for (int k = threadIdx.x; k < K; k += TILE_THREADS) h[(int64_t)k * nt + t_in_box] = (uint32_t)s_h[k];
}

// list starts and counts of every (box, key) from the scanned histogram
__global__ void k_par_lists(BoxDesc *boxes, PW W, int n_box) {
  const int box = blockIdx.y;
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= W.K_stride) return;
  const int nt = W.tiles_in_box[box];
  const int K = W.K_hist;
  const int64_t o = (int64_t)K * W.tile_first[box];
  const uint32_t base = W.hist_scan[o];
  // keys beyond the step's grid hold no particle: their lists start at the
  // end of the box's, as the last key's list ends
  const int kk = k < K ? k : K - 1;
  const uint32_t e = W.hist_scan[o + (int64_t)kk * nt + nt - 1] + W.hist[o + (int64_t)kk * nt + nt - 1];
  const uint32_t s = k < K ? W.hist_scan[o + (int64_t)k * nt] : e;
  W.lstart[(int64_t)box * W.K_stride + k] = (int)(s - base);
  W.lcnt[(int64_t)box * W.K_stride + k] = (int)(e - s);
}

// Stable scatter.  The tile's keys are loaded coalesced and transposed to a
// blocked arrangement (item order = index order), radix-sorted (stable), and
// each item's rank among equal keys of the tile is its sorted position minus
// that of the key's first item.  The first position and the list offset of
// every key present are kept in shared memory (dynamic, 2 * K_stride ints),
// written by the first item of each run.
__global__ void __launch_bounds__(TILE_THREADS)
k_par_scatter(BoxDesc *boxes, PW W, int key_bits) {
  using BLD = rocprim::block_load<int, TILE_THREADS, TILE_ITEMS,
                                  rocprim::block_load_method::block_load_transpose>;
  using BRS = rocprim::block_radix_sort<int, TILE_THREADS, TILE_ITEMS, int>;
  __shared__ union {
    typename BLD::storage_type load;
    typename BRS::storage_type sort;
  } tmp;
  __shared__ int s_keys[TILE];
  extern __shared__ int s_run[];  // [0, K): first sorted position; [K, 2K): list offset
  const int tile = blockIdx.x;
  const int box = W.tile_box[tile];
  BoxDesc *B = &boxes[box];
  const int n = B->error ? 0 : B->n_part;
  const int64_t base = B->off;
  const int sentinel = (1 << key_bits) - 1;
  const int start = W.tile_start[tile];
  const int valid = max(0, min(TILE, n - start));
  int keys[TILE_ITEMS], vals[TILE_ITEMS];
  // coalesced loads, exchanged so that each thread holds TILE_ITEMS
  // consecutive keys
  BLD().load(W.key + base + start, keys, (unsigned)valid, sentinel, tmp.load);
  for (int it = 0; it < TILE_ITEMS; ++it) vals[it] = start + threadIdx.x * TILE_ITEMS + it;
  __syncthreads();
  BRS().sort(keys, vals, tmp.sort, 0, key_bits);
  __syncthreads();
  for (int it = 0; it < TILE_ITEMS; ++it) s_keys[threadIdx.x * TILE_ITEMS + it] = keys[it];
  __syncthreads();
  const int K = W.K_stride;
  const int t_in_box = start / TILE;
  const int nt = W.tiles_in_box[box];
  const int64_t o = (int64_t)W.K_hist * W.tile_first[box];
  const uint32_t bbase = W.hist_scan[o];
  for (int it = 0; it < TILE_ITEMS; ++it) {
    const int pos = threadIdx.x * TILE_ITEMS + it;
    const int k = keys[it];
    if (k == sentinel) continue;
    if (pos == 0 || s_keys[pos - 1] != k) {
      s_run[k] = pos;
      s_run[K + k] = (int)(W.hist_scan[o + (int64_t)k * nt + t_in_box] - bbase);
    }
  }
  __syncthreads();
  for (int it = 0; it < TILE_ITEMS; ++it) {
    const int pos = threadIdx.x * TILE_ITEMS + it;
    const int k = keys[it];
    if (k == sentinel) continue;
    const int dst = s_run[K + k] + (pos - s_run[k]);
    W.perm[base + dst] = (int)(base + vals[it]);
  }
}

__device__ inline void decode_pair(int64_t q, int tc, int mb, int *c1, int *c2, int *b1, int *b2) {
  int64_t mb2 = (int64_t)mb * mb;
  int cls = (int)(q / mb2);
  int rem = (int)(q % mb2);
  *c1 = cls / tc + 1;
  *c2 = cls % tc + 1;
  *b1 = rem / mb + 1;
  *b2 = rem % mb + 1;
}

// the record without its components (they are read in a second pass
// that depends on ncomp, so a reader that may not need them defers it)
__device__ inline void load_rec_head(const Arrays &A, int64_t p, Rec &r) {
  for (int s = 0; s < A.n_spec; ++s) r.vol[s] = A.vol[(int64_t)s * A.total_cap + p];
  r.id = A.id[p];
  r.group = A.group[p];
  r.klass = A.klass[p];
  r.nprim = A.nprim[p];
  r.ncomp = A.ncomp[p];
  r.least = A.least[p];
  r.greatest = A.greatest[p];
}

__device__ inline void load_rec_comps(const Arrays &A, int64_t p, Rec &r) {
  for (int c = 0; c < r.ncomp; ++c) {
    r.csrc[c] = A.csrc[comp_at(p, c)];
    r.ctime[c] = A.ctime[comp_at(p, c)];
  }
}

__device__ inline void load_rec(const Arrays &A, int64_t p, Rec &r) {
  load_rec_head(A, p, r);
  load_rec_comps(A, p, r);
}

// aero_particle_coagulate for the parallel path (same rules as rec_coagulate)
__device__ inline int par_rec_coagulate(int n_spec, Stream &s, const Rec &p1, const Rec &p2,
                                        Rec &o) {
  for (int k = 0; k < n_spec; ++k) o.vol[k] = p1.vol[k] + p2.vol[k];
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
    int n1_new = prob_round(s, (double)p1.nprim / (double)(p1.nprim + p2.nprim) * (double)MAX_COMP);
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
  return 0;
}

// try_per_particle_coag's tests (src/coagulation.F90:150-215): monotone
// decreasing weights, target weight strictly below the source weight with
// the destination class the target's, at least one expected source per
// target, and at most 1.5x growth.  Returns 0 (per-set path), 1 (targets in
// (b1, c1)) or 2 (targets in (b2, c2)).
__device__ inline int accel_kind(const Weights &w, const double *edges, int b1, int b2, int c1,
                                 int c2, int cc, double k_max, double del_t, const int *cnt, int nb,
                                 double *n_spt_out) {
  for (int i = 0; i < w.n_group * w.n_class; ++i)
    if (w.type[i] == WEIGHT_POWER && w.expo[i] > 0.0) return 0;
  double a1 = nc_at_radius(w, c1, edges[b1 - 1]), a2 = nc_at_radius(w, c1, edges[b1]);
  double d1 = nc_at_radius(w, c2, edges[b2 - 1]), d2 = nc_at_radius(w, c2, edges[b2]);
  double nc1_min = fmin(a1, a2), nc1_max = fmax(a1, a2);
  double nc2_min = fmin(d1, d2), nc2_max = fmax(d1, d2);
  int kind = 0, bt = 0, bs = 0, cs = 0;
  if (nc1_max < nc2_min && cc == c1) {
    kind = 1; bt = b1; bs = b2; cs = c2;
  }
  if (nc2_max < nc1_min && cc == c2) {
    kind = 2; bt = b2; bs = b1; cs = c1;
  }
  if (!kind) return 0;
  double n_spt = k_max * del_t * (double)cnt[par_key(bs, cs, nb)];
  if (n_spt < 1.0) return 0;
  double min_tv = rad2vol(edges[bt - 1]);
  double max_sv = rad2vol(edges[bs]);
  if ((min_tv + max_sv * n_spt) / min_tv > 1.5) return 0;
  *n_spt_out = n_spt;
  return kind;
}

// n_samp of every (box, c1, c2, b1, b2), compute_n_samp
// src/coagulation.F90:648-702 on the pair's own Poisson stream
__global__ void k_par_pairs(BoxDesc *boxes, Arrays A, PW W, double del_t, uint64_t seed,
                            uint32_t step) {
  const int box = blockIdx.y;
  const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= W.P_stride) return;
  BoxDesc *B = &boxes[box];
  uint32_t ns = 0;
  int acc = 0;
  uint32_t na = 0;  // target events of an accelerated pair
  if (!B->error) {
    int c1, c2, b1, b2;
    decode_pair(q, W.tab_classes, W.max_bin, &c1, &c2, &b1, &b2);
    const int nb = B->n_bin, nc = B->w.n_class;
    if (c1 <= nc && c2 <= c1 && b1 <= nb && b2 <= nb && (c1 != c2 || b2 >= b1)) {
      const int *cnt = W.lcnt + (int64_t)box * W.K_stride;
      int n1 = cnt[par_key(b1, c1, nb)], n2 = cnt[par_key(b2, c2, nb)];
      const int64_t ti = table_index(A, box, nb, nc, b1, b2, c1, c2);
      double k_max = A.kw[ti];
      const double *edges = A.edges + (int64_t)box * (A.max_bin + 1);
      double n_spt;
      // eligibility needs only the source list: a target that grows into
      // an empty bin during the step still visits that bin's pairs
      acc = accel_kind(B->w, edges, b1, b2, c1, c2, A.cc[ti], k_max, del_t, cnt, nb, &n_spt);
      if (acc) {
        // one event per entry of the target list when the loop reaches the
        // pair: its reserved arrival entries, then its step-start members;
        // none when the source list is empty
        if (n1 > 0 && n2 > 0) atomicAdd((unsigned long long *)&B->n_accel_pairs, 1ull);
        const int kt = acc == 1 ? par_key(b1, c1, nb) : par_key(b2, c2, nb);
        const int n_src = acc == 1 ? n2 : n1;
        const int *acp = W.acap + (int64_t)box * W.K_stride;
        if (n_src > 0) na = (uint32_t)(cnt[kt] + acp[kt]);
        ns = na;
      } else {
        // the bound on the lists when the loop reaches the pair: step-start
        // members plus reserved arrival entries (draws thin to the members)
        const int *acp = W.acap + (int64_t)box * W.K_stride;
        const double N1 = (double)(n1 + acp[par_key(b1, c1, nb)]);
        const double N2 = (double)(n2 + acp[par_key(b2, c2, nb)]);
        bool same = (b1 == b2) && (c1 == c2);
        double n_possible = same ? N1 * (N2 - 1.0) / 2.0 : N1 * N2;
        double mean = k_max * del_t * n_possible;
        Stream ps;
        ps.open(seed, box, step, pair_site(nb, nc, b1, b2, c1, c2), 0, RNG_POISSON);
        ns = n_possible > 0.0 ? (uint32_t)rand_poisson(ps, mean) : 0u;
        if (ns) {
          atomicAdd((unsigned long long *)&B->n_samp, (unsigned long long)ns);
          atomicAdd((unsigned long long *)&B->step_samp, (unsigned long long)ns);
        }
      }
    }
  }
  W.pair_n[(int64_t)box * W.P_stride + q] = ns;
  W.pair_acc[(int64_t)box * W.P_stride + q] = (uint8_t)acc;
  W.pair_accn[(int64_t)box * W.P_stride + q] = na;
}

__device__ inline int64_t find_pair(const uint32_t *exc, int64_t n, uint32_t t) {
  // last index k with exc[k] <= t  (exc has n + 1 entries, exc[n] = total)
  int64_t lo = 0, hi = n;
  while (lo < hi) {
    int64_t mid = (lo + hi + 1) >> 1;
    if (exc[mid] <= t) lo = mid; else hi = mid - 1;
  }
  return lo;
}

// the pair of test t: from the step's test-to-pair table when it covers
// the step, else by binary search
__device__ inline int64_t test_pair(const PW &W, int64_t NP, uint32_t t) {
  return W.tpair_ok ? (int64_t)W.tpair[t] : find_pair(W.pair_exc, NP, t);
}

// the first test of every non-empty pair marks that pair; a max-scan then
// gives every test its pair
__global__ void k_mark_pair_starts(PW W, int64_t NP) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k < NP && W.pair_n[k] > 0) W.tp_in[W.pair_exc[k]] = (uint32_t)k;
}

__device__ inline void atomic_max_double(double *addr, double v) {
  unsigned long long *a = (unsigned long long *)addr;
  unsigned long long nv = (unsigned long long)__double_as_longlong(v);
  atomicMax(a, nv);  // v >= 0
}

// A particle's status this step.  Statuses carry the epoch of the step that
// set them, so a new step starts with every particle live without clearing
// the array (it is cleared when the epoch wraps).
__device__ inline uint8_t st_get(const PW &W, int64_t p) {
  const uint8_t v = W.status[p];
  return (int)(v >> 2) == W.epoch ? (uint8_t)(v & 3) : ST_LIVE;
}
__device__ inline void st_set(const PW &W, int64_t p, uint8_t s) {
  W.status[p] = (uint8_t)((W.epoch << 2) | s);
}

// a slot of a step-start list can be drawn while it holds a particle of
// that list's (bin, class): untouched, or changed in place by an earlier
// merge without leaving the list
__device__ inline bool drawable(const PW &W, int p, int key) {
  return st_get(W, p) != ST_HOLE && W.key[p] == key;
}

// ---- list membership within a step
// In the reference a bin pair draws its test count from the lists as they
// are when the loop reaches it (src/coagulation.F90:615-626) and each test
// draws among the members at that moment (:766-803).  Here the count comes
// from an upper bound, each list's n step-start members plus acap reserved
// arrival entries, and a draw thins itself: a member that left before the
// pair began, or a reserved entry with no arrival before then, drops the
// test.  Poisson thinning keeps the surviving count Poisson with the mean
// of the members present when the pair begins.  Within the pair, a member
// that left after the pair began is drawn again among the present ones.

// the first change of slot p in this step: remember its step-start key
// The fields of slot p that the list bookkeeping reads, loaded together
// (one memory round trip instead of a chain of them); a slot first touched
// this step records its step-start list and no departure yet.
struct SlotView {
  int key, key0, e;
  uint32_t dep;
};
__device__ inline SlotView slot_touch(const PW &W, int64_t p) {
  SlotView v;
  const uint8_t stv = W.status[p];
  v.key = W.key[p];
  v.key0 = W.key0[p];
  v.dep = W.depart_t[p];
  v.e = W.arr_of[p];
  if ((int)(stv >> 2) != W.epoch || (stv & 3) == ST_LIVE) {
    v.key0 = v.key;
    v.dep = UNCLAIMED;
    W.key0[p] = v.key;
    W.depart_t[p] = UNCLAIMED;
  }
  return v;
}

// An arrival takes the next entry of its list's array, or past A_MAX of
// them an overflow entry linked into the list's chain.  Which entries the
// arrivals take depends on the order of the atomics, so readers use only
// the set: counts and choices by (sequence index, slot) or by a hash of the
// slot.  The entry is complete before the readers of a later round (a
// later kernel, or a tail round behind a barrier) look at it.
__device__ inline void record_arrival(const PW &W, BoxDesc *B, int box, int key, int64_t p,
                                      uint32_t t) {
  const int64_t L = (int64_t)box * W.K_stride + key;
  const int i = atomicAdd(&W.arr_n[L], 1);
  int64_t e;
  if (i < A_MAX) {
    e = L * A_MAX + i;
  } else {
    const uint32_t o = atomicAdd(W.ap_top, 1u);
    if (o >= W.ap_cap) {
      atomicExch(&B->error, ERR_CAPACITY);
      return;
    }
    e = W.ap_ovf + o;
  }
  W.ap_p[e] = (int)p;
  W.ap_t[e] = t;
  W.ap_dep[e] = UNCLAIMED;
  W.arr_of[p] = (int)e;
  atomicAdd((unsigned long long *)&B->n_arrivals, 1ull);
  if (i < A_MAX) return;
  int old = __hip_atomic_load(&W.arr_head[L], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  for (;;) {
    W.ap_next[e] = old;
    __threadfence();
    const int prev = atomicCAS(&W.arr_head[L], old, (int)e);
    if (prev == old) break;
    old = prev;
  }
}

// visit every arrival entry of list L: the array, then the overflow chain
template <class F>
__device__ inline void arrivals_each(const PW &W, int64_t L, F f) {
  const int n = W.arr_n[L];
  const int na = n < A_MAX ? n : A_MAX;
  for (int i = 0; i < na; ++i) f(L * A_MAX + i);
  if (n > A_MAX)
    for (int e = W.arr_head[L]; e >= 0; e = W.ap_next[e]) f((int64_t)e);
}

// the particle in slot p (as v shows it) leaves the list it is in, at
// sequence index t
__device__ inline void slot_leave_view(const PW &W, int64_t p, uint32_t t, const SlotView &v) {
  if (v.key < 0) return;
  if (v.key == v.key0 && v.dep == UNCLAIMED) {
    W.depart_t[p] = t;  // the slot's step-start member
    return;
  }
  const int e = v.e;  // an arrival of list v.key
  if (e >= 0 && e < W.ap_ovf + W.ap_cap && W.ap_p[e] == (int)p && W.ap_dep[e] == UNCLAIMED)
    W.ap_dep[e] = t;
}

__device__ inline void slot_leave(const PW &W, int64_t p, uint32_t t) {
  slot_leave_view(W, p, t, slot_touch(W, p));
}

// slot p receives a particle of list new_key in place of the one it held
// (a merge into a parent's slot, a target merged in place): the old one
// leaves its list at t_leave, the new one joins at t_join
__device__ inline void slot_write(const PW &W, BoxDesc *B, int box, int64_t p, uint32_t t_leave,
                                  uint32_t t_join, int new_key) {
  const SlotView v = slot_touch(W, p);
  if (v.key == new_key) return;  // changed in place, still in its list
  slot_leave_view(W, p, t_leave, v);
  if (new_key != v.key0) record_arrival(W, B, box, new_key, p, t_join);
}

// an appended slot receives a particle of list new_key at sequence index t
__device__ inline void slot_append(const PW &W, BoxDesc *B, int box, int64_t p, uint32_t t,
                                   int new_key) {
  W.key0[p] = -1;
  W.depart_t[p] = 0;
  record_arrival(W, B, box, new_key, p, t);
}

constexpr int DRAW_THIN = -2, DRAW_ATHIN = -3, DRAW_AGAIN = -4;

// 32-bit mixing (the finalizer of MurmurHash3)
__device__ inline uint32_t mix32(uint32_t h) {
  h ^= h >> 16;
  h *= 0x85ebca6bu;
  h ^= h >> 13;
  h *= 0xc2b2ae35u;
  h ^= h >> 16;
  return h;
}

// A draw on reserved entry j of list L for a pair that begins at tau.  The
// entry holds an arrival when the list had at least j arrivals before tau;
// which one is drawn uniformly among them (not `exclude`, the first draw of
// a test on the same list) as the one whose slot hashes lowest under u.  So
// each arrival present at tau is drawn at the rate of a member.  Returns
// the pool entry, or -1.
constexpr int ACHUNK = 8;  // arrival entries loaded together

__device__ inline int64_t arrival_draw(const PW &W, int64_t L, int j, uint32_t tau, uint32_t u,
                                       int exclude) {
  int n_before = 0;
  int64_t best = -1;
  uint32_t best_h = 0;
  int best_p = 0;
  auto consider = [&](int64_t e, uint32_t te, int pe) {
    if (te >= tau) return;
    ++n_before;
    if (pe == exclude) return;
    const uint32_t h = mix32(u ^ ((uint32_t)pe * 0x9E3779B9u));
    if (best < 0 || h < best_h || (h == best_h && pe < best_p)) {
      best = e;
      best_h = h;
      best_p = pe;
    }
  };
  // the array part a chunk at a time, its loads issued before any is used
  // (index clamped to the last entry), entries in order as before
  const int n = W.arr_n[L];
  const int na = n < A_MAX ? n : A_MAX;
  const int64_t e0 = L * A_MAX;
  for (int i0 = 0; i0 < na; i0 += ACHUNK) {
    uint32_t te[ACHUNK];
    int pe[ACHUNK];
#pragma unroll
    for (int q = 0; q < ACHUNK; ++q) {
      const int64_t e = e0 + (i0 + q < na ? i0 + q : na - 1);
      te[q] = W.ap_t[e];
      pe[q] = W.ap_p[e];
    }
#pragma unroll
    for (int q = 0; q < ACHUNK; ++q)
      if (i0 + q < na) consider(e0 + i0 + q, te[q], pe[q]);
  }
  if (n > A_MAX)
    for (int e = W.arr_head[L]; e >= 0; e = W.ap_next[e]) consider((int64_t)e, W.ap_t[e], W.ap_p[e]);
  return j <= n_before ? best : -1;
}

// The arrival of rank r (1-based, in (sequence index, slot) order) among
// list L's arrivals, if it came before tau: the pool entry, or -1.
__device__ inline int64_t arrival_rank(const PW &W, int64_t L, int r, uint32_t tau) {
  int64_t found = -1;
  arrivals_each(W, L, [&](int64_t e) {
    if (found >= 0) return;
    const uint32_t te = W.ap_t[e];
    const int pe = W.ap_p[e];
    int rank = 0;
    arrivals_each(W, L, [&](int64_t f) {
      const uint32_t tf = W.ap_t[f];
      if (tf < te || (tf == te && W.ap_p[f] < pe)) ++rank;
    });
    if (rank == r - 1) found = te < tau ? e : -2;
  });
  return found >= 0 ? found : -1;
}

// entry c (1-based) of list `key` (list index L, n step-start members
// starting at s) for a pair that begins at tau: the particle, or
// DRAW_THIN (a member that left before tau), DRAW_ATHIN (no arrival there
// before tau), DRAW_AGAIN (a member that left after tau: draw again).  A
// reserved entry takes one more draw from `ts` to pick its arrival.
__device__ inline int draw_entry(const PW &W, int64_t base, int64_t L, int key, int n, int s,
                                 int c, uint32_t tau, Stream &ts, int exclude) {
  if (c <= n) {
    const int p = W.perm[base + s + c - 1];
    if (drawable(W, p, key)) return p;
    return W.depart_t[p] < tau ? DRAW_THIN : DRAW_AGAIN;
  }
  const uint32_t u = (uint32_t)(ts.uniform() * 4294967296.0);
  const int64_t e = arrival_draw(W, L, c - n, tau, u, exclude);
  if (e < 0 || W.ap_dep[e] < tau) return DRAW_ATHIN;
  const int p = W.ap_p[e];
  return drawable(W, p, key) ? p : DRAW_AGAIN;
}

// reserved arrival entries and a fresh arrival log for every list: lists
// at most ARR_REACH bins above an occupied one (particles only grow)
// reserve 2 + n/64, and a list that received arrivals in the previous step
// (on the same grid) at least twice as many as it received then
__global__ void k_par_acap(BoxDesc *boxes, PW W) {
  const int box = blockIdx.y;
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= W.K_stride) return;
  const BoxDesc *B = &boxes[box];
  const int64_t L = (int64_t)box * W.K_stride + k;
  const int nb = B->n_bin, nc = B->w.n_class;
  const int prev = (B->regridded || B->rekey) ? 0 : W.arr_n[L];
  int cap = 0;
  if (!B->error && nb > 0 && k < nb * nc) {
    const int b = k % nb + 1;
    const int *cnt = W.lcnt + (int64_t)box * W.K_stride;
    bool near = false;
    for (int c = 1; c <= nc && !near; ++c)
      for (int bb = max(1, b - ARR_REACH); bb <= b && !near; ++bb) near = cnt[par_key(bb, c, nb)] > 0;
    if (near) cap = 2 + cnt[k] / 64;
    if (prev > 0) cap = max(cap, 2 * prev + 2);
    cap = min(A_MAX, cap);
  }
  W.acap[L] = cap;
  W.arr_n[L] = 0;
  W.arr_head[L] = -1;
  if (box == 0 && k == 0) *W.ap_top = 0u;
}

// arrivals of the step beyond their list's reserved entries (not seen by
// the list's later pairs)
__global__ void k_par_arr_over(BoxDesc *boxes, PW W) {
  const int box = blockIdx.y;
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  int over = 0;
  if (k < W.K_stride) {
    const int64_t L = (int64_t)box * W.K_stride + k;
    over = max(0, W.arr_n[L] - W.acap[L]);
  }
  for (int off = 32; off > 0; off >>= 1) over += __shfl_xor(over, off);
  if ((threadIdx.x & 63) == 0 && over)
    atomicAdd((unsigned long long *)&boxes[box].n_arr_over, (unsigned long long)over);
}

// find_rand_pair + maybe_coag_pair (src/coagulation.F90:712-803) for test t
// of pair k (global pair index): the two entries, drawn from the step's
// lists skipping slots whose particle is gone or has left the list's
// (bin, class) -- with every particle present the first draw of each
// stands -- and the accept decision.  A re-run uses the test's own stream:
// on an unchanged state it repeats the earlier outcome, on a changed one it
// is what the sequential chain would draw there.
struct TestOut {
  int p1, p2;
  bool accept;
  bool warn;
  int thin;  // 0: drawn; 1: thinned on a member that left before the pair; 2 (3): on an empty
            // arrival entry of the first (second) list
};

__device__ __forceinline__ void perset_test(BoxDesc *B, const Arrays &A, const PW &W, int box, int64_t k,
                                   uint32_t t, int type, uint64_t seed, uint32_t step,
                                   int precision, double band, int validation, TestOut &o,
                                        const Weights *wl = nullptr) {
  o.p1 = o.p2 = -1;
  o.accept = false;
  o.warn = false;
  o.thin = 0;
  int c1, c2, b1, b2;
  decode_pair(k % W.P_stride, W.tab_classes, W.max_bin, &c1, &c2, &b1, &b2);
  const uint32_t j = t - W.pair_exc[k];
  const uint32_t tau = W.pair_exc[k];  // where the pair begins in the sequence
  const Weights &w = wl ? *wl : B->w;
  const int nb = B->n_bin, nc = w.n_class;
  const int64_t ti = table_index(A, box, nb, nc, b1, b2, c1, c2);
  const int cc = A.cc[ti];
  const double accept_factor = 1.0 / A.kw[ti];
  const bool same = (b1 == b2) && (c1 == c2);
  const int64_t lb = (int64_t)box * W.K_stride;
  const int key1 = par_key(b1, c1, nb), key2 = par_key(b2, c2, nb);
  const int n1 = W.lcnt[lb + key1], n2 = W.lcnt[lb + key2];
  const int s1 = W.lstart[lb + key1], s2 = W.lstart[lb + key2];
  const int N1 = n1 + W.acap[lb + key1], N2 = n2 + W.acap[lb + key2];
  const int64_t base = B->off;
  Stream ts;
  ts.open(seed, box, step, pair_site(nb, nc, b1, b2, c1, c2), j, RNG_TEST);
  // each side's first draw thins the test or stands; a member that left
  // within the pair is drawn again among those present
  int p1 = -1, p2 = -1;
  int i1 = 0;
  for (int a = 0; a < MAX_REDRAW; ++a) {
    const int c = rand_int(ts, N1);
    const int p = draw_entry(W, base, lb + key1, key1, n1, s1, c, tau, ts, -1);
    if (p >= 0) {
      i1 = c;
      p1 = p;
      break;
    }
    if (a == 0 && p != DRAW_AGAIN) {
      o.thin = p == DRAW_THIN ? 1 : 2;
      return;
    }
  }
  if (p1 >= 0 && (!same || N2 > 1)) {
    for (int a = 0; a < MAX_REDRAW; ++a) {
      int c;
      if (same) {
        c = rand_int(ts, N2 - 1);
        if (c == i1) c = N2;
      } else {
        c = rand_int(ts, N2);
      }
      const int p = draw_entry(W, base, lb + key2, key2, n2, s2, c, tau, ts, same ? p1 : -1);
      if (p >= 0 && p != p1) {
        p2 = p;
        break;
      }
      if (a == 0 && p < 0 && p != DRAW_AGAIN) {
        o.thin = p == DRAW_THIN ? 1 : 3;
        return;
      }
    }
  }
  if (p1 < 0 || p2 < 0) {
    atomicAdd(&B->n_dropped, 1);
    return;
  }
  o.p1 = p1;
  o.p2 = p2;
  const Env env = B->env;
  double u = ts.uniform();
  double pr;
  if (precision == 0) {
    double k64;
    int err = weighted_kernel64(A, w, env, type, p1, p2, c1, c2, cc, &k64);
    if (err) atomicExch(&B->error, err);
    pr = k64 * accept_factor;
    o.accept = u < pr;
  } else {
    float k32 = weighted_kernel32(type, w, env, W.terms[p1], W.terms[p2], c1, c2, cc);
    double p32 = (double)k32 * accept_factor;
    bool need64 = !(fabs(u - p32) >= band) || validation;
    double k64 = 0.0;
    if (need64) {
      int err = weighted_kernel64(A, w, env, type, p1, p2, c1, c2, cc, &k64);
      if (err) atomicExch(&B->error, err);
    }
    if (!(fabs(u - p32) >= band)) {
      atomicAdd((unsigned long long *)&B->n_recheck, 1ull);
      pr = k64 * accept_factor;
      o.accept = u < pr;
    } else {
      pr = p32;
      o.accept = u < p32;
      if (validation && (o.accept != (u < k64 * accept_factor)))
        atomicAdd((unsigned long long *)&B->n_flip, 1ull);
    }
    if (validation && k64 > 0.0) {
      atomic_max_double(&B->max_rel_err, fabs((double)k32 - k64) / k64);
      atomic_max_double(&B->max_p_err, fabs(p32 - k64 * accept_factor));
    }
  }
  o.warn = !(pr <= 1.0);
}

// outcome bits of ev_out
constexpr int O_WIN = 1, O_RM1 = 2, O_RM2 = 4, O_NEW = 8, O_SLOT1 = 16, O_SLOT2 = 32,
              O_APPEND = 64, O_ID1 = 128, O_ID2 = 256, O_NEWID = 512, O_INFO1 = 1024,
              O_INFO2 = 2048;

struct MergeCtx {
  int box, c1, c2, cc, b1, b2;
  uint32_t site, j;
};

__device__ inline MergeCtx merge_ctx(BoxDesc *boxes, const PW &W, int n_box, uint32_t t) {
  MergeCtx m;
  const int64_t NP = (int64_t)n_box * W.P_stride;
  const int64_t k = test_pair(W, NP, t);
  m.box = (int)(k / W.P_stride);
  decode_pair(k % W.P_stride, W.tab_classes, W.max_bin, &m.c1, &m.c2, &m.b1, &m.b2);
  m.j = t - W.pair_exc[k];
  return m;
}

// coagulate_weighting (src/coagulation.F90:810-922): decisions only
// wl: the box's weights staged in LDS by the caller, or null to read them
// from the box
__device__ inline int merge_decide(const Arrays &A, const PW &W, const BoxDesc *B, Stream &s,
                                   int64_t p1, int64_t p2, int c1, int c2, int cc,
                                   int *new_group, const Weights *wl = nullptr) {
  const Weights &w = wl ? *wl : B->w;
  // the parents' volumes and lists, read together up front
  const double v1 = particle_volume(A, p1), v2 = particle_volume(A, p2);
  const int key_p1 = W.key[p1], key_p2 = W.key[p2];
  double r1 = vol2rad(v1);
  double r2 = vol2rad(v2);
  double rc = vol2rad(rad2vol(r1) + rad2vol(r2));
  double nc1 = nc_at_radius(w, c1, r1);
  double nc2 = nc_at_radius(w, c2, r2);
  double ncc = nc_at_radius(w, cc, rc);
  double pdf[MAX_GROUP];
  for (int g = 0; g < w.n_group; ++g) {
    int idx = g + w.n_group * (cc - 1);
    pdf[g] = 1.0 / weight_nc(w.type[idx], w.mag[idx], w.expo[idx], rc);
  }
  *new_group = sample_cts_pdf(s, w.n_group, pdf);
  double nc_min = fmin(nc1, fmin(nc2, ncc));
  bool rm1 = s.uniform() < nc_min / nc1;
  bool rm2 = s.uniform() < nc_min / nc2;
  bool cr = s.uniform() < nc_min / ncc;
  int o = 0;
  if (rm1) o |= O_RM1;
  if (rm2) o |= O_RM2;
  if (cr) o |= O_NEW;
  bool id1_lost, id2_lost;
  if (cr) {
    id1_lost = false;
    id2_lost = false;
    if (rm1 && rm2) {
      if (v1 > v2) id2_lost = true;
      else id1_lost = true;
    }
    if (rm1 && !id1_lost) o |= O_ID1;
    else if (rm2 && !id2_lost) o |= O_ID2;
    else o |= O_NEWID;
    // the new particle takes a freed slot when there is one: the slot of
    // the parent whose list it stays in, if any, so that it can still be
    // drawn from that list in this step (in the reference it joins its
    // bin's list)
    const int nb = B->n_bin;
    const int b_new = bin_find_log(nb, B->log_lo, B->log_hi, rc);
    const int k_new = (b_new >= 1 && b_new <= nb) ? par_key(b_new, cc, nb) : -1;
    if (rm1 && rm2) o |= (key_p2 == k_new && key_p1 != k_new) ? O_SLOT2 : O_SLOT1;
    else if (rm1) o |= O_SLOT1;
    else if (rm2) o |= O_SLOT2;
    else o |= O_APPEND;
  } else {
    id1_lost = rm1;
    id2_lost = rm2;
  }
  if (rm1 && id1_lost) o |= O_INFO1;
  if (rm2 && id2_lost) o |= O_INFO2;
  return o;
}

// A claim on a particle the test changes (removes, or a target it merges
// into): recorded as the earliest accessor and the earliest writer.  A
// particle a test only reads is claimed as accessor only: tests that read
// it win together, a test that changes it waits for earlier readers, and a
// reader waits for earlier writers.
__device__ inline void claim_write(const PW &W, int p, uint32_t t) {
  atomicMin(&W.claim[p], t);
  atomicMin(&W.claim_wr[p], t);
}

// the merge decisions of an accepted per-set test (which parents it
// removes), on the particles as they are before the round's winners apply
__device__ inline int perset_decide(BoxDesc *B, const Arrays &A, const PW &W, int box, int64_t k,
                                    uint32_t t, int p1, int p2, uint64_t seed, uint32_t step,
                                    Stream &ms) {
  int c1, c2, b1, b2;
  decode_pair(k % W.P_stride, W.tab_classes, W.max_bin, &c1, &c2, &b1, &b2);
  const int nb = B->n_bin, nc = B->w.n_class;
  const int cc = A.cc[table_index(A, box, nb, nc, b1, b2, c1, c2)];
  ms.open(seed, box, step, pair_site(nb, nc, b1, b2, c1, c2), t - W.pair_exc[k], RNG_MERGE);
  int g;
  return merge_decide(A, W, B, ms, p1, p2, c1, c2, cc, &g) | (g << 16);
}

// The kernels of a claim round read their item counts from device memory
// (d.rc), and their grids are sized on the host from a bound or an
// estimate: every thread strides over the items.
#define PAR_FOR(i, n) \
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < (n); i += gridDim.x * blockDim.x)

// test i of the round: draw, evaluate, accept, claim
__device__ __forceinline__ void par_test(BoxDesc *boxes, const Arrays &A, const PW &W, int n_box, int type,
                                uint64_t seed, uint32_t step, uint32_t t0, const uint32_t *pend,
                                int precision, double band, int validation, int i) {
  const uint32_t t = pend ? pend[i] : t0 + (uint32_t)i;
  const int64_t NP = (int64_t)n_box * W.P_stride;
  const int64_t k = test_pair(W, NP, t);
  if (W.pair_acc[k]) return;  // a target slot: k_par_acc_tests
  W.acc[i] = 0;
  W.tid[i] = t;
  W.tp1[i] = -1;
  W.tp2[i] = -1;
  W.trec[i] = -1;
  W.tns[i] = 0;
  const int box = (int)(k / W.P_stride);
  BoxDesc *B = &boxes[box];
  if (B->error) return;
  TestOut o;
  perset_test(B, A, W, box, k, t, type, seed, step, precision, band, validation, o);
  W.tp1[i] = o.p1;
  W.tp2[i] = o.p2;
  W.thin[t - t0] = o.thin ? 1 : 0;
  // counted once the test is final; a test thinned on an arrival entry waits
  // for the tail, where earlier events are settled first
  W.tns[i] = (o.warn ? ACC_WARN : 0) | (o.thin >= 2 ? TNS_ATHIN : 0) | (o.thin == 3 ? TNS_SIDE2 : 0);
  if (o.accept) {
    W.acc[i] = 1;
    Stream ms;
    const int d = perset_decide(B, A, W, box, k, t, o.p1, o.p2, seed, step, ms);
    if (d & O_RM1) claim_write(W, o.p1, t);
    else atomicMin(&W.claim[o.p1], t);
    if (d & O_RM2) claim_write(W, o.p2, t);
    else atomicMin(&W.claim[o.p2], t);
  }
}

// STRIDE: the grid may be smaller than the count (a later round, sized from
// an estimate).  The first round's grid covers its tests exactly, and
// without the loop the kernel keeps half the registers.
template <bool STRIDE>
__global__ void __launch_bounds__(256)
k_par_tests(BoxDesc *boxes, Arrays A, PW W, int n_box, int type, uint64_t seed, uint32_t step,
            uint32_t t0, const uint32_t *pend, int precision, double band, int validation) {
  const int n = W.rc[RC_N];
  if (STRIDE) {
    PAR_FOR(i, n) par_test(boxes, A, W, n_box, type, seed, step, t0, pend, precision, band, validation, i);
  } else {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) par_test(boxes, A, W, n_box, type, seed, step, t0, pend, precision, band, validation, i);
  }
}

// ------------------------------------------------ accelerated target slots
// try_per_particle_coag (src/coagulation.F90:150-241).  In the reference
// loop a target particle is visited by every accelerated pair that has its
// current (bin, class) on the target side, in loop order; a target that
// grows into the next bin is visited again by the later pairs of that bin.
// Here each target particle is one test slot, placed in the first
// accelerated pair of its step-start list (targets e = n .. 1, :220-237).
// Its thread walks the remaining pairs in loop order, following the
// target's current bin; in each it samples a virtual source particle
// (sample_source_particle, :333-517) and merges it in
// (coag_target_with_source, :522-582, with aero_state_dup_particle,
// src/aero_state.F90:488-539).  The final target, its copies and the
// removed sources are applied when the event wins all of its claims.

// aero_weight_array_rand_group (src/aero_weight_array.F90:496-514)
__device__ inline int rand_group_w(Stream &s, const Weights &w, int cls, double r) {
  double pdf[MAX_GROUP];
  for (int g = 0; g < w.n_group; ++g) {
    int idx = g + w.n_group * (cls - 1);
    pdf[g] = 1.0 / weight_nc(w.type[idx], w.mag[idx], w.expo[idx], r);
  }
  return sample_cts_pdf(s, w.n_group, pdf);
}

__device__ inline double rec_vol(int n_spec, const Rec &r) {
  double v = 0.0;
  for (int s = 0; s < n_spec; ++s) v += r.vol[s];
  return v;
}

__device__ inline double rec_density(const Arrays &A, const Rec &r) {
  double m = 0.0;
  for (int s = 0; s < A.n_spec; ++s) m += r.vol[s] * A.density[s];
  return m / rec_vol(A.n_spec, r);
}

// weighted kernel of two detached records (as weighted_kernel64)
__device__ inline int weighted_kernel_rec(const Arrays &A, const Weights &w, const Env &env,
                                          int type, const Rec &a, const Rec &b, int c1, int c2,
                                          int cc, double *k_out) {
  double v1 = rec_vol(A.n_spec, a), v2 = rec_vol(A.n_spec, b);
  double k = 0.0;
  int err = 0;
  switch (type) {
    case KERNEL_ADDITIVE: k = env.beta * (v1 + v2); break;
    case KERNEL_CONSTANT: k = C_BETA_0; break;
    case KERNEL_ZERO: k = 0.0; break;
    case KERNEL_BROWN:
      k = kernel_brown_helper(v1, rec_density(A, a), v2, rec_density(A, b), env.temp,
                              env.pressure);
      break;
    case KERNEL_SEDI: err = kernel_sedi_helper(v1, v2, &k); break;
    default: err = 200724934;
  }
  *k_out = k * coag_num_conc_factor(w, vol2rad(v1), vol2rad(v2), c1, c2, cc);
  return err;
}

// the sources a target removes, in a chain of RM_CHUNK-int chunks
struct RmChain {
  int head, cur, fill, n;
  uint64_t bloom;
};

__device__ inline uint64_t rm_bit(int slot) {
  return 1ull << ((((uint32_t)slot * 2654435761u) >> 26) & 63u);
}

__device__ inline bool rm_push(const PW &W, RmChain &c, int slot) {
  if (c.fill == RM_CHUNK) {
    int nc = atomicAdd(W.rm_top, RM_CHUNK);
    if (nc + RM_CHUNK > W.rm_cap) return false;
    W.rm_buf[nc] = -1;
    if (c.cur >= 0) W.rm_buf[c.cur] = nc;
    else c.head = nc;
    c.cur = nc;
    c.fill = 1;
  }
  W.rm_buf[c.cur + c.fill] = slot;
  c.fill++;
  c.n++;
  if (slot >= 0) c.bloom |= rm_bit(slot);
  return true;
}

__device__ inline bool rm_has(const PW &W, const RmChain &c, int slot) {
  if (slot < 0 || !(c.bloom & rm_bit(slot))) return false;
  int left = c.n;
  for (int ch = c.head; ch >= 0 && left > 0; ch = W.rm_buf[ch]) {
    int m = left < RM_CHUNK - 1 ? left : RM_CHUNK - 1;
    for (int q = 1; q <= m; ++q)
      if (W.rm_buf[ch + q] == slot) return true;
    left -= m;
  }
  return false;
}

// visit the chain of record r in order: removed sources (>= 0) and the
// owner markers (-1 - k: the removals that follow belong to copy k)
template <class F>
__device__ inline void chain_walk(const PW &W, int r, F f) {
  int left = W.r_nchain[r];
  for (int ch = W.r_head[r]; ch >= 0 && left > 0; ch = W.rm_buf[ch]) {
    int m = left < RM_CHUNK - 1 ? left : RM_CHUNK - 1;
    for (int q = 1; q <= m; ++q) f(W.rm_buf[ch + q]);
    left -= m;
  }
}

// visit the removed sources of record r in removal order
template <class F>
__device__ inline void rm_walk(const PW &W, int r, F f) {
  chain_walk(W, r, [&](int q) {
    if (q >= 0) f(q);
  });
}

// visit the sources record r read from small lists
template <class F>
__device__ inline void rd_walk(const PW &W, int r, F f) {
  int left = W.r_nrd[r];
  for (int ch = W.r_rdhead[r]; ch >= 0 && left > 0; ch = W.rm_buf[ch]) {
    int m = left < RM_CHUNK - 1 ? left : RM_CHUNK - 1;
    for (int q = 1; q <= m; ++q) f(W.rm_buf[ch + q]);
    left -= m;
  }
}

__device__ inline void pool_store(const PW &W, int n_spec, int i, const Rec &s) {
  for (int k = 0; k < n_spec; ++k) W.q_vol[(int64_t)k * W.pool_cap + i] = s.vol[k];
  W.q_group[i] = s.group;
  W.q_nprim[i] = s.nprim;
  W.q_ncomp[i] = s.ncomp;
  for (int c = 0; c < s.ncomp; ++c) {
    W.q_csrc[(int64_t)c * W.pool_cap + i] = s.csrc[c];
    W.q_ctime[(int64_t)c * W.pool_cap + i] = s.ctime[c];
  }
  W.q_least[i] = s.least;
  W.q_greatest[i] = s.greatest;
}

// a pool entry as a particle of class `klass` with identifier `id`
__device__ inline void pool_load(const PW &W, int n_spec, int i, int klass, int64_t id, Rec &s) {
  for (int k = 0; k < n_spec; ++k) s.vol[k] = W.q_vol[(int64_t)k * W.pool_cap + i];
  s.id = id;
  s.group = W.q_group[i];
  s.klass = klass;
  s.nprim = W.q_nprim[i];
  s.ncomp = W.q_ncomp[i];
  for (int c = 0; c < s.ncomp; ++c) {
    s.csrc[c] = W.q_csrc[(int64_t)c * W.pool_cap + i];
    s.ctime[c] = W.q_ctime[(int64_t)c * W.pool_cap + i];
  }
  s.least = W.q_least[i];
  s.greatest = W.q_greatest[i];
}

// sample_source_particle (src/coagulation.F90:333-517) against the running
// target, drawing from the step's source list (list start ss, n_s entries)
// the particles still present that this target has not removed; returns
// an error code, 0 on success.  Sources it removes are pushed to `chain`
// and claimed with index t.
// The target's side of weighted_kernel_rec, the same for every source drawn
// for one pair: its Brownian terms, volume, radius and weight.
struct TargetSide {
  BrownAir air;
  BrownTerms terms;
  double vol, rad, nc;
};

__device__ inline TargetSide target_side(const Arrays &A, const Weights &w, const Env &env,
                                         int type, const Rec &b, int c2) {
  TargetSide ts;
  ts.vol = rec_vol(A.n_spec, b);
  if (type == KERNEL_BROWN) {
    ts.air = brown_air(env.temp, env.pressure);
    ts.terms = brown_terms(ts.air, ts.vol, rec_density(A, b), env.temp);
  }
  ts.rad = vol2rad(ts.vol);
  ts.nc = nc_at_radius(w, c2, ts.rad);
  return ts;
}

// weighted_kernel_rec(a, b, c1, c2, cc) with b's side from target_side(b, c2)
__device__ inline int weighted_kernel_tgt(const Arrays &A, const Weights &w, const Env &env,
                                          int type, const Rec &a, const TargetSide &ts,
                                          int c1, int cc, double *k_out) {
  double v1 = rec_vol(A.n_spec, a), v2 = ts.vol;
  double k = 0.0;
  int err = 0;
  switch (type) {
    case KERNEL_ADDITIVE: k = env.beta * (v1 + v2); break;
    case KERNEL_CONSTANT: k = C_BETA_0; break;
    case KERNEL_ZERO: k = 0.0; break;
    case KERNEL_BROWN:
      k = brown_combine(brown_terms(ts.air, v1, rec_density(A, a), env.temp), ts.terms);
      break;
    case KERNEL_SEDI: err = kernel_sedi_helper(v1, v2, &k); break;
    default: err = 200724934;
  }
  *k_out = k * coag_num_conc_factor_j(w, vol2rad(v1), rad2vol(ts.rad), ts.nc, c1, cc);
  return err;
}

__device__ inline int acc_sample_source(const Arrays &A, const PW &W, BoxDesc *B, int type,
                                        const Rec &target, int ct, int cs, double nc_src_min,
                                        int64_t sbase, int skey, int n_s, double n_spt,
                                        double accept_factor, Stream &aps, Stream &smp,
                                        Stream &fin, uint32_t tau, uint32_t t, bool claim,
                                        RmChain &chain,
                                        RmChain &reads, int *n_samp_out,
                                        int *n_coag_out, int *warn, Rec &source) {
  const int ns = A.n_spec;
  const Weights &w = B->w;
  const Env env = B->env;
  *n_samp_out = 0;
  *n_coag_out = 0;
  const double nc_target = nc_at_radius(w, target.klass, vol2rad(rec_vol(ns, target)));
  const double prm = nc_target / nc_src_min;
  if (!(prm <= 1.0)) return 653606684;
  const int n_remove = rand_poisson(aps, prm * n_spt);
  const int n_extra = rand_poisson(aps, (1.0 - prm) * n_spt);
  const int n_total = n_remove + n_extra;
  int n_avg = 0, n_samp = 0, n_rm_here = 0, n_thin = 0, n_samp_rm = 0;
  double prob_coag_tot = 0.0;
  double vol_sq[MAX_SPEC], vol_cv[MAX_SPEC];
  for (int q = 0; q < ns; ++q) {
    vol_sq[q] = 0.0;
    vol_cv[q] = 0.0;
    source.vol[q] = 0.0;
  }
  source.id = 0;
  source.group = source.klass = source.nprim = source.ncomp = 0;
  for (int c = 0; c < MAX_COMP; ++c) {
    source.csrc[c] = 0;
    source.ctime[c] = 0.0;
  }
  source.least = source.greatest = 0.0;
  Rec pi;
  TargetSide ts;  // formed at the first sample: many pairs draw none
  bool have_ts = false;
  int i_samp = 1;
  for (; i_samp <= n_total; ++i_samp) {
    if (n_rm_here >= n_s) break;  // every source removed
    // the reference's n_samp > n_samp_remove (:413) over the samples kept
    if (i_samp > n_remove && n_samp > n_samp_rm && n_avg >= 2) {
      double cvmax = 0.0;
      for (int q = 0; q < ns; ++q) {
        double vm = source.vol[q] / (double)n_avg;
        if (vm > 0.0) vol_cv[q] = sqrt(vol_sq[q] / (double)n_avg - vm * vm) / vm;
        if (q == 0 || vol_cv[q] > cvmax) cvmax = vol_cv[q];
      }
      double m95 = cvmax / sqrt((double)n_avg) * student_t_95_coeff(n_avg);
      if (m95 < 0.1) break;
    }
    // a uniform draw among the sources still present.  The counts above
    // come from the step-start list; a source that left before the pair
    // began (tau) was not in the reference's count (compute_n_source,
    // src/coagulation.F90:288-328, reads the list as it is), so a first
    // draw that lands on one thins the sample.  Sources that earlier
    // targets of the pair, or this one, removed are drawn again, as in the
    // reference's loop.
    int ps = -1;
    bool thinned = false;
    for (int a = 0; a < MAX_REDRAW; ++a) {
      int c = rand_int(smp, n_s);
      int q = W.perm[sbase + c - 1];
      if (drawable(W, q, skey) && !rm_has(W, chain, q)) {
        ps = q;
        break;
      }
      if (a == 0 && !drawable(W, q, skey) && W.depart_t[q] < tau) {
        thinned = true;
        break;
      }
    }
    if (thinned) {
      n_thin++;
      continue;
    }
    if (ps < 0) {
      atomicAdd(&B->n_dropped, 1);
      break;
    }
    // in a small list a source another target removes changes this
    // target's draws: remember what was read
    if (n_s <= READ_TRACK && !rm_has(W, reads, ps) && !rm_push(W, reads, ps)) return ERR_CAPACITY;
    n_samp++;
    if (i_samp <= n_remove) n_samp_rm++;
    load_rec_head(A, ps, pi);
    double kk;
    if (!have_ts) {
      ts = target_side(A, w, env, type, target, ct);
      have_ts = true;
    }
    int err = weighted_kernel_tgt(A, w, env, type, pi, ts, cs, ct, &kk);
    if (err) return err;
    double prob_coag = kk * accept_factor;
    prob_coag_tot += prob_coag;
    if (smp.uniform() < prob_coag) {
      n_avg++;
      load_rec_comps(A, ps, pi);
      err = par_rec_coagulate(ns, smp, source, pi, source);
      if (err) return err;
      for (int q = 0; q < ns; ++q) vol_sq[q] += pi.vol[q] * pi.vol[q];
      if (i_samp <= n_remove) {
        double nc_i = nc_at_radius(w, pi.klass, vol2rad(rec_vol(ns, pi)));
        double prob_remove_i = nc_target / nc_i;
        if (smp.uniform() < prob_remove_i / prm) {
          if (!rm_push(W, chain, ps)) return ERR_CAPACITY;
          n_rm_here++;
          if (claim) claim_write(W, ps, t);
        }
      }
    }
  }
  *n_samp_out = n_samp;
  if (n_avg == 0) return 0;
  // the samples the early exit left undrawn thin as the drawn ones did
  for (int i = i_samp; i <= n_total; ++i) {
    const int q = W.perm[sbase + rand_int(smp, n_s) - 1];
    if (!drawable(W, q, skey) && W.depart_t[q] < tau) n_thin++;
  }
  double prob_coag_mean = prob_coag_tot / (double)n_samp;
  if (!(prob_coag_mean <= 1.0)) *warn = 1;
  int n_coag = rand_binomial(fin, n_total - n_thin, prob_coag_mean);
  for (int q = 0; q < ns; ++q) source.vol[q] = source.vol[q] * ((double)n_coag / (double)n_avg);
  source.nprim = prob_round(fin, (double)source.nprim * ((double)n_coag / (double)n_avg));
  const int n_comp = source.ncomp;
  const int n_comp_new = source.nprim < MAX_COMP ? source.nprim : MAX_COMP;
  int sample[MAX_COMP];
  int cs_src[MAX_COMP];
  double ct_src[MAX_COMP];
  for (int c = 0; c < MAX_COMP; ++c) {
    cs_src[c] = source.csrc[c];
    ct_src[c] = source.ctime[c];
  }
  if (n_comp_new < n_comp) {
    sample_without_replacement(fin, n_comp_new, n_comp, sample);
    for (int c = 0; c < n_comp_new; ++c) {
      source.csrc[c] = cs_src[sample[c] - 1];
      source.ctime[c] = ct_src[sample[c] - 1];
    }
    source.ncomp = n_comp_new;
  } else if (n_comp < MAX_COMP) {
    int i_comp = 1;
    while (i_comp <= n_comp_new) {
      if (i_comp + n_comp <= n_comp_new) {
        for (int c = 1; c <= n_comp; ++c) {
          source.csrc[i_comp + c - 2] = cs_src[c - 1];
          source.ctime[i_comp + c - 2] = ct_src[c - 1];
        }
        i_comp += n_comp;
      } else {
        int sample_size = n_comp_new - i_comp + 1;
        sample_without_replacement(fin, sample_size, n_comp, sample);
        for (int c = 1; c <= sample_size; ++c) {
          source.csrc[i_comp + c - 2] = cs_src[sample[c - 1] - 1];
          source.ctime[i_comp + c - 2] = ct_src[sample[c - 1] - 1];
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

// One target of an accelerated pair: the body of try_per_particle_coag's
// target loop (src/coagulation.F90:216-237) as the event at sequence index
// t of pair k (record rec).  A pair's events are laid out as the reference
// loops over its target list at the pair's start: the list's arrivals of
// the step first, the latest first (they were appended last), then its
// step-start members from the last entry down.  Each visit of a target to
// a pair is its own event, so a visit sees the sources as the events
// before the pair left them, and a target that merged at an earlier pair
// is read as that pair's event left it.  Returns 1 when the event has an
// effect (record rec and the pool filled; the removed sources are claimed
// with index t when `claim`), 0 when it has none (*tns = its samples +
// ACC_WARN, or TNS_ATHIN for an arrival entry with no arrival yet), -1 on
// a box error.  *p_out is the target's slot, or -1 when there is none.
__device__ __noinline__ int acc_visit(BoxDesc *boxes, const Arrays &A, const PW &W, int64_t k,
                                      uint32_t t, int rec, int type, uint64_t seed, uint32_t step,
                                      double del_t, bool claim, int *p_out, int *tns) {
  *p_out = -1;
  *tns = 0;
  const int box = (int)(k / W.P_stride);
  BoxDesc *B = &boxes[box];
  if (B->error) return -1;
  const int nb = B->n_bin, nc = B->w.n_class;
  const int64_t lb = (int64_t)box * W.K_stride;
  const int64_t pb = (int64_t)box * W.P_stride;
  const double *edges = A.edges + (int64_t)box * (A.max_bin + 1);
  int c1, c2, b1, b2;
  const int64_t q = k - pb;
  decode_pair(q, W.tab_classes, W.max_bin, &c1, &c2, &b1, &b2);
  const int kind = W.pair_acc[k];
  const int bt = kind == 1 ? b1 : b2, ct = kind == 1 ? c1 : c2;
  const int bs = kind == 1 ? b2 : b1, cs = kind == 1 ? c2 : c1;
  const int kt = par_key(bt, ct, nb), ks = par_key(bs, cs, nb);
  const uint32_t tau = W.pair_exc[k];
  const int j = (int)(t - tau);
  const int acap_t = W.acap[lb + kt];
  int p;
  if (j < acap_t) {
    const int64_t e = arrival_rank(W, lb + kt, acap_t - j, tau);
    if (e < 0 || W.ap_dep[e] < tau) {
      *tns = TNS_ATHIN;  // an earlier event not yet settled may still fill it
      return 0;
    }
    p = W.ap_p[e];
  } else {
    const int m = W.lcnt[lb + kt] - (j - acap_t);
    p = W.perm[B->off + W.lstart[lb + kt] + m - 1];
  }
  if (!drawable(W, p, kt)) return 0;  // left the list before the pair
  *p_out = p;
  // the record's chains start empty, so an attempt cut short by an error
  // leaves nothing to walk
  W.r_nrd[rec] = 0;
  W.r_rdhead[rec] = -1;
  W.r_nchain[rec] = 0;
  W.r_nrm[rec] = 0;
  W.r_head[rec] = -1;
  const int ns = A.n_spec;
  const Weights &w = B->w;
  const uint32_t ordinal = (uint32_t)j + 1;
  const int64_t ti = table_index(A, box, nb, nc, b1, b2, c1, c2);
  const double k_max = A.kw[ti];
  const int cc = A.cc[ti];
  const int n_s = W.lcnt[lb + ks];
  const uint32_t site = pair_site(nb, nc, b1, b2, c1, c2);
  Stream aps, smp, fin;
  aps.open(seed, box, step, site, ordinal, RNG_ACC_POISSON);
  smp.open(seed, box, step, site, ordinal, RNG_ACC_SAMPLE);
  fin.open(seed, box, step, site, ordinal, RNG_ACC_FINISH);
  Rec target, source;
  load_rec(A, p, target);
  RmChain chain{-1, -1, RM_CHUNK, 0, 0ull};
  RmChain reads{-1, -1, RM_CHUNK, 0, 0ull};
  int n_samp = 0, n_coag = 0, warn = 0, flags = 0, err_code = 0;
  int err = acc_sample_source(A, W, B, type, target, ct, cs, nc_at_radius(w, cs, edges[bs - 1]),
                              B->off + W.lstart[lb + ks], ks, n_s, k_max * del_t * (double)n_s,
                              1.0 / k_max, aps, smp, fin, tau, t, claim, chain, reads, &n_samp,
                              &n_coag, &warn, source);
  if (err) {
    atomicExch(&B->error, err);
    return -1;
  }
  const int n_rm = chain.n;
  int n_dup = 0, dup_head = -1, dup_tail = -1;
  Rec np = target;
  if (n_coag > 0) {
    // coag_target_with_source (src/coagulation.F90:522-582)
    Stream ft;
    ft.open(seed, box, step, site, ordinal, RNG_ACC_TARGET);
    const double old_nc = nc_at_radius(w, target.klass, vol2rad(rec_vol(ns, target)));
    err = par_rec_coagulate(ns, ft, target, source, np);
    if (err) {
      atomicExch(&B->error, err);
      return -1;
    }
    np.id = target.id;
    const double r_new = vol2rad(rec_vol(ns, np));
    np.group = rand_group_w(ft, w, cc, r_new);
    np.klass = cc;
    flags |= RF_MERGED;
    const int new_bin = bin_find_log(nb, B->log_lo, B->log_hi, r_new);
    if (new_bin < 1 || new_bin > nb) {
      err_code = 765620746;
    } else {
      // aero_state_dup_particle (src/aero_state.F90:488-539): the copies
      // join the list of the target's new bin
      const int n_copies = prob_round(ft, old_nc / nc_at_radius(w, np.klass, r_new));
      if (n_copies == 0) flags |= RF_DIED;
      for (int c = 1; c < n_copies; ++c) {
        const int d = W.acc_cap + atomicAdd(W.dup_top, 1);
        if (d >= W.pool_cap) {
          atomicExch(&B->error, ERR_CAPACITY);
          return -1;
        }
        Rec cp = np;
        cp.group = rand_group_w(ft, w, np.klass, r_new);
        pool_store(W, ns, d, cp);
        W.q_next[d] = -1;
        if (dup_tail >= 0) W.q_next[dup_tail] = d;
        else dup_head = d;
        dup_tail = d;
        n_dup++;
      }
    }
  }
  if (warn) flags |= RF_WARN;
  W.r_rdhead[rec] = reads.head;
  W.r_nrd[rec] = reads.n;
  if (n_coag == 0 && n_rm == 0 && err_code == 0) {
    *tns = n_samp | (warn ? ACC_WARN : 0);
    return 0;
  }
  if (flags & RF_MERGED) pool_store(W, ns, rec, np);
  W.r_nsamp[rec] = n_samp;
  W.r_ncoag[rec] = n_coag;
  W.r_nrm[rec] = n_rm;
  W.r_nchain[rec] = chain.n;
  W.r_head[rec] = chain.head;
  W.r_ndup[rec] = n_dup;
  W.r_dhead[rec] = dup_head;
  W.r_flags[rec] = flags;
  W.r_err[rec] = err_code;
  if (claim) claim_write(W, p, t);
  return 1;
}

// visits per wavefront: each is one lane's serial sample loop, and fewer a
// wavefront spread the first round's visits over more compute units
constexpr int ACC_LANES = 16;
// Round 0 runs one thread per target slot of the batch (record a maps to
// its slot through acc_exc); later rounds run over the pending slots.
__device__ inline void acc_test(BoxDesc *boxes, const Arrays &A, const PW &W, int n_box, int type,
                                uint64_t seed, uint32_t step, uint32_t t0, const uint32_t *pend,
                                double del_t, uint32_t acc_base, int idx) {
  const int64_t NP = (int64_t)n_box * W.P_stride;
  uint32_t t;
  int i;
  int64_t k;
  if (pend) {
    i = idx;
    t = pend[i];
    k = test_pair(W, NP, t);
    if (!W.pair_acc[k]) return;
  } else {
    const uint32_t a = acc_base + (uint32_t)idx;
    k = find_pair(W.acc_exc, NP, a);
    t = W.pair_exc[k] + (a - W.acc_exc[k]);
    i = (int)(t - t0);
  }
  const int rec = (int)(W.acc_exc[k] + (t - W.pair_exc[k]) - acc_base);
  W.acc[i] = 0;
  W.tid[i] = t;
  W.tp2[i] = -1;
  W.trec[i] = rec;
  int p, tns;
  const int r = acc_visit(boxes, A, W, k, t, rec, type, seed, step, del_t, true, &p, &tns);
  W.tp1[i] = p;
  W.tns[i] = tns;
  if (r == 1) W.acc[i] = 1;
  // the visit's samples, for the algorithmic traffic of this kernel
  const int ns = (r == 1) ? W.r_nsamp[rec] : (tns & (TNS_SIDE2 - 1));
  if (ns > 0) atomicAdd(W.acc_samp, (unsigned long long)ns);
}

__global__ void __launch_bounds__(64)
k_par_acc_tests(BoxDesc *boxes, Arrays A, PW W, int n_box, int type, uint64_t seed,
                uint32_t step, uint32_t t0, const uint32_t *pend, double del_t,
                uint32_t acc_base, int n_acc) {
  // over a pending list most slots are per-set tests that return at once:
  // every lane takes one; over the first round's target slots, ACC_LANES a
  // wavefront
  const int lanes = pend ? 64 : ACC_LANES;
  if ((int)threadIdx.x >= lanes) return;
  const int count = pend ? W.rc[RC_N] : n_acc;
  for (int idx = blockIdx.x * lanes + threadIdx.x; idx < count; idx += gridDim.x * lanes)
    acc_test(boxes, A, W, n_box, type, seed, step, t0, pend, del_t, acc_base, idx);
}

// the largest batch end t1 <= min(T, t0 + T_cap) whose target slots fit the
// records; out[0] = t1, out[1] = acc_count(t0), out[2] = acc_count(t1)
__device__ inline uint32_t acc_count(const PW &W, int64_t NP, uint32_t T, uint32_t t) {
  if (t >= T) return W.acc_exc[NP];
  int64_t k = test_pair(W, NP, t);
  return W.acc_exc[k] + (W.pair_acc[k] ? t - W.pair_exc[k] : 0u);
}

__global__ void k_batch_end(PW W, int64_t NP, uint32_t t0, uint32_t t_cap, uint32_t a_cap,
                            uint32_t *out) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  const uint32_t T = W.pair_exc[NP];
  const uint32_t a0 = acc_count(W, NP, T, t0);
  uint32_t lo = t0 + 1, hi = (T - t0 < t_cap) ? T : t0 + t_cap;
  if (acc_count(W, NP, T, hi) - a0 <= a_cap) lo = hi;  // the usual case: all of it fits
  while (lo < hi) {  // largest t1 in [lo, hi] with acc_count(t1) - a0 <= a_cap
    uint32_t mid = lo + (hi - lo + 1) / 2;
    if (acc_count(W, NP, T, mid) - a0 <= a_cap) lo = mid; else hi = mid - 1;
  }
  out[0] = lo;
  out[1] = a0;
  out[2] = acc_count(W, NP, T, lo);
  out[3] = T;
  out[4] = W.acc_exc[NP];
}

// compact accepted tests into the event list (in test order)
__global__ void k_par_gather(PW W) {
  const int n = W.rc[RC_N];
  PAR_FOR(i, n) {
    if (!W.acc[i]) continue;
    int e = W.acc_scan[i];
    W.ev_t[e] = W.tid[i];
    W.ev_p1[e] = W.tp1[i];
    W.ev_p2[e] = W.tp2[i];
    W.ev_rec[e] = W.trec[i];
  }
}

__device__ inline void resolve_event(BoxDesc *boxes, const Arrays &A, const PW &W, int n_box,
                                     uint64_t seed, uint32_t step, int e) {
  uint32_t t = W.ev_t[e];
  int p1 = W.ev_p1[e], p2 = W.ev_p2[e];
  MergeCtx m = merge_ctx(boxes, W, n_box, t);
  W.ev_box[e] = m.box;
  const int r = W.ev_rec[e];
  if (r >= 0) {
    // a target event wins when it holds the target and every removed source
    // ... and no earlier event claimed a source it read
    bool win = W.claim[p1] == t;
    rm_walk(W, r, [&](int q) { win = win && W.claim[q] == t; });
    rd_walk(W, r, [&](int q) { win = win && W.claim_wr[q] >= t; });
    const int died = (W.r_flags[r] & RF_DIED) ? 1 : 0;
    W.ev_out[e] = win ? O_WIN : 0;
    W.f_app[e] = win ? W.r_ndup[r] : 0;
    W.f_id[e] = win ? W.r_ndup[r] : 0;
    W.f_info[e] = win ? W.r_nrm[r] + died : 0;
    W.f_hole[e] = win ? W.r_nrm[r] + died : 0;
    W.f_lose[e] = win ? 0 : 1;
    return;
  }
  // the decisions say which parents the merge changes (removes); a parent
  // it keeps is only read
  BoxDesc *B = &boxes[m.box];
  const int nb = B->n_bin, nc = B->w.n_class;
  int cc = A.cc[table_index(A, m.box, nb, nc, m.b1, m.b2, m.c1, m.c2)];
  Stream s;
  s.open(seed, m.box, step, pair_site(nb, nc, m.b1, m.b2, m.c1, m.c2), m.j, RNG_MERGE);
  int g;
  const int d = merge_decide(A, W, B, s, p1, p2, m.c1, m.c2, cc, &g);
  const bool win = ((d & O_RM1) ? W.claim[p1] == t : W.claim_wr[p1] > t) &&
                   ((d & O_RM2) ? W.claim[p2] == t : W.claim_wr[p2] > t);
  int o = 0;
  if (win) o = d | O_WIN | (g << 16);
  W.ev_out[e] = o;
  W.f_app[e] = (o & O_APPEND) ? 1 : 0;
  W.f_id[e] = (o & O_NEWID) ? 1 : 0;
  W.f_info[e] = ((o & O_INFO1) ? 1 : 0) + ((o & O_INFO2) ? 1 : 0);
  W.f_hole[e] = ((o & O_RM1) && !(o & O_SLOT1) ? 1 : 0) + ((o & O_RM2) && !(o & O_SLOT2) ? 1 : 0);
  W.f_lose[e] = win ? 0 : 1;
}

__global__ void k_par_resolve(BoxDesc *boxes, Arrays A, PW W, int n_box, uint64_t seed,
                              uint32_t step) {
  const int E = W.rc[RC_E];
  PAR_FOR(e, E) resolve_event(boxes, A, W, n_box, seed, step, e);
}

// first event of every box (events are in test order, so box-major)
__global__ void k_par_box_first(PW W) {
  const int E = W.rc[RC_E];
  PAR_FOR(e, E) {
    if (e == 0 || W.ev_box[e - 1] != W.ev_box[e]) W.box_first[W.ev_box[e]] = e;
  }
}

__device__ inline void write_new_particle(const Arrays &A, const PW &W, BoxDesc *B, int box,
                                          uint32_t t_leave, uint32_t t_join, int64_t slot,
                                          bool appended, const Rec &r, int nb, int type,
                                          int want_terms) {
  double v = 0.0;
  double m = 0.0;
  for (int s = 0; s < A.n_spec; ++s) {
    A.vol[(int64_t)s * A.total_cap + slot] = r.vol[s];
    v += r.vol[s];
    m += r.vol[s] * A.density[s];
  }
  A.vtot[slot] = v;
  A.id[slot] = r.id;
  A.group[slot] = r.group;
  A.klass[slot] = r.klass;
  A.nprim[slot] = r.nprim;
  A.ncomp[slot] = r.ncomp;
  for (int c = 0; c < r.ncomp; ++c) {
    A.csrc[comp_at(slot, c)] = r.csrc[c];
    A.ctime[comp_at(slot, c)] = r.ctime[c];
  }
  A.least[slot] = r.least;
  A.greatest[slot] = r.greatest;
  int b = bin_find_log(nb, B->log_lo, B->log_hi, vol2rad(v));
  if (b < 1 || b > nb) {
    atomicExch(&B->error, 134572570);
    st_set(W, slot, ST_NEW);
    W.key[slot] = 0;
    return;
  }
  int k = par_key(b, r.klass, nb);
  if (appended) slot_append(W, B, box, slot, t_join, k);
  else slot_write(W, B, box, slot, t_leave, t_join, k);
  st_set(W, slot, ST_NEW);
  W.key[slot] = k;
  if (want_terms) W.terms[slot] = particle_terms(type, v, m / v, B->env);
}

// Where a merge puts what it creates.  In a claim round the positions come
// from the round's scans over the events (RankAlloc, one per event); in the
// sequential tail they come straight from the box's counters (SeqAlloc).
struct RankAlloc {
  BoxDesc *B;
  int64_t app, id, info, hole;  // this event's next ranks
  int *live;                    // the box's live counts per list key
  __device__ int64_t append_slot() {
    int64_t r = app++;
    return B->n_part + r < B->cap ? B->off + B->n_part + r : -1;
  }
  __device__ int64_t new_id() {
    return (B->next_id + id++ - 1) * B->id_stride + B->id_offset + 1;
  }
  __device__ int64_t info_index() { return B->n_info + info++; }
  __device__ int64_t hole_index() { return hole++; }
  __device__ void live_add(int key, int d) { atomicAdd(&live[key], d); }
  __device__ void count(int64_t samp, int64_t coag, int warn) {
    if (samp) {
      atomicAdd((unsigned long long *)&B->n_samp, (unsigned long long)samp);
      atomicAdd((unsigned long long *)&B->step_samp, (unsigned long long)samp);
    }
    if (coag) {
      atomicAdd((unsigned long long *)&B->n_coag, (unsigned long long)coag);
      atomicAdd((unsigned long long *)&B->step_coag, (unsigned long long)coag);
    }
    if (warn) atomicAdd((unsigned long long *)&B->n_warn, (unsigned long long)warn);
  }
};

// The round's allocator: counters kept in the lane and summed over the
// wavefront per box before they reach the box (thousands of winners of one
// box would otherwise serialise on the same counter words)
struct ApplyAlloc : RankAlloc {
  int64_t n_samp, n_coag, n_warn;
  int lk[3], ld[3], nl;  // the first live-count changes, summed per wavefront later
  __device__ void count(int64_t samp, int64_t coag, int warn) {
    n_samp += samp;
    n_coag += coag;
    n_warn += warn;
  }
  __device__ void live_add(int key, int d) {
    if (nl < 3) {
      lk[nl] = key;
      ld[nl] = d;
      ++nl;
    } else {
      atomicAdd(&live[key], d);
    }
  }
};

// Sum live-count changes over the wavefront per (list, key) and add each
// sum once; key < 0: nothing from this lane.  Every lane takes part.
__device__ inline void wave_flush_live(int *live, int key, int d) {
  const int lane = threadIdx.x & 63;
  const uint64_t addr = key >= 0 ? (uint64_t)(live + key) : 0ull;
  uint64_t pending = __ballot(key >= 0 && d != 0);
  while (pending) {
    const int leader = __ffsll((long long)pending) - 1;
    const uint64_t la = __shfl(addr, leader);
    const bool mine = ((pending >> lane) & 1ull) && addr == la;
    int td = mine ? d : 0;
    for (int off = 32; off > 0; off >>= 1) td += __shfl_xor(td, off);
    if (lane == leader && td) atomicAdd((int *)la, td);
    pending &= ~__ballot(mine);
  }
}

// Sum each box's counter changes over the wavefront and add them with one
// set of atomics per box (box < 0: nothing from this lane).  Every lane of
// the wavefront takes part.
__device__ inline void wave_flush_counts(BoxDesc *boxes, int box, int64_t s, int64_t c,
                                         int64_t w) {
  const int lane = threadIdx.x & 63;
  uint64_t pending = __ballot(box >= 0 && (s | c | w));
  while (pending) {
    const int leader = __ffsll((long long)pending) - 1;
    const int lb = __shfl(box, leader);
    const bool mine = ((pending >> lane) & 1ull) && box == lb;
    long long ts = mine ? s : 0, tc = mine ? c : 0, tw = mine ? w : 0;
    for (int off = 32; off > 0; off >>= 1) {
      ts += __shfl_xor(ts, off);
      tc += __shfl_xor(tc, off);
      tw += __shfl_xor(tw, off);
    }
    if (lane == leader) {
      BoxDesc *B = &boxes[lb];
      if (ts) {
        atomicAdd((unsigned long long *)&B->n_samp, (unsigned long long)ts);
        atomicAdd((unsigned long long *)&B->step_samp, (unsigned long long)ts);
      }
      if (tc) {
        atomicAdd((unsigned long long *)&B->n_coag, (unsigned long long)tc);
        atomicAdd((unsigned long long *)&B->step_coag, (unsigned long long)tc);
      }
      if (tw) atomicAdd((unsigned long long *)&B->n_warn, (unsigned long long)tw);
    }
    pending &= ~__ballot(mine);
  }
}

// The tail's allocator: ranks as in a round, live-count changes gathered
// in the workgroup's shared memory and counters in the lane, both flushed
// once (many winners of one box would otherwise contend on the same words)
struct TailAlloc : RankAlloc {
  int *s_live;
  int64_t n_samp, n_coag, n_warn;
  __device__ void live_add(int key, int d) { atomicAdd(&s_live[key], d); }
  __device__ void count(int64_t samp, int64_t coag, int warn) {
    n_samp += samp;
    n_coag += coag;
    n_warn += warn;
  }
};

template <class Alloc>
__device__ inline void put_info(const Arrays &A, BoxDesc *B, Alloc &al, int64_t id, int action,
                                int64_t other) {
  int64_t k = al.info_index();
  if (k < B->info_cap) {
    A.info_id[B->info_off + k] = id;
    A.info_action[B->info_off + k] = action;
    A.info_other[B->info_off + k] = other;
  }
}

// Both parents' components, read before any is written (dst may be either
// parent's slot) and all in flight at once.  The loads are unconditional
// (every one of a particle's MAX_COMP entries, which sit together): a load
// under `if (c < n)` becomes a branch, and the compiler waits for every
// load in flight at each one, one memory round trip per component.
__device__ inline void load_comps2(const Arrays &A, int64_t p1, int64_t p2, int n1, int n2,
                                   int *s1, double *t1, int *s2, double *t2) {
  (void)n1;
  (void)n2;
#pragma unroll
  for (int c = 0; c < MAX_COMP; ++c) {
    s1[c] = A.csrc[comp_at(p1, c)];
    t1[c] = A.ctime[comp_at(p1, c)];
    s2[c] = A.csrc[comp_at(p2, c)];
    t2[c] = A.ctime[comp_at(p2, c)];
  }
}

// the merged particle's components when the parents hold more than
// MAX_COMP together (aero_particle_coagulate, src/aero_particle.F90:
// 917-1033): prob_round of the first parent's share, then two selection
// samples (sample_without_replacement keeps sorted order) from the parents'
// components, which are read into registers first; the loops are unrolled
// so the arrays stay in registers.
__device__ inline void resample_components(const Arrays &A, Stream &s, int64_t p1, int64_t p2,
                                           int64_t dst, int n1, int n2, int np1, int np2) {
  int s1[MAX_COMP], s2[MAX_COMP];
  double t1[MAX_COMP], t2[MAX_COMP];
  load_comps2(A, p1, p2, n1, n2, s1, t1, s2, t2);
  const int n1_new = prob_round(s, (double)np1 / (double)(np1 + np2) * (double)MAX_COMP);
  const int n2_new = MAX_COMP - n1_new;
  // first parent: n1_new of its n1 components into positions 0 ..
  {
    int m = 0;
#pragma unroll
    for (int c = 0; c < MAX_COMP; ++c) {
      if (c < n1 && m < n1_new) {
        const double r = s.uniform();
        if (!((double)(n1 - c) * r >= (double)(n1_new - m))) {
          A.csrc[comp_at(dst, m)] = s1[c];
          A.ctime[comp_at(dst, m)] = t1[c];
          m++;
        }
      }
    }
  }
  // second parent: n2_new of its n2 components after them
  {
    int m = 0;
#pragma unroll
    for (int c = 0; c < MAX_COMP; ++c) {
      if (c < n2 && m < n2_new) {
        const double r = s.uniform();
        if (!((double)(n2 - c) * r >= (double)(n2_new - m))) {
          A.csrc[comp_at(dst, n1_new + m)] = s2[c];
          A.ctime[comp_at(dst, n1_new + m)] = t2[c];
          m++;
        }
      }
    }
  }
}

// the merged particle of an accepted per-set test written straight from
// its parents' slots into dst (which may be either parent's slot):
// aero_particle_coagulate (src/aero_particle.F90:917-1033).  Components
// are concatenated, or resampled on `rs` (the merge stream after the merge
// decisions) when the parents hold more than MAX_COMP together.
__device__ inline int direct_merge(const Arrays &A, const PW &W, BoxDesc *B, int box, uint32_t t,
                                   int64_t p1, int64_t p2, int64_t dst, bool appended, int64_t id,
                                   int group, int klass, int type, int want_terms, Stream *rs) {
  const int64_t tc = A.total_cap;
  const int n1 = A.ncomp[p1], n2 = A.ncomp[p2];
  const int np1 = A.nprim[p1], np2 = A.nprim[p2];
  if (np1 < n1) return 465791384;
  if (np2 < n2) return 465791385;
  const double least = fmin(A.least[p1], A.least[p2]);
  const double greatest = fmax(A.greatest[p1], A.greatest[p2]);
  // dst may be a parent's slot: each species' sum depends only on its own
  // row, so it is written once formed; a chunk's loads go out together
  double v = 0.0, m = 0.0;
  const int ns = A.n_spec;
  for (int s0 = 0; s0 < ns; s0 += SPEC_CHUNK) {
    double x[SPEC_CHUNK], rho[SPEC_CHUNK];
#pragma unroll
    for (int u = 0; u < SPEC_CHUNK; ++u) {
      const int s = s0 + u < ns ? s0 + u : ns - 1;
      x[u] = A.vol[(int64_t)s * tc + p1] + A.vol[(int64_t)s * tc + p2];
      rho[u] = A.density[s];
    }
#pragma unroll
    for (int u = 0; u < SPEC_CHUNK; ++u) {
      if (s0 + u < ns) {
        A.vol[(int64_t)(s0 + u) * tc + dst] = x[u];
        v += x[u];
        m += x[u] * rho[u];
      }
    }
  }
  // components: the first parent's, then the second's, all loaded before
  // any is written (dst may be a parent's slot), so the loads go out
  // together
  int n_out = n1 + n2;
  if (n1 + n2 > MAX_COMP) {
    resample_components(A, *rs, p1, p2, dst, n1, n2, np1, np2);
    n_out = MAX_COMP;
  } else {
    int s1[MAX_COMP], s2[MAX_COMP];
    double t1[MAX_COMP], t2[MAX_COMP];
    load_comps2(A, p1, p2, n1, n2, s1, t1, s2, t2);
#pragma unroll
    for (int c = 0; c < MAX_COMP; ++c) {
      if (c < n1) {
        A.csrc[comp_at(dst, c)] = s1[c];
        A.ctime[comp_at(dst, c)] = t1[c];
      }
      if (c < n2) {
        A.csrc[comp_at(dst, n1 + c)] = s2[c];
        A.ctime[comp_at(dst, n1 + c)] = t2[c];
      }
    }
  }
  A.vtot[dst] = v;
  A.id[dst] = id;
  A.group[dst] = group;
  A.klass[dst] = klass;
  A.nprim[dst] = np1 + np2;
  A.ncomp[dst] = n_out;
  A.least[dst] = least;
  A.greatest[dst] = greatest;
  const int nb = B->n_bin;
  const int b = bin_find_log(nb, B->log_lo, B->log_hi, vol2rad(v));
  if (b < 1 || b > nb) {
    st_set(W, dst, ST_NEW);
    W.key[dst] = 0;
    return 134572570;
  }
  const int k_new = par_key(b, klass, nb);
  if (appended) slot_append(W, B, box, dst, t, k_new);
  else slot_write(W, B, box, dst, t, t, k_new);
  st_set(W, dst, ST_NEW);
  W.key[dst] = k_new;
  if (want_terms) W.terms[dst] = particle_terms(type, v, m / v, B->env);
  return 0;
}

// coagulate_weighting + coagulate (src/coagulation.F90:810-985) of an
// accepted per-set test: removals become holes (or the slot the merged
// particle takes), the merged particle is written, removal records follow
// the reference's order.  o_known >= 0 passes the outcome of merge_decide
// already taken on this test's stream (with the new group in bits 16+),
// and post, if given, that stream after them.
template <class Alloc>
__device__ inline void perset_merge(BoxDesc *B, const Arrays &A, const PW &W, int box, int64_t k,
                                    uint32_t t, int64_t p1, int64_t p2, uint64_t seed,
                                    uint32_t step, int type, int want_terms, int rid, Alloc &al,
                                    int o_known = -1, const Stream *post = nullptr) {
  int c1, c2, b1, b2;
  decode_pair(k % W.P_stride, W.tab_classes, W.max_bin, &c1, &c2, &b1, &b2);
  const uint32_t j = t - W.pair_exc[k];
  al.count(0, 1, 0);
  const int nb = B->n_bin, nc = B->w.n_class;
  const int cc = A.cc[table_index(A, box, nb, nc, b1, b2, c1, c2)];
  const int key1 = W.key[p1], key2 = W.key[p2];
  const int64_t id1 = A.id[p1], id2 = A.id[p2];
  int o, new_group;
  Stream s;
  const bool resample = (A.ncomp[p1] + A.ncomp[p2] > MAX_COMP);
  if (o_known >= 0 && (!resample || post)) {
    // the merge decisions are known; a resampling continues the stream
    // where they left it
    o = o_known & 0xFFFF;
    new_group = o_known >> 16;
    if (post) s = *post;
  } else {
    s.open(seed, box, step, pair_site(nb, nc, b1, b2, c1, c2), j, RNG_MERGE);
    o = merge_decide(A, W, B, s, p1, p2, c1, c2, cc, &new_group);
  }
  int64_t info_other = 0;
  int64_t slot = -1;
  if (o & O_NEW) {
    const int64_t nid = (o & O_ID1) ? id1 : (o & O_ID2) ? id2 : al.new_id();
    info_other = nid;
    if (o & O_SLOT1) slot = p1;
    else if (o & O_SLOT2) slot = p2;
    else {
      slot = al.append_slot();
      if (slot < 0) {
        atomicExch(&B->error, ERR_CAPACITY);
        return;
      }
    }
    const bool appended = !(o & (O_SLOT1 | O_SLOT2));
    const int err = direct_merge(A, W, B, box, t, p1, p2, slot, appended, nid, new_group, cc, type,
                                 want_terms, &s);
    if (err) atomicExch(&B->error, err);
  }
  if (o & O_RM1) {
    W.dead_t[p1] = t;
    W.dead_r[p1] = rid;
    al.live_add(key1, -1);
    if (!(o & O_SLOT1)) {
      slot_leave(W, p1, t);
      st_set(W, p1, ST_HOLE);
      W.holes[al.hole_index()] = ((uint64_t)box << 32) | (uint64_t)(p1 - B->off);
    }
  }
  if (o & O_RM2) {
    W.dead_t[p2] = t;
    W.dead_r[p2] = rid;
    al.live_add(key2, -1);
    if (!(o & O_SLOT2)) {
      slot_leave(W, p2, t);
      st_set(W, p2, ST_HOLE);
      W.holes[al.hole_index()] = ((uint64_t)box << 32) | (uint64_t)(p2 - B->off);
    }
  }
  if (slot >= 0) al.live_add(W.key[slot], 1);
  if (o & O_INFO1) put_info(A, B, al, id1, 2, info_other);
  if (o & O_INFO2) put_info(A, B, al, id2, 2, info_other);
}

// a winning target event (coag_target_with_source and
// aero_state_dup_particle, src/coagulation.F90:522-582,
// src/aero_state.F90:488-539): removed sources become holes with a
// coagulation record naming the particle that took them (the target or one
// of its copies), the merged target replaces the old one in place, copies
// are appended with new identifiers in creation order, and a target or
// copy that kept no copy of itself is removed with a weighting record.
template <class Alloc>
__device__ __noinline__ void acc_apply(BoxDesc *B, const Arrays &A, const PW &W, int box, uint32_t t,
                                 int r, int64_t p, int type, int want_terms, int rid, Alloc &al,
                                 bool chain_elsewhere = false) {
  const int ns = A.n_spec;
  const int nb = B->n_bin;
  const int64_t target_id = A.id[p];
  const int klass = A.klass[p];
  const int flags = W.r_flags[r];
  // copy identifiers in creation order, drawn before any record names them
  int64_t id_first = 0;
  for (int kk = 0; kk < W.r_ndup[r]; ++kk) {
    int64_t v = al.new_id();
    if (kk == 0) id_first = v;
  }
  auto copy_id = [&](int kk) { return id_first + (int64_t)(kk - 1) * B->id_stride; };
  if (chain_elsewhere) {
    // the removed sources are applied by a wavefront (k_par_apply_chains)
    // at this event's first hole and record ranks
    al.hole += W.r_nrm[r];
    al.info += W.r_nrm[r];
  } else {
    int64_t owner = target_id;
    chain_walk(W, r, [&](int q) {
      if (q < 0) {
        owner = copy_id(-1 - q);
        return;
      }
      W.dead_t[q] = t;
      W.dead_r[q] = rid;
      al.live_add(W.key[q], -1);
      slot_leave(W, q, t);
      st_set(W, q, ST_HOLE);
      W.holes[al.hole_index()] = ((uint64_t)box << 32) | (uint64_t)(q - B->off);
      put_info(A, B, al, A.id[q], 2, owner);
    });
  }
  const unsigned long long n_samp = (unsigned long long)W.r_nsamp[r];
  const unsigned long long n_coag = (unsigned long long)W.r_ncoag[r];
  al.count((int64_t)n_samp, (int64_t)n_coag, (flags & RF_WARN) ? 1 : 0);
  atomicAdd((unsigned long long *)&B->n_acc_samp, n_samp);
  if (n_coag) atomicAdd((unsigned long long *)&B->n_acc_coag, n_coag);
  if (W.r_err[r]) atomicExch(&B->error, W.r_err[r]);
  if (!(flags & RF_MERGED)) return;
  Rec np;
  pool_load(W, ns, r, klass, target_id, np);
  W.dead_t[p] = t;
  W.dead_r[p] = rid;
  al.live_add(W.key[p], -1);
  write_new_particle(A, W, B, box, t, t, p, false, np, nb, type, want_terms);
  if (flags & RF_DIED) {
    slot_leave(W, p, t);
    st_set(W, p, ST_HOLE);
    W.holes[al.hole_index()] = ((uint64_t)box << 32) | (uint64_t)(p - B->off);
    put_info(A, B, al, target_id, 4, 0);
  } else {
    al.live_add(W.key[p], 1);
  }
  int kk = 0;
  for (int d = W.r_dhead[r]; d >= 0; d = W.q_next[d]) {
    ++kk;
    pool_load(W, ns, d, klass, copy_id(kk), np);
    const int64_t slot = al.append_slot();
    if (slot < 0) {
      atomicExch(&B->error, ERR_CAPACITY);
      return;
    }
    write_new_particle(A, W, B, box, t, t, slot, true, np, nb, type, want_terms);
    al.live_add(W.key[slot], 1);
  }
}

__device__ inline void apply_event(BoxDesc *boxes, const Arrays &A, const PW &W, int n_box,
                                   uint64_t seed, uint32_t step, int type, int want_terms, int rid,
                                   int E, int64_t hole_base, int e) {
  int box = -1;
  int64_t cs = 0, cc = 0, cw = 0;
  int nl = 0, lk[3] = {0, 0, 0}, ld[3] = {0, 0, 0};
  int *live = W.live;
  if (e < E && (W.ev_out[e] & O_WIN)) {
    const uint32_t t = W.ev_t[e];
    box = W.ev_box[e];
    BoxDesc *B = &boxes[box];
    const int first = W.box_first[box];
    ApplyAlloc al;
    al.B = B;
    al.app = W.s_app[e] - W.s_app[first];
    al.id = W.s_id[e] - W.s_id[first];
    al.info = W.s_info[e] - W.s_info[first];
    al.hole = hole_base + W.s_hole[e];
    al.live = W.live + (int64_t)box * W.K_stride;
    al.n_samp = al.n_coag = al.n_warn = 0;
    al.nl = 0;
    if (W.ev_rec[e] >= 0) {
      acc_apply(B, A, W, box, t, W.ev_rec[e], W.ev_p1[e], type, want_terms, rid, al,
                W.r_nchain[W.ev_rec[e]] > CHAIN_WAVE_MIN);
    } else {
      const int64_t k = test_pair(W, (int64_t)n_box * W.P_stride, t);
      perset_merge(B, A, W, box, k, t, W.ev_p1[e], W.ev_p2[e], seed, step, type, want_terms, rid,
                   al, W.ev_out[e] & ~O_WIN);
    }
    cs = al.n_samp;
    cc = al.n_coag;
    cw = al.n_warn;
    nl = al.nl;
    for (int i = 0; i < 3; ++i) {
      lk[i] = al.lk[i];
      ld[i] = al.ld[i];
    }
    live = al.live;
  }
  wave_flush_counts(boxes, box, cs, cc, cw);
  for (int i = 0; i < 3; ++i) wave_flush_live(live, i < nl ? lk[i] : -1, i < nl ? ld[i] : 0);
}

// apply the round's winners, each at the positions its ranks give
__global__ void __launch_bounds__(256)
k_par_apply(BoxDesc *boxes, Arrays A, PW W, int n_box, uint64_t seed, uint32_t step,
            int type, int want_terms, int rid) {
  const int E = W.rc[RC_E];
  const int64_t hole_base = (int64_t)*W.hole_top;
  // the loop bound is uniform over the workgroup: every lane reaches the
  // wavefront flushes of each pass
  for (int base = blockIdx.x * blockDim.x; base < E; base += gridDim.x * blockDim.x)
    apply_event(boxes, A, W, n_box, seed, step, type, want_terms, rid, E, hole_base,
                base + (int)threadIdx.x);
}

// The removed sources of the round's winning target events with long
// removal chains, one wavefront per event: each chunk's entries go to
// lanes, the ranks of holes and removal records follow the chain order (a
// ballot count of the entries before a lane), and each record names the
// particle that took the source: the target, or the copy named by the last
// copy marker before it in the chain (a ballot of the markers).
__device__ inline void apply_chain(BoxDesc *boxes, const Arrays &A, const PW &W, int e, int lane,
                                   int64_t hole_base, int rid) {
  if (!(W.ev_out[e] & O_WIN)) return;
  const int r = W.ev_rec[e];
  if (r < 0 || W.r_nchain[r] <= CHAIN_WAVE_MIN) return;
  const uint32_t t = W.ev_t[e];
  const int box = W.ev_box[e];
  BoxDesc *B = &boxes[box];
  const int first = W.box_first[box];
  const int64_t hole0 = hole_base + W.s_hole[e];
  const int64_t info0 = B->n_info + (W.s_info[e] - W.s_info[first]);
  const int64_t id_first =
      (B->next_id + (W.s_id[e] - W.s_id[first]) - 1) * B->id_stride + B->id_offset + 1;
  int *live = W.live + (int64_t)box * W.K_stride;
  int64_t owner = A.id[W.ev_p1[e]];
  int done = 0;  // removed sources before this chunk
  int left = W.r_nchain[r];
  for (int ch = W.r_head[r]; ch >= 0 && left > 0; ch = W.rm_buf[ch]) {
    const int m = left < RM_CHUNK - 1 ? left : RM_CHUNK - 1;
    const bool here = lane < m;
    const int q = here ? W.rm_buf[ch + 1 + lane] : 0;
    const uint64_t below = (1ull << lane) - 1ull;
    const uint64_t marks = __ballot(here && q < 0);
    const uint64_t rems = __ballot(here && q >= 0);
    // the owner of this lane's entry: the last marker at or before it
    const uint64_t mk = marks & (below | (1ull << lane));
    // every lane shuffles (no divergence around the cross-lane read)
    const int src = mk ? 63 - __clzll((long long)mk) : lane;
    const int qm = __shfl(q, src);
    const int64_t own = mk ? id_first + (int64_t)(-1 - qm - 1) * B->id_stride : owner;
    if (here && q >= 0) {
      const int ord = done + __popcll(rems & below);
      W.dead_t[q] = t;
      W.dead_r[q] = rid;
      atomicAdd(&live[W.key[q]], -1);
      slot_leave(W, q, t);
      st_set(W, q, ST_HOLE);
      W.holes[hole0 + ord] = ((uint64_t)box << 32) | (uint64_t)(q - B->off);
      const int64_t ki = info0 + ord;
      if (ki < B->info_cap) {
        A.info_id[B->info_off + ki] = A.id[q];
        A.info_action[B->info_off + ki] = 2;
        A.info_other[B->info_off + ki] = own;
      }
    }
    const int last = marks ? 63 - __clzll((long long)marks) : 0;
    const int ql = __shfl(q, last);
    if (marks) owner = id_first + (int64_t)(-1 - ql - 1) * B->id_stride;
    done += __popcll(rems);
    left -= m;
  }
}

__global__ void __launch_bounds__(256)
k_par_apply_chains(BoxDesc *boxes, Arrays A, PW W, int rid) {
  const int lane = threadIdx.x & 63;
  const int E = W.rc[RC_E];
  const int64_t hole_base = (int64_t)*W.hole_top;
  // one wavefront per event, striding over the round's events
  for (int e = blockIdx.x * 4 + (threadIdx.x >> 6); e < E; e += gridDim.x * 4)
    apply_chain(boxes, A, W, e, lane, hole_base, rid);
}

// per-box totals of a round: appended slots, new ids, records, winners
__global__ void k_par_box_totals(BoxDesc *boxes, PW W, int n_box) {
  int box = blockIdx.x * blockDim.x + threadIdx.x;
  if (box >= n_box) return;
  const int E = W.rc[RC_E];
  int first = W.box_first[box];
  if (first < 0) return;
  int last = first;
  {
    int lo = first, hi = E - 1;  // last event of the box (ev_box sorted)
    while (lo < hi) {
      int mid = (lo + hi + 1) >> 1;
      if (W.ev_box[mid] == box) lo = mid; else hi = mid - 1;
    }
    last = lo;
  }
  BoxDesc *B = &boxes[box];
  int app = W.s_app[last] + W.f_app[last] - W.s_app[first];
  int ids = W.s_id[last] + W.f_id[last] - W.s_id[first];
  int info = W.s_info[last] + W.f_info[last] - W.s_info[first];
  B->n_part += app;
  B->next_id += ids;
  B->n_info += info;
}

__device__ inline void release_event(const PW &W, int e) {
  W.claim[W.ev_p1[e]] = UNCLAIMED;
  W.claim_wr[W.ev_p1[e]] = UNCLAIMED;
  const int r = W.ev_rec[e];
  if (r >= 0) {
    rm_walk(W, r, [&](int q) {
      W.claim[q] = UNCLAIMED;
      W.claim_wr[q] = UNCLAIMED;
    });
  } else {
    W.claim[W.ev_p2[e]] = UNCLAIMED;
    W.claim_wr[W.ev_p2[e]] = UNCLAIMED;
  }
}

__global__ void k_par_release(PW W) {
  const int E = W.rc[RC_E];
  PAR_FOR(e, E) release_event(W, e);
}

// Tests of the round that must be drawn again: accepted tests that lost a
// claim, and any other non-winning test that drew a particle which an
// earlier winner (smaller sequential index) of the same round removed or
// changed after the draw -- in the sequential chain that test would have
// seen the particles as that winner left them.  A particle changed in an
// earlier round was drawn in its changed state.
__device__ inline bool consumed_before(const PW &W, int p, uint32_t t, int rid) {
  return p >= 0 && st_get(W, p) != ST_LIVE && W.dead_t[p] < t && W.dead_r[p] == rid;
}

// A test on an empty arrival entry of list L (bin bL) of a pair that
// begins at tau is final once no event with a smaller index than tau that
// could add a particle to L was pending at the start of the round that ran
// it: every arrival into L before tau had been applied then.  On the
// pair's second list the bound is the test's own index, since earlier
// tests of the pair can change the first list's draws (a member they
// removed is drawn again).  Such tests change no list until an earlier
// event fills their entry, so they do not hold back others.
__device__ inline void pending_test(BoxDesc *boxes, const PW &W, int n_box, int rid, int i) {
  uint32_t t = W.tid[i];
  const int64_t NP = (int64_t)n_box * W.P_stride;
  const int64_t k = test_pair(W, NP, t);
  const int box = (int)(k / W.P_stride);
  int c1, c2, b1, b2;
  decode_pair(k % W.P_stride, W.tab_classes, W.max_bin, &c1, &c2, &b1, &b2);
  const int kind = W.pair_acc[k];
  // the lowest bin this event's results can land in
  const int b_lo = kind ? (kind == 1 ? b1 : b2) : max(b1, b2);
  const int64_t mb = (int64_t)box * W.K_stride;
  const bool athin = (W.tns[i] & TNS_ATHIN) && !W.acc[i];
  int f = 0;
  if (athin) {
    // an empty entry of the second list was reached after the first
    // list's draws, which earlier tests of the same pair can change
    const bool side2 = !kind && (W.tns[i] & TNS_SIDE2);
    const int b_list = kind ? b_lo : (side2 ? b2 : b1);
    f = (side2 ? t : W.pair_exc[k]) <= W.minpend[mb + b_list - 1] ? 0 : 1;
  } else if (W.acc[i]) {
    f = (W.ev_out[W.acc_scan[i]] & O_WIN) ? 0 : 1;
  } else if (W.trec[i] >= 0) {
    // a target event without effect: final unless an earlier winner of the
    // round changed its target; its samples count once final
    bool changed = consumed_before(W, W.tp1[i], t, rid);
    if (W.tp1[i] >= 0)
      rd_walk(W, W.trec[i], [&](int q) { changed = changed || consumed_before(W, q, t, rid); });
    f = changed ? 1 : 0;
    int v = W.tns[i];
    if (!f && v) {
      BoxDesc *B = &boxes[box];
      unsigned long long ns = (unsigned long long)(v & (ACC_WARN - 1));
      atomicAdd((unsigned long long *)&B->n_samp, ns);
      atomicAdd((unsigned long long *)&B->step_samp, ns);
      atomicAdd((unsigned long long *)&B->n_acc_samp, ns);
      if (v & ACC_WARN) atomicAdd((unsigned long long *)&B->n_warn, 1ull);
    }
  } else if (W.tp1[i] >= 0) {
    f = (consumed_before(W, W.tp1[i], t, rid) || consumed_before(W, W.tp2[i], t, rid)) ? 1 : 0;
  }
  if (!f && W.trec[i] < 0 && (W.tns[i] & ACC_WARN)) {
    // a per-set test's k_max warning (src/coagulation.F90:749-751), once final
    atomicAdd((unsigned long long *)&boxes[box].n_warn, 1ull);
  }
  // values only fall: most tests see a smaller one already and skip the atomic
  if (f && !athin && W.minpend_next[mb + b_lo - 1] > t) atomicMin(&W.minpend_next[mb + b_lo - 1], t);
  W.f_pend[i] = f;
}

__global__ void k_par_pending(BoxDesc *boxes, PW W, int n_box, int rid) {
  const int n = W.rc[RC_N];
  PAR_FOR(i, n) pending_test(boxes, W, n_box, rid, i);
}

// the running minimum over each box's bins of the next round's minima
// (one workgroup per box), and this round's cleared for the round after
constexpr int PREFIX_THREADS = 256;
__global__ void __launch_bounds__(PREFIX_THREADS) k_minpend_prefix(PW W) {
  __shared__ scan_storage<PREFIX_THREADS, uint32_t> tmp;
  __shared__ uint32_t carry;
  const int tid = threadIdx.x;
  uint32_t *m = W.minpend_next + (int64_t)blockIdx.x * W.K_stride;
  uint32_t *c = W.minpend + (int64_t)blockIdx.x * W.K_stride;
  for (int base = 0; base < W.K_stride; base += PREFIX_THREADS) {
    const int q = base + tid;
    uint32_t v = q < W.K_stride ? m[q] : UNCLAIMED;
    if (base > 0 && tid == 0) v = min(v, carry);
    uint32_t tot;
    const uint32_t r = block_scan_incl<PREFIX_THREADS>(v, OpMin(), tmp, &tot);
    if (q < W.K_stride) {
      m[q] = r;
      c[q] = UNCLAIMED;
    }
    if (tid == PREFIX_THREADS - 1) carry = r;
    __syncthreads();
  }
}

__global__ void k_par_gather_pending(PW W) {
  const int n = W.rc[RC_N];
  PAR_FOR(i, n) {
    if (W.f_pend[i]) W.pend[W.s_pend[i]] = W.tid[i];
  }
}

// The tests of a batch's first round: all of them.
__global__ void k_batch_begin(PW W, int n_batch) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    W.rc[RC_N] = n_batch;
    W.rc[RC_E] = 0;
    W.rc[RC_TAIL] = 0;
    for (int q = 0; q < 4; ++q) W.rstat[q] = 0ull;
  }
}

// The end of a device-wide round that ran: its holes join the step's count,
// and its pending tests go to another device-wide round while that settles
// most of the batch (at most half of it still pending, and not the last
// round), else to the tail.  A round without events leaves nothing
// pending: with no winner no test's inputs changed, and a test on an empty
// arrival entry stays thinned, as no particle arrived.
__global__ void k_round_end(PW W, int n_batch, int round) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  const int n = W.rc[RC_N];
  if (n == 0) return;  // the round did not run
  const int E = W.rc[RC_E];
  const int p = E > 0 ? W.s_pend[n] : 0;
  *W.hole_top += (unsigned long long)W.s_hole[E];
  W.pstat[0] += 1ull;
  W.rstat[round == 0 ? 0 : 2] = (unsigned long long)E;
  if (p > 0 && round < PAR_ROUNDS - 1 && p <= n_batch / 2) {
    W.rc[RC_N] = p;
    W.rstat[1] = (unsigned long long)p;
  } else {
    W.rc[RC_N] = 0;
    W.rc[RC_TAIL] = p;
    W.rstat[3] = (unsigned long long)p;
    W.pstat[1] += (unsigned long long)p;
  }
}

// first index of the sorted list v[0..n) holding a value >= x
__device__ inline int lower_bound_u32(const uint32_t *v, int n, uint32_t x) {
  int lo = 0, hi = n;
  while (lo < hi) {
    int mid = (lo + hi) >> 1;
    if (v[mid] < x) lo = mid + 1; else hi = mid;
  }
  return lo;
}

constexpr int TAIL_THREADS = 512;
constexpr int TF_EFFECT = 1, TF_ACC = 2, TF_RUN = 4, TF_WARN = 8, TF_ATHIN = 16, TF_SIDE2 = 32,
              TF_WAIT = 64;

// claims inside one kernel are read and released with device-scope atomics,
// so they need no cache maintenance between the phases of a round
__device__ inline uint32_t claim_load(const PW &W, int p) {
  return __hip_atomic_load(&W.claim[p], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ inline uint32_t claim_wr_load(const PW &W, int p) {
  return __hip_atomic_load(&W.claim_wr[p], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ inline void claim_release(const PW &W, int p) {
  __hip_atomic_store(&W.claim[p], UNCLAIMED, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  __hip_atomic_store(&W.claim_wr[p], UNCLAIMED, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}


// a phase boundary of the tail: earlier writes of the box's workgroup
// become visible to its next phase.  Only this workgroup touches the box,
// so a workgroup-scope fence suffices; claims go through device-scope
// atomics.
__device__ inline void tail_barrier() {
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
  __syncthreads();
}

// The tail of a batch: the tests still pending after the device-wide claim
// rounds, mostly chains of tests on one particle (a large particle drawn by
// many tests of its bin pairs).  One wavefront per box runs claim rounds
// over a window of the box's next TAIL_THREADS pending tests, in sequential
// order, with barriers in place of launches.  A test is run again only when
// a particle it read or claimed was changed by an earlier winner; a test
// that lost with its inputs unchanged keeps its outcome and claims again.
// The window's earliest test always wins or is final, so every round makes
// progress.
// Claims distinguish reading from changing.  Every particle of an accepted
// test is claimed in W.claim (earliest accessor); the particles the test
// changes are also claimed in W.claim_wr (earliest writer).  A per-set
// merge changes only the parents it removes: a parent that stays is read
// and left as it was.  A test wins when it is the earliest accessor of
// every particle it changes and no earlier test changes a particle it only
// reads.  Tests that only read a particle therefore win together (the
// large-weight particle a bin pair draws over and over, which the
// weighting keeps), and a later test that changes it waits for them.
__global__ void __launch_bounds__(TAIL_THREADS)
k_par_tail(BoxDesc *boxes, Arrays A, PW W, int n_box, const uint32_t *pend, int type,
           uint64_t seed, uint32_t step, int precision, double band, double del_t,
           uint32_t acc_base, int want_terms, int rid0, unsigned long long *rounds_out,
           uint32_t batch_t0) {
  (void)n_box;
  const int box = blockIdx.x;
  const int tid = threadIdx.x;
  BoxDesc *B = &boxes[box];
  const int64_t pb = (int64_t)box * W.P_stride;
  const int n = W.rc[RC_TAIL];
  if (n == 0) return;
  const int lo = lower_bound_u32(pend, n, W.pair_exc[pb]);
  const int hi = lower_bound_u32(pend, n, W.pair_exc[pb + W.P_stride]);
  if (lo >= hi) return;
  __shared__ scan_storage<TAIL_THREADS, int> scan_tmp;
  __shared__ scan_storage<TAIL_THREADS, uint32_t> scan_tmp3;
  __shared__ uint32_t s_carry;
  __shared__ scan_storage<TAIL_THREADS, long long> scan_tmp2;
  __shared__ uint32_t w_t[TAIL_THREADS];
  __shared__ int w_k[TAIL_THREADS], w_p1[TAIL_THREADS], w_p2[TAIL_THREADS];
  __shared__ int w_tns[TAIL_THREADS], w_fl[TAIL_THREADS];
  __shared__ int s_tot[4];
  __shared__ long long s_hole_base;
  __shared__ int s_err;
  // the box's weights, read by every test and merge decision of the tail
  __shared__ Weights s_w;
  for (int q = tid; q < (int)(sizeof(Weights) / sizeof(int)); q += TAIL_THREADS)
    ((int *)&s_w)[q] = ((const int *)&B->w)[q];
  extern __shared__ int s_live[];  // live-count changes of the box's lists
  // then, per bin, the lowest index of an unsettled test of the window whose
  // results land in that bin or a lower one (a running minimum)
  uint32_t *s_pm = (uint32_t *)(s_live + W.K_stride);
  for (int q = tid; q < W.K_stride; q += TAIL_THREADS) s_live[q] = 0;
  int64_t c_samp = 0, c_coag = 0, c_warn = 0;  // this lane's counter changes
  int wn = 0, head = lo;
  int rid = rid0;
  unsigned long long n_rounds = 0;
  for (;;) {
    if (tid == 0) s_err = B->error;
    tail_barrier();
    if (s_err || (wn == 0 && head >= hi)) break;
    // refill the window from the queue, in order
    const int add = min(TAIL_THREADS - wn, hi - head);
    if (tid < add) {
      const uint32_t t = pend[head + tid];
      const int kl = (int)(W.pend_k[head + tid] - pb);
      w_t[wn + tid] = t;
      w_k[wn + tid] = kl;
      w_fl[wn + tid] = TF_RUN | (W.pair_acc[pb + kl] ? TF_ACC : 0);
      w_p1[wn + tid] = -1;
      w_p2[wn + tid] = -1;
      w_tns[wn + tid] = 0;
    }
    __syncthreads();
    wn += add;
    head += add;
    const bool active = tid < wn;
    uint32_t t = 0;
    int64_t k = pb;
    int fl = 0, p1 = -1, p2 = -1, tns = 0;
    if (active) {
      t = w_t[tid];
      k = pb + w_k[tid];
      fl = w_fl[tid];
      p1 = w_p1[tid];
      p2 = w_p2[tid];
      tns = w_tns[tid];
    }
    const bool is_acc = fl & TF_ACC;
    const int rec = is_acc ? (int)(W.acc_exc[k] + (t - W.pair_exc[k]) - acc_base) : -1;
    // 1. run the tests whose inputs are new or changed
    if (active && (fl & TF_RUN)) {
      fl &= ~(TF_RUN | TF_EFFECT | TF_WARN | TF_ATHIN | TF_SIDE2 | TF_WAIT);
      if (is_acc) {
        const int r = acc_visit(boxes, A, W, k, t, rec, type, seed, step, del_t, false, &p1, &tns);
        if (r == 1) fl |= TF_EFFECT;
        if (tns & TNS_ATHIN) fl |= TF_ATHIN;
        p2 = -1;
      } else {
        TestOut o;
        perset_test(B, A, W, box, k, t, type, seed, step, precision, band, 0, o, &s_w);
        p1 = o.p1;
        p2 = o.p2;
        W.thin[t - batch_t0] = o.thin ? 1 : 0;
        if (o.accept) fl |= TF_EFFECT;
        if (o.warn) fl |= TF_WARN;
        if (o.thin >= 2) fl |= TF_ATHIN;
        if (o.thin == 3) fl |= TF_SIDE2;
      }
    }
    const bool effect = fl & TF_EFFECT;
    // the merge decisions of an accepted per-set test, from the particles as
    // they are now: they say which parents it changes
    int o_win = -1;
    Stream ms;  // a per-set test's merge stream, after its decisions
    if (effect && !is_acc) {
      int c1, c2, b1, b2;
      decode_pair(k % W.P_stride, W.tab_classes, W.max_bin, &c1, &c2, &b1, &b2);
      const int nb = B->n_bin, nc = B->w.n_class;
      const int cc = A.cc[table_index(A, box, nb, nc, b1, b2, c1, c2)];
      ms.open(seed, box, step, pair_site(nb, nc, b1, b2, c1, c2), t - W.pair_exc[k], RNG_MERGE);
      int g;
      o_win = merge_decide(A, W, B, ms, p1, p2, c1, c2, cc, &g, &s_w) | (g << 16);
    }
    // 2. claims
    if (effect) {
      if (is_acc) {
        claim_write(W, p1, t);
        rm_walk(W, rec, [&](int q) { claim_write(W, q, t); });
      } else {
        if (o_win & O_RM1) claim_write(W, p1, t);
        else atomicMin(&W.claim[p1], t);
        if (o_win & O_RM2) claim_write(W, p2, t);
        else atomicMin(&W.claim[p2], t);
      }
    }
    __syncthreads();
    // 3. resolve, and what each winner creates
    bool win = false;
    int c_app = 0, c_id = 0, c_info = 0, c_hole = 0;
    if (effect) {
      if (is_acc) {
        win = claim_load(W, p1) == t;
        rm_walk(W, rec, [&](int q) { win = win && claim_load(W, q) == t; });
        rd_walk(W, rec, [&](int q) { win = win && claim_wr_load(W, q) >= t; });
        if (win) {
          const int died = (W.r_flags[rec] & RF_DIED) ? 1 : 0;
          c_app = W.r_ndup[rec];
          c_id = W.r_ndup[rec];
          c_info = W.r_nrm[rec] + died;
          c_hole = W.r_nrm[rec] + died;
        }
      } else {
        const int o = o_win & 0xFFFF;
        win = ((o & O_RM1) ? claim_load(W, p1) == t : claim_wr_load(W, p1) > t) &&
              ((o & O_RM2) ? claim_load(W, p2) == t : claim_wr_load(W, p2) > t);
        if (win) {
          c_app = (o & O_APPEND) ? 1 : 0;
          c_id = (o & O_NEWID) ? 1 : 0;
          c_info = ((o & O_INFO1) ? 1 : 0) + ((o & O_INFO2) ? 1 : 0);
          c_hole = ((o & O_RM1) && !(o & O_SLOT1) ? 1 : 0) + ((o & O_RM2) && !(o & O_SLOT2) ? 1 : 0);
        }
      }
    }
    // ranks of what the winners create: two scans of packed counts
    long long ai = ((long long)c_id << 32) | (long long)c_app;
    long long ih = ((long long)c_hole << 32) | (long long)c_info;
    long long r_ai, r_ih, t_ai, t_ih;
    r_ai = block_sum_excl<TAIL_THREADS>(ai, scan_tmp2, &t_ai);
    r_ih = block_sum_excl<TAIL_THREADS>(ih, scan_tmp2, &t_ih);
    const int r_app = (int)(r_ai & 0xFFFFFFFFll), r_id = (int)(r_ai >> 32);
    const int r_info = (int)(r_ih & 0xFFFFFFFFll), r_hole = (int)(r_ih >> 32);
    int tot;
    if (tid == 0) {
      s_tot[0] = (int)(t_ai & 0xFFFFFFFFll);
      s_tot[1] = (int)(t_ai >> 32);
      s_tot[2] = (int)(t_ih & 0xFFFFFFFFll);
      s_tot[3] = (int)(t_ih >> 32);
      const int th = s_tot[3];
      s_hole_base = th ? (long long)atomicAdd(W.hole_top, (unsigned long long)th) : 0ll;
    }
    __syncthreads();
    // 4. apply the winners at their ranks
    if (win) {
      TailAlloc al;
      al.B = B;
      al.app = r_app;
      al.id = r_id;
      al.info = r_info;
      al.hole = (int64_t)s_hole_base + r_hole;
      al.live = nullptr;
      al.s_live = s_live;
      al.n_samp = al.n_coag = al.n_warn = 0;
      if (is_acc) acc_apply(B, A, W, box, t, rec, p1, type, want_terms, rid, al);
      else perset_merge(B, A, W, box, k, t, p1, p2, seed, step, type, want_terms, rid, al, o_win,
                        &ms);
      c_samp += al.n_samp;
      c_coag += al.n_coag;
      c_warn += al.n_warn;
    }
    tail_barrier();
    if (tid == 0) {
      B->n_part += s_tot[0];
      B->next_id += s_tot[1];
      B->n_info += s_tot[2];
    }
    // 5. release the claims
    if (effect) {
      claim_release(W, p1);
      if (is_acc) rm_walk(W, rec, [&](int q) { claim_release(W, q); });
      else claim_release(W, p2);
    }
    // 6. what stays in the window: losers (run again only if an earlier
    // winner changed what they read) and tests whose inputs changed
    int keep = 0;
    bool athin = false;
    if (active && !win) {
      bool changed = consumed_before(W, p1, t, rid) || consumed_before(W, p2, t, rid);
      if (is_acc && effect && !changed)
        rm_walk(W, rec, [&](int q) { changed = changed || consumed_before(W, q, t, rid); });
      if (is_acc && p1 >= 0 && !changed)
        rd_walk(W, rec, [&](int q) { changed = changed || consumed_before(W, q, t, rid); });
      if (effect || changed) {
        keep = 1;
        if (changed) fl |= TF_RUN;
      } else if (fl & TF_ATHIN) {
        athin = true;  // decided below
      } else {
        // final without effect: its counts
        if (is_acc && tns) {
          c_samp += tns & (ACC_WARN - 1);
          atomicAdd((unsigned long long *)&B->n_acc_samp,
                    (unsigned long long)(tns & (ACC_WARN - 1)));
          if (tns & ACC_WARN) c_warn += 1;
        }
        if (!is_acc && (fl & TF_WARN)) c_warn += 1;
      }
    }
    if (win && !is_acc && (fl & TF_WARN)) c_warn += 1;
    // A test thinned on an empty arrival entry of list L (bin bL) is final
    // once no earlier test of the window that won this round (its arrivals
    // show from the next round on) or waits for another round could add a
    // particle to L: particles only grow, so a test on bins b1, b2 adds
    // them to bins from max(b1, b2) up, and a target event to bins from its
    // target's.  On the pair's second list the bound is the test's own
    // index (earlier tests of the pair can change the first list's draws).
    // Until then it waits, and it runs once more when it is no longer
    // blocked.
    for (int q = tid; q < W.K_stride; q += TAIL_THREADS) s_pm[q] = UNCLAIMED;
    __syncthreads();
    int b1 = 1, b2 = 1, kind = 0;
    if (active) {
      int c1, c2;
      decode_pair(k - pb, W.tab_classes, W.max_bin, &c1, &c2, &b1, &b2);
      kind = W.pair_acc[k];
    }
    const int b_lo = kind ? (kind == 1 ? b1 : b2) : max(b1, b2);
    if (active && (win || keep)) atomicMin(&s_pm[b_lo - 1], t);
    __syncthreads();
    // running minimum over the bins, TAIL_THREADS bins at a time
    const int nbin = B->n_bin;
    for (int base = 0; base < nbin; base += TAIL_THREADS) {
      const int q = base + tid;
      uint32_t v = q < nbin ? s_pm[q] : UNCLAIMED;
      if (base > 0 && tid == 0) v = min(v, s_carry);
      uint32_t tot_m;
      const uint32_t r = block_scan_incl<TAIL_THREADS>(v, OpMin(), scan_tmp3, &tot_m);
      if (q < nbin) s_pm[q] = r;
      if (tid == TAIL_THREADS - 1) s_carry = r;
      __syncthreads();
    }
    if (athin) {
      const bool side2 = !kind && (fl & TF_SIDE2);
      const int b_list = kind ? b_lo : (side2 ? b2 : b1);
      const bool blocked = (side2 ? t : W.pair_exc[k]) > s_pm[b_list - 1];
      if (blocked) {
        // wait, without running, until the events that block it settle
        keep = 1;
        fl |= TF_WAIT;
      } else if (fl & TF_WAIT) {
        // unblocked since its last run: run once more, on the settled lists
        keep = 1;
        fl = (fl & ~TF_WAIT) | TF_RUN;
      }
    }
    int r_keep;
    r_keep = block_sum_excl<TAIL_THREADS>(keep, scan_tmp, &tot);
    if (keep) {
      w_t[r_keep] = t;
      w_k[r_keep] = (int)(k - pb);
      w_fl[r_keep] = fl;
      w_p1[r_keep] = p1;
      w_p2[r_keep] = p2;
      w_tns[r_keep] = tns;
    }
    wn = tot;
    rid++;
    n_rounds++;
  }
  __syncthreads();
  // flush the gathered live-count and counter changes
  int *live = W.live + (int64_t)box * W.K_stride;
  for (int q = tid; q < W.K_stride; q += TAIL_THREADS)
    if (s_live[q]) atomicAdd(&live[q], s_live[q]);
  if (c_samp) {
    atomicAdd((unsigned long long *)&B->n_samp, (unsigned long long)c_samp);
    atomicAdd((unsigned long long *)&B->step_samp, (unsigned long long)c_samp);
  }
  if (c_coag) {
    atomicAdd((unsigned long long *)&B->n_coag, (unsigned long long)c_coag);
    atomicAdd((unsigned long long *)&B->step_coag, (unsigned long long)c_coag);
  }
  if (c_warn) atomicAdd((unsigned long long *)&B->n_warn, (unsigned long long)c_warn);
  if (tid == 0 && n_rounds) {
    atomicAdd(rounds_out, n_rounds);
    B->n_rounds += (int)n_rounds;
    B->n_tail_tests += hi - lo;
  }
}

// the pair of every pending test, found in parallel before the tail
__global__ void k_pend_pairs(PW W, int64_t NP, const uint32_t *pend) {
  const int n = W.rc[RC_TAIL];
  PAR_FOR(i, n) W.pend_k[i] = (int)test_pair(W, NP, pend[i]);
}

// the largest bin count among the boxes (their pair space for this step);
// also clears the two entries after that pair space, which the pair scans
// read as their last input
__global__ void __launch_bounds__(256) k_max_nbin(const BoxDesc *boxes, int n_box, uint32_t *out) {
  __shared__ scan_storage<256, int> tr;
  int m = 1;
  for (int b = threadIdx.x; b < n_box; b += 256)
    if (!boxes[b].error) m = max(m, boxes[b].n_bin);
  const int r = block_reduce<256>(m, OpMax(), tr);
  if (threadIdx.x == 0) out[0] = (uint32_t)r;
}

__global__ void k_clear_pair_ends(PW W, int64_t NP) {
  if (threadIdx.x == 0) {
    W.pair_n[NP] = 0;
    W.pair_accn[NP] = 0;
  }
}

__global__ void k_set_u64(unsigned long long *p, unsigned long long v) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *p = v;
}

__global__ void k_fill_int(int *p, int64_t n, int v) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}
__global__ void k_fill_u32(uint32_t *p, int64_t n, uint32_t v) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}
__global__ void k_fill_u8(uint8_t *p, int64_t n, uint8_t v) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

__global__ void k_step_begin(BoxDesc *boxes, int n_box) {
  int box = blockIdx.x * blockDim.x + threadIdx.x;
  if (box >= n_box) return;
  BoxDesc *B = &boxes[box];
  B->step_samp = 0;
  B->step_coag = 0;
  B->part_steps += B->n_part;
}

// Tail-fill compaction over all holes at once.  Per box: the range of its
// holes in the sorted list and how many lie below the final count.  Then
// the j-th such head hole receives the j-th live slot of the tail
// [n_final, n) in slot order: v = n_final + j + m with m the number of the
// tail's holes below v, found by binary search over those holes.  Sources
// lie at or above n_final and destinations below it, so the moves run in
// place.  Live particles of the tail keep their slot order.
__global__ void k_compact_ranges(BoxDesc *boxes, PW W, int n_box, int n_holes) {
  const int box = blockIdx.x * blockDim.x + threadIdx.x;
  if (box >= n_box) return;
  auto lower = [&](uint64_t key, int lo, int hi) {
    while (lo < hi) {
      const int mid = (lo + hi) >> 1;
      if (W.holes_sorted[mid] < key) lo = mid + 1; else hi = mid;
    }
    return lo;
  };
  const int lo = lower((uint64_t)box << 32, 0, n_holes);
  const int hi = lower((uint64_t)(box + 1) << 32, lo, n_holes);
  const int n_final = boxes[box].n_part - (hi - lo);
  W.cmp_lo[box] = lo;
  W.cmp_hi[box] = hi;
  W.cmp_head[box] = lower(((uint64_t)box << 32) | (uint64_t)(uint32_t)n_final, lo, hi) - lo;
}

__global__ void k_compact_move(BoxDesc *boxes, Arrays A, PW W, int n_holes) {
  const int jg = blockIdx.x * blockDim.x + threadIdx.x;
  if (jg >= n_holes) return;
  const uint64_t h = W.holes_sorted[jg];
  const int box = (int)(h >> 32);
  const int lo = W.cmp_lo[box], hi = W.cmp_hi[box], head = W.cmp_head[box];
  const int j = jg - lo;
  if (j >= head) return;
  const BoxDesc *B = &boxes[box];
  const int n_final = B->n_part - (hi - lo);
  // tail holes: sorted slots at [lo + head, hi); the smallest m with
  // n_final + j + m < tail_hole[m] (or m = their count)
  int a = 0, b = hi - lo - head;
  while (a < b) {
    const int m = (a + b) >> 1;
    const int slot = (int)(W.holes_sorted[lo + head + m] & 0xFFFFFFFFu);
    if (n_final + j + m < slot) b = m; else a = m + 1;
  }
  const int from = n_final + j + a;
  const int hole = (int)(h & 0xFFFFFFFFu);
  const int64_t pf = B->off + from, pt = B->off + hole;
  copy_particle(A, pf, pt);
  W.key[pt] = W.key[pf];
  W.terms[pt] = W.terms[pf];
  W.status[pt] = W.status[pf];
}

__global__ void k_compact_finish(BoxDesc *boxes, PW W, int n_box) {
  const int box = blockIdx.x * blockDim.x + threadIdx.x;
  if (box >= n_box) return;
  boxes[box].n_part -= W.cmp_hi[box] - W.cmp_lo[box];
}

__global__ void k_par_invalidate(BoxDesc *boxes, int n_box) {
  int box = blockIdx.x * blockDim.x + threadIdx.x;
  if (box >= n_box) return;
  boxes[box].pkeys_valid = 0;
  boxes[box].terms_kernel = 0;
}

__global__ void k_par_end(BoxDesc *boxes, int n_box) {
  int box = blockIdx.x * blockDim.x + threadIdx.x;
  if (box >= n_box) return;
  BoxDesc *B = &boxes[box];
  B->valid_sort = 1;
  B->xlists_valid = 0;
}

// ------------------------------------------------------------ host side
// The reference's tot_n_samp counts each pair's Poisson draw over the
// members present when the loop reaches it (src/coagulation.F90:627): the
// batch's thinned tests come off the counts drawn over the bound.
__global__ void __launch_bounds__(256) k_thin_counts(BoxDesc *boxes, PW W, int64_t NP, uint32_t t0,
                                                     int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  int box = -1;
  if (i < n && W.thin[i]) box = (int)(test_pair(W, NP, t0 + (uint32_t)i) / W.P_stride);
  wave_flush_counts(boxes, box, -1, 0, 0);
}

inline int parallel_init(ParallelWork &W, int n_box, int max_bin) {
  W.n_box = n_box;
  W.d.max_bin = max_bin;
  W.d.tab_classes = 1;
  if (hipHostMalloc((void **)&W.h_u32, 64 * sizeof(uint32_t)) != hipSuccess) return ERR_HIP;
  if (hipHostMalloc((void **)&W.h_i32, 64 * sizeof(int)) != hipSuccess) return ERR_HIP;
  if (hipHostMalloc((void **)&W.h_u64, 8 * sizeof(unsigned long long)) != hipSuccess) return ERR_HIP;
  if (hipStreamCreate(&W.st2) != hipSuccess) return ERR_HIP;
  if (hipMalloc((void **)&W.pstat, 2 * sizeof(unsigned long long)) != hipSuccess) return ERR_HIP;
  if (hipMemset(W.pstat, 0, 2 * sizeof(unsigned long long)) != hipSuccess) return ERR_HIP;
  if (hipEventCreateWithFlags(&W.ev_fork, hipEventDisableTiming) != hipSuccess) return ERR_HIP;
  if (hipEventCreateWithFlags(&W.ev_join, hipEventDisableTiming) != hipSuccess) return ERR_HIP;
  return 0;
}

// the accelerated path's records, allocated when a step first has target slots
inline void parallel_free_records(ParallelWork &W) {
  PW &d = W.d;
  void *ptrs[] = {d.r_rdhead, d.r_nrd, d.r_nsamp, d.r_ncoag, d.r_nrm, d.r_nchain, d.r_head, d.r_ndup,
                  d.r_dhead, d.r_flags, d.r_err, d.q_vol, d.q_group, d.q_nprim, d.q_ncomp,
                  d.q_next, d.q_csrc, d.q_ctime, d.q_least,
                  d.q_greatest, d.dup_top, d.rm_buf, d.rm_top};
  for (void *p : ptrs)
    if (p) (void)hipFree(p);
  d.r_nsamp = d.r_ncoag = d.r_nrm = d.r_nchain = d.r_head = d.r_ndup = nullptr;
  d.r_dhead = d.r_flags = d.r_err = d.r_rdhead = d.r_nrd = nullptr;
  d.q_vol = d.q_ctime = d.q_least = d.q_greatest = nullptr;
  d.q_group = d.q_nprim = d.q_ncomp = d.q_next = d.q_csrc = nullptr;
  d.dup_top = d.rm_buf = d.rm_top = nullptr;
  d.acc_cap = d.pool_cap = d.rm_cap = 0;
}

inline int parallel_alloc_records(ParallelWork &W, int n_spec) {
  PW &d = W.d;
  const int64_t c = ACC_CAP, pc = (int64_t)ACC_CAP + DUP_CAP;
  int **rv[] = {&d.r_nsamp, &d.r_ncoag, &d.r_nrm, &d.r_nchain, &d.r_head, &d.r_ndup,
                &d.r_dhead, &d.r_flags, &d.r_err, &d.r_rdhead, &d.r_nrd};
  for (int **q : rv) PCHECK(pmalloc(q, c));
  PCHECK(pmalloc(&d.q_vol, (size_t)n_spec * pc));
  PCHECK(pmalloc(&d.q_csrc, (size_t)MAX_COMP * pc));
  PCHECK(pmalloc(&d.q_ctime, (size_t)MAX_COMP * pc));
  int **qv[] = {&d.q_group, &d.q_nprim, &d.q_ncomp, &d.q_next};
  for (int **q : qv) PCHECK(pmalloc(q, pc));
  PCHECK(pmalloc(&d.q_least, pc));
  PCHECK(pmalloc(&d.q_greatest, pc));
  PCHECK(pmalloc(&d.dup_top, 1));
  PCHECK(pmalloc(&d.rm_buf, RM_CAP));
  PCHECK(pmalloc(&d.rm_top, 1));
  d.acc_cap = (int)c;
  d.pool_cap = (int)pc;
  d.rm_cap = RM_CAP;
  return 0;
}

inline void parallel_free_arrays(ParallelWork &W) {
  PW &d = W.d;
  void *ptrs[] = {d.key,  d.status, d.claim, d.claim_wr, d.terms, d.perm, d.tile_box, d.tile_start,
                  d.tiles_in_box, d.tile_first, d.hist, d.hist_scan, d.lstart, d.lcnt, d.live,
                  d.pair_n, d.pair_exc, d.acc, d.acc_scan, d.tp1, d.tp2, d.tid, d.pend, d.ev_t,
                  d.ev_p1, d.ev_p2, d.ev_box, d.ev_out, d.f_app, d.f_id, d.f_info, d.f_hole,
                  d.f_lose, d.f_pend, d.s_pend, d.dead_t, d.dead_r, d.s_app, d.s_id, d.s_info, d.s_hole,
                  d.s_lose, d.box_first,
                  d.holes, d.holes_sorted, d.cmp_lo, d.cmp_hi, d.cmp_head, d.acc_samp,
                  d.scan_part, W.sort_tmp, d.rc,
                  d.pair_acc, d.pair_accn, d.acc_exc, d.trec, d.tns, d.ev_rec, d.misc,
                  d.hole_top, d.pend_k, d.tp_in, d.tpair,
                  d.acap, d.arr_n, d.arr_head, d.ap_p, d.ap_t, d.ap_dep, d.ap_next, d.ap_top,
                  d.arr_of, d.key0, d.depart_t, d.thin, d.minpend, d.minpend_next};
  for (void *p : ptrs)
    if (p) (void)hipFree(p);
  parallel_free_records(W);
  int ks = d.K_stride;
  int64_t ps = d.P_stride;
  int tc = d.tab_classes, mb = d.max_bin;
  d = PW{};
  d.K_stride = ks;
  d.P_stride = ps;
  d.tab_classes = tc;
  d.max_bin = mb;
  W.sort_tmp = nullptr;
  W.sort_tmp_bytes = 0;
}

inline void parallel_free(ParallelWork &W) {
  parallel_free_arrays(W);
  if (W.h_u32) (void)hipHostFree(W.h_u32);
  if (W.h_i32) (void)hipHostFree(W.h_i32);
  if (W.h_u64) (void)hipHostFree(W.h_u64);
  W.h_u64 = nullptr;
  W.h_u32 = nullptr;
  W.h_i32 = nullptr;
  if (W.ev_fork) (void)hipEventDestroy(W.ev_fork);
  if (W.ev_join) (void)hipEventDestroy(W.ev_join);
  if (W.st2) (void)hipStreamDestroy(W.st2);
  if (W.pstat) (void)hipFree(W.pstat);
  W.ev_fork = W.ev_join = nullptr;
  W.st2 = nullptr;
  W.pstat = nullptr;
}

// (re)allocate everything sized by particle capacities and the class count
inline int parallel_layout(ParallelWork &W, const Arrays &A, int n_box,
                           const std::vector<int64_t> &cap, int max_bin) {
  parallel_free_arrays(W);
  PW &d = W.d;
  d.max_bin = max_bin;
  d.K_stride = max_bin * d.tab_classes;
  d.P_stride = (int64_t)d.tab_classes * d.tab_classes * max_bin * max_bin;
  // the sort and the tail keep two ints per list of a box in LDS beside
  // their static LDS: refuse grids and class counts that do not fit
  {
    int dev = 0, lds_max = 0;
    PCHECK(hipGetDevice(&dev));
    PCHECK(hipDeviceGetAttribute(&lds_max, hipDeviceAttributeMaxSharedMemoryPerBlock, dev));
    hipFuncAttributes fa_tail, fa_scatter;
    PCHECK(hipFuncGetAttributes(&fa_tail, reinterpret_cast<const void *>(&k_par_tail)));
    PCHECK(hipFuncGetAttributes(&fa_scatter, reinterpret_cast<const void *>(&k_par_scatter)));
    const size_t dyn = 2 * (size_t)d.K_stride * sizeof(int);
    if (fa_tail.sharedSizeBytes + dyn > (size_t)lds_max ||
        fa_scatter.sharedSizeBytes + dyn > (size_t)lds_max) {
      g_parallel_error = "parallel mode: " + std::to_string(max_bin) + " bins x " +
                         std::to_string(d.tab_classes) + " weight classes need " +
                         std::to_string(dyn) + " bytes of LDS per box list set, above the " +
                         std::to_string(lds_max) + " a workgroup has beside the kernels' own";
      return ERR_CAPACITY;
    }
  }
  const int64_t tc = A.total_cap;
  PCHECK(pmalloc(&d.key, tc));
  PCHECK(pmalloc(&d.status, tc));
  PCHECK(pmalloc(&d.claim, tc));
  PCHECK(pmalloc(&d.claim_wr, tc));
  PCHECK(pmalloc(&d.dead_t, tc));
  PCHECK(pmalloc(&d.dead_r, tc));
  PCHECK(pmalloc(&d.terms, tc));
  PCHECK(pmalloc(&d.perm, tc));
  PCHECK(pmalloc(&d.key0, tc));
  PCHECK(pmalloc(&d.depart_t, tc));
  PCHECK(pmalloc(&d.arr_of, tc));
  // tiles over each box's capacity
  std::vector<int> tb, ts;
  std::vector<int64_t> hoff(n_box);
  W.tiles_in_box.assign(n_box, 0);
  int64_t H = 0;
  for (int b = 0; b < n_box; ++b) {
    int nt = (int)((cap[b] + TILE - 1) / TILE);
    if (nt < 1) nt = 1;
    W.tiles_in_box[b] = nt;
    for (int t = 0; t < nt; ++t) {
      tb.push_back(b);
      ts.push_back(t * TILE);
    }
    hoff[b] = (int64_t)tb.size() - nt;
    H += (int64_t)d.K_stride * nt;
  }
  W.n_tiles = (int)tb.size();
  W.H = H;
  PCHECK(pmalloc(&d.tile_box, tb.size()));
  PCHECK(pmalloc(&d.tile_start, ts.size()));
  PCHECK(pmalloc(&d.tiles_in_box, n_box));
  PCHECK(pmalloc(&d.tile_first, n_box));
  PCHECK(hipMemcpy(d.tile_box, tb.data(), tb.size() * sizeof(int), hipMemcpyHostToDevice));
  PCHECK(hipMemcpy(d.tile_start, ts.data(), ts.size() * sizeof(int), hipMemcpyHostToDevice));
  PCHECK(hipMemcpy(d.tiles_in_box, W.tiles_in_box.data(), n_box * sizeof(int), hipMemcpyHostToDevice));
  PCHECK(hipMemcpy(d.tile_first, hoff.data(), n_box * sizeof(int64_t), hipMemcpyHostToDevice));
  d.K_hist = d.K_stride;
  PCHECK(pmalloc(&d.hist, H));
  PCHECK(pmalloc(&d.hist_scan, H + 1));
  PCHECK(pmalloc(&d.lstart, (size_t)n_box * d.K_stride));
  PCHECK(pmalloc(&d.lcnt, (size_t)n_box * d.K_stride));
  PCHECK(pmalloc(&d.live, (size_t)n_box * d.K_stride));
  PCHECK(hipMemset(d.live, 0, sizeof(int) * (size_t)n_box * d.K_stride));
  PCHECK(pmalloc(&d.acap, (size_t)n_box * d.K_stride));
  PCHECK(pmalloc(&d.arr_n, (size_t)n_box * d.K_stride));
  PCHECK(hipMemset(d.arr_n, 0, sizeof(int) * (size_t)n_box * d.K_stride));
  PCHECK(pmalloc(&d.arr_head, (size_t)n_box * d.K_stride));
  W.NP = (int64_t)n_box * d.P_stride;
  PCHECK(pmalloc(&d.pair_n, W.NP + 1));
  PCHECK(pmalloc(&d.pair_exc, W.NP + 1));
  PCHECK(hipMemset(d.pair_n, 0, sizeof(uint32_t) * (W.NP + 1)));
  PCHECK(pmalloc(&d.pair_acc, W.NP));
  PCHECK(pmalloc(&d.pair_accn, W.NP + 1));
  PCHECK(pmalloc(&d.acc_exc, W.NP + 1));
  PCHECK(hipMemset(d.pair_accn, 0, sizeof(uint32_t) * (W.NP + 1)));
  PCHECK(pmalloc(&d.misc, 8));
  PCHECK(pmalloc(&d.rc, RC_SIZE));
  PCHECK(hipMemset(d.rc, 0, sizeof(int) * RC_SIZE));
  d.pstat = W.pstat;
  PCHECK(pmalloc(&d.hole_top, 8));
  d.tail_rounds = d.hole_top + 1;
  d.rstat = d.hole_top + 2;
  const int64_t T = W.T_cap;
  PCHECK(pmalloc(&d.acc, T + 1));
  PCHECK(pmalloc(&d.acc_scan, T + 1));
  PCHECK(pmalloc(&d.tp1, T));
  PCHECK(pmalloc(&d.tp2, T));
  PCHECK(pmalloc(&d.tid, T));
  PCHECK(pmalloc(&d.pend, T));
  PCHECK(pmalloc(&d.ev_t, T));
  PCHECK(pmalloc(&d.ev_p1, T));
  PCHECK(pmalloc(&d.ev_p2, T));
  PCHECK(pmalloc(&d.ev_box, T));
  PCHECK(pmalloc(&d.ev_out, T));
  PCHECK(pmalloc(&d.trec, T));
  PCHECK(pmalloc(&d.pend_k, T));
  PCHECK(pmalloc(&d.tp_in, T));
  PCHECK(pmalloc(&d.tpair, T));
  PCHECK(pmalloc(&d.tns, T));
  PCHECK(pmalloc(&d.thin, T));
  PCHECK(pmalloc(&d.minpend, (size_t)n_box * d.K_stride));
  PCHECK(pmalloc(&d.minpend_next, (size_t)n_box * d.K_stride));
  // arrivals of one step: at most one per merge or copy, over every batch
  d.ap_cap = (uint32_t)std::min<int64_t>(2 * T + DUP_CAP, (int64_t)INT32_MAX / 2);
  d.ap_ovf = (int64_t)n_box * d.K_stride * A_MAX;
  if (d.ap_ovf + d.ap_cap > (int64_t)INT32_MAX) {
    g_parallel_error = "arrival pool larger than 2^31 entries";
    return ERR_CAPACITY;
  }
  PCHECK(pmalloc(&d.ap_p, d.ap_ovf + d.ap_cap));
  PCHECK(pmalloc(&d.ap_t, d.ap_ovf + d.ap_cap));
  PCHECK(pmalloc(&d.ap_dep, d.ap_ovf + d.ap_cap));
  PCHECK(pmalloc(&d.ap_next, d.ap_ovf + d.ap_cap));
  PCHECK(pmalloc(&d.ap_top, 1));
  PCHECK(pmalloc(&d.ev_rec, T));
  int **fl[] = {&d.f_app, &d.f_id, &d.f_info, &d.f_hole, &d.f_lose, &d.f_pend,
                &d.s_app, &d.s_id, &d.s_info, &d.s_hole, &d.s_lose, &d.s_pend};
  for (int **f : fl) PCHECK(pmalloc(f, T + 1));
  PCHECK(pmalloc(&d.box_first, n_box));
  int64_t total_cap = 0;
  for (int b = 0; b < n_box; ++b) total_cap += cap[b];
  W.hole_cap = total_cap + 1024;
  PCHECK(pmalloc(&d.holes, W.hole_cap));
  PCHECK(pmalloc(&d.holes_sorted, W.hole_cap));
  PCHECK(pmalloc(&d.cmp_lo, n_box));
  PCHECK(pmalloc(&d.acc_samp, 1));
  PCHECK(hipMemset(d.acc_samp, 0, sizeof(unsigned long long)));
  PCHECK(pmalloc(&d.cmp_hi, n_box));
  PCHECK(pmalloc(&d.cmp_head, n_box));
  // the chunk totals of the device-wide scans
  PCHECK(pmalloc(&d.scan_part, (size_t)SCAN_COLS * SCAN_MAX_BLOCKS));
  PCHECK(rocprim::radix_sort_keys(nullptr, W.sort_tmp_bytes, d.holes, d.holes_sorted,
                                  (unsigned)W.hole_cap, 0u, 64u, (hipStream_t)0));
  PCHECK(hipMalloc(&W.sort_tmp, W.sort_tmp_bytes));
  hipLaunchKernelGGL(k_fill_u32, dim3((tc + 255) / 256), dim3(256), 0, 0, d.claim, tc, UNCLAIMED);
  hipLaunchKernelGGL(k_fill_u32, dim3((tc + 255) / 256), dim3(256), 0, 0, d.claim_wr, tc, UNCLAIMED);
  hipLaunchKernelGGL(k_fill_u8, dim3((tc + 255) / 256), dim3(256), 0, 0, d.status, tc, ST_LIVE);
  W.epoch = 0;
  PCHECK(hipGetLastError());
  PCHECK(hipDeviceSynchronize());
  return 0;
}

inline int parallel_classes(ParallelWork &W, int n_box, int classes) {
  (void)n_box;
  W.d.tab_classes = classes;
  // arrays are re-laid out at the next parallel_layout call; mark stale
  W.d.K_stride = W.d.max_bin * classes;
  W.d.P_stride = (int64_t)classes * classes * W.d.max_bin * W.d.max_bin;
  g_parallel_error = "";
  return 0;
}

// the whole parallel step; the host reads back the test and event counts
inline int parallel_step(ParallelWork &W, const Arrays &A, BoxDesc *d_box, int n_box,
                         const std::vector<int64_t> &cap, int max_bin, hipStream_t st, int type,
                         double del_t, uint64_t seed, uint32_t step, int precision, double band,
                         int validation) {
  PW &d = W.d;
  if (d.key == nullptr || d.K_stride != max_bin * d.tab_classes || W.NP != (int64_t)n_box * d.P_stride) {
    int e = parallel_layout(W, A, n_box, cap, max_bin);
    if (e) return e;
    hipLaunchKernelGGL(k_par_invalidate, dim3((n_box + 63) / 64), dim3(64), 0, st, d_box, n_box);
  }
  const int want_terms = (precision != 0) ? 1 : 0;
  const dim3 gb((n_box + 63) / 64), bb(64);
  const dim3 gt(W.n_tiles), bt(TILE_THREADS);
  d.tpair_ok = 0;  // built once the step's tests are laid out
  hipLaunchKernelGGL(k_step_begin, gb, bb, 0, st, d_box, n_box);
  // 1. keys for boxes with a valid grid but stale keys, radius extremes for
  // unsorted boxes, the remake test, and keys again after a new grid
  hipLaunchKernelGGL(k_par_zero_live, dim3(1, n_box), dim3(256), 0, st, d_box, d, n_box, 0);
  hipLaunchKernelGGL(k_par_rekey, gt, bt, 0, st, d_box, A, d, type, want_terms, 0);
  hipLaunchKernelGGL(k_par_finish_rekey, gb, bb, 0, st, d_box, n_box, type, want_terms, 0);
  hipLaunchKernelGGL(k_par_reset_minmax, gb, bb, 0, st, d_box, n_box);
  hipLaunchKernelGGL(k_par_minmax, gt, bt, 0, st, d_box, A, d);
  hipLaunchKernelGGL(k_par_remake, gb, bb, 0, st, d_box, A, d, n_box);
  hipLaunchKernelGGL(k_par_zero_live, dim3(1, n_box), dim3(256), 0, st, d_box, d, n_box, 1);
  hipLaunchKernelGGL(k_par_rekey, gt, bt, 0, st, d_box, A, d, type, 0, 1);
  hipLaunchKernelGGL(k_par_finish_rekey, gb, bb, 0, st, d_box, n_box, type, 0, 1);
  // The largest grid among the boxes this step: the pair space spans it
  // (not the allocation's max_bin) and the sort keys need only its bits.
  hipLaunchKernelGGL(k_max_nbin, dim3(1), dim3(256), 0, st, d_box, n_box, d.misc);
  PCHECK(hipMemcpyAsync(W.h_u32 + 8, d.misc, sizeof(uint32_t), hipMemcpyDeviceToHost, st));
  PCHECK(hipStreamSynchronize(st));
  const int mb_step = std::max(1, std::min((int)W.h_u32[8], max_bin));
  d.max_bin = mb_step;
  d.P_stride = (int64_t)d.tab_classes * d.tab_classes * mb_step * mb_step;
  W.NP = (int64_t)n_box * d.P_stride;
  hipLaunchKernelGGL(k_clear_pair_ends, dim3(1), dim3(64), 0, st, d, W.NP);
  // 2. majorant and pair tables (shared with the exact mode)
  {
    const int mb = max_bin;
    hipLaunchKernelGGL(k_kbounds, dim3((mb * mb + 255) / 256, n_box), dim3(256), 0, st, d_box, A, type);
    hipLaunchKernelGGL(k_mark_bounds, gb, bb, 0, st, d_box, n_box, type);
    hipLaunchKernelGGL(k_pair_tables,
                       dim3((d.tab_classes * d.tab_classes * mb * mb + 255) / 256, n_box), dim3(256),
                       0, st, d_box, A);
    hipLaunchKernelGGL(k_mark_tables, gb, bb, 0, st, d_box, n_box);
  }
  // 3. stable counting sort of the keys
  // the histogram spans the step's grid: keys are below mb_step * classes
  d.K_hist = mb_step * d.tab_classes;
  PLAUNCH(W, "k_par_hist", k_par_hist, gt, bt, d.K_hist * sizeof(int), st, d_box, d);
  {
    const int64_t nh = (int64_t)d.K_hist * W.n_tiles;
    scan_excl_sum(scan_job(d.hist, d.hist_scan, d.scan_part, nh), nh, st);
  }
  hipLaunchKernelGGL(k_par_lists, dim3((d.K_stride + 255) / 256, n_box), dim3(256), 0, st, d_box, d,
                     n_box);
  hipLaunchKernelGGL(k_par_acap, dim3((d.K_stride + 255) / 256, n_box), dim3(256), 0, st, d_box, d);
  // keys are below mb_step * classes; the sentinel 2^key_bits - 1 above them
  int key_bits = 1;
  while ((1 << key_bits) <= mb_step * d.tab_classes) key_bits++;
  PLAUNCH(W, "k_par_scatter", k_par_scatter, gt, bt, 2 * d.K_stride * sizeof(int), st, d_box, d,
          key_bits);
  // 4. Poisson test counts per bin pair, laid out in reference order
  PLAUNCH(W, "k_par_pairs", k_par_pairs, dim3((unsigned)((d.P_stride + 255) / 256), n_box),
          dim3(256), 0, st, d_box, A, d, del_t, seed, step);
  {
    // both in one pass; each writes its total after the pair space
    ScanJob J = scan_job(d.pair_n, d.pair_exc, d.scan_part, W.NP);
    J.in[1] = d.pair_accn;
    J.out[1] = d.acc_exc;
    J.cols = 2;
    scan_excl_sum(J, W.NP, st);
  }
  // the step's test count, its target slots and the first batch's bounds,
  // in one read
  hipLaunchKernelGGL(k_batch_end, dim3(1), dim3(64), 0, st, d, W.NP, 0u, (uint32_t)W.T_cap,
                     (uint32_t)ACC_CAP, d.misc);
  PCHECK(hipMemcpyAsync(W.h_u32 + 4, d.misc, 5 * sizeof(uint32_t), hipMemcpyDeviceToHost, st));
  PCHECK(hipStreamSynchronize(st));
  const uint32_t T = W.h_u32[7];
  const uint32_t T_acc = W.h_u32[8];
  // the test-to-pair table, when the step is one batch
  d.tpair_ok = 0;
  if (T > 0 && (int64_t)T <= W.T_cap) {
    hipLaunchKernelGGL(k_fill_u32, dim3((T + 255) / 256), dim3(256), 0, st, d.tp_in, (int64_t)T,
                       0u);
    hipLaunchKernelGGL(k_mark_pair_starts, dim3((unsigned)((W.NP + 255) / 256)), dim3(256), 0, st,
                       d, W.NP);
    scan_incl_max(scan_job(d.tp_in, d.tpair, d.scan_part, (int64_t)T), (int64_t)T, st);
    d.tpair_ok = 1;
  }
  if (T_acc > 0 && d.r_nsamp == nullptr) {
    int e = parallel_alloc_records(W, A.n_spec);
    if (e) return e;
  }
  // every particle starts the step live: a fresh status epoch, and a clear
  // array when the 6-bit epoch wraps
  if (++W.epoch > 63) {
    W.epoch = 1;
    hipLaunchKernelGGL(k_fill_u8, dim3((unsigned)((A.total_cap + 255) / 256)), dim3(256), 0, st,
                       d.status, A.total_cap, ST_LIVE);
  }
  d.epoch = W.epoch;
  // 5. tests, claims and merges, in batches of T_cap tests.  The rounds of
  // a batch are launched without reading their counts back: each kernel
  // reads them from d.rc and a round that does not run finds a zero count.
  // Grids are sized from the previous batch's counts, and the kernels stride
  // over any excess.
  int64_t n_holes = 0;
  hipLaunchKernelGGL(k_set_u64, dim3(1), dim3(64), 0, st, d.hole_top, 0ull);
  for (uint32_t t0 = 0; t0 < T;) {
    uint32_t t1, a0, a1;
    if (t0 == 0) {
      t1 = W.h_u32[4];
      a0 = W.h_u32[5];
      a1 = W.h_u32[6];
    } else {
      // the batch ends where its target slots would overflow the records
      hipLaunchKernelGGL(k_batch_end, dim3(1), dim3(64), 0, st, d, W.NP, t0, (uint32_t)W.T_cap,
                         (uint32_t)d.acc_cap, d.misc);
      PCHECK(hipMemcpyAsync(W.h_u32 + 4, d.misc, 3 * sizeof(uint32_t), hipMemcpyDeviceToHost, st));
      PCHECK(hipStreamSynchronize(st));
      t1 = W.h_u32[4];
      a0 = W.h_u32[5];
      a1 = W.h_u32[6];
    }
    const int n_acc = (int)(a1 - a0);
    const int n_batch = (int)(t1 - t0);
    const uint32_t batch_t0 = t0;
    t0 = t1;
    if (n_batch > 0)
      hipLaunchKernelGGL(k_fill_u8, dim3((n_batch + 255) / 256), dim3(256), 0, st, d.thin,
                         (int64_t)n_batch, (uint8_t)0);
    // at the start of the first round every event of the batch is pending
    hipLaunchKernelGGL(k_fill_u32, dim3((unsigned)(((int64_t)n_box * d.K_stride + 255) / 256)),
                       dim3(256), 0, st, d.minpend, (int64_t)n_box * d.K_stride, 0u);
    hipLaunchKernelGGL(k_fill_u32, dim3((unsigned)(((int64_t)n_box * d.K_stride + 255) / 256)),
                       dim3(256), 0, st, d.minpend_next, (int64_t)n_box * d.K_stride, UNCLAIMED);
    hipLaunchKernelGGL(k_batch_begin, dim3(1), dim3(64), 0, st, d, n_batch);
    const uint32_t *pend = nullptr;
    // a grid for at most `bound` items, `per` a workgroup, sized from the
    // last batch's count with a margin (the kernels stride over any excess)
    auto grid = [](int64_t bound, int64_t last, int per) -> unsigned {
      int64_t want = last >= 0 ? last + last / 4 + 1024 : bound;
      if (want > bound) want = bound;
      return (unsigned)std::max<int64_t>(1, (want + per - 1) / per);
    };
    for (int round = 0; round < PAR_ROUNDS; ++round) {
      // a later round runs only while at most half of the batch is pending
      const int nb = round == 0 ? n_batch : n_batch / 2;  // its largest test count
      if (nb <= 0) break;  // a batch of one test has no second round
      const unsigned g_n = round == 0 ? (unsigned)((nb + 255) / 256) : grid(nb, W.est[1], 256);
      const unsigned g_e = grid(nb, W.est[round == 0 ? 0 : 2], 256);
      W.rid++;
      if (n_acc > 0) {
        hipLaunchKernelGGL(k_fill_int, dim3(1), dim3(64), 0, st, d.rm_top, 1, 0);
        hipLaunchKernelGGL(k_fill_int, dim3(1), dim3(64), 0, st, d.dup_top, 1, 0);
      }
      // the target walks run on the second stream beside the per-set tests
      // (disjoint slots; both only claim with atomics)
      if (n_acc > 0) {
        PCHECK(hipEventRecord(W.ev_fork, st));
        PCHECK(hipStreamWaitEvent(W.st2, W.ev_fork, 0));
        W.tst = W.st2;
        const unsigned g_a = pend ? grid(nb, W.est[1], 64) : (unsigned)((n_acc + ACC_LANES - 1) / ACC_LANES);
        PLAUNCH(W, "k_par_acc_tests", k_par_acc_tests, dim3(g_a), dim3(64), 0, W.st2,
                d_box, A, d, n_box, type, seed, step, batch_t0, pend, del_t, a0, n_acc);
        W.tst = nullptr;
        PCHECK(hipEventRecord(W.ev_join, W.st2));
      }
      if (round == 0)
        PLAUNCH(W, "k_par_tests", k_par_tests<false>, dim3(g_n), dim3(256), 0, st, d_box, A, d,
                n_box, type, seed, step, batch_t0, pend, precision, band, validation);
      else
        PLAUNCH(W, "k_par_tests", k_par_tests<true>, dim3(g_n), dim3(256), 0, st, d_box, A, d,
                n_box, type, seed, step, batch_t0, pend, precision, band, validation);
      if (n_acc > 0) PCHECK(hipStreamWaitEvent(st, W.ev_join, 0));
      scan_excl_sum(scan_job(d.acc, d.acc_scan, d.scan_part, 0, d.rc + RC_N, d.rc + RC_E), nb, st);
      hipLaunchKernelGGL(k_par_gather, dim3(g_n), dim3(256), 0, st, d);
      hipLaunchKernelGGL(k_fill_int, dim3((n_box + 255) / 256), dim3(256), 0, st, d.box_first,
                         (int64_t)n_box, -1);
      hipLaunchKernelGGL(k_par_resolve, dim3(g_e), dim3(256), 0, st, d_box, A, d, n_box, seed, step);
      hipLaunchKernelGGL(k_par_box_first, dim3(g_e), dim3(256), 0, st, d);
      {
        // the winners' ranks: the five flag arrays in one pass
        const int *fs[5] = {d.f_app, d.f_id, d.f_info, d.f_hole, d.f_lose};
        int *ss[5] = {d.s_app, d.s_id, d.s_info, d.s_hole, d.s_lose};
        ScanJob J = scan_job(fs[0], ss[0], d.scan_part, 0, d.rc + RC_E);
        for (int q = 1; q < 5; ++q) {
          J.in[q] = (const uint32_t *)fs[q];
          J.out[q] = (uint32_t *)ss[q];
        }
        J.cols = 5;
        scan_excl_sum(J, nb, st);
      }
      // one wavefront per event, a few events each: most have no long chain
      hipLaunchKernelGGL(k_par_apply_chains, dim3(std::min(grid(nb, W.est[round == 0 ? 0 : 2], 16), 4096u)),
                         dim3(256), 0, st, d_box, A, d, (int)W.rid);
      PLAUNCH(W, "k_par_apply", k_par_apply, dim3(g_e), dim3(256), 0, st, d_box, A, d,
              n_box, seed, step, type, want_terms, (int)W.rid);
      hipLaunchKernelGGL(k_par_box_totals, gb, bb, 0, st, d_box, d, n_box);
      hipLaunchKernelGGL(k_par_release, dim3(g_e), dim3(256), 0, st, d);
      hipLaunchKernelGGL(k_par_pending, dim3(g_n), dim3(256), 0, st, d_box, d, n_box, (int)W.rid);
      hipLaunchKernelGGL(k_minpend_prefix, dim3(n_box), dim3(PREFIX_THREADS), 0, st, d);
      std::swap(d.minpend, d.minpend_next);
      scan_excl_sum(scan_job(d.f_pend, d.s_pend, d.scan_part, 0, d.rc + RC_N), nb, st);
      hipLaunchKernelGGL(k_par_gather_pending, dim3(g_n), dim3(256), 0, st, d);
      hipLaunchKernelGGL(k_round_end, dim3(1), dim3(64), 0, st, d, n_batch, round);
      // the next round and the tail read the pending tests from d.pend: the
      // tests kernels copy them into tid before the round writes d.pend
      pend = d.pend;
    }
    // the rest per box, one workgroup each
    if (n_acc > 0) {
      hipLaunchKernelGGL(k_fill_int, dim3(1), dim3(64), 0, st, d.rm_top, 1, 0);
      hipLaunchKernelGGL(k_fill_int, dim3(1), dim3(64), 0, st, d.dup_top, 1, 0);
    }
    hipLaunchKernelGGL(k_set_u64, dim3(1), dim3(64), 0, st, d.tail_rounds, 0ull);
    hipLaunchKernelGGL(k_pend_pairs, dim3(grid(n_batch, W.est[3], 256)), dim3(256), 0, st, d,
                       W.NP, d.pend);
    PLAUNCH(W, "k_par_tail", k_par_tail, dim3(n_box), dim3(TAIL_THREADS),
            2 * d.K_stride * sizeof(int), st, d_box, A, d,
            n_box, d.pend, type, seed, step, precision, band, del_t, a0,
            want_terms, (int)W.rid + 1, d.tail_rounds, batch_t0);
    if (n_batch > 0)
      hipLaunchKernelGGL(k_thin_counts, dim3((n_batch + 255) / 256), dim3(256), 0, st, d_box, d,
                         W.NP, batch_t0, n_batch);
    // the holes, the tail's rounds and the batch's round sizes in one read
    PCHECK(hipMemcpyAsync(W.h_u64, d.hole_top, 6 * sizeof(unsigned long long),
                          hipMemcpyDeviceToHost, st));
    PCHECK(hipStreamSynchronize(st));
    n_holes = (int64_t)W.h_u64[0];
    W.n_tail_rounds += (int64_t)W.h_u64[1];
    for (int q = 0; q < 4; ++q) W.est[q] = (int64_t)W.h_u64[2 + q];
    W.rid += (int64_t)W.h_u64[1] + 2;  // past every round id the tail used
  }
  // 6. compaction
  if (n_holes > 0) {
    PCHECK(rocprim::radix_sort_keys(W.sort_tmp, W.sort_tmp_bytes, d.holes, d.holes_sorted,
                                    (unsigned)n_holes, 0u, 64u, st));
    hipLaunchKernelGGL(k_compact_ranges, gb, bb, 0, st, d_box, d, n_box, (int)n_holes);
    PLAUNCH(W, "k_par_compact", k_compact_move, dim3((unsigned)((n_holes + 255) / 256)), dim3(256),
            0, st, d_box, A, d, (int)n_holes);
    hipLaunchKernelGGL(k_compact_finish, gb, bb, 0, st, d_box, d, n_box);
  }
  hipLaunchKernelGGL(k_par_arr_over, dim3((d.K_stride + 255) / 256, n_box), dim3(256), 0, st, d_box,
                     d);
  hipLaunchKernelGGL(k_par_end, gb, bb, 0, st, d_box, n_box);
  PCHECK(hipGetLastError());
  return 0;
}

}  // namespace pmc
