// pmc_coag.hip -- C ABI and host orchestration of the gfx950 coagulation
// step (PartMC mc_coag, src/coagulation.F90:43-96).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdarg>
#include <cstddef>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/partmc_coag.h"
#include "pmc_device.h"
#include "pmc_state.h"

namespace pmc {
// Hall (1980) collision efficiencies as tabulated in the reference's
// effic (src/coag_kernel_sedi.F90:190-245), in its DATA (column-major) order.
__constant__ double HALL_R0[15] = {6.0,  8.0,  10.0, 15.0,  20.0,  25.0,  30.0, 40.0,
                                   50.0, 60.0, 70.0, 100.0, 150.0, 200.0, 300.0};
__constant__ double HALL_RAT[21] = {0.0,  0.05, 0.1,  0.15, 0.2,  0.25, 0.3,
                                    0.35, 0.4,  0.45, 0.5,  0.55, 0.6,  0.65,
                                    0.7,  0.75, 0.8,  0.85, 0.9,  0.95, 1.0};
__constant__ double HALL_ECOLL[315] = {
    0.001, 0.001, 0.001, 0.001, 0.001, 0.001, 0.001, 0.001, 0.001, 0.001, 0.001, 0.001,
    0.001, 0.001, 0.001, 0.003, 0.003, 0.003, 0.004, 0.005, 0.005, 0.005, 0.010, 0.100,
    0.050, 0.200, 0.500, 0.770, 0.870, 0.970, 0.007, 0.007, 0.007, 0.008, 0.009, 0.010,
    0.010, 0.070, 0.400, 0.430, 0.580, 0.790, 0.930, 0.960, 1.000, 0.009, 0.009, 0.009,
    0.012, 0.015, 0.010, 0.020, 0.280, 0.600, 0.640, 0.750, 0.910, 0.970, 0.980, 1.000,
    0.014, 0.014, 0.014, 0.015, 0.016, 0.030, 0.060, 0.500, 0.700, 0.770, 0.840, 0.950,
    0.970, 1.000, 1.000, 0.017, 0.017, 0.017, 0.020, 0.022, 0.060, 0.100, 0.620, 0.780,
    0.840, 0.880, 0.950, 1.000, 1.000, 1.000, 0.030, 0.030, 0.024, 0.022, 0.032, 0.062,
    0.200, 0.680, 0.830, 0.870, 0.900, 0.950, 1.000, 1.000, 1.000, 0.025, 0.025, 0.025,
    0.036, 0.043, 0.130, 0.270, 0.740, 0.860, 0.890, 0.920, 1.000, 1.000, 1.000, 1.000,
    0.027, 0.027, 0.027, 0.040, 0.052, 0.200, 0.400, 0.780, 0.880, 0.900, 0.940, 1.000,
    1.000, 1.000, 1.000, 0.030, 0.030, 0.030, 0.047, 0.064, 0.250, 0.500, 0.800, 0.900,
    0.910, 0.950, 1.000, 1.000, 1.000, 1.000, 0.040, 0.040, 0.033, 0.037, 0.068, 0.240,
    0.550, 0.800, 0.900, 0.910, 0.950, 1.000, 1.000, 1.000, 1.000, 0.035, 0.035, 0.035,
    0.055, 0.079, 0.290, 0.580, 0.800, 0.900, 0.910, 0.950, 1.000, 1.000, 1.000, 1.000,
    0.037, 0.037, 0.037, 0.062, 0.082, 0.290, 0.590, 0.780, 0.900, 0.910, 0.950, 1.000,
    1.000, 1.000, 1.000, 0.037, 0.037, 0.037, 0.060, 0.080, 0.290, 0.580, 0.770, 0.890,
    0.910, 0.950, 1.000, 1.000, 1.000, 1.000, 0.037, 0.037, 0.037, 0.041, 0.075, 0.250,
    0.540, 0.760, 0.880, 0.920, 0.950, 1.000, 1.000, 1.000, 1.000, 0.037, 0.037, 0.037,
    0.052, 0.067, 0.250, 0.510, 0.770, 0.880, 0.930, 0.970, 1.000, 1.000, 1.000, 1.000,
    0.037, 0.037, 0.037, 0.047, 0.057, 0.250, 0.490, 0.770, 0.890, 0.950, 1.000, 1.000,
    1.000, 1.000, 1.000, 0.036, 0.036, 0.036, 0.042, 0.048, 0.230, 0.470, 0.780, 0.920,
    1.000, 1.020, 1.020, 1.020, 1.020, 1.020, 0.040, 0.040, 0.035, 0.033, 0.040, 0.112,
    0.450, 0.790, 1.010, 1.030, 1.040, 1.040, 1.040, 1.040, 1.040, 0.033, 0.033, 0.033,
    0.033, 0.033, 0.119, 0.470, 0.950, 1.300, 1.700, 2.300, 2.300, 2.300, 2.300, 2.300,
    0.027, 0.027, 0.027, 0.027, 0.027, 0.125, 0.520, 1.400, 2.300, 3.000, 4.000, 4.000,
    4.000, 4.000, 4.000};
}  // namespace pmc

#include "pmc_exact.h"
#include "pmc_parallel.h"
#include "pmc_stats.h"
#include "pmc_rebalance.h"

using namespace pmc;

// ------------------------------------------------------------------ errors
// text of the last error on this thread, written in place (an error path
// allocates nothing)
static thread_local char g_last_error[512];

static int fail(int code, const char *fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_last_error, sizeof(g_last_error), fmt, ap);
  va_end(ap);
  return code;
}

#define HIPCHECK(expr)                                                          \
  do {                                                                          \
    hipError_t e_ = (expr);                                                     \
    if (e_ != hipSuccess)                                                       \
      return fail(ERR_HIP, "%s failed: %s", #expr, hipGetErrorString(e_));      \
  } while (0)

// ------------------------------------------------------------------ handle
struct pmc_gpu {
  int device = 0;
  hipStream_t stream = nullptr;
  int n_box = 0, n_spec = 0, max_bin = 256, tab_classes = 1;
  std::vector<double> density;
  std::vector<int64_t> cap;     // per-box particle capacity
  std::vector<int64_t> info_cap;
  BoxDesc *d_box = nullptr;
  Arrays A{};
  int mode = PMC_MODE_EXACT;
  int precision = PMC_PREC_FP64;
  double band = 1e-5;
  int validation = 0;
  uint64_t seed = 0;
  uint32_t step = 0;
  ParallelWork work{};
  // population control: n_part_ideal per box (< 0: not set) and scratch
  std::vector<double> n_part_ideal;
  RebalWork rw{};
  int64_t rw_slots = 0;  // slots of rw.rm / rw.rp
  // per-kernel HIP-event timing
  int timing = 0;
  struct Timed {
    std::string name;
    hipEvent_t start, stop;
  };
  std::vector<Timed> pending;
  std::vector<std::pair<std::string, std::pair<double, int64_t>>> totals;
  std::vector<double *> tape_buf;  // per box: device values of its tape, or null
};

static void timing_begin(pmc_gpu *h, const char *name, hipEvent_t *start) {
  (void)name;
  if (!h->timing) return;
  (void)hipEventCreate(start);
  (void)hipEventRecord(*start, h->work.tst ? h->work.tst : h->stream);
}

static void timing_end(pmc_gpu *h, const char *name, hipEvent_t start) {
  if (!h->timing) return;
  hipEvent_t stop;
  (void)hipEventCreate(&stop);
  (void)hipEventRecord(stop, h->work.tst ? h->work.tst : h->stream);
  h->pending.push_back({name, start, stop});
}

static void timing_collect(pmc_gpu *h) {
  if (h->pending.empty()) return;
  (void)hipStreamSynchronize(h->stream);
  for (auto &t : h->pending) {
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, t.start, t.stop);
    bool found = false;
    for (auto &e : h->totals)
      if (e.first == t.name) {
        e.second.first += ms;
        e.second.second += 1;
        found = true;
      }
    if (!found) h->totals.push_back({t.name, {(double)ms, 1}});
    (void)hipEventDestroy(t.start);
    (void)hipEventDestroy(t.stop);
  }
  h->pending.clear();
}

#define LAUNCH_TIMED(h, name, ...)                 \
  do {                                             \
    hipEvent_t ev_start_;                          \
    timing_begin(h, name, &ev_start_);             \
    hipLaunchKernelGGL(__VA_ARGS__);               \
    timing_end(h, name, ev_start_);                \
  } while (0)

static void free_particles(pmc_gpu *h) {
  Arrays &A = h->A;
  void *ptrs[] = {A.vol,   A.vtot,    A.id,      A.group,   A.klass,    A.nprim,
                  A.ncomp, A.csrc,    A.ctime,   A.least,   A.greatest, A.fwd_bin,
                  A.fwd_class, A.fwd_idx, A.pool, A.info_id, A.info_other, A.info_action};
  for (void *p : ptrs)
    if (p) (void)hipFree(p);
  A.vol = A.vtot = A.ctime = A.least = A.greatest = nullptr;
  A.id = nullptr;
  A.group = A.klass = A.nprim = A.ncomp = A.csrc = nullptr;
  A.fwd_bin = A.fwd_class = A.fwd_idx = A.pool = nullptr;
  A.info_id = A.info_other = nullptr;
  A.info_action = nullptr;
  A.total_cap = 0;
}

template <class T>
static hipError_t dmalloc(T **p, size_t n) {
  return hipMalloc((void **)p, std::max<size_t>(n, 1) * sizeof(T));
}

static int alloc_tables(pmc_gpu *h, int classes) {
  Arrays &A = h->A;
  if (A.kw) (void)hipFree(A.kw);
  if (A.cc) (void)hipFree(A.cc);
  size_t n = (size_t)h->n_box * classes * classes * h->max_bin * h->max_bin;
  A.tab_classes = classes;
  HIPCHECK(dmalloc(&A.kw, n));
  HIPCHECK(dmalloc(&A.cc, n));
  h->tab_classes = classes;
  return 0;
}

// Lay out particle storage for the given per-box capacities, keeping the
// particles already uploaded, their bin lists (the list order is state the
// next step depends on) and the removal logs.
static int layout(pmc_gpu *h, const std::vector<int64_t> &new_cap) {
  std::vector<BoxDesc> hb(h->n_box);
  HIPCHECK(hipMemcpy(hb.data(), h->d_box, sizeof(BoxDesc) * h->n_box, hipMemcpyDeviceToHost));
  Arrays old = h->A;
  int64_t total = 0, pool_total = 0, info_total = 0;
  std::vector<int64_t> off(h->n_box), poff(h->n_box), ioff(h->n_box), phalf(h->n_box);
  for (int b = 0; b < h->n_box; ++b) {
    off[b] = total;
    total += new_cap[b];
    phalf[b] = 2 * new_cap[b] + 16 * (int64_t)h->A.max_keys;
    poff[b] = pool_total;
    pool_total += 2 * phalf[b];
    ioff[b] = info_total;
    info_total += new_cap[b] + 64;
  }
  Arrays A = h->A;
  A.total_cap = total;
  HIPCHECK(dmalloc(&A.vol, (size_t)total * h->n_spec));
  HIPCHECK(dmalloc(&A.vtot, total));
  HIPCHECK(dmalloc(&A.id, total));
  HIPCHECK(dmalloc(&A.group, total));
  HIPCHECK(dmalloc(&A.klass, total));
  HIPCHECK(dmalloc(&A.nprim, total));
  HIPCHECK(dmalloc(&A.ncomp, total));
  HIPCHECK(dmalloc(&A.csrc, (size_t)total * MAX_COMP));
  HIPCHECK(dmalloc(&A.ctime, (size_t)total * MAX_COMP));
  HIPCHECK(dmalloc(&A.least, total));
  HIPCHECK(dmalloc(&A.greatest, total));
  HIPCHECK(dmalloc(&A.fwd_bin, total));
  HIPCHECK(dmalloc(&A.fwd_class, total));
  HIPCHECK(dmalloc(&A.fwd_idx, total));
  HIPCHECK(dmalloc(&A.pool, pool_total));
  HIPCHECK(dmalloc(&A.info_id, info_total));
  HIPCHECK(dmalloc(&A.info_other, info_total));
  HIPCHECK(dmalloc(&A.info_action, info_total));
  // move existing particles
  for (int b = 0; b < h->n_box; ++b) {
    int64_t n = hb[b].n_part;
    if (old.total_cap > 0 && n > 0) {
      int64_t o = hb[b].off, d = off[b];
      for (int s = 0; s < h->n_spec; ++s)
        HIPCHECK(hipMemcpy(A.vol + s * total + d, old.vol + s * old.total_cap + o,
                           n * sizeof(double), hipMemcpyDeviceToDevice));
      HIPCHECK(hipMemcpy(A.csrc + d * MAX_COMP, old.csrc + o * MAX_COMP,
                         n * MAX_COMP * sizeof(int), hipMemcpyDeviceToDevice));
      HIPCHECK(hipMemcpy(A.ctime + d * MAX_COMP, old.ctime + o * MAX_COMP,
                         n * MAX_COMP * sizeof(double), hipMemcpyDeviceToDevice));
#define MOVE(F, T)                                                               \
  HIPCHECK(hipMemcpy(A.F + d, old.F + o, n * sizeof(T), hipMemcpyDeviceToDevice))
      MOVE(vtot, double);
      MOVE(id, int64_t);
      MOVE(group, int);
      MOVE(klass, int);
      MOVE(nprim, int);
      MOVE(ncomp, int);
      MOVE(least, double);
      MOVE(greatest, double);
      MOVE(fwd_bin, int);
      MOVE(fwd_class, int);
      MOVE(fwd_idx, int);
#undef MOVE
    }
    if (old.total_cap > 0) {
      // the live half of the list pool, with list starts shifted to match
      const int64_t o = hb[b].pool_off + (int64_t)hb[b].pool_cur * hb[b].pool_half;
      const int64_t d = poff[b] + (int64_t)hb[b].pool_cur * phalf[b];
      if (hb[b].pool_top > 0)
        HIPCHECK(hipMemcpy(A.pool + d, old.pool + o, hb[b].pool_top * sizeof(int),
                           hipMemcpyDeviceToDevice));
      std::vector<int> ls(A.max_keys);
      int *lsd = A.lstart + (int64_t)b * A.max_keys;
      HIPCHECK(hipMemcpy(ls.data(), lsd, sizeof(int) * A.max_keys, hipMemcpyDeviceToHost));
      for (int &v : ls) v = (int)(v - o + d);
      HIPCHECK(hipMemcpy(lsd, ls.data(), sizeof(int) * A.max_keys, hipMemcpyHostToDevice));
      // the removal log
      const int64_t ni = std::min<int64_t>(hb[b].n_info, hb[b].info_cap);
      if (ni > 0) {
        const int64_t io = hb[b].info_off, id_ = ioff[b];
        HIPCHECK(hipMemcpy(A.info_id + id_, old.info_id + io, ni * sizeof(int64_t), hipMemcpyDeviceToDevice));
        HIPCHECK(hipMemcpy(A.info_other + id_, old.info_other + io, ni * sizeof(int64_t), hipMemcpyDeviceToDevice));
        HIPCHECK(hipMemcpy(A.info_action + id_, old.info_action + io, ni * sizeof(int), hipMemcpyDeviceToDevice));
      }
    }
    hb[b].off = off[b];
    hb[b].cap = new_cap[b];
    hb[b].pool_off = poff[b];
    hb[b].pool_half = phalf[b];
    hb[b].info_off = ioff[b];
    hb[b].info_cap = new_cap[b] + 64;
    // parallel-mode keys and terms live in per-particle work arrays that are
    // laid out again by the next parallel step
    hb[b].pkeys_valid = 0;
    hb[b].terms_kernel = 0;
  }
  h->A = old;
  free_particles(h);
  h->A = A;
  h->cap = new_cap;
  HIPCHECK(hipMemcpy(h->d_box, hb.data(), sizeof(BoxDesc) * h->n_box, hipMemcpyHostToDevice));
  parallel_free_arrays(h->work);  // re-laid out lazily by the next parallel step
  return 0;
}

template <class T>
static int set_field(pmc_gpu *h, int box, size_t offset, const T &v) {
  HIPCHECK(hipMemcpy((char *)(h->d_box + box) + offset, &v, sizeof(T), hipMemcpyHostToDevice));
  return 0;
}
#define SET_FIELD(h, box, field, v) set_field(h, box, offsetof(BoxDesc, field), v)

static int check_box(pmc_gpu *h, int box) {
  if (!h) return fail(ERR_ARG, "null handle");
  if (box < 0 || box >= h->n_box) return fail(ERR_ARG, "box %d out of range [0, %d)", box, h->n_box);
  return 0;
}

// Size distributions (pmc_stats.h): argument checks shared by the entry
// points, and the passes of every bin window enqueued on the handle's
// stream.  partials holds DIST_BLOCKS * min(n_bin, DIST_WIN) doubles and is
// fully rewritten by each window's partial pass before its reduce reads it.
static int check_size_dist(pmc_gpu *h, int box, int quantity, int n_bin, double d_min,
                           double d_max, const double *hist) {
  int e = check_box(h, box);
  if (e) return e;
  if (!hist || n_bin < 1 || !(d_min > 0.0 && d_min < d_max))
    return fail(ERR_ARG, "size_dist: need an output array, n_bin >= 1 and 0 < d_min < d_max");
  if (quantity != DIST_NUM && quantity != DIST_MASS && quantity != DIST_VOL)
    return fail(ERR_ARG, "size_dist: unknown quantity %d", quantity);
  return 0;
}

static hipError_t size_dist_enqueue(pmc_gpu *h, int box, int quantity, int n_bin, double d_min,
                                    double d_max, double *partials, double *d_hist) {
  // the grid of bin_grid_make (log) over diameters: its end points
  const double log_lo = std::log(d_min), log_hi = std::log(d_max);
  for (int lo = 0; lo < n_bin; lo += DIST_WIN) {
    const int win = std::min(DIST_WIN, n_bin - lo);
    int key_bits = 1;
    while ((1 << key_bits) <= win) key_bits++;  // keys 0..win, win for "not in this window"
    hipLaunchKernelGGL(k_dist_partial, dim3(DIST_BLOCKS), dim3(DIST_THREADS), 0, h->stream,
                       h->d_box, h->A, box, quantity, n_bin, log_lo, log_hi, lo, win, key_bits,
                       partials);
    hipLaunchKernelGGL(k_dist_reduce, dim3((win + 255) / 256), dim3(256), 0, h->stream, partials,
                       DIST_BLOCKS, lo, win, n_bin, d_hist);
  }
  return hipGetLastError();
}

// Scratch of pmc_gpu_size_dist_noalloc: the partial histograms of its
// DIST_BLOCKS workgroups over one window of DIST_WIN bins, 8 MiB.  Every
// device the library runs on holds its own copy until the process ends.
// On the host the name only stands for the symbol: kernels get the device
// address that scratch_setup records per device.
static __device__ __align__(256) double scratch[DIST_BLOCKS * DIST_WIN];
constexpr int SCRATCH_MAX_DEVICES = 64;
static double *g_scratch[SCRATCH_MAX_DEVICES];  // device address of `scratch`, per device
static std::mutex g_scratch_mu;

// Called with `device` current, once per handle; the address is looked up
// once per device.
static int scratch_setup(int device) {
  if (device < 0 || device >= SCRATCH_MAX_DEVICES)
    return fail(ERR_ARG, "device %d out of range [0, %d)", device, SCRATCH_MAX_DEVICES);
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (g_scratch[device]) return 0;
  void *p = nullptr;
  hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !p)
    return fail(ERR_HIP, "hipGetSymbolAddress(scratch) on device %d: %s", device,
                hipGetErrorString(e));
  g_scratch[device] = (double *)p;
  return 0;
}

extern "C" {

const char *pmc_gpu_last_error(void) { return g_last_error; }

int pmc_gpu_create(pmc_gpu **out, int device, int n_box, int n_spec, const double *density,
                   int max_bin) {
  if (!out || n_box < 1 || n_spec < 1 || n_spec > MAX_SPEC || !density)
    return fail(ERR_ARG, "pmc_gpu_create: bad arguments (n_box %d, n_spec %d)", n_box, n_spec);
  if (max_bin <= 0) max_bin = 256;
  pmc_gpu *h = new pmc_gpu();
  h->device = device;
  h->n_box = n_box;
  h->n_spec = n_spec;
  h->max_bin = max_bin;
  h->density.assign(density, density + n_spec);
  HIPCHECK(hipSetDevice(device));
  HIPCHECK(hipStreamCreate(&h->stream));
  if (int e = scratch_setup(device)) return e;
  Arrays &A = h->A;
  A.n_spec = n_spec;
  A.max_bin = max_bin;
  A.max_keys = max_bin * MAX_CLASS;
  A.d_min = *std::min_element(density, density + n_spec);
  A.d_max = *std::max_element(density, density + n_spec);
  HIPCHECK(dmalloc(&A.density, n_spec));
  HIPCHECK(hipMemcpy(A.density, density, n_spec * sizeof(double), hipMemcpyHostToDevice));
  HIPCHECK(dmalloc(&A.edges, (size_t)n_box * (max_bin + 1)));
  HIPCHECK(dmalloc(&A.centers, (size_t)n_box * max_bin));
  HIPCHECK(dmalloc(&A.widths, (size_t)n_box * max_bin));
  HIPCHECK(dmalloc(&A.kmin, (size_t)n_box * max_bin * max_bin));
  HIPCHECK(dmalloc(&A.kmax, (size_t)n_box * max_bin * max_bin));
  HIPCHECK(dmalloc(&A.lstart, (size_t)n_box * A.max_keys));
  HIPCHECK(dmalloc(&A.llen, (size_t)n_box * A.max_keys));
  HIPCHECK(dmalloc(&A.lcap, (size_t)n_box * A.max_keys));
  HIPCHECK(hipMemset(A.llen, 0, sizeof(int) * (size_t)n_box * A.max_keys));
  HIPCHECK(dmalloc(&A.sort_scratch, (size_t)n_box * SORT_THREADS * A.max_keys));
  int err = alloc_tables(h, 1);
  if (err) return err;
  std::vector<BoxDesc> hb(n_box);
  for (int b = 0; b < n_box; ++b) {
    BoxDesc &B = hb[b];
    memset(&B, 0, sizeof(B));
    B.next_id = 1;
    B.id_stride = 1;
    B.grid_req_n_bin = -1;
    B.w.n_group = 1;
    B.w.n_class = 1;
    B.w.type[0] = WEIGHT_NONE;
    B.w.mag[0] = 1.0;
    B.w.expo[0] = 0.0;
    B.env = make_env(288.0, 1e5, 1000.0);
  }
  HIPCHECK(dmalloc(&h->d_box, n_box));
  HIPCHECK(hipMemcpy(h->d_box, hb.data(), sizeof(BoxDesc) * n_box, hipMemcpyHostToDevice));
  h->cap.assign(n_box, 0);
  err = parallel_init(h->work, n_box, max_bin);
  if (err) return fail(err, "parallel work allocation failed");
  h->work.tbegin = [](void *c, const char *n, hipEvent_t *e) { timing_begin((pmc_gpu *)c, n, e); };
  h->work.tend = [](void *c, const char *n, hipEvent_t e) { timing_end((pmc_gpu *)c, n, e); };
  h->work.tctx = h;
  *out = h;
  return 0;
}

int pmc_gpu_destroy(pmc_gpu *h) {
  if (!h) return 0;
  (void)hipSetDevice(h->device);
  (void)hipStreamSynchronize(h->stream);
  free_particles(h);
  Arrays &A = h->A;
  void *ptrs[] = {A.density, A.edges, A.centers, A.widths, A.kmin, A.kmax, A.kw, A.cc,
                  A.lstart,  A.llen,  A.lcap,    A.sort_scratch, h->d_box,
                  h->rw.gc,  h->rw.plan, h->rw.tile_a, h->rw.tile_off, h->rw.tile_kmax,
                  h->rw.box_tot, h->rw.box_imax, h->rw.rm, h->rw.rp, h->rw.bval};
  for (void *p : ptrs)
    if (p) (void)hipFree(p);
  parallel_free(h->work);
  for (double *t : h->tape_buf)
    if (t) (void)hipFree(t);
  (void)hipStreamDestroy(h->stream);
  delete h;
  return 0;
}

int pmc_gpu_set_env(pmc_gpu *h, int box, double temp, double pressure, double additive_coeff) {
  int e = check_box(h, box);
  if (e) return e;
  Env env = make_env(temp, pressure, additive_coeff);
  HIPCHECK(hipSetDevice(h->device));
  HIPCHECK(hipStreamSynchronize(h->stream));
  if ((e = SET_FIELD(h, box, env, env))) return e;
  int zero = 0;
  return SET_FIELD(h, box, terms_kernel, zero);
}

int pmc_gpu_set_weights(pmc_gpu *h, int box, int n_group, int n_class, const int *type,
                        const double *magnitude, const double *exponent) {
  int e = check_box(h, box);
  if (e) return e;
  if (n_group < 1 || n_group > MAX_GROUP || n_class < 1 || n_class > MAX_CLASS || !type ||
      !magnitude || !exponent)
    return fail(ERR_ARG, "set_weights: n_group %d / n_class %d outside 1..%d/1..%d", n_group,
                n_class, MAX_GROUP, MAX_CLASS);
  Weights w{};
  w.n_group = n_group;
  w.n_class = n_class;
  for (int i = 0; i < n_group * n_class; ++i) {
    if (type[i] < WEIGHT_NONE || type[i] > WEIGHT_MFA)
      return fail(700421478, "unknown aero_weight type: %d", type[i]);
    w.type[i] = type[i];
    w.mag[i] = magnitude[i];
    w.expo[i] = exponent[i];
  }
  HIPCHECK(hipSetDevice(h->device));
  HIPCHECK(hipStreamSynchronize(h->stream));
  e = SET_FIELD(h, box, w, w);
  if (e) return e;
  int zero = 0;
  e = SET_FIELD(h, box, tables_valid, zero);
  if (e) return e;
  // class count changes the key space of the lists
  if ((e = SET_FIELD(h, box, xlists_valid, zero))) return e;
  if ((e = SET_FIELD(h, box, pkeys_valid, zero))) return e;
  if (n_class > h->tab_classes) {
    if ((e = alloc_tables(h, n_class))) return e;
    if ((e = parallel_classes(h->work, h->n_box, n_class))) return fail(e, "%s", parallel_last_error());
    std::vector<BoxDesc> hb(h->n_box);
    HIPCHECK(hipMemcpy(hb.data(), h->d_box, sizeof(BoxDesc) * h->n_box, hipMemcpyDeviceToHost));
    for (auto &B : hb) B.tables_valid = 0;
    HIPCHECK(hipMemcpy(h->d_box, hb.data(), sizeof(BoxDesc) * h->n_box, hipMemcpyHostToDevice));
  }
  return SET_FIELD(h, box, valid_sort, zero);
}

int pmc_gpu_set_rng(pmc_gpu *h, uint64_t seed, uint32_t step) {
  if (!h) return fail(ERR_ARG, "null handle");
  h->seed = seed;
  h->step = step;
  return 0;
}

int pmc_gpu_set_next_id(pmc_gpu *h, int box, int64_t next_id, int64_t id_stride,
                        int64_t id_offset) {
  int e = check_box(h, box);
  if (e) return e;
  if (id_stride < 1) return fail(ERR_ARG, "id_stride must be >= 1");
  HIPCHECK(hipSetDevice(h->device));
  if ((e = SET_FIELD(h, box, next_id, next_id))) return e;
  if ((e = SET_FIELD(h, box, id_stride, id_stride))) return e;
  return SET_FIELD(h, box, id_offset, id_offset);
}

int pmc_gpu_get_next_id(pmc_gpu *h, int box, int64_t *next_id) {
  int e = check_box(h, box);
  if (e) return e;
  HIPCHECK(hipSetDevice(h->device));
  HIPCHECK(hipMemcpy(next_id, (char *)(h->d_box + box) + offsetof(BoxDesc, next_id),
                     sizeof(int64_t), hipMemcpyDeviceToHost));
  return 0;
}

__global__ void k_default_components(Arrays A, int64_t base, int n, int set_comp, int set_ct) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  int64_t p = base + i;
  if (set_comp) {
    A.ncomp[p] = 1;
    for (int c = 0; c < MAX_COMP; ++c) {
      A.csrc[comp_at(p, c)] = 0;
      A.ctime[comp_at(p, c)] = 0.0;
    }
  }
  if (set_ct) {
    A.least[p] = 0.0;
    A.greatest[p] = 0.0;
  }
}

// Particle records n_part .. into the box's slots [first, first + n_part):
// volumes particle-major vol[i * n_spec + s] as aero_particle_t%vol, the
// optional component and create-time arrays defaulted when absent.
static int put_particles(pmc_gpu *h, const BoxDesc &hb, int64_t first, int n_part,
                         const double *vol, const int64_t *id, const int *group, const int *klass,
                         const int *n_primary, const int *n_comp, const int *comp_source,
                         const double *comp_time, const double *least, const double *greatest) {
  Arrays &A = h->A;
  const int64_t base = hb.off + first, tc = A.total_cap;
  const int ns = h->n_spec;
  if (n_part > 0) {
    std::vector<double> row(n_part);
    for (int s = 0; s < ns; ++s) {
      for (int i = 0; i < n_part; ++i) row[i] = vol[(size_t)i * ns + s];
      HIPCHECK(hipMemcpy(A.vol + s * tc + base, row.data(), n_part * sizeof(double),
                         hipMemcpyHostToDevice));
    }
    for (int i = 0; i < n_part; ++i) {
      double v = 0.0;
      for (int s = 0; s < ns; ++s) v += vol[(size_t)i * ns + s];
      row[i] = v;
    }
    HIPCHECK(hipMemcpy(A.vtot + base, row.data(), n_part * sizeof(double), hipMemcpyHostToDevice));
    if (id) HIPCHECK(hipMemcpy(A.id + base, id, n_part * sizeof(int64_t), hipMemcpyHostToDevice));
    HIPCHECK(hipMemcpy(A.group + base, group, n_part * sizeof(int), hipMemcpyHostToDevice));
    HIPCHECK(hipMemcpy(A.klass + base, klass, n_part * sizeof(int), hipMemcpyHostToDevice));
    if (n_primary) {
      HIPCHECK(hipMemcpy(A.nprim + base, n_primary, n_part * sizeof(int), hipMemcpyHostToDevice));
    } else {
      std::vector<int> ones(n_part, 1);
      HIPCHECK(hipMemcpy(A.nprim + base, ones.data(), n_part * sizeof(int), hipMemcpyHostToDevice));
    }
    if (n_comp) {
      HIPCHECK(hipMemcpy(A.ncomp + base, n_comp, n_part * sizeof(int), hipMemcpyHostToDevice));
      // the caller's [n_part][MAX_COMP] arrays are the device layout
      HIPCHECK(hipMemcpy(A.csrc + comp_at(base, 0), comp_source,
                         (size_t)n_part * MAX_COMP * sizeof(int), hipMemcpyHostToDevice));
      HIPCHECK(hipMemcpy(A.ctime + comp_at(base, 0), comp_time,
                         (size_t)n_part * MAX_COMP * sizeof(double), hipMemcpyHostToDevice));
    }
    if (least && greatest) {
      HIPCHECK(hipMemcpy(A.least + base, least, n_part * sizeof(double), hipMemcpyHostToDevice));
      HIPCHECK(hipMemcpy(A.greatest + base, greatest, n_part * sizeof(double),
                         hipMemcpyHostToDevice));
    }
    if (!n_comp || !(least && greatest)) {
      hipLaunchKernelGGL(k_default_components, dim3((n_part + 255) / 256), dim3(256), 0, h->stream,
                         A, base, n_part, n_comp ? 0 : 1, (least && greatest) ? 0 : 1);
      HIPCHECK(hipGetLastError());
      HIPCHECK(hipStreamSynchronize(h->stream));
    }
  }
  return 0;
}

int pmc_gpu_upload(pmc_gpu *h, int box, int n_part, int capacity, const double *vol,
                   const int64_t *id, const int *group, const int *klass, const int *n_primary,
                   const int *n_comp, const int *comp_source, const double *comp_time,
                   const double *least, const double *greatest) {
  int e = check_box(h, box);
  if (e) return e;
  if (n_part < 0 || (n_part > 0 && (!vol || !id || !group || !klass)))
    return fail(ERR_ARG, "upload: bad arguments");
  if (n_comp && (!comp_source || !comp_time))
    return fail(ERR_ARG, "upload: n_comp given without comp_source/comp_time");
  HIPCHECK(hipSetDevice(h->device));
  HIPCHECK(hipStreamSynchronize(h->stream));
  int64_t want = std::max<int64_t>(capacity, n_part);
  if (want < 16) want = 16;
  if (want > h->cap[box]) {
    std::vector<int64_t> nc = h->cap;
    nc[box] = want;
    if ((e = layout(h, nc))) return e;
  }
  BoxDesc hb;
  HIPCHECK(hipMemcpy(&hb, h->d_box + box, sizeof(BoxDesc), hipMemcpyDeviceToHost));
  if ((e = put_particles(h, hb, 0, n_part, vol, id, group, klass, n_primary, n_comp, comp_source,
                         comp_time, least, greatest)))
    return e;
  hb.n_part = n_part;
  hb.valid_sort = 0;
  hb.xlists_valid = 0;
  hb.pkeys_valid = 0;
  hb.terms_kernel = 0;
  hb.n_info = 0;
  hb.error = 0;
  HIPCHECK(hipMemcpy(h->d_box + box, &hb, sizeof(BoxDesc), hipMemcpyHostToDevice));
  return 0;
}

static int check_box_errors(pmc_gpu *h, std::vector<BoxDesc> &hb) {
  for (int b = 0; b < h->n_box; ++b)
    if (hb[b].tape.mode != TAPE_OFF && hb[b].tape.err)
      return fail(ERR_CAPACITY, "box %d: the %s tape ran out after %lld uniforms", b,
                  hb[b].tape.mode == TAPE_REPLAY ? "replayed" : "recording",
                  (long long)hb[b].tape.n);
  for (int b = 0; b < h->n_box; ++b)
    if (hb[b].error) {
      int code = hb[b].error;
      const char *what = "";
      switch (code) {
        case 765620746: case 134572570: what = "particle outside of bin_grid: try reducing the timestep del_t"; break;
        case 99: what = "sedimentation collision efficiency below 1e-20 (pmc_stop(99))"; break;
        case ERR_CAPACITY: what = "particle, list or grid capacity exceeded"; break;
        case 653606684: what = "prob_remove_source_max > 1"; break;
        case 538534122: what = "bin_grid requires a non-negative n_bin"; break;
        case 708651144: what = "particle outside the sort grid"; break;
        default: what = "device error";
      }
      return fail(code, "box %d: %s (code %d)", b, what, code);
    }
  return 0;
}

int pmc_gpu_set_bin_grid(pmc_gpu *h, int box, int n_bin, double r_min, double r_max) {
  int e = check_box(h, box);
  if (e) return e;
  if (n_bin < 0 || n_bin > h->max_bin)
    return fail(ERR_ARG, "set_bin_grid: n_bin %d outside 0..%d", n_bin, h->max_bin);
  if (n_bin > 0 && !(r_min > 0.0 && r_min < r_max))
    return fail(711537859, "bin_grid requires 0 < min < max");
  if (h->cap[box] == 0) return fail(ERR_STATE, "set_bin_grid: upload particles first");
  HIPCHECK(hipSetDevice(h->device));
  if ((e = SET_FIELD(h, box, grid_req_n_bin, n_bin))) return e;
  if ((e = SET_FIELD(h, box, grid_req_min, r_min))) return e;
  if ((e = SET_FIELD(h, box, grid_req_max, r_max))) return e;
  hipLaunchKernelGGL(k_prepare, dim3(1), dim3(SORT_THREADS), 0, h->stream, h->d_box, h->A,
                     box);
  HIPCHECK(hipGetLastError());
  HIPCHECK(hipStreamSynchronize(h->stream));
  std::vector<BoxDesc> hb(h->n_box);
  HIPCHECK(hipMemcpy(hb.data(), h->d_box, sizeof(BoxDesc) * h->n_box, hipMemcpyDeviceToHost));
  return check_box_errors(h, hb);
}

int pmc_gpu_get_n_part(pmc_gpu *h, int box, int *n_part) {
  int e = check_box(h, box);
  if (e) return e;
  HIPCHECK(hipSetDevice(h->device));
  HIPCHECK(hipStreamSynchronize(h->stream));
  HIPCHECK(hipMemcpy(n_part, (char *)(h->d_box + box) + offsetof(BoxDesc, n_part), sizeof(int),
                     hipMemcpyDeviceToHost));
  return 0;
}

int pmc_gpu_download(pmc_gpu *h, int box, int max_part, int *n_part, double *vol, int64_t *id,
                     int *group, int *klass, int *n_primary, int *n_comp, int *comp_source,
                     double *comp_time, double *least, double *greatest) {
  int e = check_box(h, box);
  if (e) return e;
  HIPCHECK(hipSetDevice(h->device));
  HIPCHECK(hipStreamSynchronize(h->stream));
  BoxDesc hb;
  HIPCHECK(hipMemcpy(&hb, h->d_box + box, sizeof(BoxDesc), hipMemcpyDeviceToHost));
  if (n_part) *n_part = hb.n_part;
  int n = hb.n_part;
  if (n > max_part) return fail(ERR_CAPACITY, "download: box holds %d particles, room for %d", n, max_part);
  if (n == 0) return 0;
  const Arrays &A = h->A;
  const int64_t base = hb.off, tc = A.total_cap;
  const int ns = h->n_spec;
  if (vol) {
    std::vector<double> row(n);
    for (int s = 0; s < ns; ++s) {
      HIPCHECK(hipMemcpy(row.data(), A.vol + s * tc + base, n * sizeof(double), hipMemcpyDeviceToHost));
      for (int i = 0; i < n; ++i) vol[(size_t)i * ns + s] = row[i];
    }
  }
  if (id) HIPCHECK(hipMemcpy(id, A.id + base, n * sizeof(int64_t), hipMemcpyDeviceToHost));
  if (group) HIPCHECK(hipMemcpy(group, A.group + base, n * sizeof(int), hipMemcpyDeviceToHost));
  if (klass) HIPCHECK(hipMemcpy(klass, A.klass + base, n * sizeof(int), hipMemcpyDeviceToHost));
  if (n_primary) HIPCHECK(hipMemcpy(n_primary, A.nprim + base, n * sizeof(int), hipMemcpyDeviceToHost));
  if (n_comp) HIPCHECK(hipMemcpy(n_comp, A.ncomp + base, n * sizeof(int), hipMemcpyDeviceToHost));
  if (comp_source || comp_time) {
    // the device keeps only the first n_comp entries of a particle; the
    // rest read as zero
    std::vector<int> nco(n);
    HIPCHECK(hipMemcpy(nco.data(), A.ncomp + base, n * sizeof(int), hipMemcpyDeviceToHost));
    // the device layout is the caller's [n][MAX_COMP]
    if (comp_source) {
      HIPCHECK(hipMemcpy(comp_source, A.csrc + comp_at(base, 0), (size_t)n * MAX_COMP * sizeof(int),
                         hipMemcpyDeviceToHost));
      for (int64_t i = 0; i < n; ++i)
        for (int c = nco[i] > 0 ? nco[i] : 0; c < MAX_COMP; ++c) comp_source[(size_t)i * MAX_COMP + c] = 0;
    }
    if (comp_time) {
      HIPCHECK(hipMemcpy(comp_time, A.ctime + comp_at(base, 0), (size_t)n * MAX_COMP * sizeof(double),
                         hipMemcpyDeviceToHost));
      for (int64_t i = 0; i < n; ++i)
        for (int c = nco[i] > 0 ? nco[i] : 0; c < MAX_COMP; ++c) comp_time[(size_t)i * MAX_COMP + c] = 0.0;
    }
  }
  if (least) HIPCHECK(hipMemcpy(least, A.least + base, n * sizeof(double), hipMemcpyDeviceToHost));
  if (greatest) HIPCHECK(hipMemcpy(greatest, A.greatest + base, n * sizeof(double), hipMemcpyDeviceToHost));
  return 0;
}

int pmc_gpu_set_mode(pmc_gpu *h, int mode, int precision, double band) {
  if (!h) return fail(ERR_ARG, "null handle");
  if (mode != PMC_MODE_EXACT && mode != PMC_MODE_PARALLEL) return fail(ERR_ARG, "unknown mode %d", mode);
  if (precision != PMC_PREC_FP64 && precision != PMC_PREC_FP32_RECHECK)
    return fail(ERR_ARG, "unknown precision %d", precision);
  if (mode == PMC_MODE_EXACT && precision != PMC_PREC_FP64)
    return fail(ERR_UNSUPPORTED, "exact mode evaluates the kernel in FP64 only");
  if (!(band >= 0.0)) return fail(ERR_ARG, "recheck band must be >= 0");
  h->mode = mode;
  h->precision = precision;
  h->band = band;
  return 0;
}

int pmc_gpu_set_validation(pmc_gpu *h, int on) {
  if (!h) return fail(ERR_ARG, "null handle");
  h->validation = on ? 1 : 0;
  return 0;
}

static int check_kernel_type(int t) {
  if (t == 6 || t == 7) return fail(ERR_UNSUPPORTED, "fractal kernel type %d is not supported", t);
  if (t < 1 || t > 5) return fail(200724934, "Unknown kernel type: %d", t);
  return 0;
}

static int enqueue_tables(pmc_gpu *h, int kernel_type) {
  const int mb = h->max_bin;
  dim3 gb((mb * mb + 255) / 256, h->n_box);
  hipLaunchKernelGGL(k_kbounds, gb, dim3(256), 0, h->stream, h->d_box, h->A, kernel_type);
  hipLaunchKernelGGL(k_mark_bounds, dim3((h->n_box + 63) / 64), dim3(64), 0, h->stream, h->d_box,
                     h->n_box, kernel_type);
  dim3 gt((h->tab_classes * h->tab_classes * mb * mb + 255) / 256, h->n_box);
  hipLaunchKernelGGL(k_pair_tables, gt, dim3(256), 0, h->stream, h->d_box, h->A);
  hipLaunchKernelGGL(k_mark_tables, dim3((h->n_box + 63) / 64), dim3(64), 0, h->stream, h->d_box,
                     h->n_box);
  HIPCHECK(hipGetLastError());
  return 0;
}

int pmc_gpu_mc_coag_async(pmc_gpu *h, int kernel_type, double del_t) {
  if (!h) return fail(ERR_ARG, "null handle");
  int e = check_kernel_type(kernel_type);
  if (e) return e;
  if (!(del_t >= 0.0)) return fail(ERR_ARG, "del_t must be >= 0");
  for (int b = 0; b < h->n_box; ++b)
    if (h->cap[b] == 0) return fail(ERR_STATE, "box %d has no particle storage: upload first", b);
  HIPCHECK(hipSetDevice(h->device));
  if (h->mode != PMC_MODE_EXACT)
    for (int b = 0; b < (int)h->tape_buf.size(); ++b)
      if (h->tape_buf[b])
        return fail(ERR_STATE, "box %d records or replays a tape: that is the exact mode's", b);
  if (h->mode == PMC_MODE_EXACT) {
    LAUNCH_TIMED(h, "k_prepare", k_prepare, dim3(h->n_box), dim3(SORT_THREADS), 0, h->stream,
                 h->d_box, h->A, -1);
    if ((e = enqueue_tables(h, kernel_type))) return e;
    LAUNCH_TIMED(h, "k_exact_sweep", k_exact_sweep, dim3(h->n_box), dim3(64), 0, h->stream,
                 h->d_box, h->A, kernel_type, del_t, h->seed, h->step);
    HIPCHECK(hipGetLastError());
  } else {
    e = parallel_step(h->work, h->A, h->d_box, h->n_box, h->cap, h->max_bin, h->stream,
                      kernel_type, del_t, h->seed, h->step, h->precision, h->band, h->validation);
    if (e) return fail(e, "parallel step enqueue failed: %s", parallel_last_error());
  }
  h->step++;
  return 0;
}

int pmc_gpu_sync(pmc_gpu *h) {
  if (!h) return fail(ERR_ARG, "null handle");
  HIPCHECK(hipSetDevice(h->device));
  HIPCHECK(hipStreamSynchronize(h->stream));
  std::vector<BoxDesc> hb(h->n_box);
  HIPCHECK(hipMemcpy(hb.data(), h->d_box, sizeof(BoxDesc) * h->n_box, hipMemcpyDeviceToHost));
  return check_box_errors(h, hb);
}

int pmc_gpu_mc_coag(pmc_gpu *h, int kernel_type, double del_t, int64_t *tot_n_samp,
                    int64_t *tot_n_coag) {
  int e = pmc_gpu_mc_coag_async(h, kernel_type, del_t);
  if (e) return e;
  HIPCHECK(hipStreamSynchronize(h->stream));
  std::vector<BoxDesc> hb(h->n_box);
  HIPCHECK(hipMemcpy(hb.data(), h->d_box, sizeof(BoxDesc) * h->n_box, hipMemcpyDeviceToHost));
  for (int b = 0; b < h->n_box; ++b) {
    if (tot_n_samp) tot_n_samp[b] = hb[b].step_samp;
    if (tot_n_coag) tot_n_coag[b] = hb[b].step_coag;
  }
  return check_box_errors(h, hb);
}

int pmc_gpu_get_counters(pmc_gpu *h, int64_t *n_samp, int64_t *n_coag, int64_t *n_warn,
                         int64_t *n_recheck, int64_t *n_flip, double *max_rel_err) {
  if (!h) return fail(ERR_ARG, "null handle");
  HIPCHECK(hipSetDevice(h->device));
  HIPCHECK(hipStreamSynchronize(h->stream));
  std::vector<BoxDesc> hb(h->n_box);
  HIPCHECK(hipMemcpy(hb.data(), h->d_box, sizeof(BoxDesc) * h->n_box, hipMemcpyDeviceToHost));
  for (int b = 0; b < h->n_box; ++b) {
    if (n_samp) n_samp[b] = hb[b].n_samp;
    if (n_coag) n_coag[b] = hb[b].n_coag;
    if (n_warn) n_warn[b] = hb[b].n_warn;
    if (n_recheck) n_recheck[b] = hb[b].n_recheck;
    if (n_flip) n_flip[b] = hb[b].n_flip;
    if (max_rel_err) max_rel_err[b] = hb[b].max_rel_err;
  }
  return 0;
}

int pmc_gpu_reset_counters(pmc_gpu *h) {
  if (!h) return fail(ERR_ARG, "null handle");
  HIPCHECK(hipSetDevice(h->device));
  HIPCHECK(hipStreamSynchronize(h->stream));
  std::vector<BoxDesc> hb(h->n_box);
  HIPCHECK(hipMemcpy(hb.data(), h->d_box, sizeof(BoxDesc) * h->n_box, hipMemcpyDeviceToHost));
  for (auto &B : hb) {
    B.n_samp = B.n_coag = B.n_warn = B.n_recheck = B.n_flip = 0;
    B.max_rel_err = 0.0;
    B.max_p_err = 0.0;
    B.n_rounds = B.n_dropped = 0;
    B.n_tail_tests = 0;
    B.n_arrivals = B.n_arr_over = 0;
    B.n_acc_samp = B.n_acc_coag = 0;
    B.n_accel_pairs = 0;
    B.part_steps = 0;
  }
  if (h->work.pstat) HIPCHECK(hipMemset(h->work.pstat, 0, 2 * sizeof(unsigned long long)));
  if (h->work.d.acc_samp) HIPCHECK(hipMemset(h->work.d.acc_samp, 0, sizeof(unsigned long long)));
  h->work.n_tail_rounds = 0;
  HIPCHECK(hipMemcpy(h->d_box, hb.data(), sizeof(BoxDesc) * h->n_box, hipMemcpyHostToDevice));
  return 0;
}

int pmc_gpu_get_error_stats(pmc_gpu *h, double *max_p_err) {
  if (!h || !max_p_err) return fail(ERR_ARG, "get_error_stats: bad arguments");
  HIPCHECK(hipSetDevice(h->device));
  HIPCHECK(hipStreamSynchronize(h->stream));
  std::vector<BoxDesc> hb(h->n_box);
  HIPCHECK(hipMemcpy(hb.data(), h->d_box, sizeof(BoxDesc) * h->n_box, hipMemcpyDeviceToHost));
  for (int b = 0; b < h->n_box; ++b) max_p_err[b] = hb[b].max_p_err;
  return 0;
}

int pmc_gpu_get_path_stats(pmc_gpu *h, int64_t *n_accel_pairs, int64_t *n_dropped,
                           int64_t *rounds, int64_t *box_tail) {
  if (!h || !n_accel_pairs || !n_dropped || !rounds)
    return fail(ERR_ARG, "get_path_stats: bad arguments");
  HIPCHECK(hipSetDevice(h->device));
  HIPCHECK(hipStreamSynchronize(h->stream));
  std::vector<BoxDesc> hb(h->n_box);
  HIPCHECK(hipMemcpy(hb.data(), h->d_box, sizeof(BoxDesc) * h->n_box, hipMemcpyDeviceToHost));
  for (int b = 0; b < h->n_box; ++b) {
    n_accel_pairs[b] = hb[b].n_accel_pairs;
    n_dropped[b] = hb[b].n_dropped;
    if (box_tail) {
      box_tail[2 * b] = hb[b].n_tail_tests;
      box_tail[2 * b + 1] = hb[b].n_rounds;
    }
  }
  unsigned long long ps[2] = {0, 0};
  if (h->work.pstat)
    HIPCHECK(hipMemcpy(ps, h->work.pstat, sizeof(ps), hipMemcpyDeviceToHost));
  rounds[0] = (int64_t)ps[0];
  rounds[1] = (int64_t)ps[1];
  rounds[2] = h->work.n_tail_rounds;
  unsigned long long acc_samp = 0;
  if (h->work.d.acc_samp)
    HIPCHECK(hipMemcpy(&acc_samp, h->work.d.acc_samp, sizeof(acc_samp), hipMemcpyDeviceToHost));
  rounds[3] = (int64_t)acc_samp;
  return 0;
}

// ------------------------------------------------------------ tapes
static int set_tape(pmc_gpu *h, int box, int mode, const double *u, int64_t n, int ref) {
  int e = check_box(h, box);
  if (e) return e;
  HIPCHECK(hipSetDevice(h->device));
  HIPCHECK(hipStreamSynchronize(h->stream));
  if ((int)h->tape_buf.size() != h->n_box) h->tape_buf.assign(h->n_box, nullptr);
  if (h->tape_buf[box]) {
    HIPCHECK(hipFree(h->tape_buf[box]));
    h->tape_buf[box] = nullptr;
  }
  Tape t{};
  t.mode = mode;
  t.ref_samplers = ref ? 1 : 0;
  if (mode != TAPE_OFF) {
    HIPCHECK(dmalloc(&h->tape_buf[box], (size_t)std::max<int64_t>(n, 1)));
    if (mode == TAPE_REPLAY && n > 0)
      HIPCHECK(hipMemcpy(h->tape_buf[box], u, sizeof(double) * n, hipMemcpyHostToDevice));
    t.u = h->tape_buf[box];
    t.n = n;
  }
  HIPCHECK(hipMemcpy((char *)(h->d_box + box) + offsetof(BoxDesc, tape), &t, sizeof(Tape),
                     hipMemcpyHostToDevice));
  return 0;
}

int pmc_gpu_record(pmc_gpu *h, int box, int64_t capacity, int ref_samplers) {
  if (capacity < 1) return fail(ERR_ARG, "record: capacity must be >= 1");
  return set_tape(h, box, TAPE_RECORD, nullptr, capacity, ref_samplers);
}

int pmc_gpu_replay(pmc_gpu *h, int box, const double *u, int64_t n, int ref_samplers) {
  if (!u || n < 1) return fail(ERR_ARG, "replay: need n >= 1 uniforms");
  for (int64_t i = 0; i < n; ++i)
    if (!(u[i] >= 0.0 && u[i] < 1.0))
      return fail(ERR_ARG, "replay: value %lld (%g) is not in [0, 1)", (long long)i, u[i]);
  return set_tape(h, box, TAPE_REPLAY, u, n, ref_samplers);
}

int pmc_gpu_trace_off(pmc_gpu *h, int box) { return set_tape(h, box, TAPE_OFF, nullptr, 0, 0); }

int pmc_gpu_get_trace(pmc_gpu *h, int box, double *u, int64_t max_n, int64_t *n_used,
                      int *ran_out) {
  int e = check_box(h, box);
  if (e) return e;
  if (!n_used) return fail(ERR_ARG, "get_trace: n_used is required");
  HIPCHECK(hipSetDevice(h->device));
  HIPCHECK(hipStreamSynchronize(h->stream));
  Tape t{};
  HIPCHECK(hipMemcpy(&t, (char *)(h->d_box + box) + offsetof(BoxDesc, tape), sizeof(Tape),
                     hipMemcpyDeviceToHost));
  *n_used = t.pos;
  if (ran_out) *ran_out = t.err;
  if (t.mode == TAPE_RECORD && u && max_n > 0) {
    const int64_t m = std::min(std::min(t.pos, t.n), max_n);
    if (m > 0) HIPCHECK(hipMemcpy(u, t.u, sizeof(double) * m, hipMemcpyDeviceToHost));
  }
  return 0;
}

int pmc_gpu_get_step_lists(pmc_gpu *h, int box, int *counts) {
  int e = check_box(h, box);
  if (e) return e;
  if (!counts) return fail(ERR_ARG, "get_step_lists: null counts");
  HIPCHECK(hipSetDevice(h->device));
  HIPCHECK(hipStreamSynchronize(h->stream));
  const PW &d = h->work.d;
  if (d.lcnt == nullptr) return fail(ERR_STATE, "get_step_lists: no parallel-mode step has run");
  BoxDesc hb;
  HIPCHECK(hipMemcpy(&hb, h->d_box + box, sizeof(BoxDesc), hipMemcpyDeviceToHost));
  const int K = hb.n_bin * hb.w.n_class;
  if (K <= 0) return 0;
  if (K > d.K_stride) return fail(ERR_STATE, "get_step_lists: lists not laid out for this grid");
  // list (bin b, class c) sits at (b - 1) + n_bin * (c - 1), as in counts
  HIPCHECK(hipMemcpy(counts, d.lcnt + (size_t)box * d.K_stride, K * sizeof(int),
                     hipMemcpyDeviceToHost));
  return 0;
}

int pmc_gpu_get_arrival_stats(pmc_gpu *h, int64_t *n_arrivals, int64_t *n_overflow) {
  if (!h || !n_arrivals || !n_overflow) return fail(ERR_ARG, "get_arrival_stats: bad arguments");
  HIPCHECK(hipSetDevice(h->device));
  HIPCHECK(hipStreamSynchronize(h->stream));
  std::vector<BoxDesc> hb(h->n_box);
  HIPCHECK(hipMemcpy(hb.data(), h->d_box, sizeof(BoxDesc) * h->n_box, hipMemcpyDeviceToHost));
  for (int b = 0; b < h->n_box; ++b) {
    n_arrivals[b] = hb[b].n_arrivals;
    n_overflow[b] = hb[b].n_arr_over;
  }
  return 0;
}

int pmc_gpu_get_accel_counts(pmc_gpu *h, int64_t *n_samp, int64_t *n_coag) {
  if (!h || !n_samp || !n_coag) return fail(ERR_ARG, "get_accel_counts: bad arguments");
  HIPCHECK(hipSetDevice(h->device));
  HIPCHECK(hipStreamSynchronize(h->stream));
  std::vector<BoxDesc> hb(h->n_box);
  HIPCHECK(hipMemcpy(hb.data(), h->d_box, sizeof(BoxDesc) * h->n_box, hipMemcpyDeviceToHost));
  for (int b = 0; b < h->n_box; ++b) {
    n_samp[b] = hb[b].n_acc_samp;
    n_coag[b] = hb[b].n_acc_coag;
  }
  return 0;
}

int pmc_gpu_size_dist(pmc_gpu *h, int box, int quantity, int n_bin, double d_min, double d_max,
                      double *hist) {
  int e = check_size_dist(h, box, quantity, n_bin, d_min, d_max, hist);
  if (e) return e;
  HIPCHECK(hipSetDevice(h->device));
  double *partials = nullptr, *d_hist = nullptr;
  hipError_t err = dmalloc(&partials, (size_t)DIST_BLOCKS * std::min(n_bin, DIST_WIN));
  if (err == hipSuccess) err = dmalloc(&d_hist, n_bin);
  if (err == hipSuccess)
    err = size_dist_enqueue(h, box, quantity, n_bin, d_min, d_max, partials, d_hist);
  if (err == hipSuccess)
    err = hipMemcpyAsync(hist, d_hist, sizeof(double) * n_bin, hipMemcpyDeviceToHost, h->stream);
  if (err == hipSuccess) err = hipStreamSynchronize(h->stream);
  if (partials) (void)hipFree(partials);
  if (d_hist) (void)hipFree(d_hist);
  if (err != hipSuccess) return fail(ERR_HIP, "size_dist: %s", hipGetErrorString(err));
  return 0;
}

int pmc_gpu_size_dist_noalloc(pmc_gpu *h, int box, int quantity, int n_bin, double d_min,
                              double d_max, double *d_hist) {
  int e = check_size_dist(h, box, quantity, n_bin, d_min, d_max, d_hist);
  if (e) return e;
  if (n_bin > DIST_WIN)
    return fail(ERR_CAPACITY,
                "size_dist_noalloc: n_bin %d exceeds the %d bins of its fixed scratch "
                "(pmc_gpu_size_dist takes any n_bin)", n_bin, DIST_WIN);
  HIPCHECK(hipSetDevice(h->device));
  double *partials = g_scratch[h->device];
  if (!partials) return fail(ERR_STATE, "size_dist_noalloc: no scratch on device %d", h->device);
  HIPCHECK(size_dist_enqueue(h, box, quantity, n_bin, d_min, d_max, partials, d_hist));
  return 0;
}

int pmc_gpu_set_n_part_ideal(pmc_gpu *h, int box, double n_part) {
  int e = check_box(h, box);
  if (e) return e;
  if (!(n_part >= 0.0)) return fail(ERR_ARG, "set_n_part_ideal: n_part must be >= 0");
  if ((int)h->n_part_ideal.size() != h->n_box) h->n_part_ideal.assign(h->n_box, -1.0);
  h->n_part_ideal[box] = n_part;
  return 0;
}

int pmc_gpu_get_weights(pmc_gpu *h, int box, int max_n, int *n_group, int *n_class,
                        int *type, double *magnitude, double *exponent) {
  int e = check_box(h, box);
  if (e) return e;
  HIPCHECK(hipSetDevice(h->device));
  HIPCHECK(hipStreamSynchronize(h->stream));
  Weights w;
  HIPCHECK(hipMemcpy(&w, (char *)(h->d_box + box) + offsetof(BoxDesc, w), sizeof(Weights),
                     hipMemcpyDeviceToHost));
  const int n = w.n_group * w.n_class;
  if (n_group) *n_group = w.n_group;
  if (n_class) *n_class = w.n_class;
  if (n > max_n) return fail(ERR_ARG, "get_weights: %d weights, room for %d", n, max_n);
  for (int i = 0; i < n; ++i) {
    if (type) type[i] = w.type[i];
    if (magnitude) magnitude[i] = w.mag[i];
    if (exponent) exponent[i] = w.expo[i];
  }
  return 0;
}

// Scratch of the population-control passes, grown to the current layout.
static int rebal_ensure(pmc_gpu *h, bool slots) {
  RebalWork &R = h->rw;
  const int nb = h->n_box;
  if (!R.gc) {
    HIPCHECK(dmalloc(&R.gc, (size_t)nb * (GC_KEYS + 1)));
    HIPCHECK(dmalloc(&R.plan, (size_t)2 * nb));
    HIPCHECK(dmalloc(&R.box_tot, (size_t)nb));
    HIPCHECK(dmalloc(&R.box_imax, (size_t)nb));
    HIPCHECK(dmalloc(&R.bval, (size_t)nb));
  }
  int64_t cap_max = 1;
  for (int64_t c : h->cap) cap_max = std::max(cap_max, c);
  const int tiles = (int)((cap_max + REBAL_TILE - 1) / REBAL_TILE);
  if (tiles > R.max_tiles) {
    for (int *p : {R.tile_a, R.tile_off, R.tile_kmax})
      if (p) (void)hipFree(p);
    R.tile_a = R.tile_off = R.tile_kmax = nullptr;
    R.max_tiles = 0;
    HIPCHECK(dmalloc(&R.tile_a, (size_t)nb * tiles));
    HIPCHECK(dmalloc(&R.tile_off, (size_t)nb * tiles));
    HIPCHECK(dmalloc(&R.tile_kmax, (size_t)nb * tiles));
    R.max_tiles = tiles;
  }
  if (slots && h->rw_slots < h->A.total_cap) {
    if (R.rm) (void)hipFree(R.rm);
    if (R.rp) (void)hipFree(R.rp);
    R.rm = nullptr;
    R.rp = nullptr;
    h->rw_slots = 0;
    HIPCHECK(dmalloc(&R.rm, (size_t)h->A.total_cap));
    HIPCHECK(dmalloc(&R.rp, (size_t)h->A.total_cap));
    h->rw_slots = h->A.total_cap;
  }
  return 0;
}

// One removal pass over the boxes with want[b] != 0.  Boxes with current
// bin lists (exact mode) remove one at a time through the lists, in the
// reference's order; the others settle every slot at once.  hb must hold
// the boxes' current sort flags.
static int run_removal(pmc_gpu *h, const std::vector<BoxDesc> &hb, const std::vector<int> &want,
                       const Removal &X, int64_t n_max) {
  int e = rebal_ensure(h, true);
  if (e) return e;
  RebalWork &R = h->rw;
  const int nb = h->n_box;
  std::vector<int> plan(2 * nb, 0);
  int any_lists = 0, any_tiles = 0;
  for (int b = 0; b < nb; ++b) {
    const bool lists = hb[b].valid_sort && hb[b].xlists_valid;
    plan[nb + b] = want[b] && lists;
    plan[b] = want[b] && !lists;
    any_lists |= plan[nb + b];
    any_tiles |= plan[b];
  }
  HIPCHECK(hipMemcpyAsync(R.plan, plan.data(), sizeof(int) * 2 * nb, hipMemcpyHostToDevice,
                          h->stream));
  const dim3 bb(64), gb((nb + 63) / 64);
  if (any_lists)
    hipLaunchKernelGGL(k_remove_lists, gb, bb, 0, h->stream, h->d_box, h->A, nb, R.plan + nb, X,
                       h->seed, h->step);
  if (any_tiles) {
    const dim3 gt((unsigned)std::max<int64_t>(1, (n_max + REBAL_TILE - 1) / REBAL_TILE), nb);
    hipLaunchKernelGGL(k_remove_mark, gt, dim3(REBAL_THREADS), 0, h->stream, h->d_box, h->A, R, X,
                       h->seed, h->step);
    hipLaunchKernelGGL(k_tile_scan, dim3(nb), dim3(REBAL_THREADS), 0, h->stream, h->d_box, R, 1);
    hipLaunchKernelGGL(k_remove_prefix, gt, dim3(REBAL_THREADS), 0, h->stream, h->d_box, h->A, R,
                       X.kind == RM_HALVE ? INFO_HALVED : INFO_DILUTION);
    hipLaunchKernelGGL(k_remove_move, gt, dim3(REBAL_THREADS), 0, h->stream, h->d_box, h->A, R);
    hipLaunchKernelGGL(k_remove_finish, gb, bb, 0, h->stream, h->d_box, R, nb, X);
  }
  HIPCHECK(hipGetLastError());
  return 0;
}

// aero_state_rebalance (src/aero_state.F90:1790-1852) for every box at once.
// The counts of each (group, class) decide the passes: doublings follow
// from the counts alone, so capacity is checked before any pass runs;
// halving passes repeat, with fresh counts, while a class holds more than
// twice its ideal number.
int pmc_gpu_rebalance(pmc_gpu *h, int allow_doubling, int allow_halving) {
  if (!h) return fail(ERR_ARG, "null handle");
  if ((int)h->n_part_ideal.size() != h->n_box)
    return fail(ERR_STATE, "rebalance: set_n_part_ideal first");
  for (int b = 0; b < h->n_box; ++b)
    if (h->n_part_ideal[b] < 0.0) return fail(ERR_STATE, "rebalance: box %d has no n_part_ideal", b);
  HIPCHECK(hipSetDevice(h->device));
  HIPCHECK(hipStreamSynchronize(h->stream));
  int e = rebal_ensure(h, false);
  if (e) return e;
  RebalWork &R = h->rw;
  const int nb = h->n_box, stride = GC_KEYS + 1;
  std::vector<BoxDesc> hb(nb);
  HIPCHECK(hipMemcpy(hb.data(), h->d_box, sizeof(BoxDesc) * nb, hipMemcpyDeviceToHost));
  e = check_box_errors(h, hb);
  if (e) return e;
  std::vector<int> gc((size_t)nb * stride), plan(2 * nb);
  auto tiles_for = [&](int64_t n) { return (int)std::max<int64_t>(1, (n + REBAL_TILE - 1) / REBAL_TILE); };
  auto n_max = [&]() {
    int64_t m = 0;
    for (int b = 0; b < nb; ++b) m = std::max<int64_t>(m, gc[(size_t)b * stride + GC_KEYS]);
    return m;
  };
  int64_t count_tiles = 1;
  for (auto &B : hb) count_tiles = std::max<int64_t>(count_tiles, tiles_for(B.n_part));
  auto count = [&]() -> hipError_t {
    const int64_t ng = (int64_t)nb * stride;
    hipLaunchKernelGGL(k_zero_int, dim3((unsigned)((ng + 255) / 256)), dim3(256), 0, h->stream,
                       R.gc, ng);
    hipLaunchKernelGGL(k_count_gc, dim3((unsigned)count_tiles, nb), dim3(REBAL_THREADS), 0,
                       h->stream, h->d_box, h->A, R.gc);
    hipError_t x = hipGetLastError();
    if (x != hipSuccess) return x;
    x = hipMemcpyAsync(gc.data(), R.gc, sizeof(int) * gc.size(), hipMemcpyDeviceToHost,
                       h->stream);
    if (x != hipSuccess) return x;
    x = hipStreamSynchronize(h->stream);
    count_tiles = tiles_for(n_max());
    return x;
  };
  HIPCHECK(count());
  int ng_max = 1, nc_max = 1;
  for (auto &B : hb) {
    ng_max = std::max(ng_max, B.w.n_group);
    nc_max = std::max(nc_max, B.w.n_class);
  }
  auto ideal = [&](int b) {
    return h->n_part_ideal[b] / (double)(hb[b].w.n_group * hb[b].w.n_class);
  };
  auto has = [&](int b, int g, int c) { return g <= hb[b].w.n_group && c <= hb[b].w.n_class; };
  auto key = [&](int b, int g, int c) { return (size_t)b * stride + (g - 1) + hb[b].w.n_group * (c - 1); };
  const dim3 bb(64), gb((nb + 63) / 64);

  if (allow_doubling) {
    // doublings per (box, g, c): while count < ideal / 2 and count > 0
    std::vector<int> nd((size_t)nb * ng_max * nc_max, 0);
    std::vector<int64_t> n_box_now(nb);
    for (int b = 0; b < nb; ++b) {
      int64_t n_tot = gc[(size_t)b * stride + GC_KEYS];
      for (int g = 1; g <= ng_max; ++g)
        for (int c = 1; c <= nc_max; ++c) {
          if (!has(b, g, c)) continue;
          const int64_t n0 = gc[key(b, g, c)];
          int64_t n = n0;
          int d = 0;
          while ((double)n < ideal(b) / 2.0 && n > 0) {
            n *= 2;
            d++;
          }
          n_tot += n - n0;
          if (n_tot > hb[b].cap)
            return fail(ERR_CAPACITY, "rebalance: box %d needs %lld particle slots to double, capacity %lld",
                        b, (long long)n_tot, (long long)hb[b].cap);
          nd[((size_t)b * ng_max + (g - 1)) * nc_max + (c - 1)] = d;
        }
      n_box_now[b] = gc[(size_t)b * stride + GC_KEYS];
    }
    bool any_pass = false;
    for (int g = 1; g <= ng_max; ++g)
      for (int c = 1; c <= nc_max; ++c) {
        int d_max = 0;
        for (int b = 0; b < nb; ++b) {
          plan[b] = nd[((size_t)b * ng_max + (g - 1)) * nc_max + (c - 1)];
          d_max = std::max(d_max, plan[b]);
        }
        if (!d_max) continue;
        any_pass = true;
        HIPCHECK(hipMemcpyAsync(R.plan, plan.data(), sizeof(int) * nb, hipMemcpyHostToDevice,
                                h->stream));
        for (int p = 0; p < d_max; ++p) {
          int64_t m = 0;
          for (int b = 0; b < nb; ++b) m = std::max(m, n_box_now[b]);
          const dim3 gt(tiles_for(m), nb);
          hipLaunchKernelGGL(k_double_count, gt, dim3(REBAL_THREADS), 0, h->stream, h->d_box,
                             h->A, R, g, c, p);
          hipLaunchKernelGGL(k_tile_scan, dim3(nb), dim3(REBAL_THREADS), 0, h->stream, h->d_box,
                             R, 0);
          hipLaunchKernelGGL(k_double_copy, gt, dim3(REBAL_THREADS), 0, h->stream, h->d_box,
                             h->A, R, g, c, p);
          hipLaunchKernelGGL(k_double_finish, gb, bb, 0, h->stream, h->d_box, R, nb, g, c, p);
          HIPCHECK(hipGetLastError());
          // the (g, c) count of a box doubles with each of its passes
          for (int b = 0; b < nb; ++b)
            if (plan[b] > p) n_box_now[b] += (int64_t)gc[key(b, g, c)] << p;
        }
      }
    if (any_pass) {
      count_tiles = tiles_for(*std::max_element(n_box_now.begin(), n_box_now.end()));
      HIPCHECK(count());
    }
  }
  if (allow_halving) {
    bool fresh = false;
    for (int g = 1; g <= ng_max; ++g)
      for (int c = 1; c <= nc_max; ++c)
        for (int k = 0;; ++k) {
          std::vector<int> want(nb, 0);
          int any = 0;
          for (int b = 0; b < nb; ++b) {
            want[b] = (has(b, g, c) && (double)gc[key(b, g, c)] > ideal(b) * 2.0) ? 1 : 0;
            any |= want[b];
          }
          if (!any) break;
          if (k >= 64)
            return fail(ERR_STATE, "rebalance: class (%d, %d) still above twice its ideal after 64 halvings", g, c);
          if (!fresh) {
            // doubling passes may have dropped the sort: flags as they are now
            HIPCHECK(hipMemcpy(hb.data(), h->d_box, sizeof(BoxDesc) * nb, hipMemcpyDeviceToHost));
            fresh = true;
          }
          Removal X{};
          X.kind = RM_HALVE;
          X.g = g;
          X.c = c;
          X.k = k;
          if ((e = run_removal(h, hb, want, X, n_max()))) return e;
          HIPCHECK(count());
        }
  }
  HIPCHECK(hipStreamSynchronize(h->stream));
  HIPCHECK(hipMemcpy(hb.data(), h->d_box, sizeof(BoxDesc) * nb, hipMemcpyDeviceToHost));
  return check_box_errors(h, hb);
}

// A removal pass of every box with a per-box parameter (dilution
// probability or mixing height) after the usual checks.
static int removal_all(pmc_gpu *h, Removal X, const double *per_box) {
  HIPCHECK(hipSetDevice(h->device));
  HIPCHECK(hipStreamSynchronize(h->stream));
  int e = rebal_ensure(h, true);
  if (e) return e;
  const int nb = h->n_box;
  std::vector<BoxDesc> hb(nb);
  HIPCHECK(hipMemcpy(hb.data(), h->d_box, sizeof(BoxDesc) * nb, hipMemcpyDeviceToHost));
  if ((e = check_box_errors(h, hb))) return e;
  if (per_box) {
    HIPCHECK(hipMemcpyAsync(h->rw.bval, per_box, sizeof(double) * nb, hipMemcpyHostToDevice,
                            h->stream));
    if (X.kind == RM_DILUTE) X.prob = h->rw.bval;
    else X.height = h->rw.bval;
  }
  std::vector<int> want(nb, 0);
  int64_t n_max = 0;
  for (int b = 0; b < nb; ++b) {
    want[b] = hb[b].n_part > 0 && (X.kind != RM_DILUTE || per_box[b] > 0.0);
    n_max = std::max<int64_t>(n_max, hb[b].n_part);
  }
  if ((e = run_removal(h, hb, want, X, n_max))) return e;
  HIPCHECK(hipStreamSynchronize(h->stream));
  HIPCHECK(hipMemcpy(hb.data(), h->d_box, sizeof(BoxDesc) * nb, hipMemcpyDeviceToHost));
  return check_box_errors(h, hb);
}

int pmc_gpu_dilute(pmc_gpu *h, const double *sample_prob) {
  if (!h || !sample_prob) return fail(ERR_ARG, "dilute: null argument");
  for (int b = 0; b < h->n_box; ++b)
    if (!(sample_prob[b] >= 0.0 && sample_prob[b] <= 1.0))
      return fail(721006962, "dilute: box %d probability %g outside [0, 1]", b, sample_prob[b]);
  Removal X{};
  X.kind = RM_DILUTE;
  return removal_all(h, X, sample_prob);
}

int pmc_gpu_particle_loss(pmc_gpu *h, int loss_function, double del_t, const double *height) {
  if (!h) return fail(ERR_ARG, "null handle");
  if (loss_function == LOSS_NONE || loss_function == LOSS_INVALID) return 0;
  if (loss_function == LOSS_CHAMBER)
    return fail(ERR_UNSUPPORTED, "particle_loss: the chamber loss function needs chamber geometry");
  if (loss_function < LOSS_INVALID || loss_function > LOSS_CHAMBER)
    return fail(201594391, "Unknown loss function id: %d", loss_function);
  if (!(del_t >= 0.0)) return fail(ERR_ARG, "particle_loss: del_t must be >= 0");
  if (loss_function == LOSS_DRYDEP) {
    if (!height) return fail(ERR_ARG, "particle_loss: dry deposition needs the mixing height");
    for (int b = 0; b < h->n_box; ++b)
      if (!(height[b] > 0.0)) return fail(ERR_ARG, "particle_loss: box %d height must be > 0", b);
  }
  Removal X{};
  X.kind = RM_LOSS;
  X.loss = loss_function;
  X.del_t = del_t;
  return removal_all(h, X, loss_function == LOSS_DRYDEP ? height : nullptr);
}

int pmc_gpu_append(pmc_gpu *h, int box, int n, const double *vol, const int64_t *id,
                   const int *group, const int *klass, const int *n_primary, const int *n_comp,
                   const int *comp_source, const double *comp_time, const double *least,
                   const double *greatest) {
  int e = check_box(h, box);
  if (e) return e;
  if (n < 0 || (n > 0 && (!vol || !group || !klass)))
    return fail(ERR_ARG, "append: bad arguments");
  if (n_comp && (!comp_source || !comp_time))
    return fail(ERR_ARG, "append: n_comp given without comp_source/comp_time");
  if (n == 0) return 0;
  HIPCHECK(hipSetDevice(h->device));
  HIPCHECK(hipStreamSynchronize(h->stream));
  BoxDesc hb;
  HIPCHECK(hipMemcpy(&hb, h->d_box + box, sizeof(BoxDesc), hipMemcpyDeviceToHost));
  if (hb.error) {
    std::vector<BoxDesc> all(h->n_box);
    HIPCHECK(hipMemcpy(all.data(), h->d_box, sizeof(BoxDesc) * h->n_box, hipMemcpyDeviceToHost));
    return check_box_errors(h, all);
  }
  for (int i = 0; i < n; ++i)
    if (group[i] < 1 || group[i] > hb.w.n_group || klass[i] < 1 || klass[i] > hb.w.n_class)
      return fail(ERR_ARG, "append: particle %d has group %d / class %d outside the weights", i,
                  group[i], klass[i]);
  const int first = hb.n_part;
  if ((int64_t)first + n > h->cap[box]) {
    std::vector<int64_t> nc = h->cap;
    nc[box] = std::max<int64_t>(2 * h->cap[box], (int64_t)first + n);
    if ((e = layout(h, nc))) return e;
    HIPCHECK(hipMemcpy(&hb, h->d_box + box, sizeof(BoxDesc), hipMemcpyDeviceToHost));
  }
  if ((e = put_particles(h, hb, first, n, vol, id, group, klass, n_primary, n_comp, comp_source,
                         comp_time, least, greatest)))
    return e;
  if ((e = rebal_ensure(h, false))) return e;
  if (!id)
    hipLaunchKernelGGL(k_assign_ids, dim3((n + 255) / 256), dim3(256), 0, h->stream, h->d_box,
                       h->A, box, first, n);
  HIPCHECK(hipMemsetAsync(h->rw.box_tot, 0, sizeof(int), h->stream));
  hipLaunchKernelGGL(k_append_check, dim3((n + 255) / 256), dim3(256), 0, h->stream, h->d_box,
                     h->A, box, first, n, h->rw.box_tot);
  hipLaunchKernelGGL(k_append_finish, dim3(1), dim3(64), 0, h->stream, h->d_box, h->A, box, first,
                     n, h->rw.box_tot, id ? 0 : n);
  HIPCHECK(hipGetLastError());
  HIPCHECK(hipStreamSynchronize(h->stream));
  HIPCHECK(hipMemcpy(&hb, h->d_box + box, sizeof(BoxDesc), hipMemcpyDeviceToHost));
  if (hb.error) {
    std::vector<BoxDesc> all(h->n_box);
    HIPCHECK(hipMemcpy(all.data(), h->d_box, sizeof(BoxDesc) * h->n_box, hipMemcpyDeviceToHost));
    return check_box_errors(h, all);
  }
  return 0;
}

int pmc_gpu_scale_weights(pmc_gpu *h, int box, double factor) {
  int e = check_box(h, box);
  if (e) return e;
  if (!(factor > 0.0)) return fail(ERR_ARG, "scale_weights: factor must be > 0");
  HIPCHECK(hipSetDevice(h->device));
  HIPCHECK(hipStreamSynchronize(h->stream));
  Weights w;
  HIPCHECK(hipMemcpy(&w, (char *)(h->d_box + box) + offsetof(BoxDesc, w), sizeof(Weights),
                     hipMemcpyDeviceToHost));
  for (int i = 0; i < w.n_group * w.n_class; ++i) w.mag[i] *= factor;
  if ((e = SET_FIELD(h, box, w, w))) return e;
  int zero = 0;
  return SET_FIELD(h, box, tables_valid, zero);
}

int pmc_gpu_get_part_steps(pmc_gpu *h, int64_t *part_steps) {
  if (!h || !part_steps) return fail(ERR_ARG, "get_part_steps: bad arguments");
  HIPCHECK(hipSetDevice(h->device));
  HIPCHECK(hipStreamSynchronize(h->stream));
  std::vector<BoxDesc> hb(h->n_box);
  HIPCHECK(hipMemcpy(hb.data(), h->d_box, sizeof(BoxDesc) * h->n_box, hipMemcpyDeviceToHost));
  for (int b = 0; b < h->n_box; ++b) part_steps[b] = hb[b].part_steps;
  return 0;
}

int pmc_gpu_set_timing(pmc_gpu *h, int on) {
  if (!h) return fail(ERR_ARG, "null handle");
  HIPCHECK(hipSetDevice(h->device));
  timing_collect(h);
  h->totals.clear();
  h->timing = on ? 1 : 0;
  return 0;
}

int pmc_gpu_get_kernel_time(pmc_gpu *h, const char *kernel, double *total_ms, int64_t *launches) {
  if (!h || !kernel) return fail(ERR_ARG, "get_kernel_time: bad arguments");
  HIPCHECK(hipSetDevice(h->device));
  timing_collect(h);
  double ms = 0.0;
  int64_t n = 0;
  for (auto &e : h->totals)
    if (e.first == kernel) {
      ms = e.second.first;
      n = e.second.second;
    }
  if (total_ms) *total_ms = ms;
  if (launches) *launches = n;
  return 0;
}

int pmc_gpu_reserve(pmc_gpu *h, const int64_t *capacity) {
  if (!h || !capacity) return fail(ERR_ARG, "reserve: bad arguments");
  HIPCHECK(hipSetDevice(h->device));
  HIPCHECK(hipStreamSynchronize(h->stream));
  std::vector<int64_t> nc = h->cap;
  bool grow = false;
  for (int b = 0; b < h->n_box; ++b) {
    int64_t want = std::max<int64_t>(capacity[b], 16);
    if (want > nc[b]) {
      nc[b] = want;
      grow = true;
    }
  }
  if (!grow) return 0;
  return layout(h, nc);
}

int pmc_gpu_get_grid(pmc_gpu *h, int box, int *n_bin, double *edges, double *centers,
                     double *widths) {
  int e = check_box(h, box);
  if (e) return e;
  HIPCHECK(hipSetDevice(h->device));
  HIPCHECK(hipStreamSynchronize(h->stream));
  BoxDesc hb;
  HIPCHECK(hipMemcpy(&hb, h->d_box + box, sizeof(BoxDesc), hipMemcpyDeviceToHost));
  if (n_bin) *n_bin = hb.n_bin;
  int nb = hb.n_bin;
  if (nb <= 0) return 0;
  if (edges) HIPCHECK(hipMemcpy(edges, h->A.edges + (size_t)box * (h->max_bin + 1), (nb + 1) * sizeof(double), hipMemcpyDeviceToHost));
  if (centers) HIPCHECK(hipMemcpy(centers, h->A.centers + (size_t)box * h->max_bin, nb * sizeof(double), hipMemcpyDeviceToHost));
  if (widths) HIPCHECK(hipMemcpy(widths, h->A.widths + (size_t)box * h->max_bin, nb * sizeof(double), hipMemcpyDeviceToHost));
  return 0;
}

int pmc_gpu_get_lists(pmc_gpu *h, int box, int *counts, int *entries) {
  int e = check_box(h, box);
  if (e) return e;
  HIPCHECK(hipSetDevice(h->device));
  HIPCHECK(hipStreamSynchronize(h->stream));
  BoxDesc hb;
  HIPCHECK(hipMemcpy(&hb, h->d_box + box, sizeof(BoxDesc), hipMemcpyDeviceToHost));
  int K = hb.n_bin * hb.w.n_class;
  if (K <= 0) return 0;
  std::vector<int> st(K), ln(K);
  HIPCHECK(hipMemcpy(st.data(), h->A.lstart + (size_t)box * h->A.max_keys, K * sizeof(int), hipMemcpyDeviceToHost));
  HIPCHECK(hipMemcpy(ln.data(), h->A.llen + (size_t)box * h->A.max_keys, K * sizeof(int), hipMemcpyDeviceToHost));
  int off = 0;
  for (int k = 0; k < K; ++k) {
    if (counts) counts[k] = ln[k];
    if (entries && ln[k] > 0) {
      HIPCHECK(hipMemcpy(entries + off, h->A.pool + st[k], ln[k] * sizeof(int), hipMemcpyDeviceToHost));
      for (int i = 0; i < ln[k]; ++i) entries[off + i] -= 1;
    }
    off += ln[k];
  }
  return 0;
}

int pmc_gpu_get_kbounds(pmc_gpu *h, int box, double *kmin, double *kmax) {
  int e = check_box(h, box);
  if (e) return e;
  HIPCHECK(hipSetDevice(h->device));
  HIPCHECK(hipStreamSynchronize(h->stream));
  BoxDesc hb;
  HIPCHECK(hipMemcpy(&hb, h->d_box + box, sizeof(BoxDesc), hipMemcpyDeviceToHost));
  size_t nn = (size_t)hb.n_bin * hb.n_bin;
  size_t o = (size_t)box * h->max_bin * h->max_bin;
  if (kmin) HIPCHECK(hipMemcpy(kmin, h->A.kmin + o, nn * sizeof(double), hipMemcpyDeviceToHost));
  if (kmax) HIPCHECK(hipMemcpy(kmax, h->A.kmax + o, nn * sizeof(double), hipMemcpyDeviceToHost));
  return 0;
}

int pmc_gpu_get_removed(pmc_gpu *h, int box, int max_n, int *n, int64_t *id, int *action,
                        int64_t *other_id) {
  int e = check_box(h, box);
  if (e) return e;
  HIPCHECK(hipSetDevice(h->device));
  HIPCHECK(hipStreamSynchronize(h->stream));
  BoxDesc hb;
  HIPCHECK(hipMemcpy(&hb, h->d_box + box, sizeof(BoxDesc), hipMemcpyDeviceToHost));
  int64_t cnt = std::min<int64_t>(hb.n_info, hb.info_cap);
  if (n) *n = (int)hb.n_info;
  if (hb.n_info > hb.info_cap) return fail(ERR_CAPACITY, "removal log overflow (%lld records)", (long long)hb.n_info);
  if (cnt > max_n) return fail(ERR_CAPACITY, "removal log holds %lld records, room for %d", (long long)cnt, max_n);
  if (cnt == 0) return 0;
  if (id) HIPCHECK(hipMemcpy(id, h->A.info_id + hb.info_off, cnt * sizeof(int64_t), hipMemcpyDeviceToHost));
  if (action) HIPCHECK(hipMemcpy(action, h->A.info_action + hb.info_off, cnt * sizeof(int), hipMemcpyDeviceToHost));
  if (other_id) HIPCHECK(hipMemcpy(other_id, h->A.info_other + hb.info_off, cnt * sizeof(int64_t), hipMemcpyDeviceToHost));
  return 0;
}

int pmc_gpu_clear_removed(pmc_gpu *h, int box) {
  int e = check_box(h, box);
  if (e) return e;
  HIPCHECK(hipSetDevice(h->device));
  HIPCHECK(hipStreamSynchronize(h->stream));
  int64_t zero = 0;
  return SET_FIELD(h, box, n_info, zero);
}

int pmc_gpu_eval_kernel(pmc_gpu *h, int kernel_type, int n, const double *vol1, const double *vol2,
                        double temp, double pressure, double beta, double *k64, float *k32) {
  if (!h || n < 0 || !vol1 || !vol2 || !k64) return fail(ERR_ARG, "eval_kernel: bad arguments");
  int e = check_kernel_type(kernel_type);
  if (e) return e;
  if (n == 0) return 0;
  HIPCHECK(hipSetDevice(h->device));
  const int ns = h->n_spec;
  double *d1, *d2, *dk;
  float *dk32;
  int *derr;
  HIPCHECK(dmalloc(&d1, (size_t)n * ns));
  HIPCHECK(dmalloc(&d2, (size_t)n * ns));
  HIPCHECK(dmalloc(&dk, n));
  HIPCHECK(dmalloc(&dk32, n));
  HIPCHECK(dmalloc(&derr, n));
  HIPCHECK(hipMemcpy(d1, vol1, (size_t)n * ns * sizeof(double), hipMemcpyHostToDevice));
  HIPCHECK(hipMemcpy(d2, vol2, (size_t)n * ns * sizeof(double), hipMemcpyHostToDevice));
  Env env = make_env(temp, pressure, beta);
  hipLaunchKernelGGL(k_eval_kernel, dim3((n + 127) / 128), dim3(128), 0, h->stream, h->A, kernel_type,
                     n, d1, d2, env, dk, dk32, derr);
  HIPCHECK(hipGetLastError());
  HIPCHECK(hipStreamSynchronize(h->stream));
  std::vector<int> herr(n);
  HIPCHECK(hipMemcpy(k64, dk, n * sizeof(double), hipMemcpyDeviceToHost));
  if (k32) HIPCHECK(hipMemcpy(k32, dk32, n * sizeof(float), hipMemcpyDeviceToHost));
  HIPCHECK(hipMemcpy(herr.data(), derr, n * sizeof(int), hipMemcpyDeviceToHost));
  (void)hipFree(d1);
  (void)hipFree(d2);
  (void)hipFree(dk);
  (void)hipFree(dk32);
  (void)hipFree(derr);
  for (int i = 0; i < n; ++i)
    if (herr[i]) return fail(herr[i], "kernel evaluation error %d at pair %d", herr[i], i);
  return 0;
}

int pmc_gpu_rng_draw(pmc_gpu *h, uint64_t seed, uint32_t box, uint32_t step, uint32_t site,
                     uint32_t ordinal, uint32_t kind, int which, double p1, double p2, int n,
                     double *out) {
  if (!h || n < 0 || !out || which < 0 || which > 3) return fail(ERR_ARG, "rng_draw: bad arguments");
  if (n == 0) return 0;
  HIPCHECK(hipSetDevice(h->device));
  double *d;
  HIPCHECK(dmalloc(&d, n));
  hipLaunchKernelGGL(k_rng_draw, dim3((n + 255) / 256), dim3(256), 0, h->stream, seed, box, step,
                     site, ordinal, kind, which, p1, p2, n, d);
  HIPCHECK(hipGetLastError());
  HIPCHECK(hipStreamSynchronize(h->stream));
  HIPCHECK(hipMemcpy(out, d, n * sizeof(double), hipMemcpyDeviceToHost));
  (void)hipFree(d);
  return 0;
}

}  // extern "C"
