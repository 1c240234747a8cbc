// Prefix scans and reductions.
//
// Workgroup scans are rocPRIM's (cross-lane DPP within a wavefront, LDS
// across wavefronts).  A device-wide scan is two launches over contiguous
// chunks (chunk totals, then each chunk rescanned from the totals before
// it).  It may read its element count from device memory, so a kernel
// earlier on the stream can size it: the host does not read the count back
// between the launches of a claim round.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <algorithm>

#include <rocprim/block/block_scan.hpp>

namespace pmc {

struct OpSum {
  template <typename T> __device__ T operator()(const T &a, const T &b) const { return a + b; }
};
struct OpMin {
  template <typename T> __device__ T operator()(const T &a, const T &b) const { return b < a ? b : a; }
};
struct OpMax {
  template <typename T> __device__ T operator()(const T &a, const T &b) const { return a < b ? b : a; }
};

// LDS of a workgroup scan or reduction over BLOCK threads
template <int BLOCK, typename T>
using scan_storage = typename rocprim::block_scan<T, BLOCK>::storage_type;

// Inclusive scan over a workgroup of BLOCK threads; *total receives the
// whole workgroup's result on every thread.  Every thread of the workgroup
// must call it; it ends with a barrier, so st may be reused at once.
template <int BLOCK, typename T, typename Op>
__device__ inline T block_scan_incl(T v, Op op, scan_storage<BLOCK, T> &st, T *total) {
  T out, red;
  rocprim::block_scan<T, BLOCK>().inclusive_scan(v, out, red, st, op);
  __syncthreads();
  *total = red;
  return out;
}

// exclusive sum over a workgroup; *total as above
template <int BLOCK, typename T>
__device__ inline T block_sum_excl(T v, scan_storage<BLOCK, T> &st, T *total) {
  return block_scan_incl<BLOCK>(v, OpSum(), st, total) - v;
}

// reduction over a workgroup, its result on every thread
template <int BLOCK, typename T, typename Op>
__device__ inline T block_reduce(T v, Op op, scan_storage<BLOCK, T> &st) {
  T tot;
  (void)block_scan_incl<BLOCK>(v, op, st, &tot);
  return tot;
}

// ------------------------------------------------------- device-wide scans
// Up to SCAN_COLS arrays of 32-bit elements of one length n are scanned in
// the same launches (a claim round's five event flags).  Sums wrap modulo
// 2^32, so int and uint32_t arrays scan alike.  An exclusive sum also
// writes its total at out[n] and, when total_dst is set, there.
constexpr int SCAN_COLS = 5;
constexpr int SCAN_THREADS = 256;
constexpr int SCAN_UNROLL = 4;            // segments of SCAN_THREADS elements loaded together
constexpr int SCAN_MIN_SEGS = 8;          // segments a workgroup handles at least
constexpr int SCAN_MAX_BLOCKS = 1024;

struct ScanJob {
  const uint32_t *in[SCAN_COLS];
  uint32_t *out[SCAN_COLS];
  int cols;
  const int *n_dev;    // element count in device memory, or null
  int64_t n_host;      // the count when n_dev is null
  uint32_t *partial;   // [SCAN_COLS * SCAN_MAX_BLOCKS] chunk totals
  int *total_dst;      // exclusive sums: the first column's total also goes here
};

__device__ inline int64_t scan_n(const ScanJob &J) { return J.n_dev ? (int64_t)*J.n_dev : J.n_host; }

// the segments [s0, s1) of SCAN_THREADS elements that workgroup b of g scans
__device__ inline void scan_chunk(int64_t n, int g, int b, int64_t *s0, int64_t *s1) {
  const int64_t segs = (n + SCAN_THREADS - 1) / SCAN_THREADS;
  const int64_t per = (segs + g - 1) / g;
  const int64_t a = (int64_t)b * per;
  *s0 = a < segs ? a : segs;
  *s1 = *s0 + per < segs ? *s0 + per : segs;
}

template <typename Op>
__global__ void __launch_bounds__(SCAN_THREADS) k_scan_chunks(ScanJob J, Op op, uint32_t ident) {
  __shared__ scan_storage<SCAN_THREADS, uint32_t> s;
  const int64_t n = scan_n(J);
  int64_t s0, s1;
  scan_chunk(n, gridDim.x, blockIdx.x, &s0, &s1);
  for (int c = 0; c < J.cols; ++c) {
    uint32_t acc = ident;
    for (int64_t sg = s0; sg < s1; ++sg) {
      const int64_t i = sg * SCAN_THREADS + threadIdx.x;
      if (i < n) acc = op(acc, J.in[c][i]);
    }
    acc = block_reduce<SCAN_THREADS>(acc, op, s);
    if (threadIdx.x == 0) J.partial[c * SCAN_MAX_BLOCKS + blockIdx.x] = acc;
  }
}

template <typename Op, bool EXCL>
__global__ void __launch_bounds__(SCAN_THREADS) k_scan_apply(ScanJob J, Op op, uint32_t ident) {
  __shared__ scan_storage<SCAN_THREADS, uint32_t> s;
  const int64_t n = scan_n(J);
  const int g = gridDim.x, b = blockIdx.x;
  int64_t s0, s1;
  scan_chunk(n, g, b, &s0, &s1);
  for (int c = 0; c < J.cols; ++c) {
    // the chunks before this one, and all of them for the total
    uint32_t before = ident, all = ident;
    if (g > 1) {
      for (int q0 = 0; q0 < g; q0 += SCAN_THREADS) {
        const int q = q0 + threadIdx.x;
        const uint32_t p = q < g ? J.partial[c * SCAN_MAX_BLOCKS + q] : ident;
        const uint32_t rb = block_reduce<SCAN_THREADS>(q < b ? p : ident, op, s);
        const uint32_t ra = block_reduce<SCAN_THREADS>(p, op, s);
        before = op(before, rb);
        all = op(all, ra);
      }
    }
    uint32_t carry = before;
    const uint32_t *in = J.in[c];
    uint32_t *out = J.out[c];
    for (int64_t sg = s0; sg < s1; sg += SCAN_UNROLL) {
      uint32_t v[SCAN_UNROLL];
#pragma unroll
      for (int u = 0; u < SCAN_UNROLL; ++u) {
        const int64_t i = (sg + u) * SCAN_THREADS + threadIdx.x;
        v[u] = (sg + u < s1 && i < n) ? in[i] : ident;
      }
#pragma unroll
      for (int u = 0; u < SCAN_UNROLL; ++u) {
        if (sg + u >= s1) break;  // uniform over the workgroup
        const int64_t i = (sg + u) * SCAN_THREADS + threadIdx.x;
        uint32_t tot;
        const uint32_t r = block_scan_incl<SCAN_THREADS>(v[u], op, s, &tot);
        if (i < n) out[i] = EXCL ? carry + (r - v[u]) : op(carry, r);
        carry = op(carry, tot);
      }
    }
    if (g == 1) all = carry;
    if (EXCL && b == 0 && threadIdx.x == 0) {
      out[n] = all;
      if (c == 0 && J.total_dst) *J.total_dst = (int)all;
    }
  }
}

// Workgroups for a scan of at most n_bound elements.
inline int scan_blocks(int64_t n_bound) {
  const int64_t segs = (n_bound + SCAN_THREADS - 1) / SCAN_THREADS;
  const int64_t g = (segs + SCAN_MIN_SEGS - 1) / SCAN_MIN_SEGS;
  return (int)std::max<int64_t>(1, std::min<int64_t>(g, SCAN_MAX_BLOCKS));
}

// exclusive sums (out[0..n]) of J's columns; n at most n_bound
inline void scan_excl_sum(const ScanJob &J, int64_t n_bound, hipStream_t st) {
  const int g = scan_blocks(n_bound);
  if (g > 1) hipLaunchKernelGGL(k_scan_chunks<OpSum>, dim3(g), dim3(SCAN_THREADS), 0, st, J, OpSum(), 0u);
  hipLaunchKernelGGL((k_scan_apply<OpSum, true>), dim3(g), dim3(SCAN_THREADS), 0, st, J, OpSum(), 0u);
}

// inclusive running maximum (out[0..n)) of J's columns
inline void scan_incl_max(const ScanJob &J, int64_t n_bound, hipStream_t st) {
  const int g = scan_blocks(n_bound);
  if (g > 1) hipLaunchKernelGGL(k_scan_chunks<OpMax>, dim3(g), dim3(SCAN_THREADS), 0, st, J, OpMax(), 0u);
  hipLaunchKernelGGL((k_scan_apply<OpMax, false>), dim3(g), dim3(SCAN_THREADS), 0, st, J, OpMax(), 0u);
}

// a job over one array
inline ScanJob scan_job(const void *in, void *out, uint32_t *partial, int64_t n_host,
                        const int *n_dev = nullptr, int *total_dst = nullptr) {
  ScanJob J{};
  J.in[0] = (const uint32_t *)in;
  J.out[0] = (uint32_t *)out;
  J.cols = 1;
  J.n_dev = n_dev;
  J.n_host = n_host;
  J.partial = partial;
  J.total_dst = total_dst;
  return J;
}

}  // namespace pmc
