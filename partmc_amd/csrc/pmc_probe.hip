// pmc_probe.hip -- throughput probes of the gfx950 units the coagulation
// kernels lean on, for the roofline of kernel evaluation: FP32 FMA (VALU),
// FP64 FMA, and the FP32 transcendental unit (v_exp_f32, v_log_f32,
// v_rcp_f32, v_sqrt_f32).  Each lane runs 8 independent dependency chains so
// the pipelines stay full; the grid holds many waves per CU.  Not part of
// the coagulation library: bench.py --probe runs it.
#include <hip/hip_runtime.h>

#include <cstdint>

namespace {

constexpr int CHAINS = 8;
constexpr int THREADS = 256;

__global__ void __launch_bounds__(THREADS) k_fma32(float *out, int iters, float a, float b) {
  float x[CHAINS];
  for (int c = 0; c < CHAINS; ++c) x[c] = (float)(threadIdx.x + c) * 1e-3f;
  for (int i = 0; i < iters; ++i)
#pragma unroll
    for (int c = 0; c < CHAINS; ++c) x[c] = __builtin_fmaf(x[c], a, b);
  float s = 0.f;
  for (int c = 0; c < CHAINS; ++c) s += x[c];
  if (s == 12345.f) out[blockIdx.x] = s;  // keeps the chains live
}

__global__ void __launch_bounds__(THREADS) k_fma64(double *out, int iters, double a, double b) {
  double x[CHAINS];
  for (int c = 0; c < CHAINS; ++c) x[c] = (double)(threadIdx.x + c) * 1e-3;
  for (int i = 0; i < iters; ++i)
#pragma unroll
    for (int c = 0; c < CHAINS; ++c) x[c] = __builtin_fma(x[c], a, b);
  double s = 0.0;
  for (int c = 0; c < CHAINS; ++c) s += x[c];
  if (s == 12345.0) out[blockIdx.x] = s;
}

// one transcendental per chain step: exp2, log2, rcp, sqrt in rotation,
// each kept in range by the next (log2 of exp2 of a value in [0.5, 2]...)
__global__ void __launch_bounds__(THREADS) k_trans32(float *out, int iters) {
  float x[CHAINS];
  for (int c = 0; c < CHAINS; ++c) x[c] = 1.0f + (float)((threadIdx.x + c) & 7) * 0.1f;
  for (int i = 0; i < iters; ++i) {
#pragma unroll
    for (int c = 0; c < CHAINS; ++c) x[c] = __builtin_amdgcn_exp2f(x[c]);
#pragma unroll
    for (int c = 0; c < CHAINS; ++c) x[c] = __builtin_amdgcn_logf(x[c]);
#pragma unroll
    for (int c = 0; c < CHAINS; ++c) x[c] = __builtin_amdgcn_rcpf(x[c]);
#pragma unroll
    for (int c = 0; c < CHAINS; ++c) x[c] = __builtin_amdgcn_sqrtf(x[c]) + 0.5f;
  }
  float s = 0.f;
  for (int c = 0; c < CHAINS; ++c) s += x[c];
  if (s == 12345.f) out[blockIdx.x] = s;
}

template <class F>
float time_ms(F launch) {
  hipEvent_t a, b;
  (void)hipEventCreate(&a);
  (void)hipEventCreate(&b);
  launch();  // warm-up
  (void)hipEventRecord(a, nullptr);
  launch();
  (void)hipEventRecord(b, nullptr);
  (void)hipEventSynchronize(b);
  float ms = 0.f;
  (void)hipEventElapsedTime(&ms, a, b);
  (void)hipEventDestroy(a);
  (void)hipEventDestroy(b);
  return ms;
}

}  // namespace

extern "C" {

// ops per second of each probe: rates[0] FP32 FMA lanes/s (x2 for flop/s),
// rates[1] FP64 FMA lanes/s, rates[2] FP32 transcendentals/s (exp2, log2,
// rcp and sqrt counted one each; the sqrt's add is not counted).  Returns 0
// or a HIP error code.
int pmc_probe_rates(int device, double *rates) {
  if (hipSetDevice(device) != hipSuccess) return 2;
  int n_cu = 0;
  if (hipDeviceGetAttribute(&n_cu, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess)
    return 2;
  const int blocks = n_cu * 32;  // 8 waves per SIMD
  float *o32 = nullptr;
  double *o64 = nullptr;
  if (hipMalloc(&o32, blocks * sizeof(float)) != hipSuccess) return 2;
  if (hipMalloc(&o64, blocks * sizeof(double)) != hipSuccess) return 2;
  const double lanes = (double)blocks * THREADS * CHAINS;
  const int it32 = 4096, it64 = 2048, itt = 512;
  float ms = time_ms([&] {
    hipLaunchKernelGGL(k_fma32, dim3(blocks), dim3(THREADS), 0, nullptr, o32, it32, 0.999f, 1e-3f);
  });
  rates[0] = lanes * it32 / (ms * 1e-3);
  ms = time_ms([&] {
    hipLaunchKernelGGL(k_fma64, dim3(blocks), dim3(THREADS), 0, nullptr, o64, it64, 0.999, 1e-3);
  });
  rates[1] = lanes * it64 / (ms * 1e-3);
  ms = time_ms([&] {
    hipLaunchKernelGGL(k_trans32, dim3(blocks), dim3(THREADS), 0, nullptr, o32, itt);
  });
  rates[2] = lanes * itt * 4.0 / (ms * 1e-3);
  const hipError_t e = hipGetLastError();
  (void)hipFree(o32);
  (void)hipFree(o64);
  return e == hipSuccess ? 0 : 2;
}

}  // extern "C"
