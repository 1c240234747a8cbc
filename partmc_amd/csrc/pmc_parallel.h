// pmc_parallel.h -- parallel-mode step and the FP32 pair-combine kernels.
#pragma once

#include "pmc_exact.h"

namespace pmc {

// ---------------------------------------------------------- FP32 terms
// Per-particle FP64 terms, stored as FP32, from which the FP32 pair
// combine evaluates the kernel.  Brownian: (radius, diffusivity, mean
// thermal speed squared, delta squared) of kernel_brown_helper
// (src/coag_kernel_brown.F90:160-189); sedi: (radius, terminal velocity);
// additive: (radius, volume).
__device__ inline float4 particle_terms(int type, double v, double den, const Env &env) {
  float4 t = make_float4(0.f, 0.f, 0.f, 0.f);
  double r = vol2rad(v);
  t.x = (float)r;
  if (type == KERNEL_BROWN) {
    const double temp = env.temp, viscosd = env.viscosd;
    double knud = env.gasfreepath / r;
    double cunning = 1.0 + knud * (1.249 + 0.42 * exp(-0.87 / knud));
    double diffus = (C_BOLTZMANN * temp * cunning) / (6.0 * C_PI * r * viscosd);
    double speedsq = 8.0 * C_BOLTZMANN * temp / (C_PI * den * v);
    double freepath = 8.0 * diffus / (C_PI * sqrt(speedsq));
    double tmp1 = powi(2.0 * r + freepath, 3);
    const double q = 4.0 * r * r + freepath * freepath;
    double tmp2 = q * sqrt(q);  // q^1.5
    double d = (tmp1 - tmp2) / (6.0 * r * freepath) - 2.0 * r;
    t.y = (float)diffus;
    t.z = (float)speedsq;
    t.w = (float)(d * d);
  } else if (type == KERNEL_SEDI) {
    t.y = (float)fall_g(r);
  } else {
    t.y = (float)v;
  }
  return t;
}

// FP32 number-concentration weight (src/aero_weight_array.F90:232-252)
__device__ inline float nc_at_radius32(const Weights &w, int klass, float r) {
  float inv_sum = 0.f;
  for (int g = 0; g < w.n_group; ++g) {
    int idx = g + w.n_group * (klass - 1);
    float m = (float)w.mag[idx];
    float e = (float)w.expo[idx];
    float nc;
    if (w.type[idx] == WEIGHT_NONE || e == 0.f) nc = m;
    else if (e == -3.f) nc = m / (r * r * r);
    else if (e == -2.f) nc = m / (r * r);
    else if (e == -1.f) nc = m / r;
    else if (e == 1.f) nc = m * r;
    else if (e == 2.f) nc = m * (r * r);
    else if (e == 3.f) nc = m * (r * r * r);
    else nc = m * powf(r, e);
    inv_sum += 1.f / nc;
  }
  return 1.f / inv_sum;
}

// FP32 Hall efficiency lookup (same table and interpolation as effic)
__device__ inline float effic32(float r1, float r2) {
  float r_small = fminf(r1, r2) * 1e6f;
  float r_big = fmaxf(r1, r2) * 1e6f;
  float rq = r_small / r_big;
  int ir = 1, iq = 1;
  for (int k = 1; k <= 15; ++k)
    if (r_big > (float)HALL_R0[k - 1]) ir = k + 1;
  for (int kk = 1; kk <= 21; ++kk)
    if (rq > (float)HALL_RAT[kk - 1]) iq = kk + 1;
  if (iq < 2) iq = 2;
#define EC(IR, IQ) ((float)HALL_ECOLL[((IQ) - 1) * 15 + ((IR) - 1)])
  float q = (rq - (float)HALL_RAT[iq - 2]) / ((float)HALL_RAT[iq - 1] - (float)HALL_RAT[iq - 2]);
  if (ir < 16) {
    if (ir >= 2) {
      float p = (r_big - (float)HALL_R0[ir - 2]) / ((float)HALL_R0[ir - 1] - (float)HALL_R0[ir - 2]);
      return (1.f - p) * (1.f - q) * EC(ir - 1, iq - 1) + p * (1.f - q) * EC(ir, iq - 1) +
             q * (1.f - p) * EC(ir - 1, iq) + p * q * EC(ir, iq);
    }
    return (1.f - q) * EC(1, iq - 1) + q * EC(1, iq);
  }
  return fminf((1.f - q) * EC(15, iq - 1) + q * EC(15, iq), 1.f);
#undef EC
}

// FP32 pair combine of the weighted kernel
__device__ inline float weighted_kernel32(int type, const Weights &w, const Env &env, float4 a,
                                          float4 b, int c1, int c2, int cc) {
  float k;
  switch (type) {
    case KERNEL_BROWN: {
      float rs = a.x + b.x;
      float ds = a.y + b.y;
      float t1 = rs / (rs + sqrtf(a.w + b.w));
      float t2 = 4.f * ds / (rs * sqrtf(a.z + b.z));
      k = 4.f * (float)C_PI * rs * ds / (t1 + t2);
      break;
    }
    case KERNEL_SEDI: {
      float rs = a.x + b.x;
      k = effic32(a.x, b.x) * (float)C_PI * (rs * rs) * fabsf(a.y - b.y);
      break;
    }
    case KERNEL_ADDITIVE: k = (float)env.beta * (a.y + b.y); break;
    case KERNEL_CONSTANT: k = (float)C_BETA_0; break;
    default: k = 0.f;
  }
  // coag_num_conc_factor with r_ij = cbrt(r_i^3 + r_j^3)
  float r12 = cbrtf(a.x * a.x * a.x + b.x * b.x * b.x);
  float n1 = nc_at_radius32(w, c1, a.x);
  float n2 = nc_at_radius32(w, c2, b.x);
  float n12 = nc_at_radius32(w, cc, r12);
  float nmin = fminf(n1, fminf(n2, n12));
  return k * (n1 * (n2 / nmin));
}

// ------------------------------------------------------ test kernels
// kernel(type, p1, p2) of src/coag_kernel.F90:83-125 for arbitrary pairs
// given as species volumes (particle-major rows), FP64 and FP32.
__global__ void k_eval_kernel(Arrays A, int type, int n, const double *vol1,
                              const double *vol2, Env env, double *k64, float *k32,
                              int *err) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int ns = A.n_spec;
  double v1 = 0.0, v2 = 0.0, m1 = 0.0, m2 = 0.0;
  for (int s = 0; s < ns; ++s) {
    double a = vol1[(int64_t)i * ns + s], b = vol2[(int64_t)i * ns + s];
    v1 += a;
    v2 += b;
    m1 += a * A.density[s];
    m2 += b * A.density[s];
  }
  double k = 0.0;
  int e = 0;
  switch (type) {
    case KERNEL_ADDITIVE: k = env.beta * (v1 + v2); break;
    case KERNEL_CONSTANT: k = C_BETA_0; break;
    case KERNEL_ZERO: k = 0.0; break;
    case KERNEL_BROWN: k = kernel_brown_helper(v1, m1 / v1, v2, m2 / v2, env.temp, env.pressure); break;
    case KERNEL_SEDI: e = kernel_sedi_helper(v1, v2, &k); break;
    default: e = 200724934;
  }
  k64[i] = k;
  err[i] = e;
  Weights w{};
  w.n_group = 1;
  w.n_class = 1;
  w.type[0] = WEIGHT_NONE;
  w.mag[0] = 1.0;
  float4 ta = particle_terms(type, v1, m1 / v1, env);
  float4 tb = particle_terms(type, v2, m2 / v2, env);
  k32[i] = weighted_kernel32(type, w, env, ta, tb, 1, 1, 1);
}

// draws of one call-site stream family (tests of the device generators)
__global__ void k_rng_draw(uint64_t seed, uint32_t box, uint32_t step, uint32_t site,
                           uint32_t ordinal, uint32_t kind, int which, double p1, double p2,
                           int n, double *out) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (which == 0) {
    if (i != 0) return;
    Stream s;
    s.open(seed, box, step, site, ordinal, kind);
    for (int j = 0; j < n; ++j) out[j] = s.uniform();
    return;
  }
  if (i >= n) return;
  Stream s;
  s.open(seed, box, step, site, ordinal + (uint32_t)i, kind);
  if (which == 1) out[i] = (double)rand_poisson(s, p1);
  else if (which == 2) out[i] = (double)rand_binomial(s, (int)p1, p2);
  else out[i] = (double)rand_int(s, (int)p1);
}

}  // namespace pmc

#include "pmc_parallel_step.h"
