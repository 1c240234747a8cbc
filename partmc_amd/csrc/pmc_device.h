// pmc_device.h -- device-side numerics for the gfx950 coagulation step.
//
// FP64 routines keep the reference's operation order (every product and
// sum separately rounded; the library is compiled with -ffp-contract=off)
// so that FP64 results differ from a glibc-based evaluation only through
// the last-bit behaviour of pow/exp/log/sqrt.  FP32 pair-combine routines
// are separate and are always rechecked in FP64 near the accept threshold.
//
// Reference citations are to PartMC's source tree.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace pmc {

constexpr int MAX_SPEC = 32;
constexpr int MAX_COMP = 20;     // MAX_AERO_COMPONENT_SIZE, src/aero_particle.F90:22
constexpr int MAX_GROUP = 4;
constexpr int MAX_CLASS = 4;

// kernel codes, src/coag_kernel.F90:29-45
constexpr int KERNEL_SEDI = 1;
constexpr int KERNEL_ADDITIVE = 2;
constexpr int KERNEL_CONSTANT = 3;
constexpr int KERNEL_BROWN = 4;
constexpr int KERNEL_ZERO = 5;

// weight codes, src/aero_weight.F90:25-29
constexpr int WEIGHT_NONE = 1;
constexpr int WEIGHT_POWER = 2;
constexpr int WEIGHT_MFA = 3;

// RNG call-site kinds
constexpr uint32_t RNG_POISSON = 1;
constexpr uint32_t RNG_TEST = 2;
constexpr uint32_t RNG_MERGE = 3;
constexpr uint32_t RNG_ACC_POISSON = 4;
constexpr uint32_t RNG_ACC_SAMPLE = 5;
constexpr uint32_t RNG_ACC_FINISH = 6;
constexpr uint32_t RNG_ACC_TARGET = 7;
constexpr uint32_t RNG_HALVE = 8;
constexpr uint32_t RNG_DILUTE = 9;
constexpr uint32_t RNG_LOSS = 10;

constexpr double C_PI = 3.14159265358979323846;
constexpr double C_BOLTZMANN = 1.3806505e-23;
constexpr double C_AVOGADRO = 6.02214179e23;
constexpr double C_UNIV_GAS = 8.314472;
constexpr double C_AIR_MW = 2.89644e-2;
constexpr double C_BETA_0 = 0.25 / (60.0 * 2e8);

// ---------------------------------------------------------------- Philox
__host__ __device__ inline void philox4x32_10(uint32_t c0, uint32_t c1,
                                              uint32_t c2, uint32_t c3,
                                              uint32_t k0, uint32_t k1,
                                              uint32_t out[4]) {
#pragma unroll
  for (int r = 0; r < 10; ++r) {
    uint64_t p0 = (uint64_t)0xD2511F53u * c0;
    uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
    uint32_t n0 = (uint32_t)(p1 >> 32) ^ c1 ^ k0;
    uint32_t n2 = (uint32_t)(p0 >> 32) ^ c3 ^ k1;
    c0 = n0;
    c1 = (uint32_t)p1;
    c2 = n2;
    c3 = (uint32_t)p0;
    k0 += 0x9E3779B9u;
    k1 += 0xBB67AE85u;
  }
  out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

// 52 random bits as (k + 1/2) 2^-52, strictly inside (0, 1)
__host__ __device__ inline double u52(uint32_t a, uint32_t b) {
  uint64_t k = ((uint64_t)a << 20) | (uint64_t)(b >> 12);
  return (double)k * 0x1p-52 + 0x1p-53;
}

// A box's uniforms in the order the sequential chain draws them, the
// values pmc_random returns (src/rand.F90:138-166) in call order: replayed
// from a recorded run in place of the call-site streams, or recorded from
// this run.  ref_samplers: Poisson counts above mean 10 and binomials with
// np and n(1-p) >= 10 use the reference's rounded normal (src/rand.F90:
// 298-302, 359-361), as a run without GSL draws them.
constexpr int TAPE_OFF = 0, TAPE_RECORD = 1, TAPE_REPLAY = 2;
struct Tape {
  double *u;         // the uniforms
  int64_t n;         // replay: values held; record: capacity
  int64_t pos;       // next value read or written
  int mode;
  int ref_samplers;
  int err;           // the tape ran out
};

// A call-site stream: key = (seed_lo, seed_hi + box), counter =
// (ordinal, site, step, kind << 24 | block).  Two uniforms per block.  With
// a tape attached its uniforms come from, or go to, the tape.
struct Stream {
  uint32_t k0, k1, c0, c1, c2, c3;
  uint32_t block;
  uint32_t buf2, buf3;
  int have;
  Tape *tape;

  __host__ __device__ inline void open(uint64_t seed, uint32_t box,
                                       uint32_t step, uint32_t site,
                                       uint32_t ordinal, uint32_t kind) {
    k0 = (uint32_t)seed;
    k1 = (uint32_t)(seed >> 32) + box;
    c0 = ordinal;
    c1 = site;
    c2 = step;
    c3 = kind << 24;
    block = 0;
    have = 0;
    tape = nullptr;
  }
  __host__ __device__ inline double uniform() {
    if (tape) {
      if (tape->mode == TAPE_REPLAY) {
        if (tape->pos < tape->n) return tape->u[tape->pos++];
        tape->err = 1;  // ran out: the step goes on with its own stream and fails after
        return philox_uniform();
      }
      const double v = philox_uniform();
      if (tape->pos < tape->n) tape->u[tape->pos] = v;
      else tape->err = 1;
      tape->pos++;
      return v;
    }
    return philox_uniform();
  }
  __host__ __device__ inline double philox_uniform() {
    if (have == 0) {
      uint32_t o[4];
      philox4x32_10(c0, c1, c2, c3 | (block & 0xFFFFFFu), k0, k1, o);
      block++;
      buf2 = o[2];
      buf3 = o[3];
      have = 1;
      return u52(o[0], o[1]);
    }
    have = 0;
    return u52(buf2, buf3);
  }
};

// pmc_rand_int (non-GSL), src/rand.F90:202
__device__ inline int rand_int(Stream &s, int n) {
  double u = s.uniform();
  int k = (int)(u * (double)n);
  return (k % n) + 1;
}

// rand_normal without GSL (Box-Muller, the first deviate), src/rand.F90:
// 452-470 with no probability threshold
__device__ inline double rand_normal_ref(Stream &s, double mean, double stddev) {
  const double u1 = s.uniform();
  const double u2 = s.uniform();
  const double r = sqrt(-2.0 * log(u1));
  const double theta = 2.0 * C_PI * u2;
  return stddev * (r * cos(theta)) + mean;
}

// prob_round, src/rand.F90:214-227
__device__ inline int prob_round(Stream &s, double val) {
  if (s.uniform() < ceil(val) - val) return (int)floor(val);
  return (int)ceil(val);
}

// rand_poisson: Knuth for mean <= 10 (src/rand.F90:287-296), Hormann's
// PTRS transformed rejection above (an exact sampler in place of the
// reference's rounded normal, :298-302).
__device__ inline int rand_poisson(Stream &s, double mean) {
  if (mean > 10.0 && s.tape && s.tape->ref_samplers) {
    const int k = (int)round(rand_normal_ref(s, mean, sqrt(mean)));
    return k > 0 ? k : 0;
  }
  if (mean <= 10.0) {
    double L = exp(-mean);
    int k = 0;
    double p = 1.0;
    for (;;) {
      k++;
      p = p * s.uniform();
      if (p < L) break;
    }
    return k - 1;
  }
  double slam = sqrt(mean);
  double loglam = log(mean);
  double b = 0.931 + 2.53 * slam;
  double a = -0.059 + 0.02483 * b;
  double invalpha = 1.1239 + 1.1328 / (b - 3.4);
  double vr = 0.9277 - 3.6224 / (b - 2.0);
  for (;;) {
    double U = s.uniform() - 0.5;
    double V = s.uniform();
    double us = 0.5 - fabs(U);
    double kf = floor((2.0 * a / us + b) * U + mean + 0.43);
    if (us >= 0.07 && V <= vr) return (int)kf;
    if (kf < 0.0 || (us < 0.013 && V > us)) continue;
    if (log(V) + log(invalpha) - log(a / (us * us) + b) <=
        -mean + kf * loglam - lgamma(kf + 1.0))
      return (int)kf;
  }
}

// rand_binomial, src/rand.F90:319-404, with Hormann's BTRS in place of the
// normal approximation for np >= 10 and n(1-p) >= 10.
__device__ inline int rand_binomial(Stream &s, int n, double p) {
  double np = (double)n * p;
  double nomp = (double)n * (1.0 - p);
  if (np >= 10.0 && nomp >= 10.0 && s.tape && s.tape->ref_samplers) {
    const int k = (int)round(rand_normal_ref(s, np, sqrt(np * (1.0 - p))));
    return k < 0 ? 0 : (k > n ? n : k);
  }
  if (np >= 10.0 && nomp >= 10.0) {
    bool flip = p > 0.5;
    double pp = flip ? 1.0 - p : p;
    double q = 1.0 - pp;
    double spq = sqrt((double)n * pp * q);
    double b = 1.15 + 2.53 * spq;
    double a = -0.0873 + 0.0248 * b + 0.01 * pp;
    double c = (double)n * pp + 0.5;
    double vr = 0.92 - 4.2 / b;
    double alpha = (2.83 + 5.1 / b) * spq;
    double lpq = log(pp / q);
    double m = floor((double)(n + 1) * pp);
    double h = lgamma(m + 1.0) + lgamma((double)n - m + 1.0);
    for (;;) {
      double u = s.uniform() - 0.5;
      double v = s.uniform();
      double us = 0.5 - fabs(u);
      double kf = floor((2.0 * a / us + b) * u + c);
      if (kf < 0.0 || kf > (double)n) continue;
      if (us >= 0.07 && v <= vr) {
        int k = (int)kf;
        return flip ? n - k : k;
      }
      v = log(v * alpha / (a / (us * us) + b));
      if (v <= h - lgamma(kf + 1.0) - lgamma((double)n - kf + 1.0) +
                   (kf - m) * lpq) {
        int k = (int)kf;
        return flip ? n - k : k;
      }
    }
  } else if (p < 1e-15) {
    return 0;
  } else if (np < 1e-200) {
    return 0;
  } else if (nomp < 1e-200) {
    return n;
  }
  double q = (p <= 0.5) ? p : 1.0 - p;
  int k = 0, sum = 0;
  for (;;) {
    double G_real = log(s.uniform()) / log(1.0 - q);
    if (G_real > (double)(n - sum)) break;
    int G = (int)ceil(G_real);
    sum = sum + G;
    if (sum > n) break;
    k = k + 1;
  }
  return (p <= 0.5) ? k : n - k;
}

// sample_without_replacement, src/rand.F90:515-542 (1-based output)
__device__ inline void sample_without_replacement(Stream &s, int sample_size,
                                                  int population_size,
                                                  int *sample) {
  int t = 0, m = 0;
  while (m < sample_size) {
    double r = s.uniform();
    if ((double)(population_size - t) * r >= (double)(sample_size - m)) {
      t = t + 1;
    } else {
      t = t + 1;
      m = m + 1;
      sample[m - 1] = t;
    }
  }
}

// sample_cts_pdf, src/rand.F90:550-576 (1-based)
__device__ inline int sample_cts_pdf(Stream &s, int n, const double *pdf) {
  double pdf_max = pdf[0];
  for (int i = 1; i < n; ++i)
    if (pdf[i] > pdf_max) pdf_max = pdf[i];
  for (;;) {
    int k = rand_int(s, n);
    if (s.uniform() < pdf[k - 1] / pdf_max) return k;
  }
}

// --------------------------------------------------------------- util
__host__ __device__ inline double powi(double x, int n) {
  double pw = 1.0;
  unsigned u;
  if (n == 0) return 1.0;
  if (n < 0) {
    u = (unsigned)(-n);
    x = 1.0 / x;
  } else {
    u = (unsigned)n;
  }
  for (;;) {
    if (u & 1u) pw *= x;
    u >>= 1;
    if (u) x *= x; else break;
  }
  return pw;
}

// sphere_vol2rad / sphere_rad2vol, src/util.F90:240-271
__device__ __forceinline__ double vol2rad(double v) {
  return pow(3.0 * v / 4.0 / C_PI, 1.0 / 3.0);
}
__device__ inline double rad2vol(double r) {
  return 4.0 * C_PI * powi(r, 3) / 3.0;
}

// fast_pow, src/util.F90:1978-2004
__device__ inline double fast_pow(double base, double e) {
  if (e == 0.0) return 1.0;
  if (e == 1.0) return base;
  if (e == 2.0) return base * base;
  if (e == 3.0) return base * base * base;
  if (e == -1.0) return 1.0 / base;
  if (e == -2.0) return 1.0 / (base * base);
  if (e == -3.0) return 1.0 / (base * base * base);
  return pow(base, e);
}

// interp_linear_disc, src/util.F90:663-682
__device__ inline double interp_linear_disc(double x1, double xn, int n, int i) {
  if (x1 == xn) return x1;
  double alpha = (double)(i - 1) / (double)(n - 1);
  return (1.0 - alpha) * x1 + alpha * xn;
}

// linspace_find, src/util.F90:532-551
__device__ inline int linspace_find(double min_x, double max_x, int n, double x) {
  if (x == max_x) return n - 1;
  double f = floor((x - min_x) / (max_x - min_x) * (double)(n - 1));
  if (f > (double)n) f = (double)n;
  if (f < -1.0) f = -1.0;
  int i = (int)f + 1;
  if (i > n) i = n;
  if (i < 0) i = 0;
  return i;
}

// bin_grid_find (log grid), src/bin_grid.F90:143-170 via logspace_find
// src/util.F90:566-581.  log(edge_1) and log(edge_N+1) are passed in so a
// caller can hoist them.
__device__ inline int bin_find_log(int n_bin, double log_lo, double log_hi,
                                   double r) {
  if (n_bin <= 0) return -1;
  return linspace_find(log_lo, log_hi, n_bin + 1, log(r));
}

// Student t 95% coefficient, src/stats.F90:510-536
__device__ inline double student_t_95_coeff(int n_sample) {
  const double limit = 1.959963984540054;
  const double values[15] = {
      12.7062047364, 4.30265272991, 3.18244630528, 2.7764451052,
      2.57058183661, 2.44691184879, 2.36462425101, 2.30600413503,
      2.26215716274, 2.22813885196, 2.20098516008, 2.17881282966,
      2.16036865646, 2.14478668792, 2.13144954556};
  int n_dof = n_sample - 1;
  if (n_dof <= 15) return values[n_dof - 1];
  if (n_dof <= 20) return (2.11990529922 - limit) * 16.0 / (double)n_dof + limit;
  return (2.07961384473 - limit) * 21.0 / (double)n_dof + limit;
}

// ------------------------------------------------------------ weights
struct Weights {
  int n_group, n_class;
  int type[MAX_GROUP * MAX_CLASS];      // index g + n_group * c (0-based)
  double mag[MAX_GROUP * MAX_CLASS];
  double expo[MAX_GROUP * MAX_CLASS];
};

// aero_weight_num_conc_at_radius, src/aero_weight.F90:127-146
__device__ inline double weight_nc(int type, double mag, double e, double r) {
  if (type == WEIGHT_NONE) return mag;
  return mag * fast_pow(r, e);
}

// aero_weight_array_num_conc_at_radius, src/aero_weight_array.F90:232-252
// (1-based class)
__device__ __forceinline__ double nc_at_radius(const Weights &w, int klass, double r) {
  double inv_sum = 0.0;
  for (int g = 0; g < w.n_group; ++g) {
    int idx = g + w.n_group * (klass - 1);
    inv_sum = inv_sum + 1.0 / weight_nc(w.type[idx], w.mag[idx], w.expo[idx], r);
  }
  return 1.0 / inv_sum;
}

// coag_num_conc_factor, src/coag_kernel.F90:346-377
__device__ __forceinline__ double coag_num_conc_factor(const Weights &w, double i_r,
                                              double j_r, int ic, int jc,
                                              int ijc) {
  double ij_r = vol2rad(rad2vol(i_r) + rad2vol(j_r));
  double i_nc = nc_at_radius(w, ic, i_r);
  double j_nc = nc_at_radius(w, jc, j_r);
  double ij_nc = nc_at_radius(w, ijc, ij_r);
  double nc_min = fmin(i_nc, fmin(j_nc, ij_nc));
  return i_nc * j_nc / nc_min;
}

// coag_num_conc_factor with the second particle's volume and weight given
__device__ inline double coag_num_conc_factor_j(const Weights &w, double i_r, double j_v,
                                                double j_nc, int ic, int ijc) {
  double ij_r = vol2rad(rad2vol(i_r) + j_v);
  double i_nc = nc_at_radius(w, ic, i_r);
  double ij_nc = nc_at_radius(w, ijc, ij_r);
  double nc_min = fmin(i_nc, fmin(j_nc, ij_nc));
  return i_nc * j_nc / nc_min;
}

// coag_dest_class, src/coag_kernel.F90:382-416 (both radii at centers(i_bin))
__device__ inline int coag_dest_class(const Weights &w, const double *centers,
                                      int i_bin, int ic, int jc) {
  double i_r = centers[i_bin - 1];
  double j_r = centers[i_bin - 1];
  double ij_r = vol2rad(rad2vol(i_r) + rad2vol(j_r));
  double a = nc_at_radius(w, ic, ij_r);
  double b = nc_at_radius(w, jc, ij_r);
  return (a < b) ? ic : jc;
}

// max_coag_num_conc_factor, src/coag_kernel.F90:422-465
__device__ inline double max_coag_num_conc_factor(const Weights &w,
                                                  const double *edges, int ib,
                                                  int jb, int ic, int jc,
                                                  int ijc) {
  double i_lo = edges[ib - 1], i_hi = edges[ib];
  double j_lo = edges[jb - 1], j_hi = edges[jb];
  double f_max = 0.0;
  for (int is = 1; is <= 5; ++is)
    for (int js = 1; js <= 5; ++js) {
      double i_r = interp_linear_disc(i_lo, i_hi, 5, is);
      double j_r = interp_linear_disc(j_lo, j_hi, 5, js);
      double f = coag_num_conc_factor(w, i_r, j_r, ic, jc, ijc);
      f_max = fmax(f_max, f);
    }
  return f_max;
}

// ------------------------------------------------------------- kernels
struct Env {
  double temp, pressure, beta;
  // air viscosity and mean free path at (temp, pressure), for the FP32
  // Brownian terms (the FP64 kernel derives its own, as the reference does)
  double viscosd, gasfreepath;
};

__host__ __device__ inline Env make_env(double temp, double pressure, double beta) {
  Env e;
  e.temp = temp;
  e.pressure = pressure;
  e.beta = beta;
  const double rhoair = (pressure * C_AIR_MW) / (C_UNIV_GAS * temp);
  e.viscosd = 1.8325e-05 * (416.16 / (temp + 120.0)) * pow(temp / 296.16, 1.5);
  const double viscosk = e.viscosd / rhoair;
  const double gasspeed = sqrt((8.0 * C_BOLTZMANN * temp * C_AVOGADRO) / (C_PI * C_AIR_MW));
  e.gasfreepath = 2.0 * viscosk / gasspeed;
  return e;
}

// kernel_brown_helper, src/coag_kernel_brown.F90:111-197 (spherical), in
// three parts: the air's properties, the terms of one particle, and their
// combination -- the same operations in the same order, so a caller that
// keeps one particle's terms across many partners gets the same values
struct BrownAir {
  double viscosd, gasfreepath;
};
struct BrownTerms {
  double rad, diffus, speedsq, deltasq;
};

__device__ inline BrownAir brown_air(double temp, double pressure) {
  double rhoair = (pressure * C_AIR_MW) / (C_UNIV_GAS * temp);
  double viscosd = 1.8325e-05 * (416.16 / (temp + 120.0)) * pow(temp / 296.16, 1.5);
  double viscosk = viscosd / rhoair;
  double gasspeed = sqrt((8.0 * C_BOLTZMANN * temp * C_AVOGADRO) / (C_PI * C_AIR_MW));
  return BrownAir{viscosd, 2.0 * viscosk / gasspeed};
}

__device__ inline BrownTerms brown_terms(const BrownAir &air, double vol, double den,
                                         double temp) {
  double rad = vol2rad(vol);
  double Rme = vol2rad(vol);
  double knud = air.gasfreepath / Rme;
  double cunning = 1.0 + knud * (1.249 + 0.42 * exp(-0.87 / knud));
  double diffus = (C_BOLTZMANN * temp * cunning) / (6.0 * C_PI * Rme * air.viscosd);
  double speedsq = 8.0 * C_BOLTZMANN * temp / (C_PI * den * vol);
  double freepath = 8.0 * diffus / (C_PI * sqrt(speedsq));
  double tmp1 = powi(2.0 * Rme + freepath, 3);
  double tmp2 = pow(4.0 * Rme * Rme + freepath * freepath, 1.5);
  double d = (tmp1 - tmp2) / (6.0 * Rme * freepath) - 2.0 * Rme;
  return BrownTerms{rad, diffus, speedsq, d * d};
}

__device__ inline double brown_combine(const BrownTerms &i, const BrownTerms &j) {
  double rad_sum = i.rad + j.rad;
  double diffus_sum = i.diffus + j.diffus;
  double tmp1 = rad_sum / (rad_sum + sqrt(i.deltasq + j.deltasq));
  double tmp2 = 4.0 * diffus_sum / (rad_sum * sqrt(i.speedsq + j.speedsq));
  return 4.0 * C_PI * rad_sum * diffus_sum / (tmp1 + tmp2);
}

__device__ inline double kernel_brown_helper(double vol_i, double den_i,
                                             double vol_j, double den_j,
                                             double temp, double pressure) {
  const BrownAir air = brown_air(temp, pressure);
  const BrownTerms ti = brown_terms(air, vol_i, den_i, temp);
  const BrownTerms tj = brown_terms(air, vol_j, den_j, temp);
  return brown_combine(ti, tj);
}

// fall_g, src/coag_kernel_sedi.F90:110-172
__device__ inline double fall_g(double r) {
  const double b[7] = {-0.318657e1, 0.992696, -0.153193e-2, -0.987059e-3,
                       -0.578878e-3, 0.855176e-4, -0.327815e-5};
  const double c[6] = {-0.500015e1, 0.523778e1, -0.204914e1, 0.475294,
                       -0.542819e-1, 0.238449e-2};
  double eta = 1.818e-4, xlamb = 6.62e-6, rhow = 1.0, rhoa = 1.225e-3;
  double grav = 980.665;
  double cunh = 1.257 * xlamb;
  double t0 = 273.15;
  double sigma = 76.1 - 0.155 * (293.15 - t0);
  double stok = 2.0 * grav * (rhow - rhoa) / (9.0 * eta);
  double stb = 32.0 * rhoa * (rhow - rhoa) * grav / (3.0 * eta * eta);
  double phy = sigma * sigma * sigma * rhoa * rhoa /
               (powi(eta, 4) * grav * (rhow - rhoa));
  double py = pow(phy, 1.0 / 6.0);
  double rr = r * 1e2;
  double w = 0.0;
  if (rr <= 1e-3) {
    w = stok * (rr * rr + cunh * rr);
  } else if (rr > 1e-3 && rr <= 5.35e-2) {
    double x = log(stb * rr * rr * rr);
    double y = 0.0;
    for (int i = 1; i <= 7; ++i) y = y + b[i - 1] * powi(x, i - 1);
    double xrey = (1.0 + cunh / rr) * exp(y);
    w = xrey * eta / (2.0 * rhoa * rr);
  } else if (rr > 5.35e-2) {
    double bond = grav * (rhow - rhoa) * (rr * rr) / sigma;
    if (rr > 0.35) bond = grav * (rhow - rhoa) * (0.35 * 0.35) / sigma;
    double x = log(16.0 * bond * py / 3.0);
    double y = 0.0;
    for (int i = 1; i <= 6; ++i) y = y + c[i - 1] * powi(x, i - 1);
    double xrey = py * exp(y);
    w = xrey * eta / (2.0 * rhoa * rr);
    if (rr > 0.35) w = xrey * eta / (2.0 * rhoa * 0.35);
  }
  return w / 100.0;
}

// Hall (1980) collision-efficiency table of the reference's effic
// (src/coag_kernel_sedi.F90:190-245), column-major as in its DATA
// statement: element (ir, iq) at (iq-1)*15 + (ir-1).
extern __constant__ double HALL_ECOLL[315];
extern __constant__ double HALL_R0[15];
extern __constant__ double HALL_RAT[21];

// effic, src/coag_kernel_sedi.F90:180-285; returns 99 if ec < 1e-20
__device__ inline int effic(double r1, double r2, double *ec_out) {
  double r_small = fmin(r1 * 1e6, r2 * 1e6);
  double r_big = fmax(r1 * 1e6, r2 * 1e6);
  double rq = r_small / r_big;
  int ir = 1, iq = 1;
  for (int k = 1; k <= 15; ++k)
    if (r_big > HALL_R0[k - 1]) ir = k + 1;
  for (int kk = 1; kk <= 21; ++kk)
    if (rq > HALL_RAT[kk - 1]) iq = kk + 1;
  if (iq < 2) {
    *ec_out = 0.0;
    return 99;
  }
  double p, q, ec;
#define ECOLL(IR, IQ) HALL_ECOLL[((IQ) - 1) * 15 + ((IR) - 1)]
  if (ir < 16) {
    if (ir >= 2) {
      p = (r_big - HALL_R0[ir - 2]) / (HALL_R0[ir - 1] - HALL_R0[ir - 2]);
      q = (rq - HALL_RAT[iq - 2]) / (HALL_RAT[iq - 1] - HALL_RAT[iq - 2]);
      ec = (1.0 - p) * (1.0 - q) * ECOLL(ir - 1, iq - 1) +
           p * (1.0 - q) * ECOLL(ir, iq - 1) +
           q * (1.0 - p) * ECOLL(ir - 1, iq) + p * q * ECOLL(ir, iq);
    } else {
      q = (rq - HALL_RAT[iq - 2]) / (HALL_RAT[iq - 1] - HALL_RAT[iq - 2]);
      ec = (1.0 - q) * ECOLL(1, iq - 1) + q * ECOLL(1, iq);
    }
  } else {
    q = (rq - HALL_RAT[iq - 2]) / (HALL_RAT[iq - 1] - HALL_RAT[iq - 2]);
    double ek = (1.0 - q) * ECOLL(15, iq - 1) + q * ECOLL(15, iq);
    ec = fmin(ek, 1.0);
  }
#undef ECOLL
  *ec_out = ec;
  return (ec < 1e-20) ? 99 : 0;
}

// kernel_sedi_helper, src/coag_kernel_sedi.F90:79-105
__device__ inline int kernel_sedi_helper(double v1, double v2, double *k) {
  double r1 = vol2rad(v1);
  double r2 = vol2rad(v2);
  double winf1 = fall_g(vol2rad(v1));
  double winf2 = fall_g(vol2rad(v2));
  double ec;
  int err = effic(r1, r2, &ec);
  *k = ec * C_PI * ((r1 + r2) * (r1 + r2)) * fabs(winf1 - winf2);
  return err;
}

// kernel_minmax, src/coag_kernel.F90:130-169 (+ src/coag_kernel_brown.F90:58-101)
__device__ inline int kernel_minmax(int type, double d_min, double d_max,
                                    double v1, double v2, const Env &env,
                                    double *kmin, double *kmax) {
  switch (type) {
    case KERNEL_SEDI: {
      double k;
      int err = kernel_sedi_helper(v1, v2, &k);
      *kmin = k;
      *kmax = k;
      return err;
    }
    case KERNEL_ADDITIVE:
      *kmin = env.beta * (v1 + v2);
      *kmax = *kmin;
      return 0;
    case KERNEL_CONSTANT:
      *kmin = C_BETA_0;
      *kmax = C_BETA_0;
      return 0;
    case KERNEL_BROWN: {
      bool first = true;
      for (int i = 1; i <= 3; ++i)
        for (int j = 1; j <= 3; ++j) {
          double d1 = interp_linear_disc(d_min, d_max, 3, i);
          double d2 = interp_linear_disc(d_min, d_max, 3, j);
          double k = kernel_brown_helper(v1, d1, v2, d2, env.temp, env.pressure);
          if (first) {
            first = false;
            *kmin = k;
            *kmax = k;
          } else {
            *kmin = fmin(*kmin, k);
            *kmax = fmax(*kmax, k);
          }
        }
      return 0;
    }
    case KERNEL_ZERO:
      *kmin = 0.0;
      *kmax = 0.0;
      return 0;
    default:
      return 330498208;
  }
}

}  // namespace pmc
