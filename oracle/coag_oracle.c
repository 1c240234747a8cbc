/*
 * coag_oracle.c -- TEST INFRASTRUCTURE ONLY: sequential FP64 restatement of
 * PartMC's mc_coag (see coag_oracle.h).  Not part of the product; the HIP
 * library never links or calls this file.
 *
 * Citations are reference file:line.  The reference's own draw order is kept
 * inside each RNG call site; the call sites are addressed by Philox counters
 * (documented at orc_stream_open).
 */
#include "coag_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

/* physical constants, src/constants.F90:29-72 */
static const double C_PI = 3.14159265358979323846;
static const double C_BOLTZMANN = 1.3806505e-23;
static const double C_AVOGADRO = 6.02214179e23;
static const double C_UNIV_GAS = 8.314472;
static const double C_AIR_MW = 2.89644e-2;
/* constant kernel coefficient, src/coag_kernel_constant.F90:21 */
static const double C_BETA_0 = 0.25 / (60.0 * 2e8);

/* ------------------------------------------------------------------ */
/* Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11)                   */
/* ------------------------------------------------------------------ */
void orc_philox4x32_10(const uint32_t ctr_in[4], const uint32_t key_in[2],
                       uint32_t out[4]) {
    uint32_t c0 = ctr_in[0], c1 = ctr_in[1], c2 = ctr_in[2], c3 = ctr_in[3];
    uint32_t k0 = key_in[0], k1 = key_in[1];
    for (int r = 0; r < 10; ++r) {
        uint64_t p0 = (uint64_t)0xD2511F53u * c0;
        uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
        uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
        uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        uint32_t n0 = hi1 ^ c1 ^ k0;
        uint32_t n2 = hi0 ^ c3 ^ k1;
        c0 = n0; c1 = lo1; c2 = n2; c3 = lo0;
        k0 += 0x9E3779B9u; k1 += 0xBB67AE85u;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

/* A call-site stream.  key = (seed_lo, seed_hi + box); counter =
 * (ordinal, site, step, kind << 24 | block).  Each Philox block yields two
 * uniforms, each built from 52 bits as (k + 1/2) 2^-52, so 0 < U < 1. */
/* the tape of the box whose mc_coag is running, attached to every stream
 * that step opens */
static orc_tape *g_tape = NULL;

void orc_stream_open(orc_stream *s, uint64_t seed, uint32_t box, uint32_t step,
                     uint32_t site, uint32_t ordinal, uint32_t kind) {
    s->tape = g_tape;
    s->key[0] = (uint32_t)seed;
    s->key[1] = (uint32_t)(seed >> 32) + box;
    s->ctr[0] = ordinal;
    s->ctr[1] = site;
    s->ctr[2] = step;
    s->ctr[3] = kind << 24;
    s->have = 0;
    s->block = 0;
}

static double u52(uint32_t a, uint32_t b) {
    uint64_t k = ((uint64_t)a << 20) | (uint64_t)(b >> 12);
    return (double)k * 0x1p-52 + 0x1p-53;
}

static double philox_uniform(orc_stream *s);

double orc_stream_uniform(orc_stream *s) {
    orc_tape *t = s->tape;
    if (t) {
        if (t->mode == 2) {
            if (t->pos < t->n) return t->u[t->pos++];
            t->err = 1;  /* ran out: go on with the stream, fail after the step */
            return philox_uniform(s);
        }
        double v = philox_uniform(s);
        if (t->pos < t->n) t->u[t->pos] = v;
        else t->err = 1;
        t->pos++;
        return v;
    }
    return philox_uniform(s);
}

static double philox_uniform(orc_stream *s) {
    if (s->have == 0) {
        uint32_t c[4] = {s->ctr[0], s->ctr[1], s->ctr[2],
                         s->ctr[3] | (s->block & 0xFFFFFFu)};
        orc_philox4x32_10(c, s->key, s->buf);
        s->block++;
        s->have = 2;
        return u52(s->buf[0], s->buf[1]);
    }
    s->have = 0;
    return u52(s->buf[2], s->buf[3]);
}

/* pmc_rand_int, non-GSL form: mod(int(U n), n) + 1  (src/rand.F90:202) */
int orc_rand_int(orc_stream *s, int n) {
    double u = orc_stream_uniform(s);
    int k = (int)(u * (double)n);
    return (k % n) + 1;
}

/* prob_round, src/rand.F90:214-227 */
int orc_prob_round(orc_stream *s, double val) {
    if (orc_stream_uniform(s) < ceil(val) - val)
        return (int)floor(val);
    return (int)ceil(val);
}

/* rand_normal, Box-Muller as in src/rand.F90:464-484 */
double orc_rand_normal(orc_stream *s, double mean, double stddev) {
    double u1 = orc_stream_uniform(s);
    double u2 = orc_stream_uniform(s);
    double r = sqrt(-2.0 * log(u1));
    double theta = 2.0 * C_PI * u2;
    double z0 = r * cos(theta);
    return stddev * z0 + mean;
}

/* rand_poisson, src/rand.F90:252-305.  Mean <= 10: Knuth's product method,
 * exactly the reference's loop (:287-296).  Mean > 10: the reference's GSL
 * build draws an exact Poisson variate (gsl_ran_poisson); the non-GSL build
 * uses a rounded normal (:298-302) which fails the reference's own pmf tests
 * (test/rand, lambda 20 and 30 at rel-tol 1e-3).  The oracle uses an exact
 * sampler there: Hormann's PTRS transformed rejection (W. Hormann, "The
 * transformed rejection method for generating Poisson random variables",
 * Insurance: Mathematics and Economics 12, 1993). */
int orc_rand_poisson(orc_stream *s, double mean) {
    if (mean > 10.0 && s->tape && s->tape->ref_samplers) {
        /* the reference without GSL (:298-302): nint of a normal variate */
        int k = (int)lround(orc_rand_normal(s, mean, sqrt(mean)));
        return k > 0 ? k : 0;
    }
    if (mean <= 10.0) {
        double L = exp(-mean);
        int k = 0;
        double p = 1.0;
        for (;;) {
            k++;
            p = p * orc_stream_uniform(s);
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
        double U = orc_stream_uniform(s) - 0.5;
        double V = orc_stream_uniform(s);
        double us = 0.5 - fabs(U);
        double kf = floor((2.0 * a / us + b) * U + mean + 0.43);
        if (us >= 0.07 && V <= vr) return (int)kf;
        if (kf < 0.0 || (us < 0.013 && V > us)) continue;
        if (log(V) + log(invalpha) - log(a / (us * us) + b) <=
            -mean + kf * loglam - lgamma(kf + 1.0))
            return (int)kf;
    }
}

/* rand_binomial, src/rand.F90:319-404.  Branches kept from the reference
 * except the normal approximation for np >= 10 and n(1-p) >= 10 (:360-363),
 * which is replaced by an exact sampler, Hormann's BTRS ("The generation of
 * binomial random variates", J. Stat. Comput. Simul. 46, 1993). */
int orc_rand_binomial(orc_stream *s, int n, double p) {
    double np = (double)n * p;
    double nomp = (double)n * (1.0 - p);
    if (np >= 10.0 && nomp >= 10.0 && s->tape && s->tape->ref_samplers) {
        /* the reference without GSL (:359-361) */
        int k = (int)lround(orc_rand_normal(s, np, sqrt(np * (1.0 - p))));
        return k < 0 ? 0 : (k > n ? n : k);
    }
    if (np >= 10.0 && nomp >= 10.0) {
        int flip = p > 0.5;
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
            double u = orc_stream_uniform(s) - 0.5;
            double v = orc_stream_uniform(s);
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
    } else {
        double q = (p <= 0.5) ? p : 1.0 - p;
        int k = 0, sum = 0;
        for (;;) {
            double G_real = log(orc_stream_uniform(s)) / log(1.0 - q);
            if (G_real > (double)(n - sum)) break;
            int G = (int)ceil(G_real);
            sum = sum + G;
            if (sum > n) break;
            k = k + 1;
        }
        return (p <= 0.5) ? k : n - k;
    }
}

/* sample_without_replacement, src/rand.F90:515-542 (1-based output) */
static void sample_without_replacement(orc_stream *s, int sample_size,
                                       int population_size, int *sample) {
    int t = 0, m = 0;
    while (m < sample_size) {
        double r = orc_stream_uniform(s);
        if ((double)(population_size - t) * r >= (double)(sample_size - m)) {
            t = t + 1;
        } else {
            t = t + 1;
            m = m + 1;
            sample[m - 1] = t;
        }
    }
}

/* sample_cts_pdf, src/rand.F90:550-576 (returns 1-based) */
static int sample_cts_pdf(orc_stream *s, int n, const double *pdf) {
    double pdf_max = pdf[0];
    for (int i = 1; i < n; ++i)
        if (pdf[i] > pdf_max) pdf_max = pdf[i];
    for (;;) {
        int k = orc_rand_int(s, n);
        if (orc_stream_uniform(s) < pdf[k - 1] / pdf_max) return k;
    }
}

/* ------------------------------------------------------------------ */
/* util.F90 numerics                                                   */
/* ------------------------------------------------------------------ */
/* sphere_vol2rad, src/util.F90:240-247: (3 v / 4 / pi)**(1/3) */
double orc_sphere_vol2rad(double v) {
    return pow(3.0 * v / 4.0 / C_PI, 1.0 / 3.0);
}

/* sphere_rad2vol, src/util.F90:264-271: 4 pi r**3 / 3 */
double orc_sphere_rad2vol(double r) {
    return 4.0 * C_PI * orc_powi(r, 3) / 3.0;
}

/* Fortran x**n for integer n: binary powering from the low bit, the
 * evaluation order of the Fortran runtime's integer power. */
double orc_powi(double x, int n) {
    double pw = 1.0;
    unsigned u;
    if (n == 0) return 1.0;
    if (n < 0) { u = (unsigned)(-n); x = 1.0 / x; } else u = (unsigned)n;
    for (;;) {
        if (u & 1u) pw *= x;
        u >>= 1;
        if (u) x *= x; else break;
    }
    return pw;
}

/* fast_pow, src/util.F90:1978-2004 */
double orc_fast_pow(double base, double exponent) {
    if (exponent == 0.0) return 1.0;
    if (exponent == 1.0) return base;
    if (exponent == 2.0) return base * base;
    if (exponent == 3.0) return base * base * base;
    if (exponent == -1.0) return 1.0 / base;
    if (exponent == -2.0) return 1.0 / (base * base);
    if (exponent == -3.0) return 1.0 / (base * base * base);
    return pow(base, exponent);
}

/* interp_linear_disc, src/util.F90:663-682 */
double orc_interp_linear_disc(double x1, double xn, int n, int i) {
    if (x1 == xn) return x1;
    double alpha = (double)(i - 1) / (double)(n - 1);
    return (1.0 - alpha) * x1 + alpha * xn;
}

/* linspace_find, src/util.F90:532-551 */
int orc_linspace_find(double min_x, double max_x, int n, double x) {
    if (x == max_x) return n - 1;
    double f = floor((x - min_x) / (max_x - min_x) * (double)(n - 1));
    int i;
    /* clamp in floating point first so huge values cannot overflow int */
    if (f > (double)n) f = (double)n;
    if (f < -1.0) f = -1.0;
    i = (int)f + 1;
    if (i > n) i = n;
    if (i < 0) i = 0;
    return i;
}

/* logspace_find, src/util.F90:566-581 */
int orc_logspace_find(double min_x, double max_x, int n, double x) {
    return orc_linspace_find(log(min_x), log(max_x), n, log(x));
}

/* linspace / logspace, src/util.F90:460-520 */
static void linspace(double min_x, double max_x, int n, double *out) {
    for (int i = 2; i <= n - 1; ++i) {
        double a = (double)(i - 1) / (double)(n - 1);
        out[i - 1] = (1.0 - a) * min_x + a * max_x;
    }
    if (n > 0) {
        out[0] = min_x;
        out[n - 1] = max_x;
    }
}

void orc_logspace(double min_x, double max_x, int n, double *out) {
    if (n == 0) return;
    linspace(log(min_x), log(max_x), n, out);
    for (int i = 0; i < n; ++i) out[i] = exp(out[i]);
    out[0] = min_x;
    out[n - 1] = max_x;
}

/* bin_grid_make (log type), src/bin_grid.F90:83-138 */
void orc_bin_grid_make_log(int n_bin, double rmin, double rmax, double *edges,
                           double *centers, double *widths) {
    if (n_bin <= 0) return;
    orc_logspace(rmin, rmax, n_bin + 1, edges);
    double c1 = exp(orc_interp_linear_disc(log(rmin), log(rmax), 2 * n_bin + 1, 2));
    double c2 = exp(orc_interp_linear_disc(log(rmin), log(rmax), 2 * n_bin + 1,
                                           2 * n_bin));
    orc_logspace(c1, c2, n_bin, centers);
    for (int i = 0; i < n_bin; ++i)
        widths[i] = (log(rmax) - log(rmin)) / (double)n_bin;
}

/* bin_grid_find, src/bin_grid.F90:143-170 (log type) */
int orc_bin_grid_find(int n_bin, const double *edges, double r) {
    if (n_bin <= 0) return -1;
    return orc_logspace_find(edges[0], edges[n_bin], n_bin + 1, r);
}

/* student_t_95_coeff, src/stats.F90:510-536 */
double orc_student_t_95_coeff(int n_sample) {
    static const double limit = 1.959963984540054;
    static const double values[15] = {
        12.7062047364, 4.30265272991, 3.18244630528, 2.7764451052,
        2.57058183661, 2.44691184879, 2.36462425101, 2.30600413503,
        2.26215716274, 2.22813885196, 2.20098516008, 2.17881282966,
        2.16036865646, 2.14478668792, 2.13144954556};
    int n_dof = n_sample - 1;
    if (n_dof <= 15) return values[n_dof - 1];
    if (n_dof <= 20)
        return (2.11990529922 - limit) * 16.0 / (double)n_dof + limit;
    return (2.07961384473 - limit) * 21.0 / (double)n_dof + limit;
}

/* histogram, bin_grid_histogram_1d src/bin_grid.F90:225-248 */
void orc_histogram_1d(int n_bin, const double *edges, const double *widths,
                      int n, const double *x, const double *w, double *out) {
    for (int i = 0; i < n_bin; ++i) out[i] = 0.0;
    for (int i = 0; i < n; ++i) {
        int b = orc_bin_grid_find(n_bin, edges, x[i]);
        if (b >= 1 && b <= n_bin) out[b - 1] += w[i] / widths[b - 1];
    }
}

/* ------------------------------------------------------------------ */
/* kernels                                                             */
/* ------------------------------------------------------------------ */
/* The air and one particle's terms of kernel_brown_helper
 * (src/coag_kernel_brown.F90:143-189), in its order of operations:
 * out = {air mean free path, radius, Knudsen number, Cunningham factor,
 *        diffusivity, mean speed squared, particle mean free path, delta^2}. */
void orc_brown_terms(double vol, double den, double temp, double pressure,
                     double out[8]) {
    double rhoair = (pressure * C_AIR_MW) / (C_UNIV_GAS * temp);
    double viscosd = 1.8325e-05 * (416.16 / (temp + 120.0)) *
                     pow(temp / 296.16, 1.5);
    double viscosk = viscosd / rhoair;
    double gasspeed = sqrt((8.0 * C_BOLTZMANN * temp * C_AVOGADRO) /
                           (C_PI * C_AIR_MW));
    double gasfreepath = 2.0 * viscosk / gasspeed;

    double Rme = orc_sphere_vol2rad(vol);
    double knud = gasfreepath / Rme;
    double cunning = 1.0 + knud * (1.249 + 0.42 * exp(-0.87 / knud));
    double diffus = (C_BOLTZMANN * temp * cunning) /
                    (6.0 * C_PI * Rme * viscosd);
    double speedsq = 8.0 * C_BOLTZMANN * temp / (C_PI * den * vol);
    double freepath = 8.0 * diffus / (C_PI * sqrt(speedsq));
    double tmp1 = orc_powi(2.0 * Rme + freepath, 3);
    double tmp2 = pow(4.0 * Rme * Rme + freepath * freepath, 1.5);
    double d = (tmp1 - tmp2) / (6.0 * Rme * freepath) - 2.0 * Rme;
    out[0] = gasfreepath;
    out[1] = Rme;
    out[2] = knud;
    out[3] = cunning;
    out[4] = diffus;
    out[5] = speedsq;
    out[6] = freepath;
    out[7] = d * d;
}

/* kernel_brown_helper, src/coag_kernel_brown.F90:111-197 (spherical
 * particles: mobility radius = geometric radius, src/fractal.F90:334-337) */
double orc_kernel_brown_helper(double vol_i, double den_i, double vol_j,
                               double den_j, double temp, double pressure) {
    double ti[8], tj[8];
    orc_brown_terms(vol_i, den_i, temp, pressure, ti);
    orc_brown_terms(vol_j, den_j, temp, pressure, tj);
    double rad_i = orc_sphere_vol2rad(vol_i);
    double rad_j = orc_sphere_vol2rad(vol_j);
    double diffus_i = ti[4], speedsq_i = ti[5], deltasq_i = ti[7];
    double diffus_j = tj[4], speedsq_j = tj[5], deltasq_j = tj[7];

    double rad_sum = rad_i + rad_j;
    double diffus_sum = diffus_i + diffus_j;
    double tmp1 = rad_sum / (rad_sum + sqrt(deltasq_i + deltasq_j));
    double tmp2 = 4.0 * diffus_sum / (rad_sum * sqrt(speedsq_i + speedsq_j));
    return 4.0 * C_PI * rad_sum * diffus_sum / (tmp1 + tmp2);
}

/* fall_g, src/coag_kernel_sedi.F90:110-172 */
void orc_fall_g(double r, double *w_inf) {
    static const double b[7] = {-0.318657e1, 0.992696, -0.153193e-2,
                                -0.987059e-3, -0.578878e-3, 0.855176e-4,
                                -0.327815e-5};
    static const double c[6] = {-0.500015e1, 0.523778e1, -0.204914e1,
                                0.475294, -0.542819e-1, 0.238449e-2};
    double eta = 1.818e-4, xlamb = 6.62e-6, rhow = 1.0, rhoa = 1.225e-3;
    double grav = 980.665;
    double cunh = 1.257 * xlamb;
    double t0 = 273.15;
    double sigma = 76.1 - 0.155 * (293.15 - t0);
    double stok = 2.0 * grav * (rhow - rhoa) / (9.0 * eta);
    double stb = 32.0 * rhoa * (rhow - rhoa) * grav / (3.0 * eta * eta);
    double phy = sigma * sigma * sigma * rhoa * rhoa /
                 (orc_powi(eta, 4) * grav * (rhow - rhoa));
    double py = pow(phy, 1.0 / 6.0);
    double rr = r * 1e2;
    double w = 0.0;
    if (rr <= 1e-3) {
        w = stok * (rr * rr + cunh * rr);
    } else if (rr > 1e-3 && rr <= 5.35e-2) {
        double x = log(stb * rr * rr * rr);
        double y = 0.0;
        for (int i = 1; i <= 7; ++i) y = y + b[i - 1] * orc_powi(x, i - 1);
        double xrey = (1.0 + cunh / rr) * exp(y);
        w = xrey * eta / (2.0 * rhoa * rr);
    } else if (rr > 5.35e-2) {
        double bond = grav * (rhow - rhoa) * (rr * rr) / sigma;
        if (rr > 0.35) bond = grav * (rhow - rhoa) * (0.35 * 0.35) / sigma;
        double x = log(16.0 * bond * py / 3.0);
        double y = 0.0;
        for (int i = 1; i <= 6; ++i) y = y + c[i - 1] * orc_powi(x, i - 1);
        double xrey = py * exp(y);
        w = xrey * eta / (2.0 * rhoa * rr);
        if (rr > 0.35) w = xrey * eta / (2.0 * rhoa * 0.35);
    }
    *w_inf = w / 100.0;
}

/* Hall (1980) collision efficiencies as tabulated in the reference's effic
 * (src/coag_kernel_sedi.F90:190-245).  The table is data that must be
 * identical; it is stored row-major as ecoll[ir][iq] (15 radii x 21
 * ratios), i.e. the transpose of the Fortran column-major DATA order. */
static const double R0[15] = {6.0, 8.0, 10.0, 15.0, 20.0, 25.0, 30.0, 40.0,
                              50.0, 60.0, 70.0, 100.0, 150.0, 200.0, 300.0};
static const double RAT[21] = {0.0, 0.05, 0.1, 0.15, 0.2, 0.25, 0.3,
                               0.35, 0.4, 0.45, 0.5, 0.55, 0.6, 0.65,
                               0.7, 0.75, 0.8, 0.85, 0.9, 0.95, 1.0};
/* column-major flat list exactly in DATA order: element (ir, iq) is at
 * (iq-1)*15 + (ir-1) */
static const double ECOLL_F[315] = {
    0.001, 0.001, 0.001, 0.001, 0.001, 0.001, 0.001, 0.001, 0.001, 0.001,
    0.001, 0.001, 0.001, 0.001, 0.001, 0.003, 0.003, 0.003, 0.004, 0.005,
    0.005, 0.005, 0.010, 0.100, 0.050, 0.200, 0.500, 0.770, 0.870, 0.970,
    0.007, 0.007, 0.007, 0.008, 0.009, 0.010, 0.010, 0.070, 0.400, 0.430,
    0.580, 0.790, 0.930, 0.960, 1.000, 0.009, 0.009, 0.009, 0.012, 0.015,
    0.010, 0.020, 0.280, 0.600, 0.640, 0.750, 0.910, 0.970, 0.980, 1.000,
    0.014, 0.014, 0.014, 0.015, 0.016, 0.030, 0.060, 0.500, 0.700, 0.770,
    0.840, 0.950, 0.970, 1.000, 1.000, 0.017, 0.017, 0.017, 0.020, 0.022,
    0.060, 0.100, 0.620, 0.780, 0.840, 0.880, 0.950, 1.000, 1.000, 1.000,
    0.030, 0.030, 0.024, 0.022, 0.032, 0.062, 0.200, 0.680, 0.830, 0.870,
    0.900, 0.950, 1.000, 1.000, 1.000, 0.025, 0.025, 0.025, 0.036, 0.043,
    0.130, 0.270, 0.740, 0.860, 0.890, 0.920, 1.000, 1.000, 1.000, 1.000,
    0.027, 0.027, 0.027, 0.040, 0.052, 0.200, 0.400, 0.780, 0.880, 0.900,
    0.940, 1.000, 1.000, 1.000, 1.000, 0.030, 0.030, 0.030, 0.047, 0.064,
    0.250, 0.500, 0.800, 0.900, 0.910, 0.950, 1.000, 1.000, 1.000, 1.000,
    0.040, 0.040, 0.033, 0.037, 0.068, 0.240, 0.550, 0.800, 0.900, 0.910,
    0.950, 1.000, 1.000, 1.000, 1.000, 0.035, 0.035, 0.035, 0.055, 0.079,
    0.290, 0.580, 0.800, 0.900, 0.910, 0.950, 1.000, 1.000, 1.000, 1.000,
    0.037, 0.037, 0.037, 0.062, 0.082, 0.290, 0.590, 0.780, 0.900, 0.910,
    0.950, 1.000, 1.000, 1.000, 1.000, 0.037, 0.037, 0.037, 0.060, 0.080,
    0.290, 0.580, 0.770, 0.890, 0.910, 0.950, 1.000, 1.000, 1.000, 1.000,
    0.037, 0.037, 0.037, 0.041, 0.075, 0.250, 0.540, 0.760, 0.880, 0.920,
    0.950, 1.000, 1.000, 1.000, 1.000, 0.037, 0.037, 0.037, 0.052, 0.067,
    0.250, 0.510, 0.770, 0.880, 0.930, 0.970, 1.000, 1.000, 1.000, 1.000,
    0.037, 0.037, 0.037, 0.047, 0.057, 0.250, 0.490, 0.770, 0.890, 0.950,
    1.000, 1.000, 1.000, 1.000, 1.000, 0.036, 0.036, 0.036, 0.042, 0.048,
    0.230, 0.470, 0.780, 0.920, 1.000, 1.020, 1.020, 1.020, 1.020, 1.020,
    0.040, 0.040, 0.035, 0.033, 0.040, 0.112, 0.450, 0.790, 1.010, 1.030,
    1.040, 1.040, 1.040, 1.040, 1.040, 0.033, 0.033, 0.033, 0.033, 0.033,
    0.119, 0.470, 0.950, 1.300, 1.700, 2.300, 2.300, 2.300, 2.300, 2.300,
    0.027, 0.027, 0.027, 0.027, 0.027, 0.125, 0.520, 1.400, 2.300, 3.000,
    4.000, 4.000, 4.000, 4.000, 4.000};

static double ecoll(int ir, int iq) { /* 1-based like the reference */
    return ECOLL_F[(iq - 1) * 15 + (ir - 1)];
}

/* effic, src/coag_kernel_sedi.F90:180-285.  Returns 99 (pmc_stop(99)) if
 * the efficiency is below 1e-20. */
int orc_effic(double r1, double r2, double *ec_out) {
    double r_small = fmin(r1 * 1e6, r2 * 1e6);
    double r_big = fmax(r1 * 1e6, r2 * 1e6);
    double rq = r_small / r_big;
    int ir = 1, iq = 1;
    double p, q, ec;
    for (int k = 1; k <= 15; ++k)
        if (r_big > R0[k - 1]) ir = k + 1;
    for (int kk = 1; kk <= 21; ++kk)
        if (rq > RAT[kk - 1]) iq = kk + 1;
    if (iq < 2) return 99; /* rat(0) would be read: rq <= 0 */
    if (ir < 16) {
        if (ir >= 2) {
            p = (r_big - R0[ir - 2]) / (R0[ir - 1] - R0[ir - 2]);
            q = (rq - RAT[iq - 2]) / (RAT[iq - 1] - RAT[iq - 2]);
            ec = (1.0 - p) * (1.0 - q) * ecoll(ir - 1, iq - 1) +
                 p * (1.0 - q) * ecoll(ir, iq - 1) +
                 q * (1.0 - p) * ecoll(ir - 1, iq) + p * q * ecoll(ir, iq);
        } else {
            q = (rq - RAT[iq - 2]) / (RAT[iq - 1] - RAT[iq - 2]);
            ec = (1.0 - q) * ecoll(1, iq - 1) + q * ecoll(1, iq);
        }
    } else {
        q = (rq - RAT[iq - 2]) / (RAT[iq - 1] - RAT[iq - 2]);
        double ek = (1.0 - q) * ecoll(15, iq - 1) + q * ecoll(15, iq);
        ec = fmin(ek, 1.0);
    }
    *ec_out = ec;
    if (ec < 1e-20) return 99;
    return 0;
}

/* kernel_sedi_helper, src/coag_kernel_sedi.F90:79-105 */
int orc_kernel_sedi_helper(double v1, double v2, double *k) {
    double r1 = orc_sphere_vol2rad(v1);
    double r2 = orc_sphere_vol2rad(v2);
    double winf1, winf2, ec;
    orc_fall_g(orc_sphere_vol2rad(v1), &winf1);
    orc_fall_g(orc_sphere_vol2rad(v2), &winf2);
    int err = orc_effic(r1, r2, &ec);
    *k = ec * C_PI * ((r1 + r2) * (r1 + r2)) * fabs(winf1 - winf2);
    return err;
}

static double particle_volume(int n_spec, const double *vol) {
    double v = 0.0;
    for (int s = 0; s < n_spec; ++s) v += vol[s];
    return v;
}

static double particle_density(int n_spec, const double *density,
                               const double *vol) {
    double m = 0.0;
    for (int s = 0; s < n_spec; ++s) m += vol[s] * density[s];
    return m / particle_volume(n_spec, vol);
}

/* kernel, src/coag_kernel.F90:83-125 */
int orc_kernel(int type, int n_spec, const double *density, const double *vol1,
               const double *vol2, double temp, double pressure, double beta,
               double *k) {
    switch (type) {
    case ORC_KERNEL_SEDI:
        return orc_kernel_sedi_helper(particle_volume(n_spec, vol1),
                                      particle_volume(n_spec, vol2), k);
    case ORC_KERNEL_ADDITIVE:
        *k = beta * (particle_volume(n_spec, vol1) +
                     particle_volume(n_spec, vol2));
        return 0;
    case ORC_KERNEL_CONSTANT:
        *k = C_BETA_0;
        return 0;
    case ORC_KERNEL_BROWN: {
        double v1 = particle_volume(n_spec, vol1);
        double v2 = particle_volume(n_spec, vol2);
        double d1 = particle_density(n_spec, density, vol1);
        double d2 = particle_density(n_spec, density, vol2);
        *k = orc_kernel_brown_helper(v1, d1, v2, d2, temp, pressure);
        return 0;
    }
    case ORC_KERNEL_ZERO:
        *k = 0.0;
        return 0;
    default:
        return 200724934;
    }
}

/* kernel_minmax, src/coag_kernel.F90:130-169 and the per-kernel minmax
 * routines (src/coag_kernel_brown.F90:58-101 samples 3x3 densities). */
int orc_kernel_minmax(int type, int n_spec, const double *density, double v1,
                      double v2, double temp, double pressure, double beta,
                      double *kmin, double *kmax) {
    switch (type) {
    case ORC_KERNEL_SEDI: {
        double k;
        int err = orc_kernel_sedi_helper(v1, v2, &k);
        *kmin = k; *kmax = k;
        return err;
    }
    case ORC_KERNEL_ADDITIVE:
        *kmin = beta * (v1 + v2);
        *kmax = *kmin;
        return 0;
    case ORC_KERNEL_CONSTANT:
        *kmin = C_BETA_0; *kmax = C_BETA_0;
        return 0;
    case ORC_KERNEL_BROWN: {
        double d_min = density[0], d_max = density[0];
        for (int s = 1; s < n_spec; ++s) {
            if (density[s] < d_min) d_min = density[s];
            if (density[s] > d_max) d_max = density[s];
        }
        int first = 1;
        for (int i = 1; i <= 3; ++i)
            for (int j = 1; j <= 3; ++j) {
                double d1 = orc_interp_linear_disc(d_min, d_max, 3, i);
                double d2 = orc_interp_linear_disc(d_min, d_max, 3, j);
                double k = orc_kernel_brown_helper(v1, d1, v2, d2, temp,
                                                   pressure);
                if (first) { first = 0; *kmin = k; *kmax = k; }
                else { *kmin = fmin(*kmin, k); *kmax = fmax(*kmax, k); }
            }
        return 0;
    }
    case ORC_KERNEL_ZERO:
        *kmin = 0.0; *kmax = 0.0;
        return 0;
    default:
        return 330498208;
    }
}

/* est_k_minmax_binned_unweighted + est_k_minmax_for_bin_unweighted,
 * src/coag_kernel.F90:253-341.  Tables are n_bin x n_bin, index
 * (b1-1) + n_bin*(b2-1). */
int orc_est_k_minmax_binned(int type, int n_spec, const double *density,
                            int n_bin, const double *edges, double temp,
                            double pressure, double beta, double *kmin_t,
                            double *kmax_t) {
    for (int b2 = 1; b2 <= n_bin; ++b2)
        for (int b1 = 1; b1 <= n_bin; ++b1) {
            double v1_low = orc_sphere_rad2vol(edges[b1 - 1]);
            double v1_high = orc_sphere_rad2vol(edges[b1]);
            double v2_low = orc_sphere_rad2vol(edges[b2 - 1]);
            double v2_high = orc_sphere_rad2vol(edges[b2]);
            double kmin = 0, kmax = 0;
            for (int i = 1; i <= 3; ++i)
                for (int j = 1; j <= 3; ++j) {
                    double v1 = orc_interp_linear_disc(v1_low, v1_high, 3, i);
                    double v2 = orc_interp_linear_disc(v2_low, v2_high, 3, j);
                    double nkmin, nkmax;
                    int err = orc_kernel_minmax(type, n_spec, density, v1, v2,
                                                temp, pressure, beta, &nkmin,
                                                &nkmax);
                    if (err) return err;
                    if (i == 1 && j == 1) { kmin = nkmin; kmax = nkmax; }
                    else { kmin = fmin(kmin, nkmin); kmax = fmax(kmax, nkmax); }
                }
            kmin_t[(b1 - 1) + n_bin * (b2 - 1)] = kmin;
            kmax_t[(b1 - 1) + n_bin * (b2 - 1)] = kmax * 2.0;
        }
    return 0;
}

/* ------------------------------------------------------------------ */
/* weights                                                             */
/* ------------------------------------------------------------------ */
/* aero_weight_num_conc_at_radius, src/aero_weight.F90:127-146 */
static double weight_nc_at_radius(int type, double mag, double expo, double r) {
    if (type == ORC_WEIGHT_NONE) return mag;
    return mag * orc_fast_pow(r, expo);
}

/* aero_weight_array_num_conc_at_radius, src/aero_weight_array.F90:232-252
 * (1-based class) */
double orc_num_conc_at_radius(int n_group, int n_class, const int *wtype,
                              const double *wmag, const double *wexp,
                              int i_class, double r) {
    (void)n_class;
    double inv_sum = 0.0;
    for (int g = 0; g < n_group; ++g) {
        int idx = g + n_group * (i_class - 1);
        inv_sum = inv_sum + 1.0 / weight_nc_at_radius(wtype[idx], wmag[idx],
                                                      wexp[idx], r);
    }
    return 1.0 / inv_sum;
}

/* coag_num_conc_factor, src/coag_kernel.F90:346-377 */
double orc_coag_num_conc_factor(int n_group, int n_class, const int *wtype,
                                const double *wmag, const double *wexp,
                                double i_r, double j_r, int i_class,
                                int j_class, int ij_class) {
    double ij_r = orc_sphere_vol2rad(orc_sphere_rad2vol(i_r) +
                                     orc_sphere_rad2vol(j_r));
    double i_nc = orc_num_conc_at_radius(n_group, n_class, wtype, wmag, wexp,
                                         i_class, i_r);
    double j_nc = orc_num_conc_at_radius(n_group, n_class, wtype, wmag, wexp,
                                         j_class, j_r);
    double ij_nc = orc_num_conc_at_radius(n_group, n_class, wtype, wmag, wexp,
                                          ij_class, ij_r);
    double nc_min = fmin(i_nc, fmin(j_nc, ij_nc));
    return i_nc * j_nc / nc_min;
}

/* coag_dest_class, src/coag_kernel.F90:382-416.  Both radii are taken at
 * centers(i_bin): the reference sets j_r = centers(i_bin) (:403). */
int orc_coag_dest_class(int n_group, int n_class, const int *wtype,
                        const double *wmag, const double *wexp,
                        const double *centers, int i_bin, int j_bin,
                        int i_class, int j_class) {
    (void)j_bin;
    double i_r = centers[i_bin - 1];
    double j_r = centers[i_bin - 1];
    double ij_r = orc_sphere_vol2rad(orc_sphere_rad2vol(i_r) +
                                     orc_sphere_rad2vol(j_r));
    double ij_nc_i = orc_num_conc_at_radius(n_group, n_class, wtype, wmag,
                                            wexp, i_class, ij_r);
    double ij_nc_j = orc_num_conc_at_radius(n_group, n_class, wtype, wmag,
                                            wexp, j_class, ij_r);
    return (ij_nc_i < ij_nc_j) ? i_class : j_class;
}

/* max_coag_num_conc_factor, src/coag_kernel.F90:422-465 */
double orc_max_coag_num_conc_factor(int n_group, int n_class, const int *wtype,
                                    const double *wmag, const double *wexp,
                                    const double *edges, int i_bin, int j_bin,
                                    int i_class, int j_class, int ij_class) {
    double i_r_min = edges[i_bin - 1], i_r_max = edges[i_bin];
    double j_r_min = edges[j_bin - 1], j_r_max = edges[j_bin];
    double f_max = 0.0;
    for (int is = 1; is <= 5; ++is)
        for (int js = 1; js <= 5; ++js) {
            double i_r = orc_interp_linear_disc(i_r_min, i_r_max, 5, is);
            double j_r = orc_interp_linear_disc(j_r_min, j_r_max, 5, js);
            double f = orc_coag_num_conc_factor(n_group, n_class, wtype, wmag,
                                                wexp, i_r, j_r, i_class,
                                                j_class, ij_class);
            f_max = fmax(f_max, f);
        }
    return f_max;
}

/* ------------------------------------------------------------------ */
/* analytic solutions                                                  */
/* ------------------------------------------------------------------ */
/* Modified Bessel function I1(x), x >= 0: ascending power series for
 * x < 30, otherwise the Hankel asymptotic expansion
 * I1(x) ~ e^x/sqrt(2 pi x) (1 - 3/(8x) - 15/(128 x^2) - ...)
 * (Abramowitz & Stegun 9.6.10 and 9.7.1).  The reference uses Cody's
 * rational approximation (src/coag_kernel_additive.F90:170-379); both are
 * checked against scipy.special.i1 by the tests. */
double orc_bessi1(double x) {
    if (x < 30.0) {
        double term = x / 2.0, sum = term, q = x * x / 4.0;
        for (int k = 1; k < 200; ++k) {
            term *= q / ((double)k * (double)(k + 1));
            sum += term;
            if (term < 1e-17 * sum) break;
        }
        return sum;
    }
    double mu = 4.0, t = 1.0, s = 1.0;
    for (int k = 1; k < 30; ++k) {
        double num = mu - (double)((2 * k - 1) * (2 * k - 1));
        t *= -num / ((double)k * 8.0 * x);
        s += t;
        if (fabs(t) < 1e-17) break;
    }
    return exp(x) / sqrt(2.0 * C_PI * x) * s;
}

/* soln_additive_exp, src/coag_kernel_additive.F90:99-165 */
void orc_soln_additive_exp(int n_bin, const double *centers, double time,
                           double num_conc, double radius_at_mean_vol,
                           double beta, double *nc_out, double *vc_out) {
    double mean_vol = orc_sphere_rad2vol(radius_at_mean_vol);
    if (time == 0.0) {
        for (int k = 0; k < n_bin; ++k)
            nc_out[k] = C_PI / 2.0 * orc_powi(2.0 * centers[k], 3) * num_conc /
                        mean_vol *
                        exp(-(orc_sphere_rad2vol(centers[k]) / mean_vol));
    } else {
        double tau = num_conc * mean_vol * beta * time;
        double T = 1.0 - exp(-tau);
        for (int k = 0; k < n_bin; ++k) {
            double rat_v = orc_sphere_rad2vol(centers[k]) / mean_vol;
            double x = 2.0 * rat_v * sqrt(T);
            double nn;
            if (x < 500.0) {
                double b = orc_bessi1(x);
                nn = num_conc / orc_sphere_rad2vol(centers[k]) * (1.0 - T) /
                     sqrt(T) * exp(-((1.0 + T) * rat_v)) * b;
            } else {
                nn = num_conc / orc_sphere_rad2vol(centers[k]) * (1.0 - T) /
                     sqrt(T) * exp((2.0 * sqrt(T) - T - 1.0) * rat_v) /
                     sqrt(4.0 * C_PI * rat_v * sqrt(T));
            }
            nc_out[k] = C_PI / 2.0 * orc_powi(2.0 * centers[k], 3) * nn;
        }
    }
    for (int k = 0; k < n_bin; ++k)
        vc_out[k] = orc_sphere_rad2vol(centers[k]) * nc_out[k];
}

/* soln_constant_exp, src/coag_kernel_constant.F90:92-143 (lambda = 1) */
void orc_soln_constant_exp(int n_bin, const double *centers, double time,
                           double num_conc, double radius_at_mean_vol,
                           double *nc_out, double *vc_out) {
    double mean_vol = orc_sphere_rad2vol(radius_at_mean_vol);
    double beta_1 = C_BETA_0;
    for (int k = 0; k < n_bin; ++k) {
        double rv = orc_sphere_rad2vol(centers[k]);
        if (time == 0.0) {
            nc_out[k] = C_PI / 2.0 * orc_powi(2.0 * centers[k], 3) * num_conc /
                        mean_vol * exp(-(rv / mean_vol));
        } else {
            double tau = num_conc * beta_1 * time;
            double T = 2.0 / (tau + 2.0);
            double nn = num_conc / mean_vol * T * T * exp(-(T * rv / mean_vol));
            nc_out[k] = C_PI / 2.0 * orc_powi(2.0 * centers[k], 3) * nn;
        }
        vc_out[k] = rv * nc_out[k];
    }
}

/* ------------------------------------------------------------------ */
/* the box: aero_state + aero_sorted                                   */
/* ------------------------------------------------------------------ */
typedef struct { int *e; int n, cap; } varray;

static void va_append(varray *v, int val) {
    if (v->n == v->cap) {
        v->cap = v->cap ? 2 * v->cap : 8;
        v->e = (int *)realloc(v->e, sizeof(int) * (size_t)v->cap);
    }
    v->e[v->n++] = val;
}

/* integer_varray_remove_entry, src/integer_varray.F90:151-171 (1-based) */
static void va_remove_entry(varray *v, int index) {
    if (index < v->n) v->e[index - 1] = v->e[v->n - 1];
    v->e[v->n - 1] = 0;
    v->n--;
}

struct orc_box {
    int n_spec;
    double density[ORC_MAX_SPEC];
    int n_group, n_class;
    int wtype[ORC_MAX_GROUP * ORC_MAX_CLASS];
    double wmag[ORC_MAX_GROUP * ORC_MAX_CLASS];
    double wexp[ORC_MAX_GROUP * ORC_MAX_CLASS];
    double temp, pressure, beta;
    uint64_t seed;
    uint32_t box_index;
    int64_t next_id;
    int64_t id_stride, id_offset;

    /* particles (aero_particle_array), 0-based storage */
    int n, cap;
    double *vol;      /* [cap][n_spec] */
    int64_t *id;
    int *group, *klass, *n_primary, *n_comp, *comp_source;
    double *comp_time, *least_ct, *greatest_ct;

    /* aero_sorted%size_class (integer_rmap2): forward1 = bin, forward2 =
     * class, index = entry position (all 1-based values) */
    int valid_sort;
    int n_bin;
    double *edges, *centers, *widths;
    int *fwd_bin, *fwd_class, *fwd_index;
    varray *inv; /* [n_bin * n_class], index (b-1) + n_bin*(c-1) */
    int bounds_valid;
    double *kmin, *kmax;

    /* aero_info_array */
    int n_info, cap_info;
    int64_t *info_id, *info_other;
    int *info_action;

    int64_t n_warn_kmax;
    int64_t n_accel;       /* bin pairs that took the per-particle path */
    int regridded;
    orc_tape tape;         /* mode 0: the call-site streams */
};

orc_box *orc_box_create(int n_spec, const double *density, int n_group,
                        int n_class, const int *wtype, const double *wmag,
                        const double *wexp) {
    if (n_spec < 1 || n_spec > ORC_MAX_SPEC || n_group < 1 ||
        n_group > ORC_MAX_GROUP || n_class < 1 || n_class > ORC_MAX_CLASS)
        return NULL;
    orc_box *b = (orc_box *)calloc(1, sizeof(orc_box));
    b->n_spec = n_spec;
    memcpy(b->density, density, sizeof(double) * (size_t)n_spec);
    b->n_group = n_group;
    b->n_class = n_class;
    memcpy(b->wtype, wtype, sizeof(int) * (size_t)(n_group * n_class));
    memcpy(b->wmag, wmag, sizeof(double) * (size_t)(n_group * n_class));
    memcpy(b->wexp, wexp, sizeof(double) * (size_t)(n_group * n_class));
    b->temp = 288.0;
    b->pressure = 1e5;
    b->beta = 1000.0;
    b->next_id = 1;
    b->id_stride = 1;
    b->id_offset = 0;
    return b;
}

static void free_sorted(orc_box *b) {
    if (b->inv) {
        for (int i = 0; i < b->n_bin * b->n_class; ++i) free(b->inv[i].e);
        free(b->inv);
        b->inv = NULL;
    }
    free(b->edges); free(b->centers); free(b->widths);
    free(b->kmin); free(b->kmax);
    b->edges = b->centers = b->widths = b->kmin = b->kmax = NULL;
}

void orc_box_destroy(orc_box *b) {
    if (!b) return;
    free_sorted(b);
    free(b->vol); free(b->id); free(b->group); free(b->klass);
    free(b->n_primary); free(b->n_comp); free(b->comp_source);
    free(b->comp_time); free(b->least_ct); free(b->greatest_ct);
    free(b->fwd_bin); free(b->fwd_class); free(b->fwd_index);
    free(b->info_id); free(b->info_other); free(b->info_action);
    free(b->tape.u);
    free(b);
}

void orc_box_set_env(orc_box *b, double temp, double pressure, double beta) {
    b->temp = temp; b->pressure = pressure; b->beta = beta;
}
void orc_box_set_rng(orc_box *b, uint64_t seed, uint32_t box_index) {
    b->seed = seed; b->box_index = box_index;
}
void orc_box_set_next_id(orc_box *b, int64_t next_id) { b->next_id = next_id; }
int64_t orc_box_get_next_id(const orc_box *b) { return b->next_id; }
int orc_box_n_part(const orc_box *b) { return b->n; }
int orc_box_n_bin(const orc_box *b) { return b->n_bin; }
int64_t orc_box_n_warn_kmax(const orc_box *b) { return b->n_warn_kmax; }
int64_t orc_box_n_accel(const orc_box *b) { return b->n_accel; }
int orc_box_regridded(const orc_box *b) { return b->regridded; }

static void ensure_cap(orc_box *b, int need) {
    if (need <= b->cap) return;
    int nc = b->cap ? b->cap : 16;
    while (nc < need) nc *= 2;
    size_t ns = (size_t)b->n_spec;
    b->vol = (double *)realloc(b->vol, sizeof(double) * ns * (size_t)nc);
    b->id = (int64_t *)realloc(b->id, sizeof(int64_t) * (size_t)nc);
    b->group = (int *)realloc(b->group, sizeof(int) * (size_t)nc);
    b->klass = (int *)realloc(b->klass, sizeof(int) * (size_t)nc);
    b->n_primary = (int *)realloc(b->n_primary, sizeof(int) * (size_t)nc);
    b->n_comp = (int *)realloc(b->n_comp, sizeof(int) * (size_t)nc);
    b->comp_source = (int *)realloc(b->comp_source,
                                    sizeof(int) * ORC_MAX_COMP * (size_t)nc);
    b->comp_time = (double *)realloc(b->comp_time,
                                     sizeof(double) * ORC_MAX_COMP * (size_t)nc);
    b->least_ct = (double *)realloc(b->least_ct, sizeof(double) * (size_t)nc);
    b->greatest_ct = (double *)realloc(b->greatest_ct, sizeof(double) * (size_t)nc);
    b->fwd_bin = (int *)realloc(b->fwd_bin, sizeof(int) * (size_t)nc);
    b->fwd_class = (int *)realloc(b->fwd_class, sizeof(int) * (size_t)nc);
    b->fwd_index = (int *)realloc(b->fwd_index, sizeof(int) * (size_t)nc);
    b->cap = nc;
}

/* a detached particle record (aero_particle_t subset on the path) */
typedef struct {
    double vol[ORC_MAX_SPEC];
    int64_t id;
    int group, klass, n_primary, n_comp;
    int comp_source[ORC_MAX_COMP];
    double comp_time[ORC_MAX_COMP];
    double least_ct, greatest_ct;
} prec;

static void load_particle(const orc_box *b, int i, prec *p) {
    memcpy(p->vol, b->vol + (size_t)i * b->n_spec, sizeof(double) * b->n_spec);
    p->id = b->id[i];
    p->group = b->group[i];
    p->klass = b->klass[i];
    p->n_primary = b->n_primary[i];
    p->n_comp = b->n_comp[i];
    memcpy(p->comp_source, b->comp_source + (size_t)i * ORC_MAX_COMP,
           sizeof(int) * ORC_MAX_COMP);
    memcpy(p->comp_time, b->comp_time + (size_t)i * ORC_MAX_COMP,
           sizeof(double) * ORC_MAX_COMP);
    p->least_ct = b->least_ct[i];
    p->greatest_ct = b->greatest_ct[i];
}

static void store_particle(orc_box *b, int i, const prec *p) {
    memcpy(b->vol + (size_t)i * b->n_spec, p->vol, sizeof(double) * b->n_spec);
    b->id[i] = p->id;
    b->group[i] = p->group;
    b->klass[i] = p->klass;
    b->n_primary[i] = p->n_primary;
    b->n_comp[i] = p->n_comp;
    memcpy(b->comp_source + (size_t)i * ORC_MAX_COMP, p->comp_source,
           sizeof(int) * ORC_MAX_COMP);
    memcpy(b->comp_time + (size_t)i * ORC_MAX_COMP, p->comp_time,
           sizeof(double) * ORC_MAX_COMP);
    b->least_ct[i] = p->least_ct;
    b->greatest_ct[i] = p->greatest_ct;
}

static void zero_particle(const orc_box *b, prec *p) {
    memset(p, 0, sizeof(*p));
    (void)b;
}

static double prec_volume(const orc_box *b, const prec *p) {
    return particle_volume(b->n_spec, p->vol);
}
static double prec_radius(const orc_box *b, const prec *p) {
    return orc_sphere_vol2rad(prec_volume(b, p));
}
static double part_radius(const orc_box *b, int i) {
    return orc_sphere_vol2rad(particle_volume(b->n_spec,
                                              b->vol + (size_t)i * b->n_spec));
}
static double nc_at(const orc_box *b, int klass, double r) {
    return orc_num_conc_at_radius(b->n_group, b->n_class, b->wtype, b->wmag,
                                  b->wexp, klass, r);
}

/* aero_particle_new_id, src/aero_particle.F90:157-165 */
static int64_t new_id(orc_box *b) {
    int64_t id = (b->next_id - 1) * b->id_stride + b->id_offset + 1;
    b->next_id++;
    return id;
}

/* aero_particle_coagulate, src/aero_particle.F90:917-1033 (on-path fields:
 * volumes, ids, groups, components, create times, primary counts) */
static int particle_coagulate(orc_box *b, orc_stream *s, const prec *p1,
                              const prec *p2, prec *pn) {
    prec out;
    memset(&out, 0, sizeof(out));
    for (int k = 0; k < b->n_spec; ++k) out.vol[k] = p1->vol[k] + p2->vol[k];
    out.group = 0;
    out.klass = 0;
    out.id = 0;
    int n1 = p1->n_comp, n2 = p2->n_comp;
    if (p1->n_primary < n1) return 465791384;
    if (p2->n_primary < n2) return 465791385;
    if (n1 + n2 > ORC_MAX_COMP) {
        int n1_new = orc_prob_round(
            s, (double)p1->n_primary / (double)(p1->n_primary + p2->n_primary) *
                   (double)ORC_MAX_COMP);
        int n2_new = ORC_MAX_COMP - n1_new;
        int sample[ORC_MAX_COMP];
        sample_without_replacement(s, n1_new, n1, sample);
        for (int i = 0; i < n1_new; ++i) {
            out.comp_source[i] = p1->comp_source[sample[i] - 1];
            out.comp_time[i] = p1->comp_time[sample[i] - 1];
        }
        sample_without_replacement(s, n2_new, n2, sample);
        for (int i = 0; i < n2_new; ++i) {
            out.comp_source[i + n1_new] = p2->comp_source[sample[i] - 1];
            out.comp_time[i + n1_new] = p2->comp_time[sample[i] - 1];
        }
        out.n_comp = ORC_MAX_COMP;
    } else {
        for (int i = 0; i < n1; ++i) {
            out.comp_source[i] = p1->comp_source[i];
            out.comp_time[i] = p1->comp_time[i];
        }
        for (int i = 0; i < n2; ++i) {
            out.comp_source[n1 + i] = p2->comp_source[i];
            out.comp_time[n1 + i] = p2->comp_time[i];
        }
        out.n_comp = n1 + n2;
    }
    out.least_ct = fmin(p1->least_ct, p2->least_ct);
    out.greatest_ct = fmax(p1->greatest_ct, p2->greatest_ct);
    out.n_primary = p1->n_primary + p2->n_primary;
    *pn = out;
    return 0;
}

static void add_info(orc_box *b, int64_t id, int action, int64_t other) {
    if (b->n_info == b->cap_info) {
        b->cap_info = b->cap_info ? 2 * b->cap_info : 64;
        b->info_id = (int64_t *)realloc(b->info_id, sizeof(int64_t) * (size_t)b->cap_info);
        b->info_other = (int64_t *)realloc(b->info_other, sizeof(int64_t) * (size_t)b->cap_info);
        b->info_action = (int *)realloc(b->info_action, sizeof(int) * (size_t)b->cap_info);
    }
    b->info_id[b->n_info] = id;
    b->info_action[b->n_info] = action;
    b->info_other[b->n_info] = other;
    b->n_info++;
}

int orc_box_n_info(const orc_box *b) { return b->n_info; }
void orc_box_get_info(const orc_box *b, int64_t *id, int *action,
                      int64_t *other) {
    for (int i = 0; i < b->n_info; ++i) {
        id[i] = b->info_id[i];
        action[i] = b->info_action[i];
        other[i] = b->info_other[i];
    }
}
void orc_box_clear_info(orc_box *b) { b->n_info = 0; }

/* aero_particle_array_remove_particle, src/aero_particle_array.F90:182-199
 * (0-based index) */
static void apa_remove(orc_box *b, int i) {
    int last = b->n - 1;
    if (i < last) {
        prec p;
        load_particle(b, last, &p);
        store_particle(b, i, &p);
    }
    b->n--;
}

/* integer_rmap2_remove, src/integer_rmap2.F90:197-246 (i_domain 1-based) */
static void rmap2_remove(orc_box *b, int i_domain) {
    int r1_old = b->fwd_bin[i_domain - 1];
    int r2_old = b->fwd_class[i_domain - 1];
    int idx_old = b->fwd_index[i_domain - 1];
    int n_dom = b->n + 1; /* the particle array has already shrunk */
    int shifted = n_dom;
    if (shifted != i_domain) {
        int r1f = b->fwd_bin[shifted - 1];
        int r2f = b->fwd_class[shifted - 1];
        int idxf = b->fwd_index[shifted - 1];
        b->inv[(r1f - 1) + b->n_bin * (r2f - 1)].e[idxf - 1] = i_domain;
    }
    /* forward arrays: swap last into slot */
    if (i_domain < n_dom) {
        b->fwd_bin[i_domain - 1] = b->fwd_bin[n_dom - 1];
        b->fwd_class[i_domain - 1] = b->fwd_class[n_dom - 1];
        b->fwd_index[i_domain - 1] = b->fwd_index[n_dom - 1];
    }
    varray *lst = &b->inv[(r1_old - 1) + b->n_bin * (r2_old - 1)];
    int idx_fix = lst->n;
    int dom_fix = lst->e[idx_fix - 1];
    va_remove_entry(lst, idx_old);
    if (idx_fix != idx_old) b->fwd_index[dom_fix - 1] = idx_old;
}

/* aero_state_remove_particle(_with_info/_no_info),
 * src/aero_state.F90:388-446 via aero_sorted_remove_particle
 * src/aero_sorted.F90:455-470 (0-based particle index) */
static void state_remove(orc_box *b, int i, int record, int64_t id,
                         int action, int64_t other) {
    apa_remove(b, i);
    rmap2_remove(b, i + 1);
    if (record) add_info(b, id, action, other);
}

/* integer_rmap2_append, src/integer_rmap2.F90:107-131 */
static void rmap2_append(orc_box *b, int dom0, int bin, int klass) {
    varray *lst = &b->inv[(bin - 1) + b->n_bin * (klass - 1)];
    va_append(lst, dom0 + 1);
    b->fwd_bin[dom0] = bin;
    b->fwd_class[dom0] = klass;
    b->fwd_index[dom0] = lst->n;
}

/* integer_rmap2_change, src/integer_rmap2.F90:136-191.  The reference's
 * early return compares the first range twice (:160), so a change that
 * keeps the bin is a no-op even when the class differs. */
static void rmap2_change(orc_box *b, int i_domain, int bin, int klass) {
    int r1_old = b->fwd_bin[i_domain - 1];
    int r2_old = b->fwd_class[i_domain - 1];
    if (r1_old == bin) return;
    int idx_old = b->fwd_index[i_domain - 1];
    varray *lst = &b->inv[(r1_old - 1) + b->n_bin * (r2_old - 1)];
    va_remove_entry(lst, idx_old);
    if (idx_old <= lst->n) {
        int shifted = lst->e[idx_old - 1];
        b->fwd_index[shifted - 1] = idx_old;
    }
    b->fwd_bin[i_domain - 1] = bin;
    b->fwd_class[i_domain - 1] = klass;
    varray *nl = &b->inv[(bin - 1) + b->n_bin * (klass - 1)];
    va_append(nl, i_domain);
    b->fwd_index[i_domain - 1] = nl->n;
}

/* aero_sorted_add_particle with valid sort, src/aero_sorted.F90:397-449.
 * allow_resort = .false. turns an out-of-grid particle into error
 * 134572570; the oracle never resorts inside mc_coag. */
static int state_add(orc_box *b, const prec *p, int allow_resort) {
    (void)allow_resort;
    int bin = orc_bin_grid_find(b->n_bin, b->edges, prec_radius(b, p));
    ensure_cap(b, b->n + 1);
    store_particle(b, b->n, p);
    b->n++;
    if (bin < 1 || bin > b->n_bin) return 134572570;
    rmap2_append(b, b->n - 1, bin, p->klass);
    return 0;
}

static void set_bin_grid(orc_box *b, int n_bin, double rmin, double rmax) {
    free_sorted(b);
    b->n_bin = n_bin;
    int nb = n_bin > 0 ? n_bin : 1;
    b->edges = (double *)calloc((size_t)nb + 1, sizeof(double));
    b->centers = (double *)calloc((size_t)nb, sizeof(double));
    b->widths = (double *)calloc((size_t)nb, sizeof(double));
    if (n_bin > 0) orc_bin_grid_make_log(n_bin, rmin, rmax, b->edges, b->centers,
                                         b->widths);
    b->inv = (varray *)calloc((size_t)nb * (size_t)b->n_class, sizeof(varray));
    b->kmin = (double *)calloc((size_t)nb * (size_t)nb, sizeof(double));
    b->kmax = (double *)calloc((size_t)nb * (size_t)nb, sizeof(double));
    b->bounds_valid = 0;
}

/* aero_sorted_sort_particles, src/aero_sorted.F90:185-209 */
static int sort_particles(orc_box *b) {
    for (int i = 0; i < b->n_bin * b->n_class; ++i) b->inv[i].n = 0;
    for (int i = 0; i < b->n; ++i) {
        int bin = orc_bin_grid_find(b->n_bin, b->edges, part_radius(b, i));
        if (bin < 1 || bin > b->n_bin) return 708651144;
        rmap2_append(b, i, bin, b->klass[i]);
    }
    return 0;
}

int orc_box_sort_with_grid(orc_box *b, int n_bin, double rmin, double rmax) {
    set_bin_grid(b, n_bin, rmin, rmax);
    /* aero_sorted_discard_outside_grid, src/aero_sorted.F90:153-180 */
    for (int i = b->n - 1; i >= 0; --i) {
        int bin = orc_bin_grid_find(b->n_bin, b->edges, part_radius(b, i));
        if (bin < 1 || bin > b->n_bin) apa_remove(b, i);
    }
    int err = sort_particles(b);
    b->valid_sort = 1;
    return err;
}

int orc_box_add_particles(orc_box *b, int n, const double *vol,
                          const int64_t *id, const int *group,
                          const int *klass, const int *n_primary,
                          const int *n_comp, const int *comp_source,
                          const double *comp_time) {
    for (int i = 0; i < n; ++i) {
        prec p;
        memset(&p, 0, sizeof(p));
        memcpy(p.vol, vol + (size_t)i * b->n_spec, sizeof(double) * b->n_spec);
        p.id = id[i];
        p.group = group[i];
        p.klass = klass[i];
        p.n_primary = n_primary ? n_primary[i] : 1;
        if (n_comp) {
            p.n_comp = n_comp[i];
            for (int k = 0; k < p.n_comp; ++k) {
                p.comp_source[k] = comp_source[(size_t)i * ORC_MAX_COMP + k];
                p.comp_time[k] = comp_time[(size_t)i * ORC_MAX_COMP + k];
            }
        } else {
            p.n_comp = 1;
            p.comp_source[0] = 0;
            p.comp_time[0] = 0.0;
        }
        if (b->valid_sort) {
            /* aero_sorted_add_particle, src/aero_sorted.F90:397-449: a
             * particle outside the grid makes the sort invalid; the remake
             * is left to the next aero_state_sort, which sees every
             * appended particle */
            int err = state_add(b, &p, 1);
            if (err == 134572570) {
                b->valid_sort = 0;
            } else if (err) {
                return err;
            }
        } else {
            ensure_cap(b, b->n + 1);
            store_particle(b, b->n, &p);
            b->n++;
        }
    }
    return 0;
}

void orc_box_get_particles(const orc_box *b, double *vol, int64_t *id,
                           int *group, int *klass, int *n_primary,
                           int *n_comp, int *comp_source, double *comp_time) {
    for (int i = 0; i < b->n; ++i) {
        if (vol) memcpy(vol + (size_t)i * b->n_spec, b->vol + (size_t)i * b->n_spec,
                        sizeof(double) * b->n_spec);
        if (id) id[i] = b->id[i];
        if (group) group[i] = b->group[i];
        if (klass) klass[i] = b->klass[i];
        if (n_primary) n_primary[i] = b->n_primary[i];
        if (n_comp) n_comp[i] = b->n_comp[i];
        if (comp_source)
            memcpy(comp_source + (size_t)i * ORC_MAX_COMP,
                   b->comp_source + (size_t)i * ORC_MAX_COMP,
                   sizeof(int) * ORC_MAX_COMP);
        if (comp_time)
            memcpy(comp_time + (size_t)i * ORC_MAX_COMP,
                   b->comp_time + (size_t)i * ORC_MAX_COMP,
                   sizeof(double) * ORC_MAX_COMP);
    }
}

void orc_box_get_grid(const orc_box *b, double *edges, double *centers,
                      double *widths) {
    if (b->n_bin <= 0) return;
    memcpy(edges, b->edges, sizeof(double) * (size_t)(b->n_bin + 1));
    memcpy(centers, b->centers, sizeof(double) * (size_t)b->n_bin);
    memcpy(widths, b->widths, sizeof(double) * (size_t)b->n_bin);
}

void orc_box_get_lists(const orc_box *b, int *counts, int *entries) {
    int off = 0;
    for (int c = 0; c < b->n_class; ++c)
        for (int bin = 0; bin < b->n_bin; ++bin) {
            const varray *v = &b->inv[bin + b->n_bin * c];
            if (counts) counts[bin + b->n_bin * c] = v->n;
            if (entries)
                for (int k = 0; k < v->n; ++k) entries[off + k] = v->e[k] - 1;
            off += v->n;
        }
}

void orc_box_get_kmax(const orc_box *b, double *kmin, double *kmax) {
    size_t nn = (size_t)b->n_bin * (size_t)b->n_bin;
    if (kmin) memcpy(kmin, b->kmin, sizeof(double) * nn);
    if (kmax) memcpy(kmax, b->kmax, sizeof(double) * nn);
}

static int n_entry(const orc_box *b, int bin, int klass) {
    return b->inv[(bin - 1) + b->n_bin * (klass - 1)].n;
}
static int entry_at(const orc_box *b, int bin, int klass, int i) {
    return b->inv[(bin - 1) + b->n_bin * (klass - 1)].e[i - 1] - 1;
}

/* aero_sorted_remake_if_needed, src/aero_sorted.F90:214-373 (no explicit
 * grid, single process) */
static int remake_if_needed(orc_box *b) {
    double r_min = 0.0, r_max = 0.0;
    int need = 0;
    b->regridded = 0;
    if (b->valid_sort) {
        int i_min = 0, i_max = 0;
        for (int bin = 1; bin <= b->n_bin; ++bin) {
            int tot = 0;
            for (int c = 1; c <= b->n_class; ++c) tot += n_entry(b, bin, c);
            if (tot > 0) {
                if (i_min == 0) i_min = bin;
                i_max = bin;
            }
        }
        if (i_min == 0) {
            if (b->n_bin > 0) {
                r_min = b->edges[b->n_bin];
                r_max = b->edges[0];
            }
        } else {
            r_min = b->edges[i_min - 1];
            r_max = b->edges[i_max];
        }
    } else {
        for (int i = 0; i < b->n; ++i) {
            double r = part_radius(b, i);
            if (i == 0) { r_min = r; r_max = r; }
            else { r_min = fmin(r_min, r); r_max = fmax(r_max, r); }
        }
    }
    if (r_max == 0.0) {
        if (b->valid_sort) return 0;
        set_bin_grid(b, 0, 0.0, 0.0);
        return 0;
    }
    if (b->n_bin < 1) {
        need = 1;
    } else {
        double g_min = b->edges[0], g_max = b->edges[b->n_bin];
        if ((r_min / g_min < 3.0) || (g_max / r_max < 3.0)) need = 1;
    }
    if (need) {
        double g_min = r_min / 10.0, g_max = r_max * 10.0;
        double nb = ceil((log10(g_max) - log10(g_min)) * 10.0);
        if (nb < 0.0) return 538534122;
        set_bin_grid(b, (int)nb, g_min, g_max);
        b->regridded = 1;
        return sort_particles(b);
    }
    if (!b->valid_sort) return sort_particles(b);
    return 0;
}

/* num_conc_weighted_kernel, src/coag_kernel.F90:175-209 */
static int weighted_kernel(orc_box *b, int type, const prec *p1,
                           const prec *p2, int c1, int c2, int cc, double *k) {
    double uk;
    int err = orc_kernel(type, b->n_spec, b->density, p1->vol, p2->vol,
                         b->temp, b->pressure, b->beta, &uk);
    if (err) return err;
    double i_r = prec_radius(b, p1);
    double j_r = prec_radius(b, p2);
    *k = uk * orc_coag_num_conc_factor(b->n_group, b->n_class, b->wtype,
                                       b->wmag, b->wexp, i_r, j_r, c1, c2, cc);
    return 0;
}

static int rand_group(orc_box *b, orc_stream *s, int klass, double r) {
    double pdf[ORC_MAX_GROUP] = {0};
    for (int g = 0; g < b->n_group; ++g) {
        int idx = g + b->n_group * (klass - 1);
        pdf[g] = 1.0 / weight_nc_at_radius(b->wtype[idx], b->wmag[idx],
                                           b->wexp[idx], r);
    }
    return sample_cts_pdf(s, b->n_group, pdf);
}

/* coagulate_weighting + coagulate, src/coagulation.F90:810-985
 * (0-based particle indices) */
static int coagulate(orc_box *b, orc_stream *s, int p1, int p2, int c1, int c2,
                     int cc) {
    prec pt1, pt2, ptc;
    load_particle(b, p1, &pt1);
    load_particle(b, p2, &pt2);
    if (pt1.id == pt2.id) return 371947172;
    double r1 = prec_radius(b, &pt1);
    double r2 = prec_radius(b, &pt2);
    double rc = orc_sphere_vol2rad(orc_sphere_rad2vol(r1) + orc_sphere_rad2vol(r2));
    double nc1 = nc_at(b, c1, r1);
    double nc2 = nc_at(b, c2, r2);
    double ncc = nc_at(b, cc, rc);
    int new_group = rand_group(b, s, cc, rc);
    double nc_min = fmin(nc1, fmin(nc2, ncc));
    double prob_remove_1 = nc_min / nc1;
    double prob_remove_2 = nc_min / nc2;
    double prob_create_new = nc_min / ncc;
    int remove_1 = orc_stream_uniform(s) < prob_remove_1;
    int remove_2 = orc_stream_uniform(s) < prob_remove_2;
    int create_new = orc_stream_uniform(s) < prob_create_new;
    int id_1_lost, id_2_lost;
    if (create_new) {
        id_1_lost = 0; id_2_lost = 0;
        if (remove_1 && remove_2) {
            if (prec_volume(b, &pt1) > prec_volume(b, &pt2)) id_2_lost = 1;
            else id_1_lost = 1;
        }
    } else {
        id_1_lost = remove_1;
        id_2_lost = remove_2;
    }
    int64_t info_other = 0;
    if (create_new) {
        int err = particle_coagulate(b, s, &pt1, &pt2, &ptc);
        if (err) return err;
        ptc.group = new_group;
        ptc.klass = cc;
        if (remove_1 && !id_1_lost) ptc.id = pt1.id;
        else if (remove_2 && !id_2_lost) ptc.id = pt2.id;
        else ptc.id = new_id(b);
        info_other = ptc.id;
    }
    if (p2 > p1) {
        if (remove_2) state_remove(b, p2, id_2_lost, pt2.id, ORC_INFO_COAG, info_other);
        if (remove_1) state_remove(b, p1, id_1_lost, pt1.id, ORC_INFO_COAG, info_other);
    } else {
        if (remove_1) state_remove(b, p1, id_1_lost, pt1.id, ORC_INFO_COAG, info_other);
        if (remove_2) state_remove(b, p2, id_2_lost, pt2.id, ORC_INFO_COAG, info_other);
    }
    if (create_new) return state_add(b, &ptc, 0);
    return 0;
}

static uint32_t pair_site(const orc_box *b, int b1, int b2, int c1, int c2) {
    return (uint32_t)((((c1 - 1) * b->n_class + (c2 - 1)) * b->n_bin + (b1 - 1)) *
                      b->n_bin + (b2 - 1));
}

/* per_set_coag + compute_n_samp + maybe_coag_pair + find_rand_pair,
 * src/coagulation.F90:587-803 */
static int per_set_coag(orc_box *b, int type, double k_max, double del_t,
                        uint32_t step, int b1, int b2, int c1, int c2, int cc,
                        int64_t *tot_n_samp, int64_t *tot_n_coag) {
    uint32_t site = pair_site(b, b1, b2, c1, c2);
    int n1 = n_entry(b, b1, c1);
    int n2 = n_entry(b, b2, c2);
    int same = (b1 == b2) && (c1 == c2);
    double n_possible = same ? (double)n1 * ((double)n2 - 1.0) / 2.0
                             : (double)n1 * (double)n2;
    double r_samp = k_max * del_t;
    double n_samp_mean = r_samp * n_possible;
    orc_stream ps;
    orc_stream_open(&ps, b->seed, b->box_index, step, site, 0, ORC_RNG_POISSON);
    int n_samp = orc_rand_poisson(&ps, n_samp_mean);
    double accept_factor = 1.0 / k_max;
    *tot_n_samp += n_samp;
    for (int i_samp = 1; i_samp <= n_samp; ++i_samp) {
        n1 = n_entry(b, b1, c1);
        n2 = n_entry(b, b2, c2);
        if ((n1 < 2 && same) || n1 < 1 || n2 < 1) break;
        orc_stream ts;
        orc_stream_open(&ts, b->seed, b->box_index, step, site,
                        (uint32_t)(i_samp - 1), ORC_RNG_TEST);
        int i1 = orc_rand_int(&ts, n1);
        int i2;
        if (same) {
            i2 = orc_rand_int(&ts, n2 - 1);
            if (i2 == i1) i2 = n2;
        } else {
            i2 = orc_rand_int(&ts, n2);
        }
        int p1 = entry_at(b, b1, c1, i1);
        int p2 = entry_at(b, b2, c2, i2);
        prec pt1, pt2;
        load_particle(b, p1, &pt1);
        load_particle(b, p2, &pt2);
        double k;
        int err = weighted_kernel(b, type, &pt1, &pt2, c1, c2, cc, &k);
        if (err) return err;
        double p = k * accept_factor;
        if (!(p <= 1.0)) b->n_warn_kmax++;
        if (orc_stream_uniform(&ts) < p) {
            orc_stream ms;
            orc_stream_open(&ms, b->seed, b->box_index, step, site,
                            (uint32_t)(i_samp - 1), ORC_RNG_MERGE);
            err = coagulate(b, &ms, p1, p2, c1, c2, cc);
            if (err) return err;
            *tot_n_coag += 1;
        }
    }
    return 0;
}

/* aero_weight_array_check_monotonicity, src/aero_weight_array.F90:304-332 */
static void check_monotonicity(const orc_box *b, int *inc, int *dec) {
    *inc = 1; *dec = 1;
    for (int i = 0; i < b->n_group * b->n_class; ++i) {
        if (b->wtype[i] == ORC_WEIGHT_POWER) {
            if (b->wexp[i] < 0.0) *inc = 0;
            if (b->wexp[i] > 0.0) *dec = 0;
        }
        if (b->wtype[i] == ORC_WEIGHT_MFA) *inc = 0;
    }
}

static void minmax_nc(const orc_box *b, int klass, double r1, double r2,
                      double *nmin, double *nmax) {
    double a = nc_at(b, klass, r1), c = nc_at(b, klass, r2);
    *nmin = fmin(a, c);
    *nmax = fmax(a, c);
}

/* aero_state_dup_particle, src/aero_state.F90:488-539 (random group) */
static int dup_particle(orc_box *b, orc_stream *s, int i_part,
                        double n_part_mean) {
    int n_copies = orc_prob_round(s, n_part_mean);
    if (n_copies == 0) {
        state_remove(b, i_part, 1, b->id[i_part], ORC_INFO_WEIGHT, 0);
    } else if (n_copies > 1) {
        for (int i_dup = 1; i_dup <= n_copies - 1; ++i_dup) {
            prec np;
            load_particle(b, i_part, &np);
            np.id = new_id(b);
            int g = rand_group(b, s, b->klass[i_part], part_radius(b, i_part));
            np.group = g;
            int err = state_add(b, &np, 1);
            if (err) return err;
        }
    }
    return 0;
}

/* sample_source_particle, src/coagulation.F90:333-517 */
static int sample_source_particle(orc_box *b, int type, uint32_t step,
                                  uint32_t site, uint32_t ordinal, int bs,
                                  int cs, const prec *coag_particle,
                                  double n_samp_mean, double accept_factor,
                                  int *n_samp_out, int *n_coag_out,
                                  prec *source) {
    int ns = b->n_spec;
    *n_samp_out = 0;
    *n_coag_out = 0;
    if (n_entry(b, bs, cs) == 0) return 0;
    orc_stream aps, smp, fin;
    orc_stream_open(&aps, b->seed, b->box_index, step, site, ordinal,
                    ORC_RNG_ACC_POISSON);
    orc_stream_open(&smp, b->seed, b->box_index, step, site, ordinal,
                    ORC_RNG_ACC_SAMPLE);
    orc_stream_open(&fin, b->seed, b->box_index, step, site, ordinal,
                    ORC_RNG_ACC_FINISH);
    double num_conc_target = nc_at(b, coag_particle->klass,
                                   prec_radius(b, coag_particle));
    int64_t target_id = coag_particle->id;
    int ct = coag_particle->klass;
    double num_conc_source_min = nc_at(b, cs, b->edges[bs - 1]);
    double prm = num_conc_target / num_conc_source_min;
    if (!(prm <= 1.0)) return 653606684;
    int n_samp_remove = orc_rand_poisson(&aps, prm * n_samp_mean);
    int n_samp_extra = orc_rand_poisson(&aps, (1.0 - prm) * n_samp_mean);
    int n_samp_total = n_samp_remove + n_samp_extra;
    int n_avg = 0, n_samp = 0;
    double prob_coag_tot = 0.0;
    double vol_sq[ORC_MAX_SPEC], vol_mean[ORC_MAX_SPEC], vol_cv[ORC_MAX_SPEC];
    for (int k = 0; k < ns; ++k) { vol_sq[k] = 0.0; vol_cv[k] = 0.0; }
    zero_particle(b, source);
    for (int i_samp = 1; i_samp <= n_samp_total; ++i_samp) {
        if (n_entry(b, bs, cs) == 0) break;
        if (n_samp > n_samp_remove && n_avg >= 2) {
            double cvmax = 0.0;
            for (int k = 0; k < ns; ++k) {
                vol_mean[k] = source->vol[k] / (double)n_avg;
                /* where(vol_mean > 0): entries with zero mean keep their
                 * previous (here: zero) value */
                if (vol_mean[k] > 0.0)
                    vol_cv[k] = sqrt(vol_sq[k] / (double)n_avg -
                                     vol_mean[k] * vol_mean[k]) / vol_mean[k];
                if (k == 0 || vol_cv[k] > cvmax) cvmax = vol_cv[k];
            }
            double mean_95 = cvmax / sqrt((double)n_avg) *
                             orc_student_t_95_coeff(n_avg);
            if (mean_95 < 0.1) break;
        }
        n_samp++;
        int i_unif = orc_rand_int(&smp, n_entry(b, bs, cs));
        int i_part = entry_at(b, bs, cs, i_unif);
        prec pi;
        load_particle(b, i_part, &pi);
        double k;
        int err = weighted_kernel(b, type, &pi, coag_particle, cs, ct, ct, &k);
        if (err) return err;
        double prob_coag = k * accept_factor;
        prob_coag_tot += prob_coag;
        if (orc_stream_uniform(&smp) < prob_coag) {
            n_avg++;
            err = particle_coagulate(b, &smp, source, &pi, source);
            if (err) return err;
            for (int kk = 0; kk < ns; ++kk) vol_sq[kk] += pi.vol[kk] * pi.vol[kk];
            if (i_samp <= n_samp_remove) {
                double num_conc_i = nc_at(b, pi.klass, prec_radius(b, &pi));
                double prob_remove_i = num_conc_target / num_conc_i;
                if (orc_stream_uniform(&smp) < prob_remove_i / prm)
                    state_remove(b, i_part, 1, pi.id, ORC_INFO_COAG, target_id);
            }
        }
    }
    *n_samp_out = n_samp;
    if (n_avg == 0) { *n_coag_out = 0; return 0; }
    double prob_coag_mean = prob_coag_tot / (double)n_samp;
    if (!(prob_coag_mean <= 1.0)) b->n_warn_kmax++;
    int n_coag = orc_rand_binomial(&fin, n_samp_total, prob_coag_mean);
    for (int kk = 0; kk < ns; ++kk)
        source->vol[kk] = source->vol[kk] * ((double)n_coag / (double)n_avg);
    source->n_primary = orc_prob_round(
        &fin, (double)source->n_primary * ((double)n_coag / (double)n_avg));
    int n_comp = source->n_comp;
    int n_comp_new = source->n_primary < ORC_MAX_COMP ? source->n_primary
                                                      : ORC_MAX_COMP;
    int sample[ORC_MAX_COMP];
    int cs_src[ORC_MAX_COMP];
    double ct_src[ORC_MAX_COMP];
    memcpy(cs_src, source->comp_source, sizeof(cs_src));
    memcpy(ct_src, source->comp_time, sizeof(ct_src));
    if (n_comp_new < n_comp) {
        sample_without_replacement(&fin, n_comp_new, n_comp, sample);
        for (int i = 0; i < n_comp_new; ++i) {
            source->comp_source[i] = cs_src[sample[i] - 1];
            source->comp_time[i] = ct_src[sample[i] - 1];
        }
        source->n_comp = n_comp_new;
    } else if (n_comp < ORC_MAX_COMP) {
        int i_comp = 1;
        while (i_comp <= n_comp_new) {
            if (i_comp + n_comp <= n_comp_new) {
                for (int i = 1; i <= n_comp; ++i) {
                    source->comp_source[i_comp + i - 2] = cs_src[i - 1];
                    source->comp_time[i_comp + i - 2] = ct_src[i - 1];
                }
                i_comp = i_comp + n_comp;
            } else {
                int sample_size = n_comp_new - i_comp + 1;
                sample_without_replacement(&fin, sample_size, n_comp, sample);
                for (int i = 1; i <= sample_size; ++i) {
                    source->comp_source[i_comp + i - 2] = cs_src[sample[i - 1] - 1];
                    source->comp_time[i_comp + i - 2] = ct_src[sample[i - 1] - 1];
                }
                i_comp = n_comp_new + 1;
            }
        }
        source->n_comp = n_comp_new;
    }
    if (source->n_primary < source->n_comp) return 808144130;
    *n_coag_out = n_coag;
    return 0;
}

/* coag_target_with_source, src/coagulation.F90:522-582 */
static int coag_target_with_source(orc_box *b, orc_stream *fin, int bt, int ct,
                                   int target_unif_entry, const prec *source,
                                   int cc) {
    int target_part = entry_at(b, bt, ct, target_unif_entry);
    int64_t target_id = b->id[target_part];
    prec tp, np;
    load_particle(b, target_part, &tp);
    double old_nc = nc_at(b, tp.klass, prec_radius(b, &tp));
    int err = particle_coagulate(b, fin, &tp, source, &np);
    if (err) return err;
    np.id = target_id;
    int new_group = rand_group(b, fin, cc, prec_radius(b, &np));
    np.group = new_group;
    np.klass = cc;
    store_particle(b, target_part, &np);
    int new_bin = orc_bin_grid_find(b->n_bin, b->edges, prec_radius(b, &np));
    if (new_bin < 1 || new_bin > b->n_bin) return 765620746;
    rmap2_change(b, target_part + 1, new_bin, cc);
    double new_nc = nc_at(b, np.klass, prec_radius(b, &np));
    err = dup_particle(b, fin, target_part, old_nc / new_nc);
    if (err) return err;
    if (b->id[target_part] != target_id) return 654300924;
    return 0;
}

/* try_per_particle_coag + determine_target_and_source + compute_n_source,
 * src/coagulation.F90:150-328 */
static int try_per_particle_coag(orc_box *b, int type, double k_max,
                                 double del_t, uint32_t step, int b1, int b2,
                                 int c1, int c2, int cc, int64_t *tot_n_samp,
                                 int64_t *tot_n_coag, int *succeeded) {
    int inc, dec;
    *succeeded = 0;
    check_monotonicity(b, &inc, &dec);
    if (!dec) return 0;
    double nc1_min, nc1_max, nc2_min, nc2_max;
    minmax_nc(b, c1, b->edges[b1 - 1], b->edges[b1], &nc1_min, &nc1_max);
    minmax_nc(b, c2, b->edges[b2 - 1], b->edges[b2], &nc2_min, &nc2_max);
    int ok = 0, bt = 0, bs = 0, ct = 0, cs = 0;
    if (nc1_max < nc2_min && cc == c1) { bt = b1; bs = b2; ct = c1; cs = c2; ok = 1; }
    if (nc2_max < nc1_min && cc == c2) { bt = b2; bs = b1; ct = c2; cs = c1; ok = 1; }
    if (!ok) return 0;
    double n_spt = k_max * del_t * (double)n_entry(b, bs, cs);
    double accept_factor = 1.0 / k_max;
    if (n_spt < 1.0) return 0;
    double min_target_vol = orc_sphere_rad2vol(b->edges[bt - 1]);
    double max_source_vol = orc_sphere_rad2vol(b->edges[bs]);
    double max_new_target_vol = min_target_vol + max_source_vol * n_spt;
    if (max_new_target_vol / min_target_vol > 1.5) return 0;
    uint32_t site = pair_site(b, b1, b2, c1, c2);
    for (int e = n_entry(b, bt, ct); e >= 1; --e) {
        int target_part = entry_at(b, bt, ct, e);
        prec target, source;
        load_particle(b, target_part, &target);
        int n_samp, n_coag;
        int err = sample_source_particle(b, type, step, site, (uint32_t)e, bs,
                                         cs, &target, n_spt, accept_factor,
                                         &n_samp, &n_coag, &source);
        if (err) return err;
        if (n_coag > 0) {
            orc_stream fin;
            orc_stream_open(&fin, b->seed, b->box_index, step, site,
                            (uint32_t)e, ORC_RNG_ACC_TARGET);
            err = coag_target_with_source(b, &fin, bt, ct, e, &source, cc);
            if (err) return err;
        }
        *tot_n_samp += n_samp;
        *tot_n_coag += n_coag;
    }
    *succeeded = 1;
    b->n_accel++;
    return 0;
}

/* mc_coag, src/coagulation.F90:43-96 */
static int mc_coag_body(orc_box *b, int type, double del_t, uint32_t step,
                        int64_t *tot_n_samp, int64_t *tot_n_coag);

int orc_mc_coag(orc_box *b, int type, double del_t, uint32_t step,
                int64_t *tot_n_samp, int64_t *tot_n_coag) {
    g_tape = b->tape.mode ? &b->tape : NULL;
    int err = mc_coag_body(b, type, del_t, step, tot_n_samp, tot_n_coag);
    g_tape = NULL;
    if (!err && b->tape.mode && b->tape.err) err = 3;  /* the tape ran out */
    return err;
}

int orc_box_set_tape(orc_box *b, int mode, const double *u, int64_t n, int ref_samplers) {
    free(b->tape.u);
    memset(&b->tape, 0, sizeof(b->tape));
    if (mode == 0) return 0;
    if (n < 1 || (mode != 1 && mode != 2)) return 1;
    b->tape.u = (double *)malloc(sizeof(double) * (size_t)n);
    if (mode == 2) memcpy(b->tape.u, u, sizeof(double) * (size_t)n);
    b->tape.n = n;
    b->tape.mode = mode;
    b->tape.ref_samplers = ref_samplers ? 1 : 0;
    return 0;
}

int64_t orc_box_tape_state(const orc_box *b, double *u, int64_t max_n, int *ran_out) {
    if (ran_out) *ran_out = b->tape.err;
    if (b->tape.mode == 1 && u) {
        int64_t m = b->tape.pos < b->tape.n ? b->tape.pos : b->tape.n;
        if (m > max_n) m = max_n;
        if (m > 0) memcpy(u, b->tape.u, sizeof(double) * (size_t)m);
    }
    return b->tape.pos;
}

static int mc_coag_body(orc_box *b, int type, double del_t, uint32_t step,
                        int64_t *tot_n_samp, int64_t *tot_n_coag) {
    *tot_n_samp = 0;
    *tot_n_coag = 0;
    int err = remake_if_needed(b);
    b->valid_sort = 1;
    if (err) return err;
    if (!b->bounds_valid) {
        err = orc_est_k_minmax_binned(type, b->n_spec, b->density, b->n_bin,
                                      b->edges, b->temp, b->pressure, b->beta,
                                      b->kmin, b->kmax);
        if (err) return err;
        b->bounds_valid = 1;
    }
    for (int c1 = 1; c1 <= b->n_class; ++c1)
        for (int c2 = 1; c2 <= c1; ++c2)
            for (int b1 = 1; b1 <= b->n_bin; ++b1) {
                if (n_entry(b, b1, c1) == 0) continue;
                int b2_start = (c1 == c2) ? b1 : 1;
                for (int b2 = b2_start; b2 <= b->n_bin; ++b2) {
                    if (n_entry(b, b2, c2) == 0) continue;
                    /* mc_coag_for_bin, src/coagulation.F90:101-145 */
                    int cc = orc_coag_dest_class(b->n_group, b->n_class,
                                                 b->wtype, b->wmag, b->wexp,
                                                 b->centers, b1, b2, c1, c2);
                    double f_max = orc_max_coag_num_conc_factor(
                        b->n_group, b->n_class, b->wtype, b->wmag, b->wexp,
                        b->edges, b1, b2, c1, c2, cc);
                    double k_max = b->kmax[(b1 - 1) + b->n_bin * (b2 - 1)] * f_max;
                    int ok = 0;
                    err = try_per_particle_coag(b, type, k_max, del_t, step, b1,
                                                b2, c1, c2, cc, tot_n_samp,
                                                tot_n_coag, &ok);
                    if (err) return err;
                    if (ok) continue;
                    err = per_set_coag(b, type, k_max, del_t, step, b1, b2, c1,
                                       c2, cc, tot_n_samp, tot_n_coag);
                    if (err) return err;
                }
            }
    return 0;
}

/* ------------------------------------------------------------------ */
/* population control: aero_state_rebalance and its doubling/halving   */
/* ------------------------------------------------------------------ */
/* aero_state_total_particles for one (group, class),
 * src/aero_state.F90:287-310 */
static int count_group_class(const orc_box *b, int g, int c) {
    int n = 0;
    for (int i = 0; i < b->n; ++i)
        if (b->group[i] == g && b->klass[i] == c) n++;
    return n;
}

/* aero_state_double, src/aero_state.F90:1722-1750: every particle of
 * (g, c) present when the pass starts is appended again with a new id, the
 * sort is invalidated and the weight magnitude halved (aero_weight_scale,
 * src/aero_weight.F90:60-69). */
int orc_box_double(orc_box *b, int g, int c) {
    int n = b->n;
    for (int i = 0; i < n; ++i) {
        if (b->group[i] != g || b->klass[i] != c) continue;
        prec p;
        load_particle(b, i, &p);
        p.id = new_id(b);
        ensure_cap(b, b->n + 1);
        store_particle(b, b->n, &p);
        b->n++;
    }
    b->valid_sort = 0;
    b->wmag[(g - 1) + b->n_group * (c - 1)] *= 0.5;
    return 0;
}

/* aero_state_halve, src/aero_state.F90:1755-1786.  Particles are visited
 * from the last to the first; a particle of (g, c) is removed when its
 * uniform is below 1/2, recorded as AERO_INFO_HALVED (other_id 0), with the
 * swap-last removal of aero_state_remove_particle (through the sorted lists
 * when the sort is valid).  The uniform of the particle at 1-based index i
 * is the first of stream (step, site, ordinal i, ORC_RNG_HALVE). */
int orc_box_halve(orc_box *b, int g, int c, uint32_t step, uint32_t site) {
    for (int i = b->n - 1; i >= 0; --i) {
        if (b->group[i] != g || b->klass[i] != c) continue;
        orc_stream s;
        orc_stream_open(&s, b->seed, b->box_index, step, site, (uint32_t)(i + 1),
                        ORC_RNG_HALVE);
        if (orc_stream_uniform(&s) < 0.5) {
            int64_t id = b->id[i];
            if (b->valid_sort) {
                state_remove(b, i, 1, id, ORC_INFO_HALVED, 0);
            } else {
                apa_remove(b, i);
                add_info(b, id, ORC_INFO_HALVED, 0);
            }
        }
    }
    b->wmag[(g - 1) + b->n_group * (c - 1)] *= 2.0;
    return 0;
}

/* aero_state_rebalance, src/aero_state.F90:1790-1852, one process: double
 * each (g, c) while it holds fewer than half its ideal count (and some
 * particles), then halve each while it holds more than twice.  ideal is
 * n_part_ideal(g, c) at ideal[(g-1) + n_group*(c-1)].  Halving pass k of
 * (g, c) draws from site (g-1) + n_group*((c-1) + n_class*k). */
int orc_box_rebalance(orc_box *b, int allow_doubling, int allow_halving,
                      const double *ideal, uint32_t step) {
    if (allow_doubling)
        for (int g = 1; g <= b->n_group; ++g)
            for (int c = 1; c <= b->n_class; ++c) {
                double id = ideal[(g - 1) + b->n_group * (c - 1)];
                int n = count_group_class(b, g, c);
                while ((double)n < id / 2.0 && n > 0) {
                    orc_box_double(b, g, c);
                    n = count_group_class(b, g, c);
                }
            }
    if (allow_halving)
        for (int g = 1; g <= b->n_group; ++g)
            for (int c = 1; c <= b->n_class; ++c) {
                double id = ideal[(g - 1) + b->n_group * (c - 1)];
                uint32_t k = 0;
                while ((double)count_group_class(b, g, c) > id * 2.0) {
                    uint32_t site = (uint32_t)(g - 1) +
                        (uint32_t)b->n_group * ((uint32_t)(c - 1) + (uint32_t)b->n_class * k);
                    orc_box_halve(b, g, c, step, site);
                    k++;
                }
            }
    return 0;
}

/* aero_weight_array_scale, src/aero_weight_array.F90:160-169 */
void orc_box_scale_weights(orc_box *b, double factor) {
    for (int i = 0; i < b->n_group * b->n_class; ++i) b->wmag[i] *= factor;
}

/* ------------------------------------------------------------------ */
/* sinks: dilution and particle loss                                   */
/* ------------------------------------------------------------------ */
/* scenario_loss_rate_dry_dep, src/scenario.F90:444-533 (spherical
 * particles: aero_data_vol2diam = 2 vol2rad) */
double orc_loss_rate_dry_dep(double vol, double density, double temp,
                             double pressure, double height) {
    const double u_mean = 5.0, z_ref = 20.0, z_rough = 0.1;
    const double A = 2.0 / 1000.0, alpha = 1.2, beta = 2.0, gamma = 0.54;
    double d_p = 2.0 * orc_sphere_vol2rad(vol);
    double density_air = (C_AIR_MW * pressure) / (C_UNIV_GAS * temp);
    double visc_d = 1.8325e-5 * (416.16 / (temp + 120.0)) * pow(temp / 296.16, 1.5);
    double visc_k = visc_d / density_air;
    double gas_speed = sqrt((8.0 * C_BOLTZMANN * temp * C_AVOGADRO) /
                            (C_PI * C_AIR_MW));
    double gas_mean_free_path = (2.0 * visc_d) / (density_air * gas_speed);
    double knud = (2.0 * gas_mean_free_path) / d_p;
    double cunning = 1.0 + knud * (1.257 + 0.4 * exp(-1.1 / knud));
    double grav = 9.81;
    double V_s = (density * pow(d_p, 2.0) * grav * cunning) / (18.0 * visc_d);
    double u_star = 0.4 * u_mean / log(z_ref / z_rough);
    double R_a = (1.0 / (0.4 * u_star)) * log(z_ref / z_rough);
    double diff_p = (C_BOLTZMANN * temp * cunning) / (3.0 * C_PI * visc_d * d_p);
    double Sc = visc_k / diff_p;
    double E_B = pow(Sc, -gamma);
    double E_IN = 0.5 * pow(d_p / A, 2.0);
    double St = (V_s * u_star) / (grav * A);
    double E_IM = pow(St / (alpha + St), beta);
    double R1 = exp(-pow(St, 0.5));
    double eps_0 = 3.0;
    double R_s = 1.0 / (eps_0 * u_star * (E_B + E_IN + E_IM) * R1);
    double V_d = V_s + (1.0 / (R_a + R_s + R_a * R_s * V_s));
    return V_d / height;
}

/* scenario_loss_rate, src/scenario.F90:398-437, loss function codes
 * :26-36.  The chamber function needs chamber geometry that is not on the
 * path: ORC_ERR_UNSUPPORTED. */
int orc_loss_rate(int type, double vol, double density, double temp,
                  double pressure, double height, double *rate) {
    switch (type) {
    case ORC_LOSS_INVALID: case ORC_LOSS_NONE: *rate = 0.0; return 0;
    case ORC_LOSS_CONSTANT: *rate = 1e-3; return 0;
    case ORC_LOSS_VOLUME: *rate = 1e15 * vol; return 0;
    case ORC_LOSS_DRYDEP:
        *rate = orc_loss_rate_dry_dep(vol, density, temp, pressure, height);
        return 0;
    case ORC_LOSS_CHAMBER: return ORC_ERR_UNSUPPORTED;
    default: return 201594391;
    }
}

/* One removal pass: the particle at 1-based index i is removed with its
 * probability prob[i-1] when the first uniform of stream (step, site 0,
 * ordinal i, kind) is below it, from the last particle down, swap-last
 * (through the lists when the sort is valid), logged with the action and
 * other id 0.  Independent per-particle draws give the same distribution
 * as the reference's binomial count plus uniform picks
 * (aero_state_sample_particles) and its geometric-skip accept-reject
 * (scenario_particle_loss); they are drawn per particle here so the pass
 * runs in parallel on the device. */
static void remove_pass(orc_box *b, const double *prob, int action,
                        uint32_t step, uint32_t kind) {
    for (int i = b->n - 1; i >= 0; --i) {
        if (!(prob[i] > 0.0)) continue;
        orc_stream s;
        orc_stream_open(&s, b->seed, b->box_index, step, 0, (uint32_t)(i + 1), kind);
        if (orc_stream_uniform(&s) < prob[i]) {
            int64_t id = b->id[i];
            if (b->valid_sort) {
                state_remove(b, i, 1, id, action, 0);
            } else {
                apa_remove(b, i);
                add_info(b, id, action, 0);
            }
        }
    }
}

/* Dilution out of the box: aero_state_sample_particles(from, to, prob,
 * AERO_INFO_DILUTION) with equal weights on both sides
 * (src/aero_state.F90:856-929, called from src/scenario.F90:316-319): each
 * particle leaves with probability sample_prob. */
int orc_box_dilute(orc_box *b, double sample_prob, uint32_t step) {
    if (!(sample_prob >= 0.0 && sample_prob <= 1.0)) return 721006962;
    double *p = (double *)malloc(sizeof(double) * (size_t)(b->n > 0 ? b->n : 1));
    for (int i = 0; i < b->n; ++i) p[i] = sample_prob;
    remove_pass(b, p, ORC_INFO_DILUTION, step, ORC_RNG_DILUTE);
    free(p);
    return 0;
}

/* scenario_particle_loss, src/scenario.F90:619-704 with
 * scenario_try_single_particle_loss :709-745: each particle is removed
 * with probability 1 - exp(-del_t rate(vol, density)), logged as
 * AERO_INFO_DILUTION. */
int orc_box_particle_loss(orc_box *b, int loss_type, double del_t,
                          double height, uint32_t step) {
    if (loss_type == ORC_LOSS_NONE || loss_type == ORC_LOSS_INVALID) return 0;
    double *p = (double *)malloc(sizeof(double) * (size_t)(b->n > 0 ? b->n : 1));
    for (int i = 0; i < b->n; ++i) {
        const double *v = b->vol + (size_t)i * b->n_spec;
        double rate = 0.0;
        int err = orc_loss_rate(loss_type, particle_volume(b->n_spec, v),
                                particle_density(b->n_spec, b->density, v),
                                b->temp, b->pressure, height, &rate);
        if (err) {
            free(p);
            return err;
        }
        p[i] = 1.0 - exp(-del_t * rate);
    }
    remove_pass(b, p, ORC_INFO_DILUTION, step, ORC_RNG_LOSS);
    free(p);
    return 0;
}

void orc_box_get_weights(const orc_box *b, double *wmag) {
    memcpy(wmag, b->wmag, sizeof(double) * (size_t)(b->n_group * b->n_class));
}

/* ------------------------------------------------------------------ */
/* batch samplers for the pmf tests (test/rand protocol)               */
/* ------------------------------------------------------------------ */
void orc_sample_pmf(uint64_t seed, int which, double p1, double p2, int64_t n,
                    int kmax, int64_t *counts) {
    for (int k = 0; k <= kmax; ++k) counts[k] = 0;
    orc_stream s;
    for (int64_t i = 0; i < n; ++i) {
        orc_stream_open(&s, seed, 0, (uint32_t)(i >> 32), 0, (uint32_t)i, 1);
        int k = (which == 0) ? orc_rand_poisson(&s, p1)
                             : orc_rand_binomial(&s, (int)p1, p2);
        if (k >= 0 && k <= kmax) counts[k]++;
    }
}

/* ------------------------------------------------------------------ */
/* single draws of rand_group and of the component part of             */
/* aero_particle_coagulate, for their statistics tests                 */
/* ------------------------------------------------------------------ */
/* aero_weight_array_rand_group (src/aero_weight_array.F90:374-393) of
 * box b's weights on stream (seed, ordinal) */
int orc_box_rand_group(orc_box *b, uint64_t seed, uint32_t ordinal, int klass,
                       double radius) {
    orc_stream s;
    orc_stream_open(&s, seed, 0, 0, 0, ordinal, 0);
    return rand_group(b, &s, klass, radius);
}

/* aero_particle_coagulate's components (src/aero_particle.F90:960-994) of
 * two particles with n1 and n2 components of sources src1 and src2 and
 * np1, np2 primary parts, on stream (seed, ordinal): n_out and src_out
 * receive the merged particle's components */
int orc_box_coagulate_components(orc_box *b, uint64_t seed, uint32_t ordinal,
                                 int n1, int np1, const int *src1, int n2,
                                 int np2, const int *src2, int *n_out,
                                 int *src_out) {
    if (n1 < 0 || n2 < 0 || n1 > ORC_MAX_COMP || n2 > ORC_MAX_COMP) return 1;
    prec p1, p2, pn;
    memset(&p1, 0, sizeof(p1));
    memset(&p2, 0, sizeof(p2));
    p1.n_comp = n1;
    p1.n_primary = np1;
    p2.n_comp = n2;
    p2.n_primary = np2;
    for (int i = 0; i < n1; ++i) p1.comp_source[i] = src1[i];
    for (int i = 0; i < n2; ++i) p2.comp_source[i] = src2[i];
    orc_stream s;
    orc_stream_open(&s, seed, 0, 0, 0, ordinal, 0);
    int err = particle_coagulate(b, &s, &p1, &p2, &pn);
    if (err) return err;
    *n_out = pn.n_comp;
    for (int i = 0; i < pn.n_comp; ++i) src_out[i] = pn.comp_source[i];
    return 0;
}
