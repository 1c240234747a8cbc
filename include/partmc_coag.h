/*
 * partmc_coag.h -- C ABI of the MI355X (gfx950) particle-resolved Monte
 * Carlo coagulation step.
 *
 * The library replaces PartMC's mc_coag (src/coagulation.F90:43-96), which
 * run_part_timestep calls once per timestep (src/run_part.F90:919-920).
 * The calling convention follows the reference's only FFI precedent, the
 * ISO_C_BINDING interfaces to its GSL wrapper (src/rand.F90:82-86 and
 * src/rand_gsl.c:23-34): plain C types, integer status return, 0 = success.
 * A positive return is either one of the reference's own error codes
 * (e.g. 765620746 "particle outside of bin_grid", src/coagulation.F90:561)
 * or one of the PMC_ERR_* codes below; pmc_gpu_last_error() gives the text.
 *
 * Ownership: the caller owns the particle state between calls (Fortran's
 * aero_state); the device holds a mirror per box.  Upload before and
 * download after mc_coag when other processes touch the particles.
 *
 * Array conventions follow Fortran: species volumes are vol(n_spec, n_part)
 * column-major (vol(:, i) is particle i), weights are
 * weight(n_group, n_class) column-major, groups/classes/bins are 1-based.
 *
 * Threading: one handle per host thread; a handle owns one HIP stream.
 */
#ifndef PARTMC_COAG_H
#define PARTMC_COAG_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct pmc_gpu pmc_gpu;

/* library error codes (reference error codes pass through unchanged) */
#define PMC_ERR_ARG 1          /* invalid argument */
#define PMC_ERR_HIP 2          /* HIP runtime failure */
#define PMC_ERR_CAPACITY 3     /* particle or list capacity exceeded */
#define PMC_ERR_UNSUPPORTED 4  /* kernel type or option not supported */
#define PMC_ERR_STATE 5        /* call out of order (e.g. no particles) */

/* execution modes */
#define PMC_MODE_EXACT 0     /* sequential-order replay of mc_coag, one
                                workgroup per box; bit-identical to the
                                sequential algorithm for the same streams */
#define PMC_MODE_PARALLEL 1  /* all bin pairs and test pairs in parallel,
                                deterministic claim protocol for conflicts */

/* precision of the test-pair kernel evaluation (parallel mode) */
#define PMC_PREC_FP64 0          /* reference-style FP64 from raw volumes */
#define PMC_PREC_FP32_RECHECK 1  /* FP32 pair combine on FP64 per-particle
                                    terms; FP64 re-evaluation whenever
                                    |u - k/k_max| < band */

/* Kernel type codes are PartMC's (src/coag_kernel.F90:29-45):
 * 1 sedi, 2 additive, 3 constant, 4 brown, 5 zero.  6 and 7 (fractal
 * brown_free / brown_cont) return PMC_ERR_UNSUPPORTED. */

/* Create a handle on `device` for `n_box` independent boxes sharing one
 * aero_data (n_spec species of the given densities, kg/m^3).  `max_bin`
 * bounds the size of any bin grid (0 = 256).  Replaces the setup that
 * aero_data_t / aero_state_t carry (src/aero_data.F90:55-97,
 * src/aero_state.F90:69-86). */
int pmc_gpu_create(pmc_gpu **out, int device, int n_box, int n_spec,
                   const double *density, int max_bin);
int pmc_gpu_destroy(pmc_gpu *h);
const char *pmc_gpu_last_error(void);

/* env_state_t fields read by the kernels: temp (K), pressure (Pa) and
 * additive_kernel_coefficient (s^-1) (src/env_state.F90:29-78). */
int pmc_gpu_set_env(pmc_gpu *h, int box, double temp, double pressure,
                    double additive_coeff);

/* aero_weight_array_t (src/aero_weight_array.F90:35-39): arrays of
 * n_group*n_class in Fortran order; type codes of src/aero_weight.F90:25-29
 * (1 none, 2 power, 3 mfa). */
int pmc_gpu_set_weights(pmc_gpu *h, int box, int n_group, int n_class,
                        const int *type, const double *magnitude,
                        const double *exponent);

/* Seed of the counter-based streams and the index of the next step.
 * pmc_srand (src/rand.F90:66-113) seeds the reference's sequential
 * generator; here every call site of a step has its own Philox stream. */
int pmc_gpu_set_rng(pmc_gpu *h, uint64_t seed, uint32_t step);

/* next_id of aero_particle_new_id (src/aero_particle.F90:72,157-165) for
 * one box; new ids are (next_id - 1) * id_stride + id_offset + 1. */
int pmc_gpu_set_next_id(pmc_gpu *h, int box, int64_t next_id,
                        int64_t id_stride, int64_t id_offset);
int pmc_gpu_get_next_id(pmc_gpu *h, int box, int64_t *next_id);

/* Reserve particle capacity for every box at once (capacity[n_box]),
 * avoiding a re-layout per upload when many boxes are loaded. */
int pmc_gpu_reserve(pmc_gpu *h, const int64_t *capacity);

/* Replace one box's particles (aero_state%apa, src/aero_particle_array.F90).
 * capacity >= n_part reserves room for particles created during steps.
 * n_comp/comp_source/comp_time may be NULL (one component of source 0 at
 * time 0); comp arrays are (20, n_part) column-major.  least/greatest
 * create times may be NULL (0).  Invalidates the box's sort. */
int pmc_gpu_upload(pmc_gpu *h, int box, int n_part, int capacity,
                   const double *vol, const int64_t *id, const int *group,
                   const int *klass, const int *n_primary, const int *n_comp,
                   const int *comp_source, const double *comp_time,
                   const double *least_create_time,
                   const double *greatest_create_time);

/* aero_state_sort with an explicit log bin grid of n_bin bins over radii
 * [r_min, r_max] (src/aero_state.F90:3291-3306, src/aero_sorted.F90:254-262):
 * particles outside the grid are discarded, the rest sorted. */
int pmc_gpu_set_bin_grid(pmc_gpu *h, int box, int n_bin, double r_min,
                         double r_max);

int pmc_gpu_get_n_part(pmc_gpu *h, int box, int *n_part);

/* Copy one box's particles back (any pointer may be NULL). */
int pmc_gpu_download(pmc_gpu *h, int box, int max_part, int *n_part,
                     double *vol, int64_t *id, int *group, int *klass,
                     int *n_primary, int *n_comp, int *comp_source,
                     double *comp_time, double *least_create_time,
                     double *greatest_create_time);

/* Execution mode and precision (defaults: PMC_MODE_EXACT, PMC_PREC_FP64,
 * band 1e-5). */
int pmc_gpu_set_mode(pmc_gpu *h, int mode, int precision, double recheck_band);

/* mc_coag (src/coagulation.F90:43-96) for every box: one timestep of
 * length del_t.  tot_n_samp / tot_n_coag (arrays of n_box, may be NULL)
 * receive the reference's counters per box (src/coagulation.F90:57-59).
 * Blocks until the step is complete. */
int pmc_gpu_mc_coag(pmc_gpu *h, int kernel_type, double del_t,
                    int64_t *tot_n_samp, int64_t *tot_n_coag);

/* The same step enqueued on the handle's stream without waiting; counters
 * accumulate on the device (read them with pmc_gpu_get_counters). */
int pmc_gpu_mc_coag_async(pmc_gpu *h, int kernel_type, double del_t);
int pmc_gpu_sync(pmc_gpu *h);
/* Per-box counters accumulated since the last reset: test pairs, accepted
 * coagulations, kmax warnings (src/coagulation.F90:749-751), FP64 rechecks,
 * decision flips and max relative FP32 kernel error (validation mode). */
int pmc_gpu_get_counters(pmc_gpu *h, int64_t *n_samp, int64_t *n_coag,
                         int64_t *n_warn, int64_t *n_recheck,
                         int64_t *n_flip, double *max_rel_err);
int pmc_gpu_reset_counters(pmc_gpu *h);
/* Largest |p32 - p64| of the acceptance probability p = k/k_max seen in
 * validation mode, per box (the quantity the recheck band must exceed). */
int pmc_gpu_get_error_stats(pmc_gpu *h, double *max_p_err);
/* Record and replay of a box's random draws (exact mode).  The trace is
 * the sequence of uniforms in [0, 1) that pmc_random returns
 * (src/rand.F90:138-166), in call order.  A run without
 * GSL draws everything from pmc_random, so a trace written at that one seam
 * by a reference run drives this library through the same decisions, and a
 * trace this library records drives a reference run.
 * pmc_gpu_record: the next steps write every uniform the box draws (up to
 *   capacity values; beyond it the step fails with PMC_ERR_CAPACITY).
 * pmc_gpu_replay: the next steps draw the box's uniforms from u[0..n) in
 *   order instead of the call-site streams.
 * ref_samplers != 0: Poisson counts above mean 10 and binomials with np and
 *   n(1-p) >= 10 use the reference's rounded normal (src/rand.F90:298-302,
 *   359-361) instead of the exact samplers, as a run without GSL draws them.
 * pmc_gpu_get_trace: the uniforms used so far (n_used) and, when recording,
 *   the values (up to max_n); ran_out != 0 if the tape ran out.
 * pmc_gpu_trace_off: back to the call-site streams.
 * The parallel mode refuses to step a box with a tape (PMC_ERR_STATE). */
int pmc_gpu_record(pmc_gpu *h, int box, int64_t capacity, int ref_samplers);
int pmc_gpu_replay(pmc_gpu *h, int box, const double *u, int64_t n, int ref_samplers);
int pmc_gpu_get_trace(pmc_gpu *h, int box, double *u, int64_t max_n, int64_t *n_used,
                      int *ran_out);
int pmc_gpu_trace_off(pmc_gpu *h, int box);
/* Parallel mode: per box, the particles that joined a (bin, class) list
 * within a step since the last counter reset (merged particles that landed
 * in another list, appended particles), and how many of them found the
 * list's reserved arrival entries full (the later pairs of that step did
 * not see them).  No reference counterpart: the reference's lists are
 * sequential (src/aero_sorted.F90:397-494). */
int pmc_gpu_get_arrival_stats(pmc_gpu *h, int64_t *n_arrivals, int64_t *n_overflow);
/* The parallel mode's list sizes of one box at the start of its last step:
 * counts[(c - 1) * n_bin + (b - 1)] particles in bin b of weight class c,
 * the integer_varray_n_entry of aero_sorted's size_class lists
 * (src/aero_sorted.F90:40-56) after the sort of that step. */
int pmc_gpu_get_step_lists(pmc_gpu *h, int box, int *counts);
/* Per box, since the last reset: the samples and coagulations that the
 * per-particle path (try_per_particle_coag, src/coagulation.F90:150-241)
 * added to tot_n_samp and tot_n_coag (:234-235), in either mode. */
int pmc_gpu_get_accel_counts(pmc_gpu *h, int64_t *n_samp, int64_t *n_coag);
/* Size distribution of one box on a log-spaced diameter grid of n_bin bins
 * over [d_min, d_max] (m): bin_grid_histogram_1d (src/bin_grid.F90:225-248)
 * of the particle diameters weighted by their number concentrations, as
 * extract_aero_size does (src/extract_aero_size.F90:132-138).  quantity:
 * PMC_DIST_NUM (m^-3 per ln D), PMC_DIST_MASS (kg m^-3 per ln D) or
 * PMC_DIST_VOL (m^3 m^-3 per ln D).  hist receives n_bin values. */
#define PMC_DIST_NUM 0
#define PMC_DIST_MASS 1
#define PMC_DIST_VOL 2
int pmc_gpu_size_dist(pmc_gpu *h, int box, int quantity, int n_bin, double d_min,
                      double d_max, double *hist);
/* pmc_gpu_size_dist for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  It has no counterpart in the reference.  It allocates
 * nothing: its partial histograms live in a fixed 8 MiB device array per
 * GPU, held until the process ends.  d_hist is caller-owned DEVICE memory
 * of n_bin doubles; the call only enqueues work on the handle's stream
 * (pmc_gpu_sync waits for it).  n_bin above 2048 returns PMC_ERR_CAPACITY.
 * Calls that overlap on one GPU share that array and race: serialise
 * them (one stream, or wait for the previous call). */
int pmc_gpu_size_dist_noalloc(pmc_gpu *h, int box, int quantity, int n_bin, double d_min,
                              double d_max, double *d_hist);
/* Population control.
 * pmc_gpu_set_n_part_ideal: aero_state_set_n_part_ideal
 * (src/aero_state.F90:241-258); the box's ideal count is split evenly over
 * its (group, class) pairs when pmc_gpu_rebalance runs.
 * pmc_gpu_rebalance: aero_state_rebalance (src/aero_state.F90:1790-1852) for
 * every box, as run_part_timestep calls it after coagulation
 * (src/run_part.F90:982-984), one process: aero_state_double (:1722-1750)
 * while a (group, class) holds fewer than half its ideal count, then
 * aero_state_halve (:1755-1786) while it holds more than twice.  Halved
 * particles are logged as AERO_INFO_HALVED (3) with other id 0.  The
 * halving uniforms are drawn from the handle's seed and current step.
 * Doubling beyond a box's particle capacity fails with PMC_ERR_CAPACITY
 * before any box is changed.
 * pmc_gpu_get_weights: the box's aero_weight_array (magnitudes change under
 * population control), as set by pmc_gpu_set_weights; max_n bounds the
 * n_group * n_class entries written. */
int pmc_gpu_set_n_part_ideal(pmc_gpu *h, int box, double n_part);
int pmc_gpu_rebalance(pmc_gpu *h, int allow_doubling, int allow_halving);
int pmc_gpu_get_weights(pmc_gpu *h, int box, int max_n, int *n_group, int *n_class, int *type,
                        double *magnitude, double *exponent);

/* Sources and sinks of scenario_update_aero_state (src/scenario.F90:259-341).
 * pmc_gpu_dilute: the dilution loss to the background,
 * aero_state_sample_particles(aero_state, delta, 1 - p, AERO_INFO_DILUTION)
 * with equal weights (src/scenario.F90:316-319, src/aero_state.F90:856-929):
 * each particle of box b leaves with probability sample_prob[b], logged as
 * AERO_INFO_DILUTION (1).
 * pmc_gpu_particle_loss: scenario_particle_loss (src/scenario.F90:619-745)
 * with loss function codes of src/scenario.F90:26-36 (1 none, 2 constant,
 * 3 volume, 4 dry deposition; 5 chamber returns PMC_ERR_UNSUPPORTED): each
 * particle leaves with probability 1 - exp(-del_t rate), logged as
 * AERO_INFO_DILUTION.  height[b] is the box's mixing height (m), read by
 * dry deposition only (may be NULL otherwise).
 * Removal draws are per particle (the same distributions as the
 * reference's binomial count with uniform picks and its geometric-skip
 * accept-reject), from the handle's seed and current step.
 * pmc_gpu_scale_weights: aero_weight_array_scale
 * (src/aero_weight_array.F90:160-169), e.g. by
 * old_temp * pressure / (temp * old_pressure) (src/scenario.F90:336-338).
 * pmc_gpu_append: aero_state_add_particle for each of n particles
 * (src/aero_state.F90:365-383), e.g. the emission and background samples
 * of aero_state_add_aero_dist_sample; arguments as pmc_gpu_upload; id NULL
 * assigns new ids with aero_particle_new_id.  A particle outside the bin
 * grid invalidates the sort (the remake runs at the next step). */
int pmc_gpu_dilute(pmc_gpu *h, const double *sample_prob);
int pmc_gpu_particle_loss(pmc_gpu *h, int loss_function, double del_t, const double *height);
int pmc_gpu_scale_weights(pmc_gpu *h, int box, double factor);
int pmc_gpu_append(pmc_gpu *h, int box, int n, const double *vol, const int64_t *id,
                   const int *group, const int *klass, const int *n_primary, const int *n_comp,
                   const int *comp_source, const double *comp_time, const double *least,
                   const double *greatest);
/* Particle-steps per box since the last reset: the sum over steps of the
 * box's particle count at the start of each step. */
int pmc_gpu_get_part_steps(pmc_gpu *h, int64_t *part_steps);
/* Path statistics since the last reset.  Per box: bin pairs that took the
 * per-particle path (try_per_particle_coag, src/coagulation.F90:150-241)
 * and parallel-mode draws that found no particle still present within the
 * re-draw limit.  For the handle (rounds[4]): rounds[0] = device-wide
 * claim rounds run by the parallel mode, rounds[1] = tests it finished in
 * the per-box tail that follows them, rounds[2] = the tail's claim rounds
 * summed over boxes, rounds[3] = source samples drawn by the device-wide
 * target walks (runs and re-runs).  box_tail (may be NULL) receives per box the tests
 * finished in the tail and its claim rounds, [2 * b] and [2 * b + 1].  No
 * counterpart in the reference, which has no parallel mode. */
int pmc_gpu_get_path_stats(pmc_gpu *h, int64_t *n_accel_pairs, int64_t *n_dropped,
                           int64_t *rounds, int64_t *box_tail);
/* Per-kernel timing with HIP events on the handle's stream: when on, every
 * launch of the named kernels is bracketed by events; totals accumulate
 * until the next pmc_gpu_set_timing call. */
int pmc_gpu_set_timing(pmc_gpu *h, int on);
int pmc_gpu_get_kernel_time(pmc_gpu *h, const char *kernel, double *total_ms,
                            int64_t *launches);
/* validation on: every FP32 test is also evaluated in FP64 to count decision
 * flips and the max relative kernel error */
int pmc_gpu_set_validation(pmc_gpu *h, int on);

/* Inspection (tests): the box's bin grid, (bin, class) list counts and
 * entries (exact mode; 0-based particle indices, class-major then bin),
 * and the cached k_min/k_max tables (n_bin x n_bin, column-major). */
int pmc_gpu_get_grid(pmc_gpu *h, int box, int *n_bin, double *edges,
                     double *centers, double *widths);
int pmc_gpu_get_lists(pmc_gpu *h, int box, int *counts, int *entries);
int pmc_gpu_get_kbounds(pmc_gpu *h, int box, double *kmin, double *kmax);
/* Removal records (aero_info_array, src/aero_info.F90:48-56) of one box
 * since the last upload; returns the count in *n. */
int pmc_gpu_get_removed(pmc_gpu *h, int box, int max_n, int *n, int64_t *id,
                        int *action, int64_t *other_id);
/* Empty the removal log of one box, as run_part_timestep does after each
 * output (aero_info_array_zero, src/aero_info_array.F90:62-73). */
int pmc_gpu_clear_removed(pmc_gpu *h, int box);

/* Kernel evaluation on the device for n particle pairs (tests and
 * golden vectors): kernel(type, p1, p2) of src/coag_kernel.F90:83-125 in
 * FP64 and, when k32 is not NULL, the FP32 pair-combine value. */
int pmc_gpu_eval_kernel(pmc_gpu *h, int kernel_type, int n,
                        const double *vol1, const double *vol2, double temp,
                        double pressure, double beta, double *k64, float *k32);
/* Uniforms / Poisson / binomial variates of one call-site stream, for
 * testing the device generators against the oracle's. which: 0 uniform,
 * 1 poisson(param1), 2 binomial(int(param1), param2), 3 rand_int(param1). */
int pmc_gpu_rng_draw(pmc_gpu *h, uint64_t seed, uint32_t box, uint32_t step,
                     uint32_t site, uint32_t ordinal, uint32_t kind, int which,
                     double param1, double param2, int n, double *out);

#ifdef __cplusplus
}
#endif
#endif
