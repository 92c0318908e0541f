/* nmlmc: MI355X (gfx950) level-l nested Monte Carlo sampler, C ABI.
 *
 * Drop-in for the reference's extern "C" entry points (the libnested.so built by
 * the reference makefile:72-73) with the reference's -DREAL_DOUBLE data layouts:
 * every real_t / prob_t field below is a double.  Sums and work are double, as in
 * the reference (common.hpp:17-18).  The reference functions are void and abort
 * on error; here they keep their void signatures and record a status that
 * nmlmc_last_error() returns (0 = success).  The nmlmc_* entry points return that
 * status directly.
 *
 * Threading: like the reference, one problem per process and not re-entrant;
 * call from one host thread.  Outputs of a sample call are overwritten, not
 * accumulated (nested_sampler.hpp:655).
 */
#ifndef NMLMC_H
#define NMLMC_H
#include <stddef.h>
#include <stdint.h>
#ifndef __cplusplus
#include <stdbool.h>
#endif

#ifdef __cplusplus
extern "C" {
#endif

/* ---- data layouts crossing the ABI ---------------------------------------- */

/* nested_sampler_params (nested_sampler.hpp:14-34), REAL_DOUBLE layout. */
typedef struct nested_sampler_params {
    uint32_t N0;
    uint32_t beta;
    bool antithetic;
    uint32_t fine_moments;
    uint32_t diff_moments;
    double adaptive_r; /* adaptive iff 1 < r < 2 */
} nested_sampler_params;

/* sample_output (common.hpp:37-55): the caller owns fsums[phi*fine_moments] and
 * diffsums[phi*diff_moments]. */
typedef struct sample_output {
    uint64_t M;
    double* fsums;
    double* diffsums;
    double total_work;
    uint64_t inner_stats[6]; /* {inner samples, outer, at max N, at base N, #uniforms, #normals} */
} sample_output;

/* model::data (model_sampler.cpp:21-39).  phi_type: 0 Indicator, 1 Max. */
typedef struct model_data {
    double stau;     /* sqrt(tau) */
    double s1tau;    /* sqrt(1 - tau) */
    double max_loss;
    int32_t phi_type;
} model_data;

/* span<T> (common.hpp:69-131): pointer + size. */
typedef struct nmlmc_span {
    void* data;
    size_t size;
} nmlmc_span;

/* portfolio::option_params / stock_params (portfolio_sampler.cpp:42-55).
 * type: 1 call, -1 put; extension: 2 / -2 cash-or-nothing digital call / put. */
typedef struct option_params {
    double strike;
    double maturity;
    double start;     /* option value at t=0 (portfolio_comp_params) */
    double w;         /* weight */
    double delta_est; /* delta estimate (portfolio_comp_params) */
    uint32_t stock_ind;
    int32_t type;
    double prob;      /* sub-sampling weight; options sorted by descending prob per method */
    double W_est;     /* work estimate */
} option_params;

typedef struct stock_params {
    double S0, mu, sigma;
} stock_params;

/* portfolio::data (portfolio_sampler.cpp:62-92).  options[m] for m = BS, Exact,
 * Unbiased; spans of option_params / stock_params.  phi_type: 0 Indicator,
 * 1 Max, 2 Moments, 3 FMoments.  subsample: 0 None, 1 Random, -1 Random_Uniform. */
typedef struct portfolio_data {
    nmlmc_span options[3];
    nmlmc_span stocks;
    double r, rho, s1_rho2;
    double max_loss;
    double tau;
    uint32_t dt_h0inv;
    uint32_t dt_beta;
    double dt_prob_expo;
    int32_t phi_type;
    uint32_t _phi_size;
    int32_t subsample;
    bool use_delta_cv;
    bool use_coarse_delta_cv;
    bool use_antithetic_risk;
    bool use_milstein;
} portfolio_data;

/* ---- reference entry points ------------------------------------------------ */

/* common.cpp:20-26.  cores = number of virtual threads of the parity keying. */
void init_nestedmc_sampler(uint32_t cores, const nested_sampler_params* params);
/* common.cpp:28-31. */
void cleanup_nestedmc_sampler(void);

/* model_sampler.cpp:94-97. */
void model_update_data(const model_data* data);
/* model_sampler.cpp:99-104. */
void model_init(const model_data* data, uint64_t seed);
/* model_sampler.cpp:106-114. */
void model_sample(uint32_t ell, uint64_t M, sample_output* out);

/* portfolio_sampler.cpp:618-629. */
void portfolio_cleanup(void);
/* portfolio_sampler.cpp:631-678: builds per-method CDFs, compute_prob, prob_norm and
 * delta_coalesce on the host and copies every table to device memory. */
void portfolio_update_data(const portfolio_data* data);
/* portfolio_sampler.cpp:680-686. */
void portfolio_init(const portfolio_data* data, uint64_t seed);
/* portfolio_sampler.cpp:687-700: start = Black-Scholes value, delta_est = BS delta. */
void portfolio_comp_params(portfolio_data* data);
/* portfolio_sampler.cpp:749-756.  override_dyn_inner > 0 forces the number of lanes
 * that share one outer sample's inner loop (rounded to a power of two <= 64). */
void portfolio_sample(uint32_t ell, uint64_t override_dyn_inner, uint64_t M, sample_output* out);

/* portfolio_sampler.cpp:758-781.  Moments [0, moments) of *M delta samples of
 * Unbiased option option_index (Milstein against exact paths, plain keying:
 * uses and advances the inner streams of the virtual threads); *M is
 * unchanged, *total_work is overwritten. */
void portfolio_sample_delta_unbiased(uint32_t option_index, uint64_t* M, double* sums, uint32_t moments,
                                     double* total_work);

/* portfolio_sampler.cpp:914-934.  Adds to each Unbiased option's delta_est the
 * mean of batches of 1000 delta samples drawn until 3 sqrt(V/M) < TOL. */
void portfolio_estimate_delta_for_unbiased(portfolio_data* data, double TOL);

/* portfolio_sampler.cpp:792-904.  A synthetic portfolio of P options on Q
 * stocks (P >= 2Q), split over BS / Exact / Unbiased by floor(probs[m] * P),
 * priced, delta-hedged, normalised and sorted; the arrays are allocated here
 * and released by portfolio_free_data. */
portfolio_data portfolio_gen_data(uint32_t P, uint32_t Q, uint64_t seed, double* probs, bool uniform_weights,
                                  bool uniform_prob);

/* portfolio_sampler.cpp:906-912. */
void portfolio_free_data(portfolio_data* data);

/* ---- extensions (no reference counterpart) -------------------------------- */

enum {
    NMLMC_OK = 0,
    NMLMC_ERR_STATE = 1,   /* call order: not initialised, or no problem staged */
    NMLMC_ERR_ARG = 2,     /* invalid argument */
    NMLMC_ERR_HIP = 3,     /* HIP runtime error */
    NMLMC_ERR_DEVICE = 4   /* device-side check failed */
};

/* Forms: PLAIN = FP64, reference keying (per-virtual-thread streams keyed
 * {seed, t}, bit-compatible with the reference OpenMP build); BATCHED = FP64,
 * per-outer-sample keys, persistent work-queue kernel; MIXED = the batched form
 * with FP32 inner arithmetic + compensated sums and FP64 re-evaluation near the
 * Heaviside threshold. */
enum { NMLMC_FORM_PLAIN = 0, NMLMC_FORM_BATCHED = 1, NMLMC_FORM_MIXED = 2 };

int nmlmc_last_error(void);
const char* nmlmc_last_error_string(void);
int nmlmc_set_form(int form);
int nmlmc_get_form(void);

/* Adaptive rule: 0 = reference (nested_sampler.hpp:219-274); 1 = start at
 * N0*beta^l and double while |mean| < 3 stderr, up to (N0*beta^l) << cap_log2. */
int nmlmc_set_adaptive_rule(uint32_t rule, uint32_t cap_log2);

/* Model problem with a variant: 0 = reference model (model_sampler.cpp:66-79);
 * 1 = shift model X = Y + Z (one normal per inner sample), threshold = max_loss. */
int nmlmc_model_init(int variant, int phi_type, double tau, double max_loss, uint64_t seed);

/* General correlation for the portfolio outer draw: W_tau = sqrt(tau) * L z with
 * L a d x d row-major lower-triangular factor (NULL restores the reference's
 * one-factor draw).  Call after portfolio_init / portfolio_update_data. */
int nmlmc_portfolio_set_correlation(const double* chol, uint32_t d);

/* Level-coupled inner paths for the options of method 2 (Unbiased) of the
 * portfolios staged after this call (portfolio_init, nmlmc_batch_stage; no
 * counterpart in the reference): instead of the Unbiased method's randomised
 * level, such an option's inner sample at MLMC level l runs the Milstein paths
 * of sample_inner_approx (portfolio_sampler.cpp:539-604) with
 * dt_h0inv 2^floor(l/2) steps over (tau, T] - the batched-risk workload's
 * arithmetic-average Asians.  An antithetic final inner sample of a level
 * l > 0 gives the fine estimator its value on those paths and the coarse
 * estimator its value on dt_h0inv 2^floor((l-1)/2) steps of the same
 * increments; refinement tests use their level's step count.  Batched and
 * mixed forms; sub-sampling mode 1. */
int nmlmc_set_level_coupled_paths(int on);

/* One level call with a status.  outer_offset (batched forms) selects the outer
 * sample ids [outer_offset, outer_offset + M); pass UINT64_MAX to continue after
 * the previous call of this level.  Optional per-outer-sample trace arrays of
 * length M (any may be NULL): phi_f, phi_c (first phi component), nf, nc. */
int nmlmc_sample(uint32_t ell, uint64_t M, uint64_t outer_offset, sample_output* out,
                 double* tr_phi_f, double* tr_phi_c, uint64_t* tr_nf, uint64_t* tr_nc);

/* A level sweep: n level calls (ells[i], Ms[i]) into outs[i], each continuing
 * its level's outer ids as nmlmc_sample(..., UINT64_MAX, ...) does and with the
 * same results.  In the batched forms the calls' kernel chains are captured as
 * one HIP graph and run with one wait (NMLMC_SWEEP_GRAPH=0: enqueued on the
 * stream directly); the plain form runs them one by one.  No counterpart in the
 * reference, whose mlmc_t::run calls its sampler level by level
 * (mlmc.hpp:306-314); the native driver uses it for each iteration's levels.
 * nmlmc_last_kernel_ms then gives the whole sweep's device time. */
int nmlmc_sample_levels(uint32_t n, const uint32_t* ells, const uint64_t* Ms, sample_output* outs);

/* ---- A batch of portfolio problems (no counterpart in the reference, which
 * holds one problem per process) ------------------------------------------
 * nmlmc_batch_begin(n) opens a batch of n <= 64 instances (dropping any earlier
 * batch); nmlmc_batch_stage(i, data, seed, chol, d) builds instance i's tables
 * as portfolio_update_data does, with its own seed and optional Cholesky
 * factor (this replaces the engine's own staged problem, which must be staged
 * again before it is sampled).  nmlmc_batch_sample(ell, n, M, outs) is one
 * level call over all instances in one launch chain per chunk (batched and
 * mixed forms): M[i] outer samples of instance i, continuing that instance's
 * level ids, results into outs[i] - equal to those of instance i staged alone
 * with portfolio_init(data, seed) and sampled with the same calls.  n is the
 * length of M and outs and must equal the batch's instance count
 * (NMLMC_ERR_ARG otherwise).  The instances share the sampler parameters and
 * need one phi type (Indicator or Max). */
int nmlmc_batch_begin(uint32_t n);
int nmlmc_batch_stage(uint32_t i, const portfolio_data* data, uint64_t seed, const double* chol, uint32_t d);
int nmlmc_batch_sample(uint32_t ell, uint32_t n, const uint64_t* M, sample_output* outs);
/* nmlmc_autotune_level for the batch level call (ell, M[0..n-1]): the same
 * candidate shapes, timed on outer ids from 2^45 of every instance (the
 * instances' own outer ids do not move), the fastest kept for later batch
 * calls at level ell; dropped by nmlmc_batch_begin.  n as for
 * nmlmc_batch_sample. */
int nmlmc_batch_autotune_level(uint32_t ell, uint32_t n, const uint64_t* M, uint32_t* segment,
                               uint32_t* blocks_per_cu, double* ms);

/* A level call of the plain FP64 form (nmlmc_set_form(NMLMC_FORM_PLAIN), the
 * form of model_sample / portfolio_sample) that allocates no memory: the same
 * kernels and results as model_sample / portfolio_sample(ell, 0, M, out) of
 * the staged problem, with their scratch (one partial record per block of
 * virtual threads, the reduced record and the reduction's staging records) in
 * a static device array of 16384 + 1 + 256 records set up at
 * init_nestedmc_sampler.  No counterpart in the reference: for callers that
 * cannot allocate (HIP graph capture, serving loops, harnesses that forbid
 * hipMalloc).  Limit: min(M, cores) x lanes / 256 <= 16384 blocks (the virtual
 * threads holding outer samples: 65536 of 64 lanes; the portfolio: min(M,
 * cores) <= 16384), else NMLMC_ERR_ARG.  Calls that overlap on one GPU share the array and race.
 * NMLMC_ERR_STATE if the library is not initialised or the form is not plain. */
int nmlmc_sample_noalloc(uint32_t ell, uint64_t M, sample_output* out);

/* model::sampler (model_sampler.cpp:42-92) written as a class of the
 * reference's sampler concept and run through the C++ template entry points
 * nmlmc::init_samplers<SAMPLER> / nmlmc::sample<SAMPLER> of
 * include/nmlmc_sampler.hpp (the replacement of the reference's
 * init_samplers / sample<SAMPLER>, nested_sampler.hpp:618-694): the same
 * arguments as model_init / model_sample (model_sampler.cpp:100-114), the
 * batched FP64 form's keying and results.  Staging it replaces the engine's
 * problem; nmlmc_sample and model_sample then refuse until a problem is
 * staged again. */
int nmlmc_sampler_model_init(const model_data* data, uint64_t seed);
int nmlmc_sampler_model_sample(uint32_t ell, uint64_t M, sample_output* out);

/* Mixed form diagnostics of the last call: outer samples re-evaluated in FP64,
 * indicator flips (FP32 decision differing from the FP64 one), and refinement
 * decisions rechecked in FP64 (their FP32 test statistic lay within its error
 * bound of the threshold).  Any pointer may be NULL. */
int nmlmc_mixed_stats(uint64_t* reevaluated, uint64_t* flips, uint64_t* rechecked);

/* Device time (ms) of the last sample call's kernels, from HIP events on the
 * library's stream, and the name of its dominant kernel. */
double nmlmc_last_kernel_ms(void);
const char* nmlmc_last_kernel_name(void);

/* First 16 hex digits of the sha256 of the library's sources (csrc/ and this
 * header, concatenated in path order) as built; loaders compare it with the
 * sources beside them to refuse a stale build.  No counterpart in the reference. */
const char* nmlmc_source_digest(void);

/* Lanes per outer sample chosen for the last call (autotuned or overridden). */
uint32_t nmlmc_last_group_lanes(void);

/* Launch-shape autotuner of the batched forms (no counterpart in the reference,
 * whose launch geometry is fixed): times the level call (ell, M) of the staged
 * problem in the current form over segments of 4, 8, 16, 32, 64 and 128 inner units
 * per lane and grids of 4, 8 and 16 blocks per CU, on outer samples set apart
 * for tuning (ids from 2^45), and keeps the fastest for later calls at level
 * ell.  Writes the choice and its kernel time (any pointer may be NULL).  The
 * choices are dropped when a problem is staged or the adaptive rule changes;
 * per-sample results never depend on them. */
int nmlmc_autotune_level(uint32_t ell, uint64_t M, uint32_t* segment, uint32_t* blocks_per_cu, double* ms);

/* ---- MLMC driver (mlmc.hpp:207-337) over the current problem -------------- */

/* Output of nmlmc_mlmc_run: per level sums as in mlmc_t::mc_diff_output_t
 * (mlmc.hpp:21-27), 4 moments. */
typedef struct nmlmc_level_stats {
    double sums_fine[4];
    double sums_diff[4];
    double work;
    double time;
    uint64_t M;
} nmlmc_level_stats;

/* Runs mlmc_t<double,4>::run with the library's sample call as sample_level.
 * levels_out has room for max_levels entries; *n_levels receives the count.
 * summary_out[5] = {starting_level, bias_est, stat_error, final TOL, wall seconds}. */
int nmlmc_mlmc_run(double targetTOL, uint64_t M0, double theta, double Ca,
                   double starting_level_coeff, int verbose, uint32_t max_levels,
                   nmlmc_level_stats* levels_out, uint32_t* n_levels, double* summary_out);

/* Same driver over a caller-supplied level sampler (mlmc_input_t::sample_level,
 * mlmc.hpp:224): fn fills *out for M fresh samples of level ell and returns 0, or
 * a non-zero code that aborts the run. */
/* The level table of the last completed nmlmc_mlmc_run / nmlmc_mlmc_run_cb:
 * mlmc_output_t::to_string (mlmc.hpp:106-183), the text the reference prints
 * after each iteration in verbose mode (and that parse_mlmc_output.py reads).
 * The string stays valid until the next run. */
const char* nmlmc_mlmc_last_table(void);

typedef int (*nmlmc_sample_level_fn)(void* user, uint32_t ell, uint64_t M, nmlmc_level_stats* out);
int nmlmc_mlmc_run_cb(nmlmc_sample_level_fn fn, void* user, double targetTOL, uint64_t M0,
                      double theta, double Ca, double starting_level_coeff, int verbose,
                      uint32_t max_levels, nmlmc_level_stats* levels_out, uint32_t* n_levels,
                      double* summary_out);

/* The MLMC drivers of the staged batch (nmlmc_batch_*) run together (no
 * counterpart in the reference): instance i runs mlmc_t::run to targetTOL[i]
 * with the other parameters given, in a host thread of its own; each step's
 * level requests of all running instances are sampled by one batch level call
 * per level.  Per instance: its levels at levels_out[i * max_levels ...], the
 * level count at n_levels[i], the summary at summary_out[i * 5 ...] as
 * nmlmc_mlmc_run gives them, with the whole batch's wall time.  n is the
 * length of targetTOL, n_levels and summary_out (levels_out: n * max_levels)
 * and must equal the batch's instance count (NMLMC_ERR_ARG otherwise). */
int nmlmc_batch_mlmc_run(uint32_t n, const double* targetTOL, uint64_t M0, double theta, double Ca,
                         double starting_level_coeff, uint32_t max_levels, nmlmc_level_stats* levels_out,
                         uint32_t* n_levels, double* summary_out);

#ifdef __cplusplus
}
#endif
#endif /* NMLMC_H */
