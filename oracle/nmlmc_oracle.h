/* CPU restatement of the reference's level-l nested sampler (TEST INFRASTRUCTURE).
 *
 * This is the oracle: a plain-C restatement of the reference algorithm used only
 * by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg, as the
 * checker.  The product library never links or calls it.
 *
 * Compiled twice by oracle/Makefile: with -DREAL_DOUBLE (real_t = double, the
 * "plain FP64" form) and without (real_t = float, the reference default).
 * Sums and work are double in both, as in the reference (common.hpp:11-18).
 */
#ifndef NMLMC_ORACLE_H
#define NMLMC_ORACLE_H
#include <stdint.h>

#ifdef REAL_DOUBLE
typedef double real_t;
#else
typedef float real_t;
#endif

#ifdef __cplusplus
extern "C" {
#endif

/* Engine parameters: nested_sampler_params (nested_sampler.hpp:14-34) plus the
 * adaptive-rule extension used by the capped 3-sigma configuration. */
typedef struct {
    uint32_t N0;
    uint32_t beta;
    int32_t antithetic;
    uint32_t fine_moments;
    uint32_t diff_moments;
    double adaptive_r;
    uint32_t adapt_rule;   /* 0: reference rule; 1: doubling while |mean| < 3 stderr, capped */
    uint32_t cap_log2;     /* rule 1: N_max = N_base << cap_log2 */
} orc_params;

enum { ORC_MODEL_REF = 0, ORC_MODEL_SHIFT = 1, ORC_PORTFOLIO = 2 };
enum { ORC_KEY_PARITY = 0, ORC_KEY_BATCHED = 1 };

typedef struct orc_ctx orc_ctx;

orc_ctx* orc_new(uint32_t cores, const orc_params* p);
void orc_free(orc_ctx* c);
int orc_real_bytes(void);

/* model::data (model_sampler.cpp:21-39).  variant: ORC_MODEL_REF or ORC_MODEL_SHIFT. */
void orc_model_init(orc_ctx* c, int variant, int phi_type, double tau, double max_loss,
                    uint64_t seed);

/* portfolio: flat arrays, options method-major (BS, Exact, Unbiased) in staged order. */
void orc_portfolio_init(orc_ctx* c, uint32_t Q, const double* S0, const double* mu,
                        const double* sigma, const uint32_t* counts, const double* strike,
                        const double* maturity, const double* start, const double* w,
                        const double* delta_est, const uint32_t* stock_ind, const int* type,
                        const double* prob, const double* W_est, double r, double rho,
                        double max_loss, double tau, uint32_t dt_h0inv, uint32_t dt_beta,
                        double dt_prob_expo, int phi_type, uint32_t phi_size, int subsample,
                        int use_delta_cv, int use_coarse_delta_cv, int use_antithetic_risk,
                        int use_milstein, const double* chol, uint64_t seed);

/* One level call.  keying = ORC_KEY_PARITY (reference OpenMP semantics with
 * `cores` virtual threads, persistent per-thread streams) or ORC_KEY_BATCHED
 * (per-outer-sample keys {seed, (ell<<46) | (outer_offset+m)}; outer draws from
 * counters {j, 0}; inner unit u of draw group `tag` from counters
 * {j, (tag << 40) | u}).
 * Optional per-outer-sample trace (indexed by m): phi_f[m], phi_c[m], nf[m], nc[m]. */
void orc_sample(orc_ctx* c, int keying, uint32_t ell, uint64_t M, uint64_t outer_offset,
                double* fsums, double* diffsums, double* work, uint64_t* stats,
                double* tr_phi_f, double* tr_phi_c, uint64_t* tr_nf, uint64_t* tr_nc);

/* Level-coupled paths for the staged portfolio's method-2 options (the
 * batched-risk workload's Asians; needs sub-sampling mode 1 and method-2
 * options, else off): at level l such an option's inner sample runs
 * dt_h0inv 2^floor(l/2) Milstein steps, fine and (antithetic finals of l > 0)
 * coarse values from the same increments. */
void orc_set_coupled(orc_ctx* c, int on);

/* Portfolio outer draws of outer samples [offset, offset + M) at level ell
 * under the batched keying: R_out[m * Q + i] = R_tau[i]; const_out[m] (may be
 * NULL) = the sampler's constant. */
void orc_outer_states(orc_ctx* c, uint32_t ell, uint64_t M, uint64_t outer_offset, double* R_out,
                      double* const_out);

/* Building blocks exposed for known-answer tests. */
void orc_threefry2x64(const uint64_t* ctr, const uint64_t* key, uint64_t* out);
double orc_u01_f64(uint64_t u);
double orc_uneg11_f64(uint64_t u);
float orc_u01_f32(uint64_t u);
void orc_boxmuller(uint64_t u0, uint64_t u1, double* out);
void orc_stream(uint64_t seed, uint64_t core, const int* kinds, int n, uint64_t* out_u,
                double* out_r);
uint32_t orc_rand_int_cdf(double u, const double* cdf, uint32_t n);
uint64_t orc_rand_int_pdf(double u, const double* pdf, uint32_t n);
uint64_t orc_uint_pow(uint64_t x, uint64_t n);
double orc_norm_cdf_inv(double p);
double orc_norm_cdf(double x);
double orc_bs_price(int type, double S0, double K, double r, double t, double sig);
double orc_bs_delta(int type, double S0, double K, double r, double t, double sig);
double orc_simulate_stock(double S0, double mu, double sig, double t, double dW);
double orc_simulate_stock_dt(int with_coarse, const double* normals, double* S, uint64_t N,
                             double mu, double sig, double t, uint32_t size, int milstein,
                             uint32_t coarse_step);
/* The same with A: path averages after each step (the Asian extension). */
double orc_simulate_stock_dt_avg(int with_coarse, const double* normals, double* S, double* A, uint64_t N,
                                 double mu, double sig, double t, uint32_t size, int milstein,
                                 uint32_t coarse_step);
uint64_t orc_samples_count(orc_ctx* c, uint32_t ell, int which);  /* 0 base, 1 max */

/* portfolio_sample_delta_unbiased (portfolio_sampler.cpp:758-781) with the
 * plain keying: sums[0..moments) of the delta samples of Unbiased option
 * option_index and their work. */
void orc_sample_delta_unbiased(orc_ctx* c, uint32_t option_index, uint64_t M, uint32_t moments,
                               double* sums, double* work);

#ifdef __cplusplus
}
#endif
#endif
