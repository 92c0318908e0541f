/* CPU restatement of the reference's level-l nested sampler.
 *
 * TEST INFRASTRUCTURE ONLY.  This oracle is imported, linked or executed solely
 * by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg, as the
 * checker.  The product library (nested-risk-mlmc_amd/) never calls it.
 *
 * Every function cites the reference file:line it restates.  Floating-point
 * expressions keep the reference's operand types and evaluation order
 * (real_t vs double promotions included) so that, compiled with plain gcc on
 * x86-64 (no FMA contraction), results are bit-identical to the reference's
 * OpenMP build on the same inputs.  Pinning (see tests/test_oracle_*.py):
 *   - threefry2x64_20: Random123 published known-answer vectors + the compiled
 *     reference (oracle/_ref);
 *   - u01/uneg11/boxmuller/streams: closed forms + compiled reference;
 *   - rand_int_cdf/pdf: brute-force linear scans + compiled reference;
 *   - samplers/engine: per-sample and per-call goldens from the compiled
 *     reference (tests/golden/), closed-form answers, invariants;
 *   - extensions absent from the reference (model shift variant, capped 3-sigma
 *     adaptive rule, Cholesky outer draw, digital options, batched keying):
 *     closed forms and reductions to the reference paths (documented per test).
 */
#define _GNU_SOURCE
#include "nmlmc_oracle.h"
#include <math.h>
#include <omp.h>
#include <stdlib.h>
#include <string.h>

#ifdef REAL_DOUBLE
#define RLOG log
#define RSQRT sqrt
#define REXP exp
#define RPOW pow
#define RCOS cos
#define RSIN sin
#define RFLOOR floor
#define RCEIL ceil
#else
#define RLOG logf
#define RSQRT sqrtf
#define REXP expf
#define RPOW powf
#define RCOS cosf
#define RSIN sinf
#define RFLOOR floorf
#define RCEIL ceilf
#endif

typedef double sum_t;
typedef double work_t;
typedef real_t prob_t;

#define MAX_PHI 4
#define MAX_MOM 4
#define NSTATS 6

int orc_real_bytes(void) { return (int)sizeof(real_t); }

/* ---------------------------------------------------------------------------
 * Threefry2x64-20 (Random123/threefry.h:190-293; rotation constants 79-93,
 * key-schedule parity 171, 20 rounds 178-180), restated from the published
 * Threefish/Threefry algorithm (Salmon et al., SC'11).
 * ------------------------------------------------------------------------- */
static inline uint64_t rotl64(uint64_t x, unsigned n) {
    return (x << (n & 63)) | (x >> ((64 - n) & 63));
}

void orc_threefry2x64(const uint64_t* ctr, const uint64_t* key, uint64_t* out) {
    static const unsigned R[8] = {16, 42, 12, 31, 16, 32, 24, 21};
    uint64_t ks[3];
    ks[0] = key[0];
    ks[1] = key[1];
    ks[2] = 0x1BD11BDAA9FC1A22ULL ^ key[0] ^ key[1];
    uint64_t x0 = ctr[0] + ks[0], x1 = ctr[1] + ks[1];
    for (int r = 0; r < 20; r++) {
        x0 += x1;
        x1 = rotl64(x1, R[r & 7]);
        x1 ^= x0;
        if ((r & 3) == 3) {
            int s = (r >> 2) + 1;
            x0 += ks[s % 3];
            x1 += ks[(s + 1) % 3] + (uint64_t)s;
        }
    }
    out[0] = x0;
    out[1] = x1;
}

/* r123::u01 / uneg11 (Random123/uniform.hpp:144-180): W=64. */
double orc_u01_f64(uint64_t u) { return (double)u * 0x1p-64 + 0x1p-65; }
double orc_uneg11_f64(uint64_t u) { return (double)(int64_t)u * 0x1p-63 + 0x1p-64; }
float orc_u01_f32(uint64_t u) { return (float)u * 0x1p-64f + 0x1p-65f; }
static inline real_t u01_real(uint64_t u) {
#ifdef REAL_DOUBLE
    return orc_u01_f64(u);
#else
    return orc_u01_f32(u);
#endif
}

/* r123::boxmuller(uint64,uint64) (Random123/boxmuller.hpp:128-137) with the
 * sincospi fallback sincos(PI*x) (boxmuller.hpp:104-107). */
void orc_boxmuller(uint64_t u0, uint64_t u1, double* out) {
    const double PI = 3.1415926535897932;
    double s, c;
    sincos(PI * orc_uneg11_f64(u0), &s, &c);
    double r = sqrt(-2. * log(orc_u01_f64(u1)));
    out[0] = s * r;
    out[1] = c * r;
}

/* Zeroed memory on 128-byte boundaries in whole 128-byte lines: arrays of
 * per-thread records written by OpenMP threads must not share cache lines. */
static void* orc_calloc_lines(size_t n, size_t size) {
    size_t bytes = (n ? n : 1) * size;
    bytes = (bytes + 127) & ~(size_t)127;
    void* p = aligned_alloc(128, bytes);
    if (p) memset(p, 0, bytes);
    return p;
}
static size_t orc_line_reals(size_t n) { return (n + 15) & ~(size_t)15; }

/* random_state (par_omp.hpp:72-121).  rand() hands out cur[1] then cur[0]
 * (par_omp.hpp:92-99); ctr.incr() is a carrying little-endian increment
 * (Random123/array.h:133-163).  randn() calls boxmuller(rand(), rand()): the
 * reference built with g++ evaluates the second argument first, so the first
 * u64 consumed is u1 and the second is u0 (pinned by the compiled reference). */
typedef struct {
    uint64_t key[2], ctr[2], cur[2];
    int cached;
    int is_saved;
    real_t saved;
} orc_rs;

static void rs_seed(orc_rs* s, uint64_t k0, uint64_t k1) {
    s->is_saved = 0;
    s->cached = 0;
    s->key[0] = k0;
    s->key[1] = k1;
    s->ctr[0] = s->ctr[1] = 0;
    s->cur[0] = s->cur[1] = 0;
    s->saved = 0;
}

static uint64_t rs_rand(orc_rs* s) {
    if (s->cached <= 0) {
        orc_threefry2x64(s->ctr, s->key, s->cur);
        if (++s->ctr[0] == 0) ++s->ctr[1];
        s->cached = 2;
    }
    return s->cur[--s->cached];
}

static real_t rs_randn(orc_rs* s) {
    if (s->is_saved) {
        s->is_saved = 0;
        return s->saved;
    }
    uint64_t first = rs_rand(s);
    uint64_t second = rs_rand(s);
    double d[2];
    orc_boxmuller(second, first, d);
    s->is_saved = 1;
    s->saved = (real_t)d[1];
    return (real_t)d[0];
}

static real_t rs_randu(orc_rs* s) { return u01_real(rs_rand(s)); }

void orc_stream(uint64_t seed, uint64_t core, const int* kinds, int n, uint64_t* out_u,
                double* out_r) {
    orc_rs s;
    rs_seed(&s, seed, core);
    for (int i = 0; i < n; i++) {
        out_u[i] = 0;
        out_r[i] = 0;
        if (kinds[i] == 0) out_u[i] = rs_rand(&s);
        else if (kinds[i] == 1) out_r[i] = rs_randu(&s);
        else out_r[i] = rs_randn(&s);
    }
}

/* ---------------------------------------------------------------------------
 * Utilities (common.hpp:146-203, par_omp.hpp:52-60, common.cpp:53-86)
 * ------------------------------------------------------------------------- */
static inline uint64_t clog2_64(uint64_t x) { return 64 - __builtin_clzl(x - 1); }
static inline uint32_t clog2_32(uint32_t x) { return 32 - __builtin_clz(x - 1); }

/* uint_pow (common.hpp:146-157).  For x == 1 the reference's power-of-two branch
 * evaluates clog2(1) = 64 - clz(0), which is undefined; 1^n = 1 is returned. */
uint64_t orc_uint_pow(uint64_t x, uint64_t n) {
    if (n == 0 || x == 1) return 1;
    if ((x & (x - 1)) == 0) return (uint64_t)1 << (n * clog2_64(x));
    uint64_t r;
    for (r = 1; n--; r *= x) {}
    return r;
}

/* rand_int_cdf (common.hpp:160-176): fixed-step branch-free binary search over
 * an exclusive-prefix CDF. */
static uint32_t rand_int_cdf_r(real_t u, const real_t* cdf, uint32_t n) {
    uint32_t pos[2] = {0, n};
    if (pos[1] == 1) return 0;
    uint32_t steps = clog2_32(pos[1]);
    for (uint32_t i = 0; i < steps; i++) {
        uint32_t mid = (pos[0] + pos[1]) / 2;
        pos[mid > 0 && u < cdf[mid]] = mid;
    }
    return pos[0];
}

/* rand_int_pdf (common.hpp:178-195). */
static uint64_t rand_int_pdf_r(real_t u, const real_t* pdf, uint64_t n) {
    real_t p = 0;
    uint64_t ret = 0;
    for (uint64_t i = 0; i < n; i++) {
        real_t c = pdf[i];
        if (c <= 0) continue;
        p += c;
        ret = i;
        if (u <= p) break;
    }
    return ret;
}

uint32_t orc_rand_int_cdf(double u, const double* cdf, uint32_t n) {
    real_t* c = (real_t*)malloc(sizeof(real_t) * n);
    for (uint32_t i = 0; i < n; i++) c[i] = (real_t)cdf[i];
    uint32_t r = rand_int_cdf_r((real_t)u, c, n);
    free(c);
    return r;
}

uint64_t orc_rand_int_pdf(double u, const double* pdf, uint32_t n) {
    real_t* c = (real_t*)malloc(sizeof(real_t) * n);
    for (uint32_t i = 0; i < n; i++) c[i] = (real_t)pdf[i];
    uint64_t r = rand_int_pdf_r((real_t)u, c, n);
    free(c);
    return r;
}

static const real_t SQRT_2 = (real_t)1.414213562373095;
static const real_t SQRT_2PI = (real_t)2.506628274631;

/* norm_cdf (common.hpp:201-202): the double overload uses erfc, the float one
 * erfcf; both return 0.5*erfc(..) computed in double, narrowed to real_t. */
static inline real_t norm_cdf_r(real_t x) {
#ifdef REAL_DOUBLE
    return 0.5 * erfc(-x / SQRT_2);
#else
    return (real_t)(0.5 * erfcf(-x / SQRT_2));
#endif
}
double orc_norm_cdf(double x) { return norm_cdf_r((real_t)x); }
static inline real_t norm_pdf_r(real_t x) { return (real_t)(exp(-0.5 * (x * x)) / SQRT_2PI); }

/* norm_cdf_inv, Acklam's rational approximation with float-literal
 * coefficients (common.cpp:53-86). */
static real_t norm_cdf_inv_r(real_t p) {
    const real_t a[] = {-39.696830f, 220.946098f, -275.928510f, 138.357751f, -30.664798f, 2.506628f};
    const real_t b[] = {-54.476098f, 161.585836f, -155.698979f, 66.801311f, -13.280681f};
    const real_t c[] = {-0.007784894002f, -0.32239645f, -2.400758f, -2.549732f, 4.374664f, 2.938163f};
    const real_t d[] = {0.007784695709f, 0.32246712f, 2.445134f, 3.754408f};
    real_t plow = 0.02425f;
    real_t phigh = 1 - plow;
    if (p < plow) {
        real_t q = RSQRT(-2 * RLOG(p));
        return (((((c[0] * q + c[1]) * q + c[2]) * q + c[3]) * q + c[4]) * q + c[5]) /
               ((((d[0] * q + d[1]) * q + d[2]) * q + d[3]) * q + 1);
    }
    if (phigh < p) {
        real_t q = RSQRT(-2 * RLOG(1 - p));
        return -(((((c[0] * q + c[1]) * q + c[2]) * q + c[3]) * q + c[4]) * q + c[5]) /
               ((((d[0] * q + d[1]) * q + d[2]) * q + d[3]) * q + 1);
    }
    {
        real_t q = p - 0.5f;
        real_t r = q * q;
        return (((((a[0] * r + a[1]) * r + a[2]) * r + a[3]) * r + a[4]) * r + a[5]) * q /
               (((((b[0] * r + b[1]) * r + b[2]) * r + b[3]) * r + b[4]) * r + 1);
    }
}
double orc_norm_cdf_inv(double p) { return norm_cdf_inv_r((real_t)p); }

/* ---------------------------------------------------------------------------
 * Portfolio pricing helpers (portfolio_sampler.cpp:136-225).  Option type
 * codes: 1 call, -1 put (reference); 2 / -2 cash-or-nothing digital call/put
 * paying 1 (extension, absent from the reference).
 * ------------------------------------------------------------------------- */
typedef struct {
    real_t strike, maturity, start, w, delta_est;
    uint32_t stock_ind;
    int type;
    prob_t prob;
    real_t W_est;
} orc_option;

typedef struct {
    real_t S0, mu, sigma;
} orc_stock;

/* simulate_stock (136-138): exponent evaluated in double (0.5 is a double). */
static inline real_t simulate_stock_r(real_t S0, real_t mu, real_t sigma, real_t t, real_t dW) {
    return (real_t)(S0 * exp((double)(sigma * dW) + ((double)mu - 0.5 * sigma * sigma) * t));
}
double orc_simulate_stock(double S0, double mu, double sig, double t, double dW) {
    return simulate_stock_r((real_t)S0, (real_t)mu, (real_t)sig, (real_t)t, (real_t)dW);
}

/* calc_option_blackscholes (140-150); digitals: e^{-rt} N(+-d2). */
static real_t bs_price_r(int type, real_t S0, real_t strike, real_t r, real_t t, real_t sigma) {
    real_t v1 = RLOG(S0 / strike);
    real_t v2 = (real_t)((r + (sigma * sigma) / 2.0) * t);
    real_t v3 = sigma * RSQRT(t);
    real_t d1 = (v1 + v2) / v3;
    real_t d2 = d1 - sigma * RSQRT(t);
    if (type == 2 || type == -2) {
        real_t mod = (real_t)(type / 2);
        return REXP(-r * t) * norm_cdf_r(mod * d2);
    }
    real_t mod = (real_t)type;
    return mod * S0 * norm_cdf_r(mod * d1) - mod * strike * REXP(-r * t) * norm_cdf_r(mod * d2);
}

/* calc_delta_blackscholes (152-160); digitals: +-e^{-rt} n(d2) / (S0 sigma sqrt t). */
static real_t bs_delta_r(int type, real_t S0, real_t strike, real_t r, real_t t, real_t sigma) {
    real_t v1 = RLOG(S0 / strike);
    real_t v2 = (real_t)((r + (sigma * sigma) / 2.0) * t);
    real_t v3 = sigma * RSQRT(t);
    real_t d1 = (v1 + v2) / v3;
    if (type == 2 || type == -2) {
        real_t d2 = d1 - sigma * RSQRT(t);
        real_t mod = (real_t)(type / 2);
        return mod * REXP(-r * t) * norm_pdf_r(d2) / (S0 * sigma * RSQRT(t));
    }
    return norm_cdf_r(d1) - (real_t)(type == -1);
}

double orc_bs_price(int type, double S0, double K, double r, double t, double sig) {
    return bs_price_r(type, (real_t)S0, (real_t)K, (real_t)r, (real_t)t, (real_t)sig);
}
double orc_bs_delta(int type, double S0, double K, double r, double t, double sig) {
    return bs_delta_r(type, (real_t)S0, (real_t)K, (real_t)r, (real_t)t, (real_t)sig);
}

/* option_diff_from_blackscholes (186-200). */
static inline real_t option_diff_from_bs(const orc_option* o, real_t R_tau, real_t r,
                                         real_t sigma, real_t tau) {
    real_t start = o->start;
    real_t finish = REXP(-r * tau) * bs_price_r(o->type, R_tau, o->strike, r, o->maturity - tau, sigma);
    return o->w * (start - finish);
}

/* option_derv_from_bs (202-210). */
static inline real_t option_derv_from_bs(const orc_option* o, real_t S0, real_t r, real_t sigma) {
    return o->w * bs_delta_r(o->type, S0, o->strike, r, o->maturity, sigma);
}

/* Terminal payoff; MAX(a, 0) as in common.hpp:135-137.  Types 3 / -3
 * (arithmetic-average Asian call / put, an extension) take the path's average
 * in place of S_T. */
static inline int is_asian(int type) { return type == 3 || type == -3; }
static inline real_t payoff_r(int type, real_t S_T, real_t strike) {
    if (type == 2) return (real_t)(S_T > strike);
    if (type == -2) return (real_t)(S_T < strike);
    real_t mod = (real_t)(is_asian(type) ? type / 3 : type);
    real_t v = mod * (S_T - strike);
    return v > 0 ? v : 0;
}

/* option_diff_from_stock (212-219). */
static inline real_t option_diff_from_stock(const orc_option* o, real_t S0_T, real_t St_T, real_t r) {
    real_t start = payoff_r(o->type, S0_T, o->strike);
    real_t finish = payoff_r(o->type, St_T, o->strike);
    return o->w * REXP(-r * o->maturity) * (start - finish);
}

/* option_derv_from_stock (221-225): pathwise delta; zero for digitals. */
static inline real_t option_derv_from_stock(const orc_option* o, real_t St_T, real_t S0, real_t r) {
    if (o->type == 2 || o->type == -2 || is_asian(o->type)) return 0;
    real_t df = (real_t)((o->type == 1) - (St_T < o->strike));
    return o->w * (St_T / S0) * REXP(-r * o->maturity) * df;
}

/* simulate_stock_dt (gbm_approx.hpp:18-57) with scalar mu/sigma and a
 * real_t-valued normal source.  With A (the Asian extension), A[i] becomes the
 * average of path i's values after each of its steps (fine paths: N steps;
 * coarse paths: N / coarse_step), summed in step order. */
typedef real_t (*randn_fn)(void* ctx);

static real_t simulate_stock_dt_r(int with_coarse, randn_fn rn, void* rctx, sum_t* S, sum_t* A, uint64_t N,
                                  real_t mu, real_t sigma, real_t t, uint32_t size, int milstein,
                                  uint32_t coarse_step) {
    real_t W = 0;
    real_t dt = t / N;
    real_t sqrt_dt = RSQRT(dt);
    real_t dW_c = 0;
    real_t dt_c = t / (N / coarse_step);
    for (uint64_t n = 0; n < N; n++) {
        real_t dW = sqrt_dt * rn(rctx);
        W += dW;
        for (uint32_t i = 0; i < size; i++) {
            real_t m = milstein ? (real_t)(0.5 * (sigma * sigma) * (dW * dW - dt)) : 0;
            S[i] += S[i] * (mu * dt + sigma * dW + m);
            if (A) A[i] += S[i];
        }
        if (with_coarse) {
            dW_c += dW;
            if (((n + 1) % coarse_step) == 0) {
                for (uint32_t i = 0; i < size; i++) {
                    real_t m = milstein ? (real_t)(0.5 * (sigma * sigma) * (dW_c * dW_c - dt_c)) : 0;
                    S[i + size] += S[i + size] * (mu * dt_c + sigma * dW_c + m);
                    if (A) A[i + size] += S[i + size];
                }
                dW_c = 0;
            }
        }
    }
    if (A) {
        for (uint32_t i = 0; i < size; i++) A[i] /= (sum_t)N;
        if (with_coarse)
            for (uint32_t i = 0; i < size; i++) A[i + size] /= (sum_t)(N / coarse_step);
    }
    return W;
}

typedef struct {
    const double* p;
} fixed_normals;
static real_t fixed_randn(void* ctx) {
    fixed_normals* f = (fixed_normals*)ctx;
    return (real_t)(*f->p++);
}

double orc_simulate_stock_dt(int with_coarse, const double* normals, double* S, uint64_t N,
                             double mu, double sig, double t, uint32_t size, int milstein,
                             uint32_t coarse_step) {
    fixed_normals f = {normals};
    return simulate_stock_dt_r(with_coarse, fixed_randn, &f, S, NULL, N, (real_t)mu, (real_t)sig,
                               (real_t)t, size, milstein, coarse_step);
}

double orc_simulate_stock_dt_avg(int with_coarse, const double* normals, double* S, double* A, uint64_t N,
                                 double mu, double sig, double t, uint32_t size, int milstein,
                                 uint32_t coarse_step) {
    fixed_normals f = {normals};
    sum_t SS[8], AA[8] = {0};
    for (uint32_t i = 0; i < 2 * size && i < 8; i++) SS[i] = (sum_t)S[i];
    real_t W = simulate_stock_dt_r(with_coarse, fixed_randn, &f, SS, AA, N, (real_t)mu, (real_t)sig,
                                   (real_t)t, size, milstein, coarse_step);
    for (uint32_t i = 0; i < 2 * size && i < 8; i++) {
        S[i] = (double)SS[i];
        A[i] = (double)AA[i];
    }
    return W;
}

/* ---------------------------------------------------------------------------
 * Problem data and samplers
 * ------------------------------------------------------------------------- */
typedef struct {
    orc_option* options[3];
    uint32_t nopt[3];
    real_t* cdfs[3];
    orc_stock* stocks;
    uint32_t Q;
    real_t r, rho, s1_rho2, max_loss, tau, stau;
    uint32_t dt_h0inv, dt_beta;
    real_t dt_prob_expo;
    int subsample;
    int use_delta_cv, use_coarse_delta_cv, use_antithetic_risk, use_milstein;
    int coupled; /* extension: method-2 options on level-coupled paths */
    prob_t compute_prob[3];
    real_t* delta_coalesce;
    uint32_t options_count;
    prob_t prob_norm, total_std;
    real_t* chol; /* Q x Q row-major lower triangle, or NULL (one-factor) */
} orc_portfolio;

/* inner_sampler_base (nested_sampler.hpp:36-85): two streams plus draw
 * counters; the streams are pointers so the batched keying can alias them. */
typedef struct {
    orc_rs own_inner, own_sync;
    orc_rs *inner, *sync;
    uint64_t randn_count, rand_count;
    real_t Y;
    real_t *W_tau, *R_tau;
    real_t constant;
    /* Level-coupled paths (extension): the current draw group's steps, whether
     * it is a two-valued final (and the fine-to-coarse step ratio), and the
     * coarse value of the last inner sample (its fine value otherwise). */
    uint64_t cpl_steps;
    int cpl_two;
    uint32_t cpl_cs;
    real_t last_vc;
} __attribute__((aligned(128))) orc_sampler; /* one cache-line pair each: no false sharing between threads */

struct orc_ctx {
    uint32_t cores;
    orc_params p;
    int problem;
    int phi_type;
    uint32_t phi_size;
    uint64_t seed;
    real_t stau, s1tau, max_loss; /* model */
    orc_portfolio pf;
    orc_sampler* samplers;
    real_t* workspace;
    orc_sampler* workers; /* batched keying: one sampler per OpenMP worker */
    real_t* worker_ws;
    int nworkers;
    uint32_t ws_reals;
};

static inline real_t s_randn(orc_sampler* s, orc_rs* st) {
    s->randn_count++;
    return rs_randn(st);
}
static inline real_t s_randu(orc_sampler* s, orc_rs* st) {
    s->rand_count++;
    return rs_randu(st);
}

/* ---- model::sampler (model_sampler.cpp:42-92) and the shift variant ---- */
static void model_sample_outer(orc_ctx* c, orc_sampler* s) { s->Y = s_randn(s, s->sync); }

static real_t model_sample_inner(orc_ctx* c, orc_sampler* s, work_t* total_work) {
    *total_work += 1;
    if (c->problem == ORC_MODEL_SHIFT) {
        /* Shift variant: X = Y + Z, one normal per inner sample. */
        real_t Z = s_randn(s, s->inner);
        return s->Y + Z;
    }
    real_t X = s_randn(s, s->inner);
    real_t Y2 = s_randn(s, s->inner);
    real_t a = c->stau * s->Y + c->s1tau * X;
    real_t r = a * a;
    real_t b = c->stau * Y2 + c->s1tau * X;
    r -= b * b;
    return r;
}

/* ---- portfolio::sampler (portfolio_sampler.cpp:227-615) ---- */
static inline real_t delta_factor(const orc_portfolio* pf, const orc_sampler* s, uint32_t i) {
    return s->R_tau[i] - pf->stocks[i].S0;
}

static real_t compute_delta_cv(const orc_portfolio* pf, const orc_sampler* s) {
    if (!pf->use_delta_cv) return 0;
    real_t sum = 0;
    for (uint32_t i = 0; i < pf->Q; i++) sum += delta_factor(pf, s, i) * pf->delta_coalesce[i];
    return sum;
}

/* sample_outer (293-311); with a Cholesky factor (extension):
 * W_i = stau * sum_{j<=i} L_ij z_j with z drawn in order 0..Q-1. */
static void portfolio_sample_outer(orc_ctx* c, orc_sampler* s) {
    const orc_portfolio* pf = &c->pf;
    if (pf->chol) {
        real_t z[512];
        for (uint32_t i = 0; i < pf->Q; i++) z[i] = s_randn(s, s->sync);
        for (uint32_t i = 0; i < pf->Q; i++) {
            real_t acc = 0;
            for (uint32_t j = 0; j <= i; j++) acc += pf->chol[(size_t)i * pf->Q + j] * z[j];
            s->W_tau[i] = pf->stau * acc;
            s->R_tau[i] = simulate_stock_r(pf->stocks[i].S0, pf->stocks[i].mu, pf->stocks[i].sigma,
                                           pf->tau, s->W_tau[i]);
        }
    } else {
        real_t dW_sys = pf->stau * s_randn(s, s->sync);
        if (pf->Q == 1) {
            s->W_tau[0] = dW_sys;
            s->R_tau[0] = simulate_stock_r(pf->stocks[0].S0, pf->stocks[0].mu, pf->stocks[0].sigma,
                                           pf->tau, s->W_tau[0]);
        } else {
            for (uint32_t i = 0; i < pf->Q; i++) {
                real_t dW_ind = pf->stau * s_randn(s, s->sync);
                s->W_tau[i] = pf->rho * dW_sys + pf->s1_rho2 * dW_ind;
                s->R_tau[i] = simulate_stock_r(pf->stocks[i].S0, pf->stocks[i].mu,
                                               pf->stocks[i].sigma, pf->tau, s->W_tau[i]);
            }
        }
    }
    s->constant = -pf->max_loss - compute_delta_cv(pf, s);
}

/* sample_inner_bs (453-468). */
static real_t sample_inner_bs(orc_ctx* c, orc_sampler* s, const orc_option* o, work_t* tw) {
    const orc_portfolio* pf = &c->pf;
    *tw += o->W_est;
    uint32_t si = o->stock_ind;
    const orc_stock* st = &pf->stocks[si];
    real_t val = option_diff_from_bs(o, s->R_tau[si], pf->r, st->sigma, pf->tau);
    real_t cv = pf->use_delta_cv ? (delta_factor(pf, s, si) * option_derv_from_bs(o, st->S0, pf->r, st->sigma)) : 0;
    return val + cv;
}

/* sample_inner_exact (470-519). */
static real_t sample_inner_exact(orc_ctx* c, orc_sampler* s, const orc_option* o, work_t* tw) {
    const orc_portfolio* pf = &c->pf;
    int anti = pf->use_antithetic_risk;
    real_t S0_T_1_i, S0_T_2_i = 0, St_T_i;
    *tw += o->W_est;
    uint32_t si = o->stock_ind;
    const orc_stock* st = &pf->stocks[si];
    real_t dW_0_tau = pf->stau * s_randn(s, s->inner);
    real_t dW_tau_T = RSQRT(o->maturity - pf->tau) * s_randn(s, s->inner);
    real_t S0_t_1 = simulate_stock_r(st->S0, pf->r, st->sigma, pf->tau, dW_0_tau);
    S0_T_1_i = simulate_stock_r(S0_t_1, pf->r, st->sigma, o->maturity - pf->tau, dW_tau_T);
    if (anti) {
        real_t S0_t_2 = simulate_stock_r(st->S0, pf->r, st->sigma, pf->tau, -dW_0_tau);
        S0_T_2_i = simulate_stock_r(S0_t_2, pf->r, st->sigma, o->maturity - pf->tau, dW_tau_T);
    }
    St_T_i = simulate_stock_r(s->R_tau[si], pf->r, st->sigma, o->maturity - pf->tau, dW_tau_T);
    real_t val1 = option_diff_from_stock(o, S0_T_1_i, St_T_i, pf->r);
    real_t val2 = val1;
    if (anti) val2 = option_diff_from_stock(o, S0_T_2_i, St_T_i, pf->r);
    real_t dv1 = 0., dv2 = 0;
    if (pf->use_delta_cv) {
        real_t diff = delta_factor(pf, s, si);
        dv1 = diff * option_derv_from_stock(o, S0_T_1_i, st->S0, pf->r);
        dv2 = dv1;
        if (anti) dv2 = diff * option_derv_from_stock(o, S0_T_2_i, st->S0, pf->r);
    }
    return (real_t)(0.5 * (val1 + val2 + dv1 + dv2));
}

typedef struct {
    orc_sampler* s;
} sampler_randn_ctx;
static real_t sampler_inner_randn(void* ctx) {
    orc_sampler* s = ((sampler_randn_ctx*)ctx)->s;
    return s_randn(s, s->inner);
}

/* sample_inner_approx (539-604). */
static real_t sample_inner_approx(orc_ctx* c, orc_sampler* s, const orc_option* o, work_t* tw,
                                  uint32_t L) {
    const orc_portfolio* pf = &c->pf;
    uint32_t si = o->stock_ind;
    const orc_stock* st = &pf->stocks[si];
    real_t dW_0_tau = pf->stau * s_randn(s, s->inner);
    real_t S0_t_1_i = simulate_stock_r(st->S0, pf->r, st->sigma, pf->tau, dW_0_tau);
    real_t S0_t_2_i = simulate_stock_r(st->S0, pf->r, st->sigma, pf->tau, -dW_0_tau);
    uint64_t N_l = orc_uint_pow((uint64_t)pf->dt_beta, (uint64_t)L);
    uint64_t N = pf->dt_h0inv * N_l;
    *tw += (real_t)(N_l * o->W_est);
    sum_t S[6] = {S0_t_1_i, s->R_tau[si], S0_t_2_i, S0_t_1_i, s->R_tau[si], S0_t_2_i};
    sum_t A[6] = {0, 0, 0, 0, 0, 0};
    sum_t* A_p = is_asian(o->type) ? A : NULL;
    sampler_randn_ctx rc = {s};
    if (L) {
        simulate_stock_dt_r(1, sampler_inner_randn, &rc, S, A_p, N, pf->r, st->sigma,
                            o->maturity - pf->tau, 3, pf->use_milstein, pf->dt_beta);
        *tw += (real_t)((N_l / pf->dt_beta) * o->W_est);
    } else {
        simulate_stock_dt_r(0, sampler_inner_randn, &rc, S, A_p, N, pf->r, st->sigma,
                            o->maturity - pf->tau, 3, pf->use_milstein, 2);
    }
    /* An Asian option's payoffs take the paths' averages over (tau, T]. */
    const sum_t* V = A_p ? A : S;
    real_t val_f = 0, val_c = 0, val_d_f = 0, val_d_c = 0;
    const int cc = 3;
    const real_t m = pf->use_antithetic_risk ? 0.5 : 1.;
    const real_t diff = delta_factor(pf, s, si);
    const int do_d_f = pf->use_delta_cv && (pf->use_coarse_delta_cv || L == 0);
    const int do_d_c = pf->use_delta_cv && pf->use_coarse_delta_cv;
    for (int i = 0; i < (pf->use_antithetic_risk ? 2 : 1); i++) {
        int ic = 2 * i;
        val_f += m * option_diff_from_stock(o, (real_t)V[ic + 0], (real_t)V[1], pf->r);
        if (do_d_f) val_d_f += m * option_derv_from_stock(o, (real_t)S[ic + 0], st->S0, pf->r);
        if (L == 0) continue;
        val_c += m * option_diff_from_stock(o, (real_t)V[ic + cc + 0], (real_t)V[cc + 1], pf->r);
        if (do_d_c) val_d_c += m * option_derv_from_stock(o, (real_t)S[ic + cc + 0], st->S0, pf->r);
    }
    return ((val_f - val_c) + diff * (val_d_f - val_d_c));
}

/* sample_inner_unbiased (521-537). */
static real_t sample_inner_unbiased(orc_ctx* c, orc_sampler* s, const orc_option* o, work_t* tw) {
    const orc_portfolio* pf = &c->pf;
    prob_t e = (prob_t)pf->dt_prob_expo;
    prob_t x = (prob_t)pf->dt_beta;
    uint32_t L = (uint32_t)RFLOOR(RLOG(s_randu(s, s->sync)) / (RLOG(x) * e));
    real_t ret = sample_inner_approx(c, s, o, tw, L);
    prob_t L_prob_1 = (prob_t)(RPOW(x, -e * L) / (1. - RPOW(x, e)));
    return (real_t)((prob_t)ret * L_prob_1);
}

/* Level-coupled inner sample of a method-2 option (extension; the batched-risk
 * workload's Asians): sample_inner_approx's paths (539-604) with the step count
 * of the draw group's level, dt_h0inv 2^floor(l/2), instead of a drawn level:
 * the S0 path and its antithetic twin exactly to tau (one normal), then
 * s->cpl_steps Milstein steps of the three paths over (tau, T]; in a two-valued
 * final also the coarse paths of cpl_steps / cpl_cs steps on the same
 * increments.  Fine value returned, coarse value (the fine one when there are
 * no coarse paths) left in s->last_vc; each value is the antithetic mean of
 * the option differences plus the delta control variate on its own paths. */
static real_t sample_inner_coupled(orc_ctx* c, orc_sampler* s, const orc_option* o, work_t* tw) {
    const orc_portfolio* pf = &c->pf;
    uint32_t si = o->stock_ind;
    const orc_stock* st = &pf->stocks[si];
    real_t dW_0_tau = pf->stau * s_randn(s, s->inner);
    real_t S0_t_1_i = simulate_stock_r(st->S0, pf->r, st->sigma, pf->tau, dW_0_tau);
    real_t S0_t_2_i = simulate_stock_r(st->S0, pf->r, st->sigma, pf->tau, -dW_0_tau);
    const uint64_t N = s->cpl_steps;
    const int coarse = s->cpl_two && s->cpl_cs > 1;
    *tw += (real_t)((N / pf->dt_h0inv) * o->W_est);
    if (coarse) *tw += (real_t)((N / s->cpl_cs / pf->dt_h0inv) * o->W_est);
    sum_t S[6] = {S0_t_1_i, s->R_tau[si], S0_t_2_i, S0_t_1_i, s->R_tau[si], S0_t_2_i};
    sum_t A[6] = {0, 0, 0, 0, 0, 0};
    sum_t* A_p = is_asian(o->type) ? A : NULL;
    sampler_randn_ctx rc = {s};
    simulate_stock_dt_r(coarse, sampler_inner_randn, &rc, S, A_p, N, pf->r, st->sigma, o->maturity - pf->tau, 3,
                        pf->use_milstein, coarse ? s->cpl_cs : 2);
    const sum_t* V = A_p ? A : S;
    real_t val_f = 0, val_c = 0, val_d_f = 0, val_d_c = 0;
    const real_t m = pf->use_antithetic_risk ? 0.5 : 1.;
    const real_t diff = delta_factor(pf, s, si);
    for (int i = 0; i < (pf->use_antithetic_risk ? 2 : 1); i++) {
        int ic = 2 * i;
        val_f += m * option_diff_from_stock(o, (real_t)V[ic], (real_t)V[1], pf->r);
        if (pf->use_delta_cv) val_d_f += m * option_derv_from_stock(o, (real_t)S[ic], st->S0, pf->r);
        if (!coarse) continue;
        val_c += m * option_diff_from_stock(o, (real_t)V[ic + 3], (real_t)V[4], pf->r);
        if (pf->use_delta_cv) val_d_c += m * option_derv_from_stock(o, (real_t)S[ic + 3], st->S0, pf->r);
    }
    real_t vf = val_f + diff * val_d_f;
    s->last_vc = coarse ? val_c + diff * val_d_c : vf;
    return vf;
}

/* sampler::sample_delta_unbiased (portfolio_sampler.cpp:397-444): pathwise
 * delta of Unbiased option `o` from an antithetic pair of paths, the Milstein
 * paths (dt_h0inv steps, sigma and -sigma) against the exact GBM paths driven by
 * the same Brownian increment. */
static real_t sample_delta_unbiased(orc_ctx* c, orc_sampler* s, const orc_option* o, work_t* tw) {
    const orc_portfolio* pf = &c->pf;
    const orc_stock* st = &pf->stocks[o->stock_ind];
    const real_t S0 = st->S0, sigma = st->sigma, r = pf->r;
    const real_t T = o->maturity - pf->tau;
    const uint64_t N = pf->dt_h0inv;
    real_t dW_0_tau = pf->stau * s_randn(s, s->inner);
    real_t S1 = simulate_stock_r(S0, r, sigma, pf->tau, dW_0_tau);
    real_t S2 = simulate_stock_r(S0, r, sigma, pf->tau, -dW_0_tau);
    sum_t ST_f[2] = {S1, S2};
    const real_t sig[2] = {sigma, -sigma};
    real_t W = 0;
    const real_t dt = T / N;
    const real_t sqrt_dt = RSQRT(dt);
    for (uint64_t n = 0; n < N; n++) {
        real_t dW = sqrt_dt * s_randn(s, s->inner);
        W += dW;
        for (int i = 0; i < 2; i++) {
            real_t m = pf->use_milstein ? (real_t)(0.5 * (sig[i] * sig[i]) * (dW * dW - dt)) : 0;
            ST_f[i] += ST_f[i] * (r * dt + sig[i] * dW + m);
        }
    }
    real_t ST0 = simulate_stock_r(S1, r, sigma, T, W);
    real_t ST1 = simulate_stock_r(S2, r, sigma, T, -W);
    *tw += (real_t)(2 * N) * o->W_est;
    real_t val = 0;
    val += option_derv_from_stock(o, (real_t)ST_f[0], S0, r);
    val += option_derv_from_stock(o, (real_t)ST_f[1], S0, r);
    val -= option_derv_from_stock(o, ST0, S0, r);
    val -= option_derv_from_stock(o, ST1, S0, r);
    return (real_t)(val / (2.0 * o->w));
}

/* sample_inner_option (382-396). */
static real_t sample_inner_option(orc_ctx* c, orc_sampler* s, work_t* tw, int method,
                                  const orc_option* o) {
    switch (method) {
    case 0: return sample_inner_bs(c, s, o, tw);
    case 1: return sample_inner_exact(c, s, o, tw);
    case 2: return c->pf.coupled ? sample_inner_coupled(c, s, o, tw) : sample_inner_unbiased(c, s, o, tw);
    default: abort();
    }
}

/* sample_inner_random_option(_ind) (346-380). */
static real_t sample_inner_random_option(orc_ctx* c, orc_sampler* s, work_t* tw) {
    const orc_portfolio* pf = &c->pf;
    prob_t prob_norm = pf->prob_norm;
    real_t uc = s_randu(s, s->sync);
    uint64_t mi = rand_int_pdf_r(uc * prob_norm, pf->compute_prob, 3);
    uint64_t idx;
    prob_t prob;
    if (pf->subsample == -1) {
        real_t u = s_randu(s, s->inner);
        idx = (uint64_t)(RCEIL(u * (real_t)pf->nopt[mi]) - 1);
        prob = (prob_t)(1. / pf->options_count);
    } else {
        real_t u = s_randu(s, s->inner) * pf->compute_prob[mi];
        idx = rand_int_cdf_r(u, pf->cdfs[mi], pf->nopt[mi]);
        prob = pf->options[mi][idx].prob / prob_norm;
    }
    orc_option o = pf->options[mi][idx];
    s->last_vc = NAN;
    real_t v = (real_t)((prob_t)sample_inner_option(c, s, tw, (int)mi, &o) / prob);
    /* Only a coupled sample has a coarse value of its own. */
    s->last_vc = (mi == 2 && pf->coupled) ? (real_t)((prob_t)s->last_vc / prob) : v;
    return v;
}

/* sample_inner_all_options (336-344). */
static real_t sample_inner_all_options(orc_ctx* c, orc_sampler* s, work_t* tw) {
    sum_t ret = 0;
    for (int i = 1; i < 3; i++)
        for (uint32_t j = 0; j < c->pf.nopt[i]; j++)
            ret += sample_inner_option(c, s, tw, i, &c->pf.options[i][j]);
    return (real_t)ret;
}

/* ---- sampler dispatch (the SAMPLER concept, nested_sampler.hpp) ---- */
static void sampler_outer(orc_ctx* c, orc_sampler* s) {
    if (c->problem == ORC_PORTFOLIO) portfolio_sample_outer(c, s);
    else model_sample_outer(c, s);
}

/* init_inner_sampler: model 64; portfolio 312-324. */
static real_t sampler_init_inner(orc_ctx* c, orc_sampler* s, work_t* tw) {
    if (c->problem != ORC_PORTFOLIO) return -c->max_loss;
    if (c->pf.subsample == 0) {
        sum_t det = 0;
        for (uint32_t j = 0; j < c->pf.nopt[0]; j++)
            det += sample_inner_option(c, s, tw, 0, &c->pf.options[0][j]);
        return (real_t)det + s->constant;
    }
    return s->constant;
}

static real_t sampler_inner(orc_ctx* c, orc_sampler* s, work_t* tw) {
    if (c->problem != ORC_PORTFOLIO) return model_sample_inner(c, s, tw);
    if (c->pf.subsample == 0) return sample_inner_all_options(c, s, tw);
    return sample_inner_random_option(c, s, tw);
}

/* add_phi: model 81-91, portfolio 261-291. */
static void sampler_add_phi(const orc_ctx* c, real_t u, real_t coeff, real_t* phi) {
    switch (c->phi_type) {
    case 1: {
        real_t m = u > 0 ? u : 0;
        *phi += coeff * m;
        break;
    }
    case 2: {
        real_t p = 1;
        for (uint32_t i = 0; i < c->phi_size; i++) {
            p *= u;
            phi[i] += coeff * p;
        }
        break;
    }
    case 3: {
        real_t tmp = (real_t)(M_PI * u);
        for (uint32_t k = 0; k < c->phi_size / 2; k++) {
            phi[2 * k] += coeff * RCOS((k + 1) * tmp);
            phi[2 * k + 1] += coeff * RSIN((k + 1) * tmp);
        }
        if (c->phi_size % 2) {
            uint32_t k = c->phi_size / 2;
            phi[2 * k] += coeff * RCOS((k + 1) * tmp);
        }
        break;
    }
    default:
        *phi += coeff * (real_t)(u >= 0.f);
    }
}

/* ---------------------------------------------------------------------------
 * Engine (nested_sampler.hpp:87-485)
 * ------------------------------------------------------------------------- */
typedef struct {
    int keying;
    orc_sampler* s;
    uint64_t key[2]; /* batched: the outer sample's key */
    int spu;         /* batched: inner samples per keyed unit (2 for the shift model) */
    uint32_t ell;    /* the level of the call */
} draw_env;

#define TAG_FINE_TEST 1
#define TAG_COARSE_TEST 33
#define TAG_FINAL 65
#define TAG_FINAL_COARSE 66

/* inner_sampler::sample_inner (131-161) over samples [first, first+count) of
 * one draw group.  Parity keying: the thread's persistent streams in order.
 * Batched keying: inner sample i draws from the keyed unit stream whose
 * blocks are ctr = {j, (tag << 40) | (i / spu)}, j = 0, 1, ..., key = the outer
 * sample's key, both the sync-role and inner-role draws of that sample coming
 * from that stream. */
static void inner_loop(orc_ctx* c, draw_env* e, uint32_t tag, uint64_t first, uint64_t count,
                       sum_t* out_s1, sum_t* out_s2, work_t* out_tw, sum_t* out_s1c) {
    sum_t s1 = 0, s2 = 0, s1c = 0;
    work_t tw = 0.;
    orc_sampler* s = e->s;
    /* Level-coupled paths: coarse tests and the plain coarse final take the
     * level below's steps; the antithetic final of a level > 0 is two-valued. */
    if (c->problem == ORC_PORTFOLIO && c->pf.coupled) {
        const int below = (tag >= TAG_COARSE_TEST && tag < TAG_FINAL) || tag == TAG_FINAL_COARSE;
        const uint32_t lv = below ? (e->ell ? e->ell - 1 : 0) : e->ell;
        s->cpl_steps = (uint64_t)c->pf.dt_h0inv << (lv / 2);
        s->cpl_two = tag == TAG_FINAL && e->ell > 0 && c->p.antithetic;
        s->cpl_cs = s->cpl_two ? (uint32_t)(s->cpl_steps / ((uint64_t)c->pf.dt_h0inv << ((e->ell - 1) / 2))) : 1;
    }
    if (e->keying == ORC_KEY_PARITY) {
        for (uint64_t i = 0; i < count; i++) {
            s->last_vc = NAN;
            real_t d = sampler_inner(c, s, &tw);
            s1 += d;
            if (out_s2) s2 += d * d;
            if (out_s1c) s1c += (c->problem == ORC_PORTFOLIO && c->pf.coupled) ? s->last_vc : d;
        }
    } else {
        orc_rs unit;
        orc_rs *save_in = s->inner, *save_sync = s->sync;
        s->inner = s->sync = &unit;
        for (uint64_t ii = 0; ii < count; ii++) {
            uint64_t i = first + ii;
            if (ii == 0 || i % e->spu == 0) {
                rs_seed(&unit, e->key[0], e->key[1]);
                unit.ctr[0] = 0;
                unit.ctr[1] = ((uint64_t)tag << 40) | (i / e->spu);
                if (i % e->spu) (void)rs_randn(&unit); /* skip the pair's first normal */
            }
            s->last_vc = NAN;
            real_t d = sampler_inner(c, s, &tw);
            s1 += d;
            if (out_s2) s2 += d * d;
            if (out_s1c) s1c += (c->problem == ORC_PORTFOLIO && c->pf.coupled) ? s->last_vc : d;
        }
        s->inner = save_in;
        s->sync = save_sync;
    }
    *out_s1 += s1;
    if (out_s2) *out_s2 += s2;
    if (out_s1c) *out_s1c += s1c;
    *out_tw += tw;
}

/* sample_inner_sum (105-111). */
static void sample_inner_sum(orc_ctx* c, draw_env* e, uint32_t tag, uint64_t first, uint64_t count,
                             real_t* start, sum_t* out_s1, work_t* tw, sum_t* out_s1c) {
    *start = sampler_init_inner(c, e->s, tw);
    inner_loop(c, e, tag, first, count, out_s1, NULL, tw, out_s1c);
}

/* sample_inner_mean (112-119). */
static real_t sample_inner_mean(orc_ctx* c, draw_env* e, uint32_t tag, uint64_t count, work_t* tw) {
    real_t start = sampler_init_inner(c, e->s, tw);
    sum_t s1 = 0;
    inner_loop(c, e, tag, 0, count, &s1, NULL, tw, NULL);
    return (real_t)(start + s1 / count);
}

/* sample_inner_delta2 (120-129). */
static real_t sample_inner_delta2(orc_ctx* c, draw_env* e, uint32_t tag, uint64_t count, work_t* tw) {
    real_t start = sampler_init_inner(c, e->s, tw);
    sum_t s1 = 0, s2 = 0;
    inner_loop(c, e, tag, 0, count, &s1, &s2, tw, NULL);
    real_t out_var = (real_t)(s2 / (count - 1) - (s1 * s1) / (double)(count * (count - 1)));
    real_t dist = start + (real_t)(s1 / count);
    return dist * dist / out_var;
}

/* sample_phi (164-169). */
static void sample_phi(orc_ctx* c, draw_env* e, uint32_t tag, uint64_t count, work_t* tw, real_t* out) {
    for (uint32_t i = 0; i < c->phi_size; i++) out[i] = 0;
    sampler_add_phi(c, sample_inner_mean(c, e, tag, count, tw), 1.f, out);
}

/* sample_phi_antithetic (171-216). */
static uint64_t sample_phi_antithetic(orc_ctx* c, draw_env* e, uint64_t count_f, uint64_t count_c,
                                      work_t* tw, real_t* out_f, real_t* out_c) {
    for (uint32_t i = 0; i < c->phi_size; i++) out_f[i] = out_c[i] = 0.f;
    uint64_t count_big, count_small;
    real_t *out_big, *out_small;
    if (count_f > count_c) {
        out_big = out_f;
        out_small = out_c;
        count_big = count_f;
        count_small = count_c;
    } else {
        out_big = out_c;
        out_small = out_f;
        count_big = count_c;
        count_small = count_f;
    }
    uint64_t groups_small = count_big / count_small;
    real_t coeff_small = 1.f / groups_small;
    real_t start_small = 0;
    sum_t s1_big = 0;
    /* Two-valued finals (level-coupled paths): the fine estimator takes the
     * fine values, the coarse one the coarse values; the big side is the fine
     * one when count_f > count_c.  Otherwise both sums are the same. */
    const int fine_big = count_f > count_c;
    for (uint64_t i = 0; i < groups_small; i++) {
        sum_t s1_f = 0, s1_c = 0;
        sample_inner_sum(c, e, TAG_FINAL, i * count_small, count_small, &start_small, &s1_f, tw, &s1_c);
        sum_t s1_small = fine_big ? s1_c : s1_f;
        real_t mean_small = (real_t)(start_small + s1_small / count_small);
        sampler_add_phi(c, mean_small, coeff_small, out_small);
        s1_big += fine_big ? s1_f : s1_c;
    }
    real_t mean_big = (real_t)(start_small + s1_big / count_big);
    sampler_add_phi(c, mean_big, 1.f, out_big);
    return count_big;
}

/* get_samples_count (219-274); rule 1 (extension) starts at N0*beta^l, caps at
 * N_base << cap_log2 and doubles while delta2 < 9/N (|mean| < 3 stderr). */
typedef struct {
    real_t tmp;
    uint64_t base_count, max_count;
    int adaptive;
} samples_count_t;

static void sc_init(const orc_ctx* c, uint32_t ell, samples_count_t* sc) {
    const uint64_t N0 = c->p.N0;
    if (c->p.adapt_rule == 1) {
        sc->base_count = N0 * orc_uint_pow((uint64_t)c->p.beta, (uint64_t)ell);
        sc->max_count = sc->base_count << c->p.cap_log2;
        sc->tmp = 9;
        sc->adaptive = 1;
        return;
    }
    sc->max_count = N0 * orc_uint_pow((uint64_t)c->p.beta, (uint64_t)ell);
    real_t r = (real_t)c->p.adaptive_r;
    sc->adaptive = r > 1 && r < 2;
    if (sc->adaptive) {
        sc->base_count = N0 * orc_uint_pow((uint64_t)(c->p.beta >> 1), (uint64_t)ell);
        const real_t c_conf2 = 9.;
        sc->tmp = sc->max_count * RPOW(c_conf2 / sc->max_count, r / 2);
    } else {
        sc->base_count = sc->max_count;
        sc->tmp = 0;
    }
}

static uint64_t sc_call(orc_ctx* c, const samples_count_t* sc, draw_env* e, uint32_t tag0,
                        work_t* tw, uint64_t* stats, int with_stats) {
    uint64_t cur = sc->base_count;
    const real_t r = (real_t)c->p.adaptive_r;
    uint32_t round = 0;
    while (1) {
        if (sc->max_count <= 2 * cur) {
            cur = sc->max_count;
            break;
        }
        stats[0] += cur;
        real_t delta2 = sample_inner_delta2(c, e, tag0 + round, cur, tw);
        round++;
        real_t new_count;
        if (c->p.adapt_rule == 1) new_count = sc->tmp / delta2;
        else new_count = RPOW(delta2, -r / 2) * sc->tmp;
        if (new_count <= (real_t)cur) break;
        cur *= 2;
    }
    if (with_stats) {
        stats[1] += 1;
        stats[2] += (cur == sc->max_count);
        stats[3] += (cur == sc->base_count);
    }
    return cur;
}

uint64_t orc_samples_count(orc_ctx* c, uint32_t ell, int which) {
    samples_count_t sc;
    sc_init(c, ell, &sc);
    return which ? sc.max_count : sc.base_count;
}

/* add_moments (87-100). */
static void add_moments(const orc_ctx* c, const real_t* phi, sum_t* out, uint32_t moments) {
    for (uint32_t i = 0; i < c->phi_size; i++) {
        real_t p = 1.f;
        for (uint32_t j = 0; j < moments; j++) {
            p *= phi[i];
            *(out++) += p;
        }
    }
}

/* One outer sample: the body of kernel_sample's m-loop (419-451). */
static void process_outer(orc_ctx* c, draw_env* e, uint32_t ell, int fine_only,
                          const samples_count_t* sc_f, const samples_count_t* sc_c,
                          sum_t* f_sums, sum_t* diff_sums, work_t* tw, uint64_t* stats,
                          real_t* f_phi, real_t* c_phi, uint64_t* nf_out, uint64_t* nc_out) {
    sampler_outer(c, e->s);
    uint64_t nf = sc_call(c, sc_f, e, TAG_FINE_TEST, tw, stats, 1), nc = 0;
    if (fine_only) {
        stats[0] += nf;
        sample_phi(c, e, TAG_FINAL, nf, tw, f_phi);
    } else {
        nc = sc_call(c, sc_c, e, TAG_COARSE_TEST, tw, stats, 0);
        if (c->p.antithetic) {
            stats[0] += sample_phi_antithetic(c, e, nf, nc, tw, f_phi, c_phi);
        } else {
            stats[0] += nf + nc;
            sample_phi(c, e, TAG_FINAL, nf, tw, f_phi);
            sample_phi(c, e, TAG_FINAL_COARSE, nc, tw, c_phi);
        }
    }
    *nf_out = nf;
    *nc_out = nc;
    add_moments(c, f_phi, f_sums, c->p.fine_moments);
    if (!fine_only) {
        real_t d[MAX_PHI];
        for (uint32_t i = 0; i < c->phi_size; i++) d[i] = f_phi[i] - c_phi[i];
        add_moments(c, d, diff_sums, c->p.diff_moments);
    }
}

static void orc_ensure_workers(orc_ctx* c);

/* Per-worker accumulators of kernel_sample (nested_sampler.hpp:379-382). */
typedef struct {
    sum_t f[MAX_PHI * MAX_MOM], d[MAX_PHI * MAX_MOM];
    work_t tw;
    uint64_t st[NSTATS];
} __attribute__((aligned(128))) orc_acc;

/* Outer samples [s0, e0) on sampler s: the m-loop of kernel_sample. */
static void run_range(orc_ctx* c, orc_sampler* s, int keying, uint32_t ell, uint64_t s0, uint64_t e0,
                      uint64_t outer_offset, const samples_count_t* sc_f, const samples_count_t* sc_c,
                      orc_acc* a, double* tr_phi_f, double* tr_phi_c, uint64_t* tr_nf, uint64_t* tr_nc) {
    const int fine_only = (ell == 0);
    real_t f_phi[MAX_PHI], c_phi[MAX_PHI];
    s->randn_count = s->rand_count = 0;
    for (uint64_t m = s0; m < e0; m++) {
        draw_env e;
        e.keying = keying;
        e.s = s;
        e.spu = (c->problem == ORC_MODEL_SHIFT) ? 2 : 1;
        e.ell = ell;
        orc_rs outer;
        if (keying == ORC_KEY_BATCHED) {
            e.key[0] = c->seed;
            e.key[1] = ((uint64_t)ell << 46) | (outer_offset + m);
            rs_seed(&outer, e.key[0], e.key[1]);
            s->sync = &outer;
            s->inner = &outer;
        }
        for (uint32_t i = 0; i < c->phi_size; i++) c_phi[i] = 0;
        uint64_t nf, nc;
        process_outer(c, &e, ell, fine_only, sc_f, sc_c, a->f, a->d, &a->tw, a->st, f_phi, c_phi, &nf,
                      &nc);
        if (tr_phi_f) tr_phi_f[m] = f_phi[0];
        if (tr_phi_c) tr_phi_c[m] = fine_only ? 0 : c_phi[0];
        if (tr_nf) tr_nf[m] = nf;
        if (tr_nc) tr_nc[m] = nc;
        if (keying == ORC_KEY_BATCHED) {
            s->sync = &s->own_sync;
            s->inner = &s->own_inner;
        }
    }
    a->st[4] += s->rand_count;
    a->st[5] += s->randn_count;
}

#define ORC_BATCH_CHUNK 256

/* kernel_sample (369-485) + sample (646-694).  Parity keying restates the
 * OpenMP build: thread t takes the libgomp static-schedule block of [0, M)
 * and consumes its own persistent streams in order.  The virtual threads (or,
 * for the batched keying, fixed chunks of outer samples) run on OpenMP
 * workers; their accumulators are added in index order. */
void orc_sample(orc_ctx* c, int keying, uint32_t ell, uint64_t M, uint64_t outer_offset,
                double* fsums, double* diffsums, double* work, uint64_t* stats,
                double* tr_phi_f, double* tr_phi_c, uint64_t* tr_nf, uint64_t* tr_nc) {
    const int fine_only = (ell == 0);
    const uint32_t nf_sums = c->phi_size * c->p.fine_moments;
    const uint32_t nd_sums = c->phi_size * c->p.diff_moments;
    for (uint32_t i = 0; i < nf_sums; i++) fsums[i] = 0;
    for (uint32_t i = 0; i < nd_sums; i++) diffsums[i] = 0;
    *work = 0;
    for (int i = 0; i < NSTATS; i++) stats[i] = 0;
    samples_count_t sc_f, sc_c;
    sc_init(c, ell, &sc_f);
    sc_init(c, fine_only ? 0 : ell - 1, &sc_c);

    const uint64_t T = (keying == ORC_KEY_PARITY) ? c->cores : (M + ORC_BATCH_CHUNK - 1) / ORC_BATCH_CHUNK;
    orc_acc* acc = (orc_acc*)orc_calloc_lines(T, sizeof(orc_acc));
    if (keying == ORC_KEY_BATCHED) orc_ensure_workers(c);
#pragma omp parallel for schedule(dynamic, 1)
    for (uint64_t t = 0; t < T; t++) {
        uint64_t s0, e0;
        orc_sampler* s;
        if (keying == ORC_KEY_PARITY) {
            uint64_t q = M / T, rr = M % T;
            if (t < rr) {
                q++;
                rr = 0;
            }
            s0 = q * t + rr;
            e0 = s0 + q;
            s = &c->samplers[t];
        } else {
            s0 = t * ORC_BATCH_CHUNK;
            e0 = s0 + ORC_BATCH_CHUNK < M ? s0 + ORC_BATCH_CHUNK : M;
            s = &c->workers[omp_get_thread_num()];
        }
        run_range(c, s, keying, ell, s0, e0, outer_offset, &sc_f, &sc_c, &acc[t], tr_phi_f, tr_phi_c,
                  tr_nf, tr_nc);
    }
    for (uint64_t t = 0; t < T; t++) {
        for (uint32_t i = 0; i < nf_sums; i++) fsums[i] += acc[t].f[i];
        if (!fine_only)
            for (uint32_t i = 0; i < nd_sums; i++) diffsums[i] += acc[t].d[i];
        *work += acc[t].tw;
        for (int i = 0; i < NSTATS; i++) stats[i] += acc[t].st[i];
    }
    free(acc);
}

/* Level-coupled paths for the staged portfolio's method-2 options (extension;
 * sub-sampling mode 1). */
void orc_set_coupled(orc_ctx* c, int on) { c->pf.coupled = on && c->pf.subsample == 1 && c->pf.nopt[2] > 0; }

/* The outer draws of a level call under the batched keying, for tests of the
 * outer draw: R_out[m * Q + i] = R_tau[i] of outer sample offset + m (the
 * state sample_outer, portfolio_sampler.cpp:293-311, leaves behind), and the
 * constant (init_inner_sampler's offset before any option term) in
 * const_out[m] when const_out is not NULL. */
void orc_outer_states(orc_ctx* c, uint32_t ell, uint64_t M, uint64_t outer_offset, double* R_out,
                      double* const_out) {
    orc_ensure_workers(c);
    orc_sampler* s = &c->workers[0];
    const uint32_t Q = c->pf.Q;
    for (uint64_t m = 0; m < M; m++) {
        orc_rs outer;
        rs_seed(&outer, c->seed, ((uint64_t)ell << 46) | (outer_offset + m));
        s->sync = &outer;
        s->inner = &outer;
        sampler_outer(c, s);
        for (uint32_t i = 0; i < Q; i++) R_out[m * Q + i] = (double)s->R_tau[i];
        if (const_out) const_out[m] = (double)s->constant;
        s->sync = &s->own_sync;
        s->inner = &s->own_inner;
    }
}

/* kernel_sample_delta + portfolio_sample_delta_unbiased (portfolio_sampler.cpp:
 * 702-746, 758-781): T virtual threads, each its libgomp static share of M
 * delta samples of Unbiased option `option_index` from its persistent inner
 * stream; `moments` moments of the sample and the work are summed in thread order. */
void orc_sample_delta_unbiased(orc_ctx* c, uint32_t option_index, uint64_t M, uint32_t moments,
                               double* sums, double* work) {
    const uint32_t T = c->cores;
    orc_acc* acc = (orc_acc*)orc_calloc_lines(T, sizeof(orc_acc));
    const orc_option* o = &c->pf.options[2][option_index];
#pragma omp parallel for schedule(dynamic, 1)
    for (uint32_t t = 0; t < T; t++) {
        uint64_t q = M / T, rr = M % T;
        if (t < rr) {
            q++;
            rr = 0;
        }
        const uint64_t s0 = q * t + rr, e0 = s0 + q;
        orc_sampler* s = &c->samplers[t];
        for (uint64_t m = s0; m < e0; m++) {
            real_t d = sample_delta_unbiased(c, s, o, &acc[t].tw);
            real_t p = 1.f; /* add_moments(&delta, 1, f_sums, moments) */
            for (uint32_t j = 0; j < moments; j++) {
                p *= d;
                acc[t].f[j] += p;
            }
        }
    }
    for (uint32_t i = 0; i < moments; i++) sums[i] = 0;
    *work = 0;
    for (uint32_t t = 0; t < T; t++) {
        for (uint32_t i = 0; i < moments; i++) sums[i] += acc[t].f[i];
        *work += acc[t].tw;
    }
    free(acc);
}

/* ---------------------------------------------------------------------------
 * Setup (init_nestedmc_sampler common.cpp:20-26; init_samplers
 * nested_sampler.hpp:618-644: OpenMP seeds both streams with {seed, t})
 * ------------------------------------------------------------------------- */
orc_ctx* orc_new(uint32_t cores, const orc_params* p) {
    orc_ctx* c = (orc_ctx*)calloc(1, sizeof(orc_ctx));
    c->cores = cores;
    c->p = *p;
    c->phi_size = 1;
    c->samplers = (orc_sampler*)orc_calloc_lines(cores, sizeof(orc_sampler));
    return c;
}

static void free_portfolio(orc_portfolio* pf) {
    for (int m = 0; m < 3; m++) {
        free(pf->options[m]);
        free(pf->cdfs[m]);
        pf->options[m] = NULL;
        pf->cdfs[m] = NULL;
    }
    free(pf->stocks);
    free(pf->delta_coalesce);
    free(pf->chol);
    memset(pf, 0, sizeof(*pf));
}

void orc_free(orc_ctx* c) {
    if (!c) return;
    free_portfolio(&c->pf);
    free(c->samplers);
    free(c->workspace);
    free(c->workers);
    free(c->worker_ws);
    free(c);
}

/* Per-worker samplers of the batched keying (their streams are re-keyed per
 * outer sample; only the portfolio workspace matters). */
static void orc_ensure_workers(orc_ctx* c) {
    int n = omp_get_max_threads();
    if (c->workers && c->nworkers >= n) return;
    free(c->workers);
    free(c->worker_ws);
    c->workers = (orc_sampler*)orc_calloc_lines(n, sizeof(orc_sampler));
    c->worker_ws = c->ws_reals ? (real_t*)orc_calloc_lines(orc_line_reals(c->ws_reals) * n, sizeof(real_t)) : NULL;
    for (int i = 0; i < n; i++) {
        orc_sampler* s = &c->workers[i];
        s->sync = &s->own_sync;
        s->inner = &s->own_inner;
        if (c->ws_reals) {
            s->W_tau = c->worker_ws + (size_t)i * orc_line_reals(c->ws_reals);
            s->R_tau = s->W_tau + c->ws_reals / 2;
        }
    }
    c->nworkers = n;
}

static void seed_samplers(orc_ctx* c, uint64_t seed, uint32_t ws_reals) {
    c->seed = seed;
    c->ws_reals = ws_reals;
    free(c->workers);
    free(c->worker_ws);
    c->workers = NULL;
    c->worker_ws = NULL;
    c->nworkers = 0;
    free(c->workspace);
    c->workspace = ws_reals ? (real_t*)orc_calloc_lines(orc_line_reals(ws_reals) * c->cores, sizeof(real_t)) : NULL;
    for (uint32_t t = 0; t < c->cores; t++) {
        orc_sampler* s = &c->samplers[t];
        memset(s, 0, sizeof(*s));
        rs_seed(&s->own_sync, seed, t);
        rs_seed(&s->own_inner, seed, t);
        s->sync = &s->own_sync;
        s->inner = &s->own_inner;
        s->Y = NAN;
        if (ws_reals) {
            s->W_tau = c->workspace + (size_t)t * orc_line_reals(ws_reals);
            s->R_tau = s->W_tau + ws_reals / 2;
        }
    }
}

/* model_init / model_update_data (model_sampler.cpp:94-104); model::data ctor 24-29. */
void orc_model_init(orc_ctx* c, int variant, int phi_type, double tau, double max_loss,
                    uint64_t seed) {
    c->problem = variant;
    c->phi_type = phi_type;
    c->phi_size = 1;
    real_t t = (real_t)tau;
    c->stau = RSQRT(t);
    c->s1tau = RSQRT(1 - t);
    c->max_loss = (real_t)max_loss;
    seed_samplers(c, seed, 0);
}

/* portfolio_update_data (portfolio_sampler.cpp:631-678) + portfolio_init (680-686). */
void orc_portfolio_init(orc_ctx* c, uint32_t Q, const double* S0, const double* mu,
                        const double* sigma, const uint32_t* counts, const double* strike,
                        const double* maturity, const double* start, const double* w,
                        const double* delta_est, const uint32_t* stock_ind, const int* type,
                        const double* prob, const double* W_est, double r, double rho,
                        double max_loss, double tau, uint32_t dt_h0inv, uint32_t dt_beta,
                        double dt_prob_expo, int phi_type, uint32_t phi_size, int subsample,
                        int use_delta_cv, int use_coarse_delta_cv, int use_antithetic_risk,
                        int use_milstein, const double* chol, uint64_t seed) {
    orc_portfolio* pf = &c->pf;
    free_portfolio(pf);
    c->problem = ORC_PORTFOLIO;
    c->phi_type = phi_type;
    c->phi_size = (phi_type == 0 || phi_type == 1) ? 1 : phi_size;
    pf->Q = Q;
    pf->stocks = (orc_stock*)malloc(sizeof(orc_stock) * Q);
    for (uint32_t i = 0; i < Q; i++) {
        pf->stocks[i].S0 = (real_t)S0[i];
        pf->stocks[i].mu = (real_t)mu[i];
        pf->stocks[i].sigma = (real_t)sigma[i];
    }
    pf->r = (real_t)r;
    pf->rho = (real_t)rho;
    pf->s1_rho2 = RSQRT(1 - pf->rho * pf->rho);
    pf->max_loss = (real_t)max_loss;
    pf->tau = (real_t)tau;
    pf->stau = RSQRT(pf->tau);
    pf->dt_h0inv = dt_h0inv;
    pf->dt_beta = dt_beta;
    pf->dt_prob_expo = (real_t)dt_prob_expo;
    pf->subsample = subsample;
    pf->use_delta_cv = use_delta_cv;
    pf->use_coarse_delta_cv = use_coarse_delta_cv;
    pf->use_antithetic_risk = use_antithetic_risk;
    pf->use_milstein = use_milstein;
    if (chol) {
        pf->chol = (real_t*)malloc(sizeof(real_t) * Q * Q);
        for (uint32_t i = 0; i < Q * Q; i++) pf->chol[i] = (real_t)chol[i];
    }
    real_t* delta = (real_t*)calloc(Q, sizeof(real_t));
    pf->total_std = 0;
    pf->prob_norm = 0;
    pf->options_count = 0;
    uint32_t off = 0;
    for (int m = 0; m < 3; m++) {
        uint32_t n = counts[m];
        pf->nopt[m] = n;
        pf->options[m] = (orc_option*)malloc(sizeof(orc_option) * (n ? n : 1));
        pf->cdfs[m] = (real_t*)malloc(sizeof(real_t) * (n ? n : 1));
        real_t cc = 0;
        pf->compute_prob[m] = 0;
        for (uint32_t j = 0; j < n; j++, off++) {
            orc_option* o = &pf->options[m][j];
            o->strike = (real_t)strike[off];
            o->maturity = (real_t)maturity[off];
            o->start = (real_t)start[off];
            o->w = (real_t)w[off];
            o->delta_est = (real_t)delta_est[off];
            o->stock_ind = stock_ind[off];
            o->type = type[off];
            o->prob = (prob_t)prob[off];
            o->W_est = (real_t)W_est[off];
            pf->cdfs[m][j] = cc;
            cc += o->prob;
            delta[o->stock_ind] += o->w * o->delta_est;
            pf->prob_norm += o->prob;
            pf->total_std += o->prob * o->W_est;
            pf->compute_prob[m] += o->prob;
        }
        pf->options_count += n;
    }
    pf->delta_coalesce = delta;
    seed_samplers(c, seed, 2 * Q);
}
