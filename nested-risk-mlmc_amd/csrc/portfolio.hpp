// Portfolio problem (portfolio_sampler.cpp:16-615): device tables and the launch
// interface of its kernels.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "batched.hpp"
#include "engine.hpp"

namespace nmlmc {

// Option record in device memory: the reference's option_params with real_t =
// double (64 bytes), one aligned record per gather.
struct alignas(16) OptionRec {
    double strike, maturity, start, w, delta_est;
    uint32_t stock_ind;
    int32_t type;
    double prob, W_est;
};
static_assert(sizeof(OptionRec) == 64, "option record must stay 64 bytes");

struct StockRec {
    double S0, mu, sigma;
};

// data_ex (portfolio_sampler.cpp:94-132) with device pointers.
// One Exact option for the FP32 inner loop of the mixed form (16 bytes).
struct MixRec {
    float K;         // strike
    float sTt;       // sqrt(T - tau)
    float discp;     // w e^{-rT} / p, p = prob / prob_norm
    uint32_t stype;  // stock index << 8 | (type + 2)
};
static_assert(sizeof(MixRec) == 16, "mixed option record must stay 16 bytes");

// Per-stock constants of the mixed form's inner loop (32 bytes).
struct MixStock {
    float sig;     // sigma
    float dcoef;   // r - sigma^2 / 2
    float sigtau;  // sigma sqrt(tau)
    float S0e;     // S0 e^{(r - sigma^2 / 2) tau}
    float S0;      // S0
    float invS0;   // 1 / S0 (the control variate's S / S0 as a product)
    float pad[2];
};

constexpr uint32_t kGuideLo = 0x7fffffffu, kGuideMulti = 0x80000000u;
// Unbiased levels whose weight beta^(-e L) / (1 - beta^e) is tabulated.
constexpr uint32_t kUnbLevels = 32;

struct PortfolioDevice {
    const OptionRec* options[3];
    const double* cdfs[3];
    // Option-search guide per method: u of bucket b = (uint32)(u * guide_scale)
    // (clamped to nb - 1) has its answer max{i : cdf[i] <= u} in [lo(b), lo(b + 1)]
    // with lo = guides[m][.] & kGuideLo; bit 31 marks a bucket holding more than
    // one answer.  nb = 4n buckets with exact FP64 edges, nb + 1 entries.
    const uint32_t* guides[3];
    double guide_scale[3];
    uint32_t guide_nb[3];
    // Mixed form, Exact options (method 1) in staged order: FP32 constants of
    // the inner loss with the stock's parameters and the weight 1/p folded in.
    const MixRec* mix;
    const MixStock* mix_stocks;
    double w_est_const;  // the common W_est of all options, or -1 when they differ
    uint32_t nopt[3];
    const StockRec* stocks;
    const double* delta_coalesce;
    const double* chol;  // d x d row-major lower triangle, or nullptr
    uint32_t d;
    double r, rho, s1_rho2, max_loss, tau, stau;
    uint32_t dt_h0inv, dt_beta;
    double dt_prob_expo;
    // The Unbiased level law's constants, evaluated on the host as the
    // reference evaluates them (sample_inner_unbiased, 521-537):
    // ln(beta) e, and the weights beta^(-e L) / (1 - beta^e) of L < kUnbLevels.
    double unb_lnbe;
    double unb_weight[kUnbLevels];
    // Level-coupled paths (batched forms, the batched-risk workload's Asians):
    // the options of method 2 are priced on Milstein paths of
    // dt_h0inv 2^floor(l / 2) steps at level l, fine and coarse from the same
    // increments, instead of the Unbiased method's randomised level.
    uint32_t coupled;
    int phi_type;
    uint32_t phi_size;
    int subsample;
    int use_delta_cv, use_coarse_delta_cv, use_antithetic_risk, use_milstein;
    double compute_prob[3];
    uint32_t options_count;
    double prob_norm, total_std;
};

struct PortfolioParityHost {
    PortfolioDevice pf;
    EngineParams ep;
    LevelCounts cf, cc;
    uint64_t seed, M;
    uint32_t T, ell;
    uint32_t blocks;  // launched: one per virtual thread that holds outer samples
    double bs_work;
    StreamState* states;
    Partial* partials;
    double* tr_phi_f;
    double* tr_phi_c;
    uint64_t* tr_nf;
    uint64_t* tr_nc;
};
hipError_t launch_portfolio_parity(const PortfolioParityHost& H, hipStream_t s);

// kernel_sample_delta (portfolio_sampler.cpp:702-746): T virtual threads, one
// lane each, delta samples of one Unbiased option from the inner streams.
struct PortfolioDeltaHost {
    PortfolioDevice pf;
    uint64_t seed, M;
    uint32_t T, option_index, moments;
    StreamState* states;
    Partial* partials;  // one per virtual thread
};
hipError_t launch_portfolio_delta(const PortfolioDeltaHost& H, hipStream_t s);

// A batch of portfolio problems in one chain: the chain's outer samples come in
// segments, each of one instance (device tables; host-built per chunk).
struct PortfolioBatchTables {
    const PortfolioDevice* inst;  // [n_inst]
    const uint64_t* inst_seed;    // [n_inst]
    const double* inst_bs_work;   // [n_inst]
    const uint64_t* seg_item0;    // [n_seg + 1]: first item of each segment, then the chunk's item count
    const uint32_t* seg_inst;     // [n_seg]
    const uint64_t* seg_outer0;   // [n_seg]: outer id of the segment's first item
    const uint32_t* seg_blk0;     // [n_seg + 1]: first block of each segment in the outer-draw kernel
    uint32_t n_seg;
    uint32_t outer_blocks;        // blocks of the outer-draw kernel (0: no instance has a Cholesky factor)
    uint32_t dmax;                // largest d of the batch (row stride of R_tau)
};

struct PortfolioBatchedHost {
    BatchedCommon c;
    PortfolioDevice pf;
    double* R_tau;
    float* R32;       // FP32 copies of R_tau for the mixed Exact loop (else NULL)
    double bs_work;
    uint32_t blocks;
    bool mixed;
    bool exact_only;  // sub-sampling mode 1 and every option Exact (method 1)
    bool multi;       // a batch of problems (tables below; pf and bs_work unused)
    PortfolioBatchTables bt;
};
hipError_t launch_portfolio_batched(const PortfolioBatchedHost& H, hipStream_t s);

}  // namespace nmlmc
