// Host-side engine state: the one-problem-per-process globals of the reference
// (gc_sampler_data, gc_phi_size, <problem>::gc_data, g_p_nestedmc_data;
// common.cpp:15-17) gathered in one object.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <functional>
#include <map>
#include <vector>

#include "../../include/nmlmc.h"
#include "batched.hpp"
#include "engine.hpp"
#include "model.hpp"
#include "portfolio.hpp"

namespace nmlmc {

// PROBLEM_SAMPLER_CLASS: a sampler class of the reference's concept staged by
// init_samplers<S> (include/nmlmc_sampler.hpp), sampled only through sample<S>.
enum { PROBLEM_MODEL = 0, PROBLEM_PORTFOLIO = 1, PROBLEM_SAMPLER_CLASS = 2 };

struct TraceBufs {
    double* phi_f = nullptr;
    double* phi_c = nullptr;
    uint64_t* nf = nullptr;
    uint64_t* nc = nullptr;
};

struct Engine {
    bool inited = false;
    uint32_t cores = 0;
    EngineParams ep{};
    int form = NMLMC_FORM_PLAIN;
    int problem = -1;
    ModelParams mp{};
    uint64_t seed = 0;
    int device = 0;

    // Portfolio tables (device) and host mirrors.
    PortfolioDevice pf{};
    uint32_t pf_phi_size = 1;
    double pf_bs_work = 0.;        // work of one init_inner_sampler call
    std::vector<void*> pf_allocs;
    double* pf_R = nullptr;        // batched: outer states [M x d]
    float* pf_R32 = nullptr;       // their FP32 copies (same buffer, after the FP64 rows)
    size_t pf_R_cap = 0;

    StreamState* d_states = nullptr;  // [2 cores]: parity streams (sync, inner)
    Partial* d_partials = nullptr;    // [cores] or [batched groups]
    size_t partial_cap = 0;
    Partial* d_out = nullptr;         // the reduced record, then kReduceBlocks staging records
    Partial* h_out = nullptr;
    double* d_ws = nullptr;           // per-group workspace (portfolio R_tau)
    size_t ws_cap = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    // Scratch of a plain-form level call: the per-block partial records (room
    // for scr_cap) and the reduced record, followed by kReduceBlocks staging
    // records of the reduction.  NULL: the engine's own buffers,
    // grown on demand; the no-allocation entry point points them at its
    // static device array for the length of its call.
    Partial* scr_partials = nullptr;
    size_t scr_cap = 0;
    Partial* scr_result = nullptr;
    float last_ms = 0;
    const char* last_kernel = "";
    uint32_t last_G = 0;
    std::vector<uint64_t> next_outer;   // batched keying: next outer id per level
    uint64_t* d_queue = nullptr;
    // Batched engine buffers.
    ItemState* bat_items = nullptr;
    TestEntry* bat_tf[2] = {nullptr, nullptr};
    TestEntry* bat_tc[2] = {nullptr, nullptr};
    TestEntry* bat_rq[2] = {nullptr, nullptr};  // mixed: decisions left to the FP64 recheck
    FinalEntry* bat_fpool = nullptr;
    uint32_t* bat_reeval = nullptr;
    uint32_t* bat_pend = nullptr;
    double* bat_redo_acc = nullptr;
    uint32_t* bat_redo_cnt = nullptr;
    BatchedCtl* bat_ctl = nullptr;
    // The control block is clear between single-problem chains: run_batched
    // copies it out after each chain and then clears it, so the next chain
    // need not start with a clearing launch (which leaves the GPU idle while
    // the host enqueues the chain's next kernel).  True when that may not hold
    // (new buffer, an enqueue that failed, a batch chain, a capture not run):
    // the next chain then clears it first.
    bool ctl_dirty = true;
    BatchedCtl* bat_ctl_h = nullptr;
    uint64_t bat_cap = 0, bat_fcap = 0;
    uint64_t mixed_reeval = 0, mixed_flips = 0, mixed_rechecks = 0;

    int ensure_device();
    int stage_portfolio(const portfolio_data* data);
    int seed_samplers(uint64_t s);
    void free_samplers();
    void free_portfolio();
    void free_batched();
    uint32_t phi_size() const;
    void copy_out(uint32_t ell, sample_output* out);
    int ensure_partials(size_t n);
    int ensure_ws(size_t n);
    int trace_alloc(uint64_t M, bool want, TraceBufs& tb);
    int trace_finish(uint64_t M, TraceBufs& tb, double* tr_phi_f, double* tr_phi_c,
                     uint64_t* tr_nf, uint64_t* tr_nc);
    int finish_call(uint32_t ell, sample_output* out, const Partial* result);
    // The plain form's partial records for n blocks and the reduced record:
    // the scratch override when set (an error if n exceeds it), else the
    // engine's own buffers.
    int plain_buffers(size_t n, Partial*& partials, Partial*& result);
    int sample(uint32_t ell, uint64_t M, uint64_t outer_offset, uint32_t lanes_override,
               sample_output* out, double* tr_phi_f, double* tr_phi_c, uint64_t* tr_nf,
               uint64_t* tr_nc);
    int sample_model_parity(uint32_t ell, uint64_t M, uint32_t lanes_override, sample_output* out,
                            double* tr_phi_f, double* tr_phi_c, uint64_t* tr_nf, uint64_t* tr_nc);
    int sample_portfolio_parity(uint32_t ell, uint64_t M, uint32_t lanes_override,
                                sample_output* out, double* tr_phi_f, double* tr_phi_c,
                                uint64_t* tr_nf, uint64_t* tr_nc);
    int cus = 0;
    int device_cus();
    int sample_delta_unbiased(uint32_t option_index, uint64_t M, uint32_t moments, double* sums, double* work);
    uint64_t batched_chunk();
    static bool fused_enabled();
    // Launch shape of a batched level call: inner units per segment (S) and
    // blocks of 256 threads per CU.
    struct Shape {
        uint32_t S = 0, bpc = 0;
    };
    std::map<uint64_t, Shape> tuned;  // autotune_level's choices, key (form << 32) | level
    // Level-coupled paths for the method-2 options of portfolios staged from
    // now on (nmlmc_set_level_coupled_paths).
    bool level_coupled = false;
    Shape force{};                    // the candidate autotune_level is timing
    Shape batched_shape(uint32_t ell, uint64_t M, uint64_t spu, const BatchedCommon& c, uint32_t lanes_override);
    int autotune_level(uint32_t ell, uint64_t M, Shape& best, double& best_ms);
    // The candidate loop of the autotuners: run(ms) makes one level call at
    // the forced shape and reports its kernel time.
    int tune_shapes(uint32_t ell, const std::function<int(double&)>& run, Shape& best, double& best_ms);
    // Level sweeps (sample_levels): several batched level calls enqueued back
    // to back, each chain's control block copied to its own pinned host slot.
    bool sweep_on = false;                    // run_batched enqueues and returns
    uint32_t sweep_level = 0;                 // index of the level call being enqueued
    std::vector<uint32_t> sweep_chain_level;  // per enqueued chain: its level call
    BatchedCtl* sweep_slots = nullptr;        // pinned, one per chain
    uint32_t sweep_slot_cap = 0;
    hipGraphExec_t sweep_exec = nullptr;      // the last captured sweep, updated in place
    size_t sweep_exec_nodes = 0;              // its node count
    int sample_levels(uint32_t n, const uint32_t* ells, const uint64_t* Ms, sample_output* outs);
    int ensure_pf_R(uint64_t M);
    int ensure_R(size_t elems);

    // A batch of portfolio problems sampled in one launch chain per level call
    // (nmlmc_batch_*): each instance keeps its tables, seed and outer ids.
    struct BatchInst {
        PortfolioDevice pf{};
        std::vector<void*> allocs;
        double bs_work = 0.;
        uint64_t seed = 0;
        bool staged = false, exact_only = false;
        std::vector<uint64_t> next_outer;
    };
    std::vector<BatchInst> batch;
    bool batch_dirty = true;                 // device copies of the instance records are stale
    PortfolioDevice* d_bpf = nullptr;        // [n] instance records
    uint64_t* d_bseed = nullptr;             // [n]
    double* d_bwork = nullptr;               // [n]
    void* d_btab = nullptr;                  // per-chunk segment tables (grown)
    void* h_btab = nullptr;                  // pinned staging of the same
    size_t btab_cap = 0;
    Partial* d_inst_out = nullptr;           // [chunks x n] per-instance accumulators (grown)
    Partial* h_inst_out = nullptr;           // pinned
    size_t inst_out_cap = 0;
    int batch_begin(uint32_t n);
    int batch_stage(uint32_t i, const portfolio_data* data, uint64_t seed, const double* chol, uint32_t d);
    int batch_sample(uint32_t ell, const uint64_t* Ms, sample_output* outs);
    // Launch-shape autotuning of a batch level call (outer ids from 2^45 while
    // batch_tuning is set; the instances' outer ids do not move).
    bool batch_tuning = false;
    int batch_autotune_level(uint32_t ell, const uint64_t* Ms, Shape& best, double& best_ms);
    void free_batch();
    int prepare_batched(uint64_t M, BatchedCommon& c);
    int batched_common(uint32_t ell, uint64_t M, uint64_t outer_offset, BatchedCommon& c);
    int run_batched(uint32_t ell, uint64_t M, BatchedCommon& c, sample_output* out,
                    const std::function<hipError_t()>& launch, double* tr_phi_f, double* tr_phi_c,
                    uint64_t* tr_nf, uint64_t* tr_nc);
    int sample_portfolio_batched(uint32_t ell, uint64_t M, uint64_t outer_offset,
                                 uint32_t lanes_override, sample_output* out, double* tr_phi_f,
                                 double* tr_phi_c, uint64_t* tr_nf, uint64_t* tr_nc);
    int sample_batched(uint32_t ell, uint64_t M, uint64_t outer_offset, uint32_t lanes_override,
                       sample_output* out, double* tr_phi_f, double* tr_phi_c, uint64_t* tr_nf,
                       uint64_t* tr_nc);
};

extern Engine g_engine;

int set_error(int code, const char* fmt, ...);
int hip_fail(hipError_t e, const char* what);
// Takes the device address of the no-allocation entry point's scratch array on
// `device` (once per device; from ensure_device, at init).  A failure leaves
// that entry point refusing calls and is not an error of the init.
int noalloc_setup(int device);
uint64_t uint_pow_u64(uint64_t x, uint64_t n);
LevelCounts level_counts(const EngineParams& ep, uint32_t ell);
uint32_t test_segment_min(bool portfolio);
uint32_t xcd_slices(bool exact_book);

}  // namespace nmlmc
