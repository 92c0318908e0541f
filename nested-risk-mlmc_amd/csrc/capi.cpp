// C ABI of the MI355X nested sampler: the reference's extern "C" entry points
// (common.cpp:20-31, model_sampler.cpp:94-114, portfolio_sampler.cpp:618-756)
// and the status-returning nmlmc_* extensions declared in include/nmlmc.h.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <algorithm>
#include <cstring>
#include <functional>
#include <vector>

#include "../../include/nmlmc.h"
#include "engine.hpp"
#include "host.hpp"
#include "model.hpp"
#include "portfolio.hpp"

namespace nmlmc {

Engine g_engine;

static thread_local int t_last_error = NMLMC_OK;
static thread_local char t_last_msg[512] = "";

int set_error(int code, const char* fmt, ...) {
    t_last_error = code;
    if (code == NMLMC_OK) {
        t_last_msg[0] = 0;
        return code;
    }
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(t_last_msg, sizeof(t_last_msg), fmt, ap);
    va_end(ap);
    return code;
}

int hip_fail(hipError_t e, const char* what) {
    return set_error(NMLMC_ERR_HIP, "%s: %s", what, hipGetErrorString(e));
}

// uint_pow (common.hpp:146-157); x == 1 returns 1 (the reference's power-of-two
// branch evaluates clz(0) there, which is undefined).
uint64_t uint_pow_u64(uint64_t x, uint64_t n) {
    if (n == 0 || x == 1) return 1;
    if ((x & (x - 1)) == 0) return (uint64_t)1 << (n * (64 - __builtin_clzll(x - 1)));
    uint64_t r = 1;
    while (n--) r *= x;
    return r;
}

// get_samples_count constructor (nested_sampler.hpp:225-239) and its rule-1
// extension, evaluated with the host libm.
LevelCounts level_counts(const EngineParams& ep, uint32_t ell) {
    LevelCounts lc;
    lc.rule = (int)ep.adapt_rule;
    const uint64_t N0 = ep.N0;
    if (ep.adapt_rule == 1) {
        lc.base = N0 * uint_pow_u64(ep.beta, ell);
        lc.max = lc.base << ep.cap_log2;
        lc.tmp = 9.0;
        lc.adaptive = 1;
        return lc;
    }
    lc.max = N0 * uint_pow_u64(ep.beta, ell);
    const double r = ep.adaptive_r;
    lc.adaptive = (r > 1 && r < 2);
    if (lc.adaptive) {
        lc.base = N0 * uint_pow_u64(ep.beta >> 1, ell);
        const double c_conf2 = 9.;
        lc.tmp = (double)lc.max * std::pow(c_conf2 / (double)lc.max, r / 2);
    } else {
        lc.base = lc.max;
        lc.tmp = 0;
    }
    return lc;
}

// Virtual threads with outer samples up to which a plain model level call whose
// inner loops fill 256 lanes runs one virtual thread per 256-lane block
// (NMLMC_PLAIN_WIDE_MAX overrides; 0 keeps at most a wave).
static uint64_t plain_wide_max() {
    if (const char* env = getenv("NMLMC_PLAIN_WIDE_MAX")) return strtoull(env, nullptr, 10);
    return 32768;
}

static uint32_t pow2_ceil_u32(uint64_t x) {
    uint32_t g = 1;
    while (g < x && g < 64) g <<= 1;
    return g;
}

int Engine::ensure_device() {
    if (stream) return NMLMC_OK;
    // NMLMC_SPIN_WAIT=1: host waits on the level calls spin rather than yield
    // (unset: the runtime's default).  Possible only before the process's first
    // HIP context exists; otherwise refused, and the refusal is cleared.
    const char* spin = getenv("NMLMC_SPIN_WAIT");
    if (spin && spin[0] == '1') {
        if (hipSetDeviceFlags(hipDeviceScheduleSpin) != hipSuccess) (void)hipGetLastError();
    }
    hipError_t e = hipGetDevice(&device);
    if (e != hipSuccess) return hip_fail(e, "hipGetDevice");
    e = hipStreamCreateWithFlags(&stream, hipStreamNonBlocking);
    if (e != hipSuccess) return hip_fail(e, "hipStreamCreate");
    if ((e = hipEventCreate(&ev0)) != hipSuccess) return hip_fail(e, "hipEventCreate");
    if ((e = hipEventCreate(&ev1)) != hipSuccess) return hip_fail(e, "hipEventCreate");
    if ((e = hipMalloc(&d_out, sizeof(Partial) * (1 + kReduceBlocks))) != hipSuccess) return hip_fail(e, "hipMalloc");
    if ((e = hipHostMalloc(&h_out, sizeof(Partial))) != hipSuccess) return hip_fail(e, "hipHostMalloc");
    // The no-allocation entry point's scratch address; if it cannot be taken,
    // that entry point reports it and everything else works on.
    noalloc_setup(device);
    return NMLMC_OK;
}

void Engine::free_samplers() {
    if (d_states) hipFree(d_states);
    if (d_partials) hipFree(d_partials);
    if (d_ws) hipFree(d_ws);
    d_states = nullptr;
    d_partials = nullptr;
    d_ws = nullptr;
    partial_cap = 0;
    ws_cap = 0;
}

// Buffers of the batched forms (kept between level calls, released here).
void Engine::free_batched() {
    for (void* p : {(void*)bat_items, (void*)bat_tf[0], (void*)bat_tf[1], (void*)bat_tc[0], (void*)bat_tc[1],
                    (void*)bat_rq[0], (void*)bat_rq[1], (void*)bat_reeval, (void*)bat_pend, (void*)bat_fpool,
                    (void*)bat_redo_acc, (void*)bat_redo_cnt, (void*)bat_ctl, (void*)pf_R})
        if (p) hipFree(p);
    if (bat_ctl_h) hipHostFree(bat_ctl_h);
    if (sweep_slots) hipHostFree(sweep_slots);
    if (sweep_exec) hipGraphExecDestroy(sweep_exec);
    sweep_slots = nullptr;
    sweep_slot_cap = 0;
    sweep_exec = nullptr;
    bat_items = nullptr;
    bat_tf[0] = bat_tf[1] = bat_tc[0] = bat_tc[1] = nullptr;
    bat_rq[0] = bat_rq[1] = nullptr;
    bat_reeval = nullptr;
    bat_pend = nullptr;
    bat_fpool = nullptr;
    bat_redo_acc = nullptr;
    bat_redo_cnt = nullptr;
    bat_ctl = nullptr;
    ctl_dirty = true;
    bat_ctl_h = nullptr;
    pf_R = nullptr;
    pf_R32 = nullptr;
    bat_cap = bat_fcap = 0;
    pf_R_cap = 0;
}

int Engine::ensure_partials(size_t n) {
    if (n <= partial_cap) return NMLMC_OK;
    if (d_partials) hipFree(d_partials);
    d_partials = nullptr;
    partial_cap = 0;
    hipError_t e = hipMalloc(&d_partials, sizeof(Partial) * n);
    if (e != hipSuccess) return hip_fail(e, "hipMalloc partials");
    partial_cap = n;
    return NMLMC_OK;
}

int Engine::ensure_ws(size_t n) {
    if (n <= ws_cap) return NMLMC_OK;
    if (d_ws) hipFree(d_ws);
    d_ws = nullptr;
    ws_cap = 0;
    hipError_t e = hipMalloc(&d_ws, sizeof(double) * n);
    if (e != hipSuccess) return hip_fail(e, "hipMalloc workspace");
    ws_cap = n;
    return NMLMC_OK;
}

// init_samplers (nested_sampler.hpp:618-644): both streams of thread t keyed
// {seed, t} at position 0 (the OpenMP seeding, 640-641).
int Engine::seed_samplers(uint64_t s) {
    seed = s;
    int rc = ensure_device();
    if (rc) return rc;
    free_samplers();
    hipError_t e = hipMalloc(&d_states, sizeof(StreamState) * 2 * (size_t)cores);
    if (e != hipSuccess) return hip_fail(e, "hipMalloc states");
    e = hipMemsetAsync(d_states, 0, sizeof(StreamState) * 2 * (size_t)cores, stream);
    if (e != hipSuccess) return hip_fail(e, "hipMemset states");
    next_outer.assign(64, 0);
    tuned.clear();
    e = hipStreamSynchronize(stream);
    if (e != hipSuccess) return hip_fail(e, "hipStreamSynchronize");
    return NMLMC_OK;
}

uint32_t Engine::phi_size() const {
    if (problem == PROBLEM_PORTFOLIO) return pf_phi_size;
    return 1;
}

// Copies the reduced record into a sample_output (device_to_host,
// nested_sampler.hpp:562-577).
void Engine::copy_out(uint32_t ell, sample_output* out) {
    const Partial& p = *h_out;
    const uint32_t ps = phi_size();
    if (out->fsums)
        for (uint32_t i = 0; i < ps * ep.fine_moments; i++) out->fsums[i] = p.f[i];
    if (out->diffsums)
        for (uint32_t i = 0; i < ps * ep.diff_moments; i++) out->diffsums[i] = ell == 0 ? 0. : p.d[i];
    out->total_work = p.work;
    for (int i = 0; i < kStats; i++) out->inner_stats[i] = p.stats[i];
}

int Engine::sample(uint32_t ell, uint64_t M, uint64_t outer_offset, uint32_t lanes_override,
                   sample_output* out, double* tr_phi_f, double* tr_phi_c, uint64_t* tr_nf,
                   uint64_t* tr_nc) {
    if (!inited) return set_error(NMLMC_ERR_STATE, "init_nestedmc_sampler has not been called");
    if (problem < 0) return set_error(NMLMC_ERR_STATE, "no problem staged (model_init / portfolio_init)");
    if (problem == PROBLEM_SAMPLER_CLASS)
        return set_error(NMLMC_ERR_STATE, "a sampler class is staged: sample it with nmlmc::sample<SAMPLER>");
    if (!out) return set_error(NMLMC_ERR_ARG, "sample_output is NULL");
    if (ell > 40) return set_error(NMLMC_ERR_ARG, "level %u out of range", ell);
    out->M = M;
    if (M == 0) {
        std::memset(h_out, 0, sizeof(Partial));
        copy_out(ell, out);
        return set_error(NMLMC_OK, "");
    }
    if (form == NMLMC_FORM_PLAIN) {
        if (problem == PROBLEM_MODEL)
            return sample_model_parity(ell, M, lanes_override, out, tr_phi_f, tr_phi_c, tr_nf, tr_nc);
        return sample_portfolio_parity(ell, M, lanes_override, out, tr_phi_f, tr_phi_c, tr_nf, tr_nc);
    }
    if (ell >= next_outer.size()) next_outer.resize(ell + 1, 0);
    const uint64_t off = (outer_offset == UINT64_MAX) ? next_outer[ell] : outer_offset;
    int rc = sample_batched(ell, M, off, lanes_override, out, tr_phi_f, tr_phi_c, tr_nf, tr_nc);
    if (rc == NMLMC_OK) next_outer[ell] = off + M;
    return rc;
}

// Device trace buffers for one call (test/diagnostic path only).
int Engine::trace_alloc(uint64_t M, bool want, TraceBufs& tb) {
    tb = TraceBufs{};
    if (!want) return NMLMC_OK;
    hipError_t e;
    if ((e = hipMalloc(&tb.phi_f, 8 * M)) != hipSuccess) return hip_fail(e, "hipMalloc trace");
    if ((e = hipMalloc(&tb.phi_c, 8 * M)) != hipSuccess) return hip_fail(e, "hipMalloc trace");
    if ((e = hipMalloc(&tb.nf, 8 * M)) != hipSuccess) return hip_fail(e, "hipMalloc trace");
    if ((e = hipMalloc(&tb.nc, 8 * M)) != hipSuccess) return hip_fail(e, "hipMalloc trace");
    return NMLMC_OK;
}

int Engine::trace_finish(uint64_t M, TraceBufs& tb, double* tr_phi_f, double* tr_phi_c,
                         uint64_t* tr_nf, uint64_t* tr_nc) {
    hipError_t e = hipSuccess;
    if (tb.phi_f) {
        if (tr_phi_f && e == hipSuccess) e = hipMemcpy(tr_phi_f, tb.phi_f, 8 * M, hipMemcpyDeviceToHost);
        if (tr_phi_c && e == hipSuccess) e = hipMemcpy(tr_phi_c, tb.phi_c, 8 * M, hipMemcpyDeviceToHost);
        if (tr_nf && e == hipSuccess) e = hipMemcpy(tr_nf, tb.nf, 8 * M, hipMemcpyDeviceToHost);
        if (tr_nc && e == hipSuccess) e = hipMemcpy(tr_nc, tb.nc, 8 * M, hipMemcpyDeviceToHost);
        hipFree(tb.phi_f);
        hipFree(tb.phi_c);
        hipFree(tb.nf);
        hipFree(tb.nc);
        tb = TraceBufs{};
    }
    if (e != hipSuccess) return hip_fail(e, "trace copy");
    return NMLMC_OK;
}

int Engine::plain_buffers(size_t n, Partial*& partials, Partial*& result) {
    if (scr_partials) {
        if (n > scr_cap)
            return set_error(NMLMC_ERR_ARG, "the level call needs %zu partial records, the scratch holds %zu "
                                            "(fewer virtual threads or lanes)", n, scr_cap);
        partials = scr_partials;
        result = scr_result;
        return NMLMC_OK;
    }
    const int rc = ensure_partials(n);
    if (rc) return rc;
    partials = d_partials;
    result = d_out;
    return NMLMC_OK;
}

int Engine::finish_call(uint32_t ell, sample_output* out, const Partial* result) {
    hipError_t e = hipMemcpyAsync(h_out, result, sizeof(Partial), hipMemcpyDeviceToHost, stream);
    if (e != hipSuccess) return hip_fail(e, "hipMemcpyAsync result");
    e = hipStreamSynchronize(stream);
    if (e != hipSuccess) return hip_fail(e, "kernel execution");
    float ms = 0;
    hipEventElapsedTime(&ms, ev0, ev1);
    last_ms = ms;
    copy_out(ell, out);
    return NMLMC_OK;
}

int Engine::sample_model_parity(uint32_t ell, uint64_t M, uint32_t lanes_override,
                                sample_output* out, double* tr_phi_f, double* tr_phi_c,
                                uint64_t* tr_nf, uint64_t* tr_nc) {
    ModelParityArgsHost H{};
    H.mp = mp;
    H.ep = ep;
    H.cf = level_counts(ep, ell);
    H.cc = level_counts(ep, ell == 0 ? 0 : ell - 1);
    H.seed = seed;
    H.M = M;
    H.T = cores;
    H.ell = ell;
    const int nps = mp.variant == MODEL_REF ? 2 : 1;
    const uint64_t pairs = (H.cf.base * nps + 1) / 2;
    // Lanes per virtual thread: the next power of two of the base count's
    // Box-Muller pairs, at most a wave; a whole block of 256 lanes when at most
    // plain_wide_max() virtual threads hold outer samples and the pairs fill
    // it (the deep levels, whose few outer samples each take up to 2^cap times
    // the base count on one virtual thread) and the partial records fit.
    uint64_t want = lanes_override;
    if (const char* s = getenv("NMLMC_GROUP_LANES")) want = strtoul(s, nullptr, 10);
    if (want) {
        H.G = want >= 256 ? 256 : pow2_ceil_u32(want);
    } else {
        H.G = pow2_ceil_u32(pairs);
        const uint64_t act = plain_active(M, cores);
        if (pairs >= 256 && act <= plain_wide_max() && (!scr_partials || parity_blocks(act, 256) <= scr_cap))
            H.G = 256;
    }
    last_G = H.G;
    const uint64_t nblocks = parity_blocks(plain_active(M, cores), H.G);
    H.blocks = (uint32_t)nblocks;
    Partial *parts, *res;
    int rc = plain_buffers(nblocks, parts, res);
    if (rc) return rc;
    H.states = d_states;
    H.partials = parts;
    TraceBufs tb;
    rc = trace_alloc(M, tr_phi_f || tr_phi_c || tr_nf || tr_nc, tb);
    if (rc) return rc;
    H.tr_phi_f = tb.phi_f;
    H.tr_phi_c = tb.phi_c;
    H.tr_nf = tb.nf;
    H.tr_nc = tb.nc;
    hipError_t e = hipEventRecord(ev0, stream);
    if (e != hipSuccess) return hip_fail(e, "hipEventRecord");
    e = launch_model_parity(H, stream);
    if (e != hipSuccess) return hip_fail(e, "launch k_model_parity");
    if ((e = hipEventRecord(ev1, stream)) != hipSuccess) return hip_fail(e, "hipEventRecord");
    e = launch_reduce_partials(parts, (uint32_t)nblocks, res, stream);
    if (e != hipSuccess) return hip_fail(e, "launch k_reduce_partials");
    last_kernel = "k_model_parity";
    rc = finish_call(ell, out, res);
    if (rc) return rc;
    rc = trace_finish(M, tb, tr_phi_f, tr_phi_c, tr_nf, tr_nc);
    if (rc) return rc;
    return set_error(NMLMC_OK, "");
}

int Engine::device_cus() {
    if (cus > 0) return cus;
    int n = 0;
    if (hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess || n <= 0) n = 256;
    cus = n;
    return cus;
}

// Buffers of the batched engine for M outer samples (grown, never shrunk).
int Engine::prepare_batched(uint64_t M, BatchedCommon& c) {
    hipError_t e;
    if (M > bat_cap) {
        uint64_t cap = bat_cap ? bat_cap : 1024;
        while (cap < M) cap *= 2;
        for (void* p : {(void*)bat_items, (void*)bat_tf[0], (void*)bat_tf[1], (void*)bat_tc[0],
                        (void*)bat_tc[1], (void*)bat_rq[0], (void*)bat_rq[1], (void*)bat_reeval, (void*)bat_pend})
            if (p) hipFree(p);
        bat_items = nullptr;
        bat_tf[0] = bat_tf[1] = bat_tc[0] = bat_tc[1] = nullptr;
        bat_rq[0] = bat_rq[1] = nullptr;
        bat_reeval = nullptr;
        bat_pend = nullptr;
        bat_cap = 0;
        if ((e = hipMalloc(&bat_items, sizeof(ItemState) * cap)) != hipSuccess) return hip_fail(e, "hipMalloc items");
        for (int i = 0; i < 2; i++) {
            if ((e = hipMalloc(&bat_tf[i], sizeof(TestEntry) * cap)) != hipSuccess) return hip_fail(e, "hipMalloc queue");
            if ((e = hipMalloc(&bat_tc[i], sizeof(TestEntry) * cap)) != hipSuccess) return hip_fail(e, "hipMalloc queue");
            if ((e = hipMalloc(&bat_rq[i], sizeof(TestEntry) * cap)) != hipSuccess) return hip_fail(e, "hipMalloc queue");
        }
        if ((e = hipMalloc(&bat_reeval, sizeof(uint32_t) * cap)) != hipSuccess) return hip_fail(e, "hipMalloc reeval");
        if ((e = hipMalloc(&bat_pend, sizeof(uint32_t) * cap)) != hipSuccess) return hip_fail(e, "hipMalloc pending");
        bat_cap = cap;
    }
    uint64_t want_f = 3 * M + 4096;
    // Test hook: a first pool of this many entries, so that a level call
    // overflows it and takes the grow-and-redo path.
    if (const char* env = getenv("NMLMC_FINAL_POOL")) want_f = std::max<uint64_t>(1, strtoull(env, nullptr, 10));
    if (want_f > bat_fcap) {
        if (bat_fpool) hipFree(bat_fpool);
        bat_fpool = nullptr;
        bat_fcap = 0;
        if ((e = hipMalloc(&bat_fpool, sizeof(FinalEntry) * want_f)) != hipSuccess) return hip_fail(e, "hipMalloc finals");
        bat_fcap = want_f;
    }
    if (!bat_redo_acc) {
        if ((e = hipMalloc(&bat_redo_acc, sizeof(double) * kRedoSlots * kRedoGroups)) != hipSuccess)
            return hip_fail(e, "hipMalloc redo sums");
        if ((e = hipMalloc(&bat_redo_cnt, sizeof(uint32_t) * kRedoSlots)) != hipSuccess)
            return hip_fail(e, "hipMalloc redo counts");
    }
    if (!bat_ctl) {
        if ((e = hipMalloc(&bat_ctl, sizeof(BatchedCtl))) != hipSuccess) return hip_fail(e, "hipMalloc ctl");
        ctl_dirty = true;
        if ((e = hipHostMalloc(&bat_ctl_h, sizeof(BatchedCtl))) != hipSuccess) return hip_fail(e, "hipHostMalloc ctl");
    }
    c.items = bat_items;
    c.tf[0] = bat_tf[0];
    c.tf[1] = bat_tf[1];
    c.tc[0] = bat_tc[0];
    c.tc[1] = bat_tc[1];
    c.rq[0] = bat_rq[0];
    c.rq[1] = bat_rq[1];
    c.fpool = bat_fpool;
    c.fcap = (uint32_t)std::min<uint64_t>(bat_fcap, 0xffffffffull);
    c.reeval = bat_reeval;
    // The fused first pass (on unless NMLMC_FUSED=0) lists its leftovers here.
    c.pend = fused_enabled() ? bat_pend : nullptr;
    c.redo_acc = bat_redo_acc;
    c.redo_cnt = bat_redo_cnt;
    c.ctl = bat_ctl;
    c.zero_ctl = 1;
    c.xcd_sliced = xcd_slices(problem == PROBLEM_PORTFOLIO && pf.subsample == 1 && pf.nopt[0] == 0 &&
                              pf.nopt[2] == 0);
    return NMLMC_OK;
}

// Items of sub-sampled books of Exact options read their outer states (d
// values each) in every inner sample: sliced per XCD (xcd_range), each L2
// holds an eighth of them (the large book 1% and the 64-book batch 5%
// faster).  The model's items carry no such data, and its uneven per-item
// work balances better unsliced (1%); so do books with Asians on coupled
// paths, whose deep levels hold few long items (the 16-book batch 4%).
// NMLMC_XCD_SLICES=0/1 overrides.
uint32_t xcd_slices(bool exact_book) {
    if (const char* xs = getenv("NMLMC_XCD_SLICES")) return xs[0] == '0' ? 0u : 1u;
    return exact_book ? 1u : 0u;
}

// Smallest segment a sparse test round may shrink to (BatchedCommon::
// s_test_min): portfolios only, since a model inner sample is too cheap for a
// short segment to pay its hand-off (measured: the model sweep 5% slower at a
// minimum of 8, the large book's 2.5% faster at 16).  NMLMC_TEST_SEGMENT_MIN
// overrides it.
uint32_t test_segment_min(bool portfolio) {
    uint32_t v = portfolio ? 16u : 0u;
    if (const char* env = getenv("NMLMC_TEST_SEGMENT_MIN")) v = (uint32_t)strtoul(env, nullptr, 10);
    return v;
}

// Common arguments of a batched level call; checks that every draw-group size
// is the smallest one times a power of two (finals are bucketed that way).
int Engine::batched_common(uint32_t ell, uint64_t M, uint64_t outer_offset, BatchedCommon& c) {
    std::memset(&c, 0, sizeof(c));
    c.ep = ep;
    c.cf = level_counts(ep, ell);
    c.cc = level_counts(ep, ell == 0 ? 0 : ell - 1);
    c.seed = seed;
    c.M = M;
    c.offset = outer_offset;
    c.ell = ell;
    c.s_test_min = test_segment_min(problem == PROBLEM_PORTFOLIO);
    // Test hook: NMLMC_MIXED_REDO_ALL=1 sends every Indicator outer sample of
    // the mixed form to the FP64 redo (more work, the same results), so that
    // tests reach the redo kernels' wide and many-sample paths at will.
    if (const char* env = getenv("NMLMC_MIXED_REDO_ALL")) c.redo_all = strtoul(env, nullptr, 10) ? 1u : 0u;
    auto rounds_of = [](const LevelCounts& lc) {
        uint32_t n = 0;
        for (uint64_t cur = lc.base; lc.max > 2 * cur; cur *= 2) n++;
        return n;
    };
    c.rounds = rounds_of(c.cf);
    uint64_t bmin = c.cf.base;
    if (ell > 0) {
        c.rounds = std::max(c.rounds, rounds_of(c.cc));
        bmin = std::min(c.cf.base, c.cc.base);
    }
    if (c.rounds > (uint32_t)kMaxRounds) return set_error(NMLMC_ERR_ARG, "too many refinement rounds");
    auto pow2_multiple = [bmin](uint64_t n) {
        if (n % bmin) return false;
        const uint64_t q = n / bmin;
        return (q & (q - 1)) == 0 && q < (1ull << (kMaxClasses - 1));
    };
    bool ok = pow2_multiple(c.cf.base) && pow2_multiple(c.cf.max);
    if (ell > 0) ok = ok && pow2_multiple(c.cc.base) && pow2_multiple(c.cc.max);
    if (!ok && !(ell > 0 && ep.antithetic))
        return set_error(NMLMC_ERR_ARG, "batched forms need inner counts that are the smallest count "
                                        "times a power of two (beta a power of two)");
    if (!ok) {
        // Antithetic finals only use the smaller count as the group size.
        bmin = std::min(c.cf.max, c.cc.max);
        if (c.cf.base != c.cf.max || c.cc.base != c.cc.max)
            return set_error(NMLMC_ERR_ARG, "batched forms need power-of-two inner count ratios");
    }
    c.base_min = bmin;
    c.fcap = 0;
    return prepare_batched(std::min(M, batched_chunk()), c);
}

int Engine::run_batched(uint32_t ell, uint64_t M, BatchedCommon& c, sample_output* out,
                        const std::function<hipError_t()>& launch, double* tr_phi_f, double* tr_phi_c,
                        uint64_t* tr_nf, uint64_t* tr_nc) {
    if (sweep_on) {
        // Level sweep: enqueue the chunks' chains, each followed by a copy of
        // its control block into its host slot; sample_levels waits once.
        const uint64_t chunk = batched_chunk();
        const uint64_t offset0 = c.offset;
        for (uint64_t start = 0; start < M; start += chunk) {
            const uint32_t k = (uint32_t)sweep_chain_level.size();
            if (k >= sweep_slot_cap) return set_error(NMLMC_ERR_STATE, "level sweep: no control slot left");
            c.M = std::min(chunk, M - start);
            c.offset = offset0 + start;
            c.zero_ctl = ctl_dirty ? 1u : 0u;
            ctl_dirty = true;
            hipError_t e = launch();
            if (e != hipSuccess) return hip_fail(e, "launch batched level");
            e = hipMemcpyAsync(sweep_slots + k, bat_ctl, sizeof(BatchedCtl), hipMemcpyDeviceToHost, stream);
            if (e != hipSuccess) return hip_fail(e, "hipMemcpyAsync ctl");
            if ((e = launch_clear_ctl(bat_ctl, stream)) != hipSuccess) return hip_fail(e, "launch k_b_zero");
            ctl_dirty = false;
            sweep_chain_level.push_back(sweep_level);
        }
        c.offset = offset0;
        return NMLMC_OK;
    }
    TraceBufs tb;
    int rc = trace_alloc(M, tr_phi_f || tr_phi_c || tr_nf || tr_nc, tb);
    if (rc) return rc;
    // The level call runs in chunks of at most batched_chunk() outer samples:
    // the queues and item states hold one chunk.  Outer samples are keyed by
    // their global index, so only the order of the FP64 level sums depends on
    // the chunk size.
    const uint64_t chunk = batched_chunk();
    const uint64_t offset0 = c.offset;
    Partial total;
    std::memset(&total, 0, sizeof(total));
    double ms_total = 0;
    for (uint64_t start = 0; start < M; start += chunk) {
        const uint64_t n = std::min(chunk, M - start);
        c.M = n;
        c.offset = offset0 + start;
        c.tr_phi_f = tb.phi_f ? tb.phi_f + start : nullptr;
        c.tr_phi_c = tb.phi_c ? tb.phi_c + start : nullptr;
        c.tr_nf = tb.nf ? tb.nf + start : nullptr;
        c.tr_nc = tb.nc ? tb.nc + start : nullptr;
        for (int attempt = 0; attempt < 4; attempt++) {
            hipError_t e = hipEventRecord(ev0, stream);
            if (e != hipSuccess) return hip_fail(e, "hipEventRecord");
            c.zero_ctl = ctl_dirty ? 1u : 0u;
            ctl_dirty = true;
            e = launch();
            if (e != hipSuccess) return hip_fail(e, "launch batched level");
            if ((e = hipEventRecord(ev1, stream)) != hipSuccess) return hip_fail(e, "hipEventRecord");
            e = hipMemcpyAsync(bat_ctl_h, bat_ctl, sizeof(BatchedCtl), hipMemcpyDeviceToHost, stream);
            if (e != hipSuccess) return hip_fail(e, "hipMemcpyAsync ctl");
            // Cleared for the next chain once copied out (outside the timed
            // chain, like the copy: one block, about 2 us).
            if ((e = launch_clear_ctl(bat_ctl, stream)) != hipSuccess) return hip_fail(e, "launch k_b_zero");
            ctl_dirty = false;
            if ((e = hipStreamSynchronize(stream)) != hipSuccess) return hip_fail(e, "batched kernels");
            if (!bat_ctl_h->overflow) break;
            // More final groups than the pool holds: grow it and redo the chunk
            // (the draws are keyed, so the redo is identical).
            const uint64_t need = (uint64_t)bat_ctl_h->final_total + 4096;
            if (bat_fpool) hipFree(bat_fpool);
            bat_fpool = nullptr;
            bat_fcap = 0;
            if ((e = hipMalloc(&bat_fpool, sizeof(FinalEntry) * need)) != hipSuccess)
                return hip_fail(e, "hipMalloc finals");
            bat_fcap = need;
            c.fpool = bat_fpool;
            c.fcap = (uint32_t)std::min<uint64_t>(need, 0xffffffffull);
        }
        if (bat_ctl_h->overflow) return set_error(NMLMC_ERR_DEVICE, "final-group pool overflow");
        float ms = 0;
        hipEventElapsedTime(&ms, ev0, ev1);
        ms_total += ms;
        const Partial& a = bat_ctl_h->acc;
        for (int i = 0; i < 16; i++) {
            total.f[i] += a.f[i];
            total.d[i] += a.d[i];
        }
        total.work += a.work;
        for (int i = 0; i < kStatsExt; i++) total.stats[i] += a.stats[i];
    }
    c.offset = offset0;
    last_ms = ms_total;
    std::memcpy(h_out, &total, sizeof(Partial));
    copy_out(ell, out);
    mixed_reeval = h_out->stats[6];
    mixed_flips = h_out->stats[7];
    mixed_rechecks = h_out->stats[8];
    rc = trace_finish(M, tb, tr_phi_f, tr_phi_c, tr_nf, tr_nc);
    if (rc) return rc;
    return set_error(NMLMC_OK, "");
}

// Several level calls of the batched forms enqueued back to back with one wait
// at the end, captured from the stream as one HIP graph (NMLMC_SWEEP_GRAPH=0:
// enqueued directly).  The graph is captured anew on every call, so it always
// holds the current arguments, and an existing executable graph of the same
// shape is updated in place instead of instantiated again.  Each call
// continues its level's outer ids as a lone call would, and its results equal
// those of the lone call.  The plain form runs the calls one by one.
int Engine::sample_levels(uint32_t n, const uint32_t* ells, const uint64_t* Ms, sample_output* outs) {
    if (!inited) return set_error(NMLMC_ERR_STATE, "init_nestedmc_sampler has not been called");
    if (problem < 0) return set_error(NMLMC_ERR_STATE, "no problem staged (model_init / portfolio_init)");
    if (n == 0) return set_error(NMLMC_OK, "");
    if (!ells || !Ms || !outs) return set_error(NMLMC_ERR_ARG, "sample_levels: NULL argument");
    for (uint32_t i = 0; i < n; i++)
        if (ells[i] > 40) return set_error(NMLMC_ERR_ARG, "level %u out of range", ells[i]);
    if (form == NMLMC_FORM_PLAIN) {
        for (uint32_t i = 0; i < n; i++) {
            const int rc = sample(ells[i], Ms[i], UINT64_MAX, 0, &outs[i], nullptr, nullptr, nullptr, nullptr);
            if (rc) return rc;
        }
        return set_error(NMLMC_OK, "");
    }
    const uint64_t chunk = batched_chunk();
    uint64_t maxM = 1;
    uint64_t chains = 0;
    for (uint32_t i = 0; i < n; i++) {
        maxM = std::max(maxM, std::min(Ms[i], chunk));
        chains += (Ms[i] + chunk - 1) / chunk;
    }
    // Buffers for the largest chunk before anything is enqueued: nothing may be
    // reallocated while chains are in flight or being captured.
    BatchedCommon tmp;
    std::memset(&tmp, 0, sizeof(tmp));
    int rc = prepare_batched(maxM, tmp);
    if (rc) return rc;
    if (problem == PROBLEM_PORTFOLIO && (rc = ensure_pf_R(maxM))) return rc;
    hipError_t e;
    if (chains > sweep_slot_cap) {
        if (sweep_slots) hipHostFree(sweep_slots);
        sweep_slots = nullptr;
        sweep_slot_cap = 0;
        if ((e = hipHostMalloc(&sweep_slots, sizeof(BatchedCtl) * chains)) != hipSuccess)
            return hip_fail(e, "hipHostMalloc sweep slots");
        sweep_slot_cap = (uint32_t)chains;
    }
    std::vector<uint64_t> off(n);
    {
        std::vector<uint64_t> nxt = next_outer;
        for (uint32_t i = 0; i < n; i++) {
            if (ells[i] >= nxt.size()) nxt.resize(ells[i] + 1, 0);
            off[i] = nxt[ells[i]];
            nxt[ells[i]] += Ms[i];
        }
    }
    const char* env = getenv("NMLMC_SWEEP_GRAPH");
    const bool graph = !(env && env[0] == '0');
    if (graph) {
        if ((e = hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal)) != hipSuccess)
            return hip_fail(e, "hipStreamBeginCapture");
    } else if ((e = hipEventRecord(ev0, stream)) != hipSuccess) {
        return hip_fail(e, "hipEventRecord");
    }
    sweep_on = true;
    sweep_chain_level.clear();
    for (uint32_t i = 0; i < n && !rc; i++) {
        if (Ms[i] == 0) continue;
        sweep_level = i;
        outs[i].M = Ms[i];
        rc = sample_batched(ells[i], Ms[i], off[i], 0, &outs[i], nullptr, nullptr, nullptr, nullptr);
    }
    sweep_on = false;
    if (graph) {
        hipGraph_t g = nullptr;
        e = hipStreamEndCapture(stream, &g);
        if (rc || e != hipSuccess) {
            // The captured chains never run: the control block's state is
            // the one before the capture, which may not be clear.
            ctl_dirty = true;
            if (g) hipGraphDestroy(g);
            return rc ? rc : hip_fail(e, "hipStreamEndCapture");
        }
        // Update the executable graph in place when the new capture has the
        // same shape (same node count; the update checks the rest), else
        // instantiate anew.  A refused update is not an error: its status is
        // cleared so that no later call reports it.
        size_t nodes = 0;
        hipGraphGetNodes(g, nullptr, &nodes);
        bool updated = false;
        if (sweep_exec && nodes == sweep_exec_nodes) {
            hipGraphNode_t bad = nullptr;
            hipGraphExecUpdateResult res = hipGraphExecUpdateError;
            updated = hipGraphExecUpdate(sweep_exec, g, &bad, &res) == hipSuccess && res == hipGraphExecUpdateSuccess;
            if (!updated) (void)hipGetLastError();
        }
        if (!updated) {
            if (sweep_exec) hipGraphExecDestroy(sweep_exec);
            sweep_exec = nullptr;
            if ((e = hipGraphInstantiate(&sweep_exec, g, nullptr, nullptr, 0)) != hipSuccess) {
                sweep_exec = nullptr;
                ctl_dirty = true;
                hipGraphDestroy(g);
                return hip_fail(e, "hipGraphInstantiate");
            }
            sweep_exec_nodes = nodes;
        }
        hipGraphDestroy(g);
        if ((e = hipEventRecord(ev0, stream)) != hipSuccess) {
            ctl_dirty = true;
            return hip_fail(e, "hipEventRecord");
        }
        if ((e = hipGraphLaunch(sweep_exec, stream)) != hipSuccess) {
            ctl_dirty = true;
            return hip_fail(e, "hipGraphLaunch");
        }
    } else if (rc) {
        hipStreamSynchronize(stream);
        return rc;
    }
    if ((e = hipEventRecord(ev1, stream)) != hipSuccess) return hip_fail(e, "hipEventRecord");
    if ((e = hipStreamSynchronize(stream)) != hipSuccess) return hip_fail(e, "level sweep kernels");
    float ms = 0;
    hipEventElapsedTime(&ms, ev0, ev1);
    double ms_total = ms;
    // Per level call: the sum of its chains' control blocks.  A call whose
    // final-group pool overflowed is redone on its own (run_batched grows the
    // pool; the draws are keyed, so the redo is identical).
    std::vector<Partial> tot(n);
    std::memset(tot.data(), 0, sizeof(Partial) * n);
    std::vector<char> redo(n, 0);
    for (size_t k = 0; k < sweep_chain_level.size(); k++) {
        const BatchedCtl& b = sweep_slots[k];
        Partial& t = tot[sweep_chain_level[k]];
        if (b.overflow) redo[sweep_chain_level[k]] = 1;
        for (int j = 0; j < kMaxPhi * kMaxMoments; j++) {
            t.f[j] += b.acc.f[j];
            t.d[j] += b.acc.d[j];
        }
        t.work += b.acc.work;
        for (int j = 0; j < kStatsExt; j++) t.stats[j] += b.acc.stats[j];
    }
    uint64_t reev = 0, flips = 0, rck = 0;
    for (uint32_t i = 0; i < n; i++) {
        outs[i].M = Ms[i];
        if (redo[i]) {
            if ((rc = sample_batched(ells[i], Ms[i], off[i], 0, &outs[i], nullptr, nullptr, nullptr, nullptr)))
                return rc;
            ms_total += last_ms;
            reev += mixed_reeval;
            flips += mixed_flips;
            rck += mixed_rechecks;
        } else {
            std::memcpy(h_out, &tot[i], sizeof(Partial));
            copy_out(ells[i], &outs[i]);
            reev += tot[i].stats[6];
            flips += tot[i].stats[7];
            rck += tot[i].stats[8];
        }
        if (ells[i] >= next_outer.size()) next_outer.resize(ells[i] + 1, 0);
        next_outer[ells[i]] = off[i] + Ms[i];
    }
    last_ms = (float)ms_total;
    mixed_reeval = reev;
    mixed_flips = flips;
    mixed_rechecks = rck;
    return set_error(NMLMC_OK, "");
}

// The batched forms' fused first pass; NMLMC_FUSED=0 turns it off (results
// are the same either way).
bool Engine::fused_enabled() {
    const char* s = getenv("NMLMC_FUSED");
    return !(s && s[0] == '0');
}

// Outer samples per batched launch chain: 2^23 (about 2.8 GB of queues and item
// states, plus 8 B x d per outer sample of stock states for the portfolio);
// NMLMC_BATCH_CHUNK overrides it.
uint64_t Engine::batched_chunk() {
    uint64_t ch = 1ull << 23;
    if (const char* s = getenv("NMLMC_BATCH_CHUNK")) {
        const uint64_t v = strtoull(s, nullptr, 10);
        if (v > 0) ch = v;
    }
    return ch;
}

// Launch shape of a batched level call.  An explicit segment size (the
// override argument, else NMLMC_SEGMENT) wins; otherwise the shape
// autotune_level chose for this form and level, else the default: the largest
// segment of 64, 32 or 16 units that still cuts the outer samples' first draw
// groups into at least 2^19 segments (two waves per SIMD over 256 CUs), with
// 8 blocks per CU.  On the adaptive benchmarks this default is within ~3% of
// the best fixed segment at every level (larger segments pay fewer hand-offs,
// smaller ones spread the few large groups of the high levels).
Engine::Shape Engine::batched_shape(uint32_t ell, uint64_t M, uint64_t spu, const BatchedCommon& c,
                                    uint32_t lanes_override) {
    if (force.S) return force;
    Shape sh;
    const auto t = tuned.find(((uint64_t)form << 32) | ell);
    if (t != tuned.end()) {
        sh = t->second;
    } else {
        sh.bpc = 8;
        const uint64_t group = c.cf.max > 2 * c.cf.base ? c.cf.base : c.cf.max;
        const uint64_t units = (group + spu - 1) / spu;
        const uint64_t m = std::min(M, batched_chunk());
        // The largest segment that still gives the grid about two segments a
        // lane; groups of at least 1024 units (the deep levels) may go down
        // to 4 units a segment, smaller ones stop at 16 (their hand-offs).
        // With 4 allowed, a level sweep of the 16-book Asian batch at 2048
        // outer samples per instance takes 39.3 instead of 45.0 ms (the
        // autotuner's shapes: 39.1); the batch drivers' calls, with more outer
        // samples, keep their segments.
        const uint32_t smin = units >= 1024 ? 4u : 16u;
        sh.S = smin;
        for (uint32_t S : {64u, 32u, 16u, 8u})
            if (S >= smin && m * ((units + S - 1) / S) >= (1ull << 19)) {
                sh.S = S;
                break;
            }
    }
    if (lanes_override) {
        sh.S = lanes_override;
    } else if (const char* env = getenv("NMLMC_SEGMENT")) {
        const uint32_t v = (uint32_t)strtoul(env, nullptr, 10);
        if (v) sh.S = v;
    }
    return sh;
}

// Times the batched level call (ell, M) of the staged problem in the current
// form over candidate launch shapes and keeps the fastest for later calls at
// this level: segments of 4, 8, 16, 32, 64 and 128 units at 8 blocks per CU,
// then 4 and 16 blocks per CU at the best segment (the small segments spread
// the few large draw groups of the high levels over more lanes; the large ones
// pay fewer hand-offs at the low levels).  Each candidate runs twice (the
// faster kernel time counts) on outer samples 2^45 onwards, far from those of
// the level's calls, so later results do not depend on whether tuning ran
// (and per-sample results never depend on the shape).
int Engine::tune_shapes(uint32_t ell, const std::function<int(double&)>& run, Shape& best, double& best_ms) {
    auto time_shape = [&](Shape s, double& ms) -> int {
        force = s;
        ms = 1e300;
        for (int rep = 0; rep < 2; rep++) {
            double t = 0.;
            const int rc = run(t);
            if (rc) {
                force = Shape{};
                return rc;
            }
            ms = std::min(ms, t);
        }
        force = Shape{};
        return NMLMC_OK;
    };
    best = Shape{};
    best_ms = 1e300;
    for (uint32_t S : {4u, 8u, 16u, 32u, 64u, 128u}) {
        double ms;
        if (int rc = time_shape(Shape{S, 8}, ms)) return rc;
        if (ms < best_ms) {
            best = Shape{S, 8};
            best_ms = ms;
        }
    }
    const uint32_t S = best.S;
    for (uint32_t bpc : {4u, 16u}) {
        double ms;
        if (int rc = time_shape(Shape{S, bpc}, ms)) return rc;
        if (ms < best_ms) {
            best = Shape{S, bpc};
            best_ms = ms;
        }
    }
    tuned[((uint64_t)form << 32) | ell] = best;
    return NMLMC_OK;
}

int Engine::autotune_level(uint32_t ell, uint64_t M, Shape& best, double& best_ms) {
    if (!inited || problem < 0) return set_error(NMLMC_ERR_STATE, "no problem staged");
    if (form == NMLMC_FORM_PLAIN) return set_error(NMLMC_ERR_ARG, "autotuning applies to the batched forms");
    if (M == 0 || M >= (1ull << 45) || ell > 40) return set_error(NMLMC_ERR_ARG, "autotune: bad level or M");
    std::vector<double> f(kMaxPhi * kMaxMoments), d(kMaxPhi * kMaxMoments);
    sample_output out{};
    out.fsums = f.data();
    out.diffsums = d.data();
    const int rc = tune_shapes(ell, [&](double& ms) {
        out.M = M;
        const int r = sample_batched(ell, M, 1ull << 45, 0, &out, nullptr, nullptr, nullptr, nullptr);
        ms = last_ms;
        return r;
    }, best, best_ms);
    if (rc) return rc;
    return set_error(NMLMC_OK, "");
}

// Batched forms: per-outer-sample keys, segment queues, one launch per round.
int Engine::sample_batched(uint32_t ell, uint64_t M, uint64_t outer_offset, uint32_t lanes_override,
                           sample_output* out, double* tr_phi_f, double* tr_phi_c, uint64_t* tr_nf,
                           uint64_t* tr_nc) {
    if (problem != PROBLEM_MODEL) return sample_portfolio_batched(ell, M, outer_offset, lanes_override, out,
                                                                   tr_phi_f, tr_phi_c, tr_nf, tr_nc);
    ModelBatchedHost H{};
    int rc = batched_common(ell, M, outer_offset, H.c);
    if (rc) return rc;
    H.mp = mp;
    H.c.phi_type = mp.phi_type;
    H.c.nps = mp.variant == MODEL_REF ? 2 : 1;
    H.c.ups = 0;
    // One unit = one Threefry block + one Box-Muller pair (two samples of the
    // shift model, one of the reference model).
    const Shape sh = batched_shape(ell, M, mp.variant == MODEL_SHIFT ? 2 : 1, H.c, lanes_override);
    H.c.S = sh.S;
    last_G = H.c.S;
    H.blocks = (uint32_t)device_cus() * sh.bpc;
    H.mixed = form == NMLMC_FORM_MIXED;
    last_kernel = "k_b_final";
    return run_batched(ell, M, H.c, out, [&]() { return launch_model_batched(H, stream); }, tr_phi_f, tr_phi_c,
                       tr_nf, tr_nc);
}

}  // namespace nmlmc

using nmlmc::g_engine;
using nmlmc::set_error;

extern "C" {

int nmlmc_last_error(void) { return nmlmc::t_last_error; }
const char* nmlmc_last_error_string(void) { return nmlmc::t_last_msg; }

// init_nestedmc_sampler (common.cpp:20-26).
void init_nestedmc_sampler(uint32_t cores, const nested_sampler_params* params) {
    if (!params || cores == 0) {
        set_error(NMLMC_ERR_ARG, "init_nestedmc_sampler: cores must be > 0 and params non-NULL");
        return;
    }
    if (params->fine_moments > nmlmc::kMaxMoments || params->diff_moments > nmlmc::kMaxMoments ||
        params->N0 == 0 || params->beta == 0) {
        set_error(NMLMC_ERR_ARG, "init_nestedmc_sampler: invalid parameters");
        return;
    }
    g_engine.free_samplers();
    g_engine.cores = cores;
    g_engine.ep.N0 = params->N0;
    g_engine.ep.beta = params->beta;
    g_engine.ep.antithetic = params->antithetic ? 1 : 0;
    g_engine.ep.fine_moments = params->fine_moments;
    g_engine.ep.diff_moments = params->diff_moments;
    g_engine.ep.adaptive_r = params->adaptive_r;
    g_engine.ep.adapt_rule = 0;
    g_engine.ep.cap_log2 = 0;
    g_engine.level_coupled = false;
    g_engine.problem = -1;
    g_engine.inited = true;
    int rc = g_engine.ensure_device();
    if (rc) return;
    set_error(NMLMC_OK, "");
}

// cleanup_nestedmc_sampler (common.cpp:28-31).
void cleanup_nestedmc_sampler(void) {
    g_engine.free_batch();
    g_engine.free_samplers();
    g_engine.free_portfolio();
    g_engine.free_batched();
    g_engine.inited = false;
    g_engine.problem = -1;
    set_error(NMLMC_OK, "");
}

int nmlmc_set_form(int form) {
    if (form < NMLMC_FORM_PLAIN || form > NMLMC_FORM_MIXED) return set_error(NMLMC_ERR_ARG, "unknown form %d", form);
    g_engine.form = form;
    return set_error(NMLMC_OK, "");
}
int nmlmc_get_form(void) { return g_engine.form; }

int nmlmc_set_adaptive_rule(uint32_t rule, uint32_t cap_log2) {
    if (!g_engine.inited) return set_error(NMLMC_ERR_STATE, "init_nestedmc_sampler has not been called");
    if (rule > 1 || cap_log2 > 24) return set_error(NMLMC_ERR_ARG, "invalid adaptive rule");
    g_engine.ep.adapt_rule = rule;
    g_engine.ep.cap_log2 = cap_log2;
    g_engine.tuned.clear();
    return set_error(NMLMC_OK, "");
}

// model_update_data (model_sampler.cpp:94-97).
void model_update_data(const model_data* data) {
    if (!data) {
        set_error(NMLMC_ERR_ARG, "model_update_data: NULL data");
        return;
    }
    g_engine.mp.stau = data->stau;
    g_engine.mp.s1tau = data->s1tau;
    g_engine.mp.max_loss = data->max_loss;
    g_engine.mp.phi_type = data->phi_type;
    g_engine.mp.variant = nmlmc::MODEL_REF;
    g_engine.problem = nmlmc::PROBLEM_MODEL;
    set_error(NMLMC_OK, "");
}

// model_init (model_sampler.cpp:99-104).
void model_init(const model_data* data, uint64_t seed) {
    if (!g_engine.inited) {
        set_error(NMLMC_ERR_STATE, "init_nestedmc_sampler has not been called");
        return;
    }
    model_update_data(data);
    if (nmlmc_last_error()) return;
    g_engine.seed_samplers(seed);
}

int nmlmc_model_init(int variant, int phi_type, double tau, double max_loss, uint64_t seed) {
    if (!g_engine.inited) return set_error(NMLMC_ERR_STATE, "init_nestedmc_sampler has not been called");
    if (variant != nmlmc::MODEL_REF && variant != nmlmc::MODEL_SHIFT)
        return set_error(NMLMC_ERR_ARG, "unknown model variant %d", variant);
    if (phi_type != 0 && phi_type != 1) return set_error(NMLMC_ERR_ARG, "model phi_type must be 0 or 1");
    if (!(tau > 0 && tau < 1)) return set_error(NMLMC_ERR_ARG, "tau must lie in (0, 1)");
    model_data d;
    d.stau = std::sqrt(tau);
    d.s1tau = std::sqrt(1 - tau);
    d.max_loss = max_loss;
    d.phi_type = phi_type;
    model_update_data(&d);
    g_engine.mp.variant = variant;
    return g_engine.seed_samplers(seed) ? nmlmc_last_error() : set_error(NMLMC_OK, "");
}

// model_sample (model_sampler.cpp:106-114).
void model_sample(uint32_t ell, uint64_t M, sample_output* out) {
    if (g_engine.problem != nmlmc::PROBLEM_MODEL) {
        set_error(NMLMC_ERR_STATE, "model_sample: no model problem staged");
        return;
    }
    g_engine.sample(ell, M, UINT64_MAX, 0, out, nullptr, nullptr, nullptr, nullptr);
}

int nmlmc_sample(uint32_t ell, uint64_t M, uint64_t outer_offset, sample_output* out,
                 double* tr_phi_f, double* tr_phi_c, uint64_t* tr_nf, uint64_t* tr_nc) {
    return g_engine.sample(ell, M, outer_offset, 0, out, tr_phi_f, tr_phi_c, tr_nf, tr_nc);
}

int nmlmc_mixed_stats(uint64_t* reevaluated, uint64_t* flips, uint64_t* rechecked) {
    if (reevaluated) *reevaluated = g_engine.mixed_reeval;
    if (flips) *flips = g_engine.mixed_flips;
    if (rechecked) *rechecked = g_engine.mixed_rechecks;
    return set_error(NMLMC_OK, "");
}

double nmlmc_last_kernel_ms(void) { return g_engine.last_ms; }
const char* nmlmc_last_kernel_name(void) { return g_engine.last_kernel; }

#ifndef NMLMC_SOURCE_DIGEST
#define NMLMC_SOURCE_DIGEST "unknown"
#endif
const char* nmlmc_source_digest(void) { return NMLMC_SOURCE_DIGEST; }
uint32_t nmlmc_last_group_lanes(void) { return g_engine.last_G; }

int nmlmc_sample_levels(uint32_t n, const uint32_t* ells, const uint64_t* Ms, sample_output* outs) {
    return g_engine.sample_levels(n, ells, Ms, outs);
}

int nmlmc_batch_autotune_level(uint32_t ell, uint32_t n, const uint64_t* M, uint32_t* segment,
                               uint32_t* blocks_per_cu, double* ms) {
    if (n != g_engine.batch.size())
        return set_error(NMLMC_ERR_ARG, "batch_autotune_level: %u array entries for a batch of %zu instances", n,
                         g_engine.batch.size());
    nmlmc::Engine::Shape best;
    double best_ms = 0.;
    const int rc = g_engine.batch_autotune_level(ell, M, best, best_ms);
    if (rc) return rc;
    if (segment) *segment = best.S;
    if (blocks_per_cu) *blocks_per_cu = best.bpc;
    if (ms) *ms = best_ms;
    return set_error(NMLMC_OK, "");
}

int nmlmc_autotune_level(uint32_t ell, uint64_t M, uint32_t* segment, uint32_t* blocks_per_cu, double* ms) {
    nmlmc::Engine::Shape best;
    double best_ms = 0.;
    const int rc = g_engine.autotune_level(ell, M, best, best_ms);
    if (rc) return rc;
    if (segment) *segment = best.S;
    if (blocks_per_cu) *blocks_per_cu = best.bpc;
    if (ms) *ms = best_ms;
    return set_error(NMLMC_OK, "");
}

}  // extern "C"
