// A batch of independent portfolio problems sampled together: every level call
// runs one launch chain over the outer samples of all instances (no
// counterpart in the reference, which holds one problem per process).  Each
// instance keeps its own option tables, Cholesky factor, seed and per-level
// outer ids, so that its results equal those of the same problem staged alone
// and sampled with the same level calls.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <vector>

#include "../../include/nmlmc.h"
#include "host.hpp"

namespace nmlmc {

void Engine::free_batch() {
    for (BatchInst& b : batch)
        for (void* p : b.allocs) hipFree(p);
    batch.clear();
    for (void* p : {(void*)d_bpf, (void*)d_bseed, (void*)d_bwork, d_btab, (void*)d_inst_out})
        if (p) hipFree(p);
    if (h_btab) hipHostFree(h_btab);
    if (h_inst_out) hipHostFree(h_inst_out);
    d_bpf = nullptr;
    d_bseed = nullptr;
    d_bwork = nullptr;
    d_btab = h_btab = nullptr;
    d_inst_out = h_inst_out = nullptr;
    btab_cap = inst_out_cap = 0;
    batch_dirty = true;
}

int Engine::batch_begin(uint32_t n) {
    if (!inited) return set_error(NMLMC_ERR_STATE, "init_nestedmc_sampler has not been called");
    if (n == 0 || n > kMaxInst) return set_error(NMLMC_ERR_ARG, "a batch holds 1..%u problems", kMaxInst);
    if (ensure_device()) return nmlmc_last_error();
    free_batch();
    tuned.clear();  // shapes tuned for another workload
    batch.resize(n);
    if (hipMalloc(&d_bpf, sizeof(PortfolioDevice) * n) != hipSuccess ||
        hipMalloc(&d_bseed, sizeof(uint64_t) * n) != hipSuccess || hipMalloc(&d_bwork, sizeof(double) * n) != hipSuccess)
        return hip_fail(hipErrorOutOfMemory, "hipMalloc batch records");
    return set_error(NMLMC_OK, "");
}

// Instance i: the tables are built as portfolio_update_data builds them (the
// engine's own problem is replaced in the process and left unstaged), then
// moved into the batch.
int Engine::batch_stage(uint32_t i, const portfolio_data* data, uint64_t s, const double* chol, uint32_t d) {
    if (i >= batch.size()) return set_error(NMLMC_ERR_ARG, "batch instance %u out of range (nmlmc_batch_begin)", i);
    if (!data) return set_error(NMLMC_ERR_ARG, "portfolio data is NULL");
    if (data->phi_type != PHI_INDICATOR && data->phi_type != PHI_MAX)
        return set_error(NMLMC_ERR_ARG, "batched forms support the Indicator and Max phi types");
    int rc = stage_portfolio(data);
    if (rc) return rc;
    if (chol && (rc = nmlmc_portfolio_set_correlation(chol, d))) return rc;
    if (pf.chol && pf.d > 256) return set_error(NMLMC_ERR_ARG, "batched forms take a correlation factor of at most 256 stocks");
    BatchInst& b = batch[i];
    for (void* p : b.allocs) hipFree(p);
    b.allocs.clear();
    b.allocs.swap(pf_allocs);
    b.pf = pf;
    b.bs_work = pf_bs_work;
    b.seed = s;
    b.staged = true;
    b.exact_only = pf.subsample == 1 && pf.nopt[0] == 0 && pf.nopt[2] == 0;
    b.next_outer.assign(64, 0);
    std::memset(&pf, 0, sizeof(pf));
    problem = -1;
    batch_dirty = true;
    return set_error(NMLMC_OK, "");
}

namespace {
// One chunk's segment tables, packed: seg_item0[n_seg + 1], seg_outer0[n_seg]
// (u64), then seg_inst[n_seg], seg_blk0[n_seg + 1] (u32).
struct Seg {
    uint32_t inst;
    uint64_t item0, count, outer0;
};
size_t tab_words(size_t n_seg) { return 2 * n_seg + 1 + (2 * n_seg + 2) / 2 + 1; }  // in u64 words
}  // namespace

// One level call over every instance of the batch: Ms[i] outer samples of
// instance i (continuing its level's outer ids), results into outs[i].
int Engine::batch_sample(uint32_t ell, const uint64_t* Ms, sample_output* outs) {
    if (batch.empty()) return set_error(NMLMC_ERR_STATE, "no batch (nmlmc_batch_begin)");
    if (!Ms || !outs) return set_error(NMLMC_ERR_ARG, "batch_sample: NULL argument");
    if (ell > 40) return set_error(NMLMC_ERR_ARG, "level %u out of range", ell);
    if (form == NMLMC_FORM_PLAIN) return set_error(NMLMC_ERR_ARG, "batches run in the batched forms");
    const uint32_t n = (uint32_t)batch.size();
    bool exact_only = true;
    uint32_t dmax = 0;
    int phi_type = -1;
    uint64_t total = 0;
    for (uint32_t i = 0; i < n; i++) {
        const BatchInst& b = batch[i];
        if (!b.staged) return set_error(NMLMC_ERR_STATE, "batch instance %u is not staged", i);
        if (phi_type >= 0 && b.pf.phi_type != phi_type)
            return set_error(NMLMC_ERR_ARG, "the instances of a batch need one phi type");
        phi_type = b.pf.phi_type;
        exact_only = exact_only && b.exact_only;
        dmax = std::max(dmax, b.pf.d);
        total += Ms[i];
        if (ell >= b.next_outer.size()) batch[i].next_outer.resize(ell + 1, 0);
    }
    hipError_t e;
    if (batch_dirty) {
        std::vector<PortfolioDevice> recs(n);
        std::vector<uint64_t> seeds(n);
        std::vector<double> works(n);
        for (uint32_t i = 0; i < n; i++) {
            recs[i] = batch[i].pf;
            seeds[i] = batch[i].seed;
            works[i] = batch[i].bs_work;
        }
        if ((e = hipMemcpy(d_bpf, recs.data(), sizeof(PortfolioDevice) * n, hipMemcpyHostToDevice)) != hipSuccess ||
            (e = hipMemcpy(d_bseed, seeds.data(), 8 * n, hipMemcpyHostToDevice)) != hipSuccess ||
            (e = hipMemcpy(d_bwork, works.data(), 8 * n, hipMemcpyHostToDevice)) != hipSuccess)
            return hip_fail(e, "hipMemcpy batch records");
        batch_dirty = false;
    }
    // Chunks of at most batched_chunk() outer samples, laid out instance after
    // instance; an instance may span chunk boundaries.
    const uint64_t chunk = batched_chunk();
    std::vector<std::vector<Seg>> chunks(1);
    uint64_t fill = 0;
    for (uint32_t i = 0; i < n; i++) {
        uint64_t left = Ms[i], done = 0;
        while (left) {
            if (fill == chunk) {
                chunks.emplace_back();
                fill = 0;
            }
            const uint64_t take = std::min(left, chunk - fill);
            const uint64_t first = batch_tuning ? (1ull << 45) : batch[i].next_outer[ell];
            chunks.back().push_back(Seg{i, fill, take, first + done});
            fill += take;
            done += take;
            left -= take;
        }
    }
    if (total == 0) chunks.clear();
    const size_t nc = chunks.size();
    // Segment tables of every chunk in one upload; per-instance accumulators
    // per chunk (a chunk redone after a pool overflow clears only its own).
    std::vector<size_t> tab_off(nc + 1, 0);
    for (size_t k = 0; k < nc; k++) tab_off[k + 1] = tab_off[k] + tab_words(chunks[k].size());
    if (tab_off[nc] > btab_cap) {
        if (d_btab) hipFree(d_btab);
        if (h_btab) hipHostFree(h_btab);
        d_btab = h_btab = nullptr;
        btab_cap = 0;
        if ((e = hipMalloc(&d_btab, 8 * tab_off[nc])) != hipSuccess) return hip_fail(e, "hipMalloc batch tables");
        if ((e = hipHostMalloc(&h_btab, 8 * tab_off[nc])) != hipSuccess) return hip_fail(e, "hipHostMalloc batch tables");
        btab_cap = tab_off[nc];
    }
    if (nc * n > inst_out_cap) {
        if (d_inst_out) hipFree(d_inst_out);
        if (h_inst_out) hipHostFree(h_inst_out);
        d_inst_out = h_inst_out = nullptr;
        inst_out_cap = 0;
        if ((e = hipMalloc(&d_inst_out, sizeof(Partial) * nc * n)) != hipSuccess)
            return hip_fail(e, "hipMalloc batch accumulators");
        if ((e = hipHostMalloc(&h_inst_out, sizeof(Partial) * nc * n)) != hipSuccess)
            return hip_fail(e, "hipHostMalloc batch accumulators");
        inst_out_cap = nc * n;
    }
    std::vector<PortfolioBatchTables> tabs(nc);
    for (size_t k = 0; k < nc; k++) {
        const std::vector<Seg>& sg = chunks[k];
        const size_t ns = sg.size();
        uint64_t* w = (uint64_t*)h_btab + tab_off[k];
        uint64_t* item0 = w;
        uint64_t* outer0 = w + ns + 1;
        uint32_t* inst = (uint32_t*)(w + 2 * ns + 1);
        uint32_t* blk0 = inst + ns;
        uint32_t blocks = 0;
        for (size_t j = 0; j < ns; j++) {
            item0[j] = sg[j].item0;
            outer0[j] = sg[j].outer0;
            inst[j] = sg[j].inst;
            blk0[j] = blocks;
            if (batch[sg[j].inst].pf.chol) blocks += (uint32_t)((sg[j].count + 63) / 64);
        }
        item0[ns] = sg[ns - 1].item0 + sg[ns - 1].count;
        blk0[ns] = blocks;
        uint64_t* dw = (uint64_t*)d_btab + tab_off[k];
        PortfolioBatchTables& t = tabs[k];
        t.inst = d_bpf;
        t.inst_seed = d_bseed;
        t.inst_bs_work = d_bwork;
        t.seg_item0 = dw;
        t.seg_outer0 = dw + ns + 1;
        t.seg_inst = (const uint32_t*)(dw + 2 * ns + 1);
        t.seg_blk0 = t.seg_inst + ns;
        t.n_seg = (uint32_t)ns;
        t.outer_blocks = blocks;
        t.dmax = dmax;
    }
    if (nc && (e = hipMemcpy(d_btab, h_btab, 8 * tab_off[nc], hipMemcpyHostToDevice)) != hipSuccess)
        return hip_fail(e, "hipMemcpy batch tables");
    // Buffers for the largest chunk, then the chains chunk by chunk.
    const uint64_t maxc = std::min(std::max<uint64_t>(total, 1), chunk);
    PortfolioBatchedHost H{};
    int rc = batched_common(ell, maxc, 0, H.c);
    if (rc) return rc;
    H.c.s_test_min = test_segment_min(true);
    H.c.xcd_sliced = xcd_slices(exact_only);
    if ((rc = ensure_R((size_t)maxc * dmax))) return rc;
    std::memset(&H.pf, 0, sizeof(H.pf));
    H.R_tau = pf_R;
    H.bs_work = 0.;
    H.mixed = form == NMLMC_FORM_MIXED;
    H.exact_only = exact_only;
    H.multi = true;
    H.R32 = (H.mixed && H.exact_only) ? pf_R32 : nullptr;
    H.c.phi_type = phi_type;
    H.c.n_inst = n;
    // Every chunk's chain clears the control block and its accumulators first
    // (zero_ctl, set by batched_common) and leaves them filled.
    ctl_dirty = true;
    float ms_total = 0;
    for (size_t k = 0; k < nc; k++) {
        const uint64_t Mk = ((const uint64_t*)h_btab)[tab_off[k] + tabs[k].n_seg];
        H.c.M = Mk;
        H.c.offset = 0;
        H.c.inst_out = d_inst_out + k * n;
        H.bt = tabs[k];
        const Shape sh = batched_shape(ell, Mk, 1, H.c, 0);
        H.c.S = sh.S;
        H.blocks = (uint32_t)device_cus() * sh.bpc;
        last_G = H.c.S;
        for (int attempt = 0; attempt < 4; attempt++) {
            if ((e = hipEventRecord(ev0, stream)) != hipSuccess) return hip_fail(e, "hipEventRecord");
            if ((e = launch_portfolio_batched(H, stream)) != hipSuccess) return hip_fail(e, "launch batch level");
            if ((e = hipEventRecord(ev1, stream)) != hipSuccess) return hip_fail(e, "hipEventRecord");
            e = hipMemcpyAsync(bat_ctl_h, bat_ctl, sizeof(BatchedCtl), hipMemcpyDeviceToHost, stream);
            if (e != hipSuccess) return hip_fail(e, "hipMemcpyAsync ctl");
            if ((e = hipStreamSynchronize(stream)) != hipSuccess) return hip_fail(e, "batch kernels");
            if (!bat_ctl_h->overflow) break;
            const uint64_t need = (uint64_t)bat_ctl_h->final_total + 4096;
            if (bat_fpool) hipFree(bat_fpool);
            bat_fpool = nullptr;
            bat_fcap = 0;
            if ((e = hipMalloc(&bat_fpool, sizeof(FinalEntry) * need)) != hipSuccess) return hip_fail(e, "hipMalloc finals");
            bat_fcap = need;
            H.c.fpool = bat_fpool;
            H.c.fcap = (uint32_t)std::min<uint64_t>(need, 0xffffffffull);
        }
        if (bat_ctl_h->overflow) return set_error(NMLMC_ERR_DEVICE, "final-group pool overflow");
        float ms = 0;
        hipEventElapsedTime(&ms, ev0, ev1);
        ms_total += ms;
    }
    if (nc && (e = hipMemcpy(h_inst_out, d_inst_out, sizeof(Partial) * nc * n, hipMemcpyDeviceToHost)) != hipSuccess)
        return hip_fail(e, "hipMemcpy batch accumulators");
    uint64_t reev = 0, flips = 0, rck = 0;
    for (uint32_t i = 0; i < n; i++) {
        Partial t;
        std::memset(&t, 0, sizeof(t));
        for (size_t k = 0; k < nc; k++) {
            const Partial& a = h_inst_out[k * n + i];
            for (int j = 0; j < kMaxPhi * kMaxMoments; j++) {
                t.f[j] += a.f[j];
                t.d[j] += a.d[j];
            }
            t.work += a.work;
            for (int j = 0; j < kStatsExt; j++) t.stats[j] += a.stats[j];
        }
        std::memcpy(h_out, &t, sizeof(Partial));
        outs[i].M = Ms[i];
        copy_out(ell, &outs[i]);
        reev += t.stats[6];
        flips += t.stats[7];
        rck += t.stats[8];
        if (!batch_tuning) batch[i].next_outer[ell] += Ms[i];
    }
    last_ms = ms_total;
    last_kernel = "k_b_final";
    mixed_reeval = reev;
    mixed_flips = flips;
    mixed_rechecks = rck;
    return set_error(NMLMC_OK, "");
}

// Times the batch level call (ell, Ms) over the candidate launch shapes, as
// autotune_level does for a single problem, on outer ids from 2^45 of every
// instance (the instances' level calls then draw what they would have drawn
// without it), and keeps the fastest for later batch calls at level ell.
int Engine::batch_autotune_level(uint32_t ell, const uint64_t* Ms, Shape& best, double& best_ms) {
    if (batch.empty()) return set_error(NMLMC_ERR_STATE, "no batch (nmlmc_batch_begin)");
    if (!Ms) return set_error(NMLMC_ERR_ARG, "batch_autotune_level: NULL argument");
    for (size_t i = 0; i < batch.size(); i++)
        if (Ms[i] >= (1ull << 45)) return set_error(NMLMC_ERR_ARG, "batch_autotune_level: bad M");
    const size_t n = batch.size();
    std::vector<double> f(n * kMaxPhi * kMaxMoments), d(n * kMaxPhi * kMaxMoments);
    std::vector<sample_output> outs(n);
    for (size_t i = 0; i < n; i++) {
        outs[i] = sample_output{};
        outs[i].fsums = f.data() + i * kMaxPhi * kMaxMoments;
        outs[i].diffsums = d.data() + i * kMaxPhi * kMaxMoments;
    }
    batch_tuning = true;
    const int rc = tune_shapes(ell, [&](double& ms) {
        const int r = batch_sample(ell, Ms, outs.data());
        ms = last_ms;
        return r;
    }, best, best_ms);
    batch_tuning = false;
    if (rc) return rc;
    return set_error(NMLMC_OK, "");
}

}  // namespace nmlmc

using nmlmc::g_engine;
using nmlmc::set_error;

extern "C" {

int nmlmc_batch_begin(uint32_t n) { return g_engine.batch_begin(n); }

int nmlmc_batch_stage(uint32_t i, const portfolio_data* data, uint64_t seed, const double* chol, uint32_t d) {
    return g_engine.batch_stage(i, data, seed, chol, d);
}

int nmlmc_batch_sample(uint32_t ell, uint32_t n, const uint64_t* M, sample_output* outs) {
    if (n != g_engine.batch.size())
        return set_error(NMLMC_ERR_ARG, "batch_sample: %u arrays entries for a batch of %zu instances", n,
                         g_engine.batch.size());
    return g_engine.batch_sample(ell, M, outs);
}

}  // extern "C"
