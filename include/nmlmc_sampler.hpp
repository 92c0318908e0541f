/* The reference's sampler extension point on the MI355X engine (C++/HIP; the
 * C ABI is include/nmlmc.h).
 *
 * The reference defines a problem as a class SAMPLER derived from
 * inner_sampler_base (nested_sampler.hpp:36-85) with the members its engine
 * calls (nested_sampler.hpp:102-275, 369-485; model::sampler,
 * model_sampler.cpp:42-92, and portfolio::sampler, portfolio_sampler.cpp:
 * 227-615, are the two it ships):
 *
 *   static uint32_t GetWorkspaceSize(const data&);   reals of workspace per sampler
 *   SAMPLER(void* workspace);
 *   void sample_outer(uint64_t ell, work_t&);         the outer draw
 *   void copy_outer(const SAMPLER&);                  take another sampler's outer state
 *   real_t init_inner_sampler(uint64_t& count, work_t&);  the inner mean's offset
 *   real_t sample_inner(uint64_t i, uint64_t count, work_t&);  one inner sample
 *   static void add_phi(real_t avg, real_t coeff, real_t* phi);
 *
 * and drives it with init_samplers<SAMPLER>(seed, workspace) and
 * sample<SAMPLER>(ell, override_dyn_inner, M, fsums, diffsums, total_work,
 * inner_stats) (nested_sampler.hpp:618-694).  The same class, written against
 * nmlmc::inner_sampler_base below with __device__ members, runs here on the
 * batched engine's FP64 form through nmlmc::init_samplers<SAMPLER> and
 * nmlmc::sample<SAMPLER>, with the reference's arguments and output layout.
 *
 * Draws: outer sample g of level ell owns the key {seed, ell << 46 | g} (ids
 * continue across the calls of a level); sample_outer draws from that key's
 * stream at counters {j, 0}; inner sample u of draw group `tag` (fine test
 * round a: 1 + a, coarse test round a: 33 + a, final: 65, plain coarse final:
 * 66) draws from counters {j, tag << 40 | u}.  Both of the reference's roles
 * (randn/randu and randn_sync/randu_sync) read the stream bound for the call,
 * in the order the sampler makes them: a sample's draws are a pure function of
 * its key and counter, so any lane can compute any sample.
 *
 * Inner samples run, as in the reference's child grids (nested_sampler.hpp:
 * 296-298), on a sampler constructed on the outer sample's workspace that took
 * the outer state with copy_outer from the sampler that drew it.
 *
 * Restrictions of the batched engine: one scalar phi (gc_phi_size = 1); the
 * sampler's state after sample_outer lives in its members and its workspace;
 * init_inner_sampler must be a deterministic function of that state (its value
 * and its work are taken once per outer sample, as both reference samplers
 * allow); sample_inner is called with the sample's index within its draw group
 * and the group's size.
 */
#pragma once
#include <algorithm>
#include <cstring>
#include <new>

#include "../nested-risk-mlmc_amd/csrc/batched_kernels.hpp"
#include "../nested-risk-mlmc_amd/csrc/host.hpp"

namespace nmlmc {

typedef double real_t;
typedef double work_t;
typedef double sum_t;

// inner_sampler_base (nested_sampler.hpp:36-85): the draws of a sampler.
class inner_sampler_base {
public:
    __device__ real_t randn() { return rs_->randn(); }
    __device__ real_t randu() { return rs_->randu(); }
    __device__ real_t randn_sync() { return rs_->randn(); }
    __device__ real_t randu_sync() { return rs_->randu(); }
    // The engine binds the stream of the call in progress.
    __device__ void bind_stream(DStream* s) { rs_ = s; }

private:
    DStream* rs_ = nullptr;
};

// The batched engine's policy for a sampler class S (FP64 form).  Per outer
// sample the engine keeps S's object (after sample_outer) and its workspace in
// one record of `words` doubles; ItemState::Y holds the work that one
// init_inner_sampler call adds (charged once per draw group, as the reference
// calls it once per sample_inner_sum / _mean / _delta2).
template <class S>
struct SamplerPolicy {
    struct Args {
        BatchedCommon c;
        double* store;       // [M x words]: S object, then its workspace
        uint32_t words;      // doubles per outer sample
        uint32_t obj_words;  // doubles of the S object
    };
    static constexpr uint64_t spu = 1;
    static constexpr bool kMulti = false;
    static constexpr bool kFuse = true;
    static constexpr bool kMixedWaves = false;
    static constexpr bool kTwoValued = false;

    __device__ static S* object(const Args& A, uint64_t m) { return reinterpret_cast<S*>(A.store + m * A.words); }
    __device__ static uint64_t key1(const Args& A, uint64_t m) {
        return ((uint64_t)A.c.ell << 46) | (A.c.offset + m);
    }
    __device__ static double group_work(const Args&, const ItemState& it) { return it.Y; }
    __device__ static double phi(const BatchedCommon&, double u, double coeff, double acc) {
        real_t p = acc;
        S::add_phi(u, coeff, &p);
        return p;
    }
    // The FP64 form only: the mixed form's bounds are never asked for.
    __device__ static double mixed_bound(double, double) { return 0.; }
    __device__ static void decision_bounds(double, double, double, double, double& e1, double& e2) {
        e1 = e2 = INFINITY;
    }
    static hipError_t launch_outer(const Args&, hipStream_t) { return hipSuccess; }

    // kernel_sample's sample_outer (nested_sampler.hpp:419-421) and
    // init_inner_sampler, on the outer sample's stream.
    template <class ACC>
    __device__ static void init_item(const Args& A, uint64_t m, ItemState& it, ACC& acc) {
        S* so = ::new (object(A, m)) S(A.store + m * A.words + A.obj_words);
        DStream st;
        st.init(A.c.seed, key1(A, m), 0, 0, 0., 0);
        so->bind_stream(&st);
        work_t w = 0.;
        so->sample_outer(A.c.ell, w);
        uint64_t count = 0;
        work_t wi = 0.;
        it.start = so->init_inner_sampler(count, wi);
        so->bind_stream(nullptr);
        it.Y = wi;
        it.inst = 0;
        acc.work(0, w);
        acc.st(0, 4, st.nu);
        acc.st(0, 5, st.nn);
    }

    // inner_sampler::sample_inner (nested_sampler.hpp:131-161) over units
    // [u0, u1) of draw group `tag` (samples [i0, i1)).
    template <bool F32>
    __device__ static void units(const Args& A, uint32_t m, const ItemState&, uint32_t tag, uint64_t u0,
                                 uint64_t u1, uint64_t i0, uint64_t i1, SegAcc& s) {
        static_assert(!F32, "sampler classes run in the FP64 form");
        S so(A.store + (uint64_t)m * A.words + A.obj_words);
        so.copy_outer(*object(A, m));
        const uint64_t k1 = key1(A, m);
        double a1 = 0., a2 = 0.;
        work_t w = 0.;
        for (uint64_t u = u0; u < u1; u++) {
            DStream st;
            st.init(A.c.seed, k1, ((uint64_t)tag << 40) | u, 0, 0., 0);
            so.bind_stream(&st);
            const double d = so.sample_inner(u - i0, i1 - i0, w);
            a1 += d;
            a2 += d * d;
            s.nu += st.nu;
            s.nn += st.nn;
        }
        s.s1 += a1;
        s.s2 += a2;
        s.work += w;
    }
};

// Engine state of a staged sampler class: which class, and the record size
// per outer sample.
struct SamplerStage {
    uint32_t words = 0, obj_words = 0;
    const void* tag = nullptr;
};
inline SamplerStage& sampler_stage() {
    static SamplerStage st;
    return st;
}
template <class S>
const void* sampler_tag() {
    static const char t = 0;
    return &t;
}

// init_samplers<SAMPLER> (nested_sampler.hpp:618-644): stages sampler class S
// as the engine's problem, with `workspace` reals (S::GetWorkspaceSize) per
// sampler and the given seed.  Needs init_nestedmc_sampler first; replaces any
// staged problem.  Returns an NMLMC_* status.
template <class S>
int init_samplers(uint64_t seed, uint32_t workspace) {
    if (!g_engine.inited) return set_error(NMLMC_ERR_STATE, "init_nestedmc_sampler has not been called");
    int rc = g_engine.ensure_device();
    if (rc) return rc;
    g_engine.free_portfolio();
    g_engine.problem = PROBLEM_SAMPLER_CLASS;
    g_engine.seed = seed;
    g_engine.next_outer.clear();
    SamplerStage& st = sampler_stage();
    st.obj_words = (uint32_t)((sizeof(S) + 7) / 8);
    st.words = st.obj_words + workspace;
    st.tag = sampler_tag<S>();
    return set_error(NMLMC_OK, "");
}

// sample<SAMPLER> (nested_sampler.hpp:646-694): one level call of M outer
// samples of the staged class S; fine-only when ell == 0 or diffsums is NULL
// (as the reference); outputs overwritten.  override_dyn_inner is accepted and
// ignored (there are no child grids to size).  The batched FP64 form.
template <class S>
int sample(uint32_t ell, uint64_t override_dyn_inner, uint64_t* M, sum_t* fsums, sum_t* diffsums,
           work_t* total_work, uint64_t* inner_stats) {
    (void)override_dyn_inner;
    const SamplerStage& st = sampler_stage();
    if (!g_engine.inited) return set_error(NMLMC_ERR_STATE, "init_nestedmc_sampler has not been called");
    if (g_engine.problem != PROBLEM_SAMPLER_CLASS || st.tag != sampler_tag<S>())
        return set_error(NMLMC_ERR_STATE, "this sampler class is not the staged one (init_samplers<SAMPLER>)");
    if (!M) return set_error(NMLMC_ERR_ARG, "M is NULL");
    if (ell > 40) return set_error(NMLMC_ERR_ARG, "level %u out of range", ell);
    sample_output out{};
    out.M = *M;
    out.fsums = fsums;
    out.diffsums = diffsums;
    const uint32_t ell_eff = diffsums ? ell : 0;
    if (ell_eff >= g_engine.next_outer.size()) g_engine.next_outer.resize(ell_eff + 1, 0);
    const uint64_t off = g_engine.next_outer[ell_eff];
    int rc = NMLMC_OK;
    if (*M == 0) {
        std::memset(g_engine.h_out, 0, sizeof(Partial));
        g_engine.copy_out(ell_eff, &out);
    } else {
        typename SamplerPolicy<S>::Args A;
        rc = g_engine.batched_common(ell_eff, *M, off, A.c);
        if (rc) return rc;
        const uint64_t chunk = std::min<uint64_t>(*M, g_engine.batched_chunk());
        if ((rc = g_engine.ensure_R((size_t)chunk * st.words))) return rc;
        A.store = g_engine.pf_R;
        A.words = st.words;
        A.obj_words = st.obj_words;
        A.c.phi_type = 0;
        const Engine::Shape sh = g_engine.batched_shape(ell_eff, *M, 1, A.c, 0);
        A.c.S = sh.S;
        const uint32_t blocks = (uint32_t)g_engine.device_cus() * sh.bpc;
        g_engine.last_kernel = "k_b_final";
        rc = g_engine.run_batched(ell_eff, *M, A.c, &out, [&]() {
            typename SamplerPolicy<S>::Args B = A;
            B.c = A.c;
            return launch_batched_level<SamplerPolicy<S>, false>(B, blocks, g_engine.stream);
        }, nullptr, nullptr, nullptr, nullptr);
        if (rc) return rc;
    }
    g_engine.next_outer[ell_eff] = off + *M;
    if (total_work) *total_work = out.total_work;
    if (inner_stats)
        for (int i = 0; i < 6; i++) inner_stats[i] = out.inner_stats[i];
    return set_error(NMLMC_OK, "");
}

}  // namespace nmlmc
