// Batched forms of the portfolio problem: the policy of the batched engine.
//
// Keying as for the model: outer sample g of level l owns the key
// {seed, (l << 46) | g}; its outer draw (portfolio::sampler::sample_outer,
// portfolio_sampler.cpp:293-311, or the Cholesky extension) reads the normal
// stream at counters {j, 0}; inner sample i of draw group `tag` is a
// random_state on counters {j, (tag << 40) | i}, which serves both the sync
// role (method choice, Unbiased level) and the inner role (option choice,
// normals) of that sample in the reference's order.
#include "batched_kernels.hpp"
#include "portfolio.hpp"
#include "portfolio_dev.hpp"

namespace nmlmc {

struct PortfolioPolicy {
    struct Args {
        BatchedCommon c;
        PortfolioDevice pf;
        double* R_tau;  // [M x d] outer state per outer sample
        double bs_work; // work of one init_inner_sampler call (sub-sampling off)
    };
    static constexpr uint64_t spu = 1;

    __device__ static double group_work(const Args& A) { return A.bs_work; }

    __device__ static uint64_t key1(const Args& A, uint64_t m) {
        return ((uint64_t)A.c.ell << 46) | (A.c.offset + m);
    }

    __device__ static void init_item(const Args& A, uint64_t m, ItemState& it, LaneAcc& acc) {
        const PortfolioDevice& pf = A.pf;
        double* R = A.R_tau + m * pf.d;
        DStream os;
        os.init(A.c.seed, key1(A, m), 0, 0, 0., 0);
        if (pf.chol) {
            // W = sqrt(tau) L z, one row at a time; z kept in R until replaced.
            for (uint32_t i = 0; i < pf.d; i++) R[i] = os.randn();
            for (int i = (int)pf.d - 1; i >= 0; i--) {
                double a = 0.;
                for (int j = 0; j <= i; j++) a += pf.chol[(size_t)i * pf.d + j] * R[j];
                const StockRec st = pf.stocks[i];
                R[i] = simulate_stock_d(st.S0, st.mu, st.sigma, pf.tau, pf.stau * a);
            }
        } else {
            const double dW_sys = pf.stau * os.randn();
            if (pf.d == 1) {
                const StockRec st = pf.stocks[0];
                R[0] = simulate_stock_d(st.S0, st.mu, st.sigma, pf.tau, dW_sys);
            } else {
                for (uint32_t i = 0; i < pf.d; i++) {
                    const double dW_ind = pf.stau * os.randn();
                    const double W = pf.rho * dW_sys + pf.s1_rho2 * dW_ind;
                    const StockRec st = pf.stocks[i];
                    R[i] = simulate_stock_d(st.S0, st.mu, st.sigma, pf.tau, W);
                }
            }
        }
        double cv = 0.;
        if (pf.use_delta_cv)
            for (uint32_t i = 0; i < pf.d; i++) cv += (R[i] - pf.stocks[i].S0) * pf.delta_coalesce[i];
        double w0 = 0.;
        it.start = det_bs_sum(pf, R, w0) + (-pf.max_loss - cv);
        it.Y = 0.;
        acc.st[5] += os.nn;
    }

    template <bool F32>
    __device__ static void units(const Args& A, uint32_t m, const ItemState& it, uint32_t tag, uint64_t u0,
                                 uint64_t u1, uint64_t, uint64_t, SegAcc& s) {
        const PortfolioDevice& pf = A.pf;
        const double* R = A.R_tau + (uint64_t)m * pf.d;
        const uint64_t k1 = key1(A, m);
        double a1 = 0., a2 = 0., ab = 0., work = 0.;
        Kahan b1, b2, bab;
        for (uint64_t u = u0; u < u1; u++) {
            DStream st;
            st.init(A.c.seed, k1, ((uint64_t)tag << 40) | u, 0, 0., 0);
            if (F32 && pf.subsample != 0) {
                uint32_t method, idx;
                const double prob = pick_option(pf, st, st, method, idx);
                const OptionRec o = pf.options[method][idx];
                if (method == 1) {
                    const uint64_t first = st.rand(), second = st.rand();
                    st.nn += 2;
                    float x, y, mag;
                    boxmuller_f32(second, first, x, y);
                    work += o.W_est;
                    const float ip = (float)(1. / prob);
                    const float d = inner_exact_f(pf, o, (float)R[o.stock_ind], x, y, mag) * ip;
                    b1.add(d);
                    b2.add(d * d);
                    bab.add(mag * ip);
                    s.nu += st.nu;
                    s.nn += st.nn;
                    continue;
                }
                double v = method == 0 ? inner_bs_d(pf, o, R, work) : inner_unbiased_d(pf, o, R, st, st, work);
                v /= prob;
                a1 += v;
                a2 += v * v;
                ab += fabs(v);
            } else {
                const double d = inner_sample_d(pf, R, st, st, work);
                a1 += d;
                a2 += d * d;
                ab += fabs(d);
            }
            s.nu += st.nu;
            s.nn += st.nn;
        }
        s.s1 += a1 + (F32 ? b1.value() : 0.);
        s.s2 += a2 + (F32 ? b2.value() : 0.);
        s.sab += F32 ? ab + bab.value() : 0.;
        s.work += work;
    }
};

hipError_t launch_portfolio_batched(const PortfolioBatchedHost& H, hipStream_t s) {
    PortfolioPolicy::Args A{H.c, H.pf, H.R_tau, H.bs_work};
    return H.mixed ? launch_batched_level<PortfolioPolicy, true>(A, H.blocks, s)
                   : launch_batched_level<PortfolioPolicy, false>(A, H.blocks, s);
}

}  // namespace nmlmc
