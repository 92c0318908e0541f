// Batched forms of the model problem: the policy of the batched engine.
//
// Keying: outer sample g of level l owns the key {seed, (l << 46) | g}.  Its
// outer draw is the first normal of the stream at counter {0, 0}; inner sample
// i of draw group `tag` takes its Box-Muller pair from counter
// {0, (tag << 40) | (i / spu)} (spu = inner samples per pair: 2 for the shift
// model, 1 for the reference model).  Tags: fine test round a -> 1 + a, coarse test
// round a -> 33 + a, final draws -> 65 (66 for the coarse final of the plain
// fine/coarse estimator).
#include "batched_kernels.hpp"
#include "model.hpp"

namespace nmlmc {

template <int VAR>
struct ModelPolicy {
    struct Args {
        BatchedCommon c;
        ModelParams mp;
    };
    static constexpr uint64_t spu = (VAR == MODEL_SHIFT) ? 2 : 1;
    static constexpr bool kMulti = false;
    static constexpr bool kFuse = true;  // k_b_fused is instantiated
    static constexpr bool kMixedWaves = true;
    static constexpr bool kTwoValued = false;

    // init_inner_sampler does no work for the model (model_sampler.cpp:64).
    __device__ static double group_work(const Args&, const ItemState&) { return 0.; }

    // add_phi (model_sampler.cpp:81-91): coeff times the Indicator or Max of u, added to acc.
    __device__ static double phi(const BatchedCommon& c, double u, double coeff, double acc) {
        return phi_scalar(c.phi_type, u, coeff, acc);
    }

    // Mixed form: an FP32 mean within this distance of the threshold is
    // re-evaluated in FP64 (mean_mag: mean of the inner terms' magnitudes).
    __device__ static double mixed_bound(double mean_mag, double start) {
        return kMixedBound * (mean_mag + fabs(start));
    }

    // Mixed form, refinement decisions: bounds on |s1 - s1_64| and |s2 - s2_64|
    // of n FP32 terms (their error scales sum to sab, squared to smm) against
    // the FP64 sums of the same draws.  Each FP32 term is within
    // e_i = 2^-19 (mag_i + 1) of its FP64 value (normals within 2^-20.5, float
    // rounding of Y and of the products 2^-23 of mag), so |s1 - s1_64| is at
    // most 2^-19 (sab + n), inside the threshold bound plus 2^-19 n, and
    // |s2 - s2_64| <= sum(2 |d_i| e_i + e_i^2) with |d_i| <= mag_i.
    __device__ static void decision_bounds(double n, double start, double sab, double smm, double& e1, double& e2) {
        e1 = n * mixed_bound(sab / n, start) + 0x1p-19 * n;
        e2 = 0x1p-17 * (smm + sab) + 0x1p-36 * n;
    }

    __device__ static uint64_t key1(const Args& A, uint64_t m) {
        return ((uint64_t)A.c.ell << 46) | (A.c.offset + m);
    }

    // The outer draw is one normal, taken in k_b_init.
    static hipError_t launch_outer(const Args&, hipStream_t) { return hipSuccess; }

    // model::sampler::sample_outer (model_sampler.cpp:57-59) and
    // init_inner_sampler (64).
    template <class ACC>
    __device__ static void init_item(const Args& A, uint64_t m, ItemState& it, ACC& acc) {
        const u64x2 w = threefry2x64_20(0, 0, A.c.seed, key1(A, m));
        double x, y;
        boxmuller_f64(w.a, w.b, x, y);
        it.Y = x;
        it.start = -A.mp.max_loss;
        it.inst = 0;
        acc.st(0, 5, 1);
    }

    // Inner samples of units [u0, u1) restricted to samples [i0, i1)
    // (model::sampler::sample_inner, model_sampler.cpp:66-79, and the shift
    // variant X = Y + Z).
    template <bool F32>
    __device__ static void units(const Args& A, uint32_t m, const ItemState& it, uint32_t tag,
                                 uint64_t u0, uint64_t u1, uint64_t i0, uint64_t i1, SegAcc& s) {
        const uint64_t k1 = key1(A, m);
        const double Y = it.Y;
        uint64_t cnt = 0;
        if (!F32) {
            double a1 = 0., a2 = 0.;
            for (uint64_t u = u0; u < u1; u++) {
                const u64x2 w = threefry2x64_20(0, ((uint64_t)tag << 40) | u, A.c.seed, k1);
                double x, y;
                boxmuller_f64(w.a, w.b, x, y);
                if (VAR == MODEL_REF) {
                    const double d = model_ref_inner(A.mp, Y, x, y);
                    a1 += d;
                    a2 += d * d;
                    cnt++;
                } else {
                    const uint64_t sm = 2 * u;
                    if (sm >= i0) {
                        const double d = Y + x;
                        a1 += d;
                        a2 += d * d;
                        cnt++;
                    }
                    if (sm + 1 < i1) {
                        const double d = Y + y;
                        a1 += d;
                        a2 += d * d;
                        cnt++;
                    }
                }
            }
            s.s1 += a1;
            s.s2 += a2;
        } else {
            const float Yf = (float)Y;
            const float sa = (float)A.mp.stau, sb = (float)A.mp.s1tau;
            Kahan b1, b2, bab;
            float m2 = 0.f;
            for (uint64_t u = u0; u < u1; u++) {
                const u64x2 w = threefry2x64_20(0, ((uint64_t)tag << 40) | u, A.c.seed, k1);
                float x, y;
                boxmuller_f32(w.a, w.b, x, y);
                if (VAR == MODEL_REF) {
                    const float p = sa * Yf + sb * x, q = sa * y + sb * x;
                    const float d = p * p - q * q;
                    b1.add(d);
                    b2.add(d * d);
                    const float mg = p * p + q * q;
                    bab.add(mg);
                    m2 += mg * mg;
                    cnt++;
                } else {
                    const uint64_t sm = 2 * u;
                    if (sm >= i0) {
                        const float d = Yf + x, mg = fabsf(Yf) + fabsf(x);
                        b1.add(d);
                        b2.add(d * d);
                        bab.add(mg);
                        m2 += mg * mg;
                        cnt++;
                    }
                    if (sm + 1 < i1) {
                        const float d = Yf + y, mg = fabsf(Yf) + fabsf(y);
                        b1.add(d);
                        b2.add(d * d);
                        bab.add(mg);
                        m2 += mg * mg;
                        cnt++;
                    }
                }
            }
            s.s1 += b1.value();
            s.s2 += b2.value();
            s.sab += bab.value();
            s.smm += m2;
        }
        s.work += (double)cnt;
        s.nn += cnt * (VAR == MODEL_REF ? 2 : 1);
    }
};

hipError_t launch_model_batched(const ModelBatchedHost& H, hipStream_t s) {
    if (H.mp.variant == MODEL_REF) {
        ModelPolicy<MODEL_REF>::Args A{H.c, H.mp};
        return H.mixed ? launch_batched_level<ModelPolicy<MODEL_REF>, true>(A, H.blocks, s)
                       : launch_batched_level<ModelPolicy<MODEL_REF>, false>(A, H.blocks, s);
    }
    ModelPolicy<MODEL_SHIFT>::Args A{H.c, H.mp};
    return H.mixed ? launch_batched_level<ModelPolicy<MODEL_SHIFT>, true>(A, H.blocks, s)
                   : launch_batched_level<ModelPolicy<MODEL_SHIFT>, false>(A, H.blocks, s);
}

}  // namespace nmlmc
