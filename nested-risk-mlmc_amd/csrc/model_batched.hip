// Batched forms of the model problem: per-outer-sample keys, persistent grid,
// device work queue.
//
// Keying: outer sample g of level l owns the key {seed, (l << 46) | g}.  Its
// outer draw is the normal-only stream at counter {j, 0}; inner sample i of
// draw group `tag` takes its Box-Muller pair from counter {i / spu, tag << 40}
// (spu = inner samples per pair: 2 for the shift model, 1 for the reference
// model).  Tags: fine test round a -> 1 + a, coarse test round a -> 33 + a,
// final draws -> 65 (and 66 for the coarse final of the non-antithetic
// estimator).  Every draw is a pure function of (key, counter), so an outer
// sample can be processed by any lanes in any order.
//
// Work distribution: a lane group of G lanes (G | 64) pulls outer samples from
// a device counter; each outer sample runs as a flat state machine
// (nested_sampler.hpp:369-485 per sample: fine adaptive count, coarse adaptive
// count, antithetic or plain final draws) in which every loop iteration gives
// each lane exactly one Box-Muller pair of the group's current draw group, so
// groups at different phases, refinement depths or outer samples still execute
// the same instruction stream.  A refined outer sample keeps its group and
// re-enters the draw loop at doubled count: the "re-enqueue at 2N" of the
// work queue without a round trip through memory.
//
// MIXED: inner arithmetic in FP32 (Box-Muller from the same 64-bit words,
// compensated sums); refinement decisions and level sums in FP64; a final draw
// whose mean lies within a bound of the Heaviside threshold is re-evaluated in
// FP64 from the same draws, and FP32/FP64 indicator disagreements are counted.
#include "engine.hpp"
#include "model.hpp"

namespace nmlmc {

enum {
    PH_NEXT = 0,   // fetch the next outer sample
    PH_FT_BEGIN,   // begin a fine test round
    PH_FT,         // fine test round drawing
    PH_CT_BEGIN,
    PH_CT,
    PH_FIN_BEGIN,  // begin the final draws
    PH_FIN,        // antithetic / fine-only final draws, group gi
    PH_NA_F,       // non-antithetic fine final
    PH_NA_C,       // non-antithetic coarse final
    PH_EXIT
};

constexpr uint32_t TAG_FINE_TEST = 1, TAG_COARSE_TEST = 33, TAG_FINAL = 65, TAG_FINAL_COARSE = 66;

// Relative bound (times the mean magnitude of the inner terms) on the FP32
// inner mean's error; 16x above the worst case seen for 2^-22-accurate normals
// and compensated sums.
constexpr double kMixedBound = 0x1p-16;

struct Kahan {
    float s = 0.f, c = 0.f;
    __device__ __forceinline__ void add(float v) {
        const float y = v - c;
        const float t = s + y;
        c = (t - s) - y;
        s = t;
    }
    __device__ __forceinline__ double value() const { return (double)s - (double)c; }
};

template <int VAR, int G, bool MIXED>
__global__ __launch_bounds__(256) void k_model_batched(ModelBatchedArgs A) {
    const uint32_t lane = threadIdx.x & 63;
    const uint32_t gl = lane % G;
    const uint32_t leader = lane - gl;
    constexpr int spu = (VAR == MODEL_SHIFT) ? 2 : 1;
    constexpr int nps = (VAR == MODEL_REF) ? 2 : 1;
    const ModelParams& mp = A.mp;
    const double start = -mp.max_loss;
    const bool fine_only = (A.ell == 0);
    const bool anti = A.ep.antithetic != 0;

    double f_sums[kMaxMoments] = {0, 0, 0, 0}, d_sums[kMaxMoments] = {0, 0, 0, 0};
    uint64_t st[kStatsExt] = {0, 0, 0, 0, 0, 0, 0, 0};

    // Outer-sample state (group-uniform).
    uint64_t m = 0, k1 = 0;
    double Y = 0.;
    float Yf = 0.f;
    int phase = PH_NEXT;
    uint32_t round = 0, tag = 0;
    uint64_t cur = 0, nf = 0, nc = 0;
    uint64_t big = 0, small = 0, groups = 0, gi = 0;
    bool f_big = true, fp64_pass = false;
    double s1_big = 0., sab_big = 0., phi_small = 0., coeff = 1., phi_f = 0., phi_c = 0.;
    double phi32_f = 0., phi32_c = 0.;
    bool near = false;
    // Current draw group: samples [i0, i1), iteration it of nit.
    uint64_t i0 = 0, i1 = 0, u0 = 0, nit = 0, it = 0;
    // Lane partial sums.
    double a1 = 0., a2 = 0., aab = 0.;
    Kahan b1, b2, bab;

    while (true) {
        if (it < nit) {
            // ---- one Box-Muller pair per lane ----
            const uint64_t u = u0 + gl + (uint64_t)G * it;
            const uint64_t s = u * spu;
            if (s < i1) {
                const u64x2 w = threefry2x64_20(u, (uint64_t)tag << 40, A.seed, k1);
                if (!MIXED || fp64_pass) {
                    double x, y;
                    boxmuller_f64(w.a, w.b, x, y);
                    if (VAR == MODEL_REF) {
                        const double d = model_ref_inner(mp, Y, x, y);
                        a1 += d;
                        a2 += d * d;
                    } else {
                        if (s >= i0) {
                            const double d = Y + x;
                            a1 += d;
                            a2 += d * d;
                        }
                        if (s + 1 < i1) {
                            const double d = Y + y;
                            a1 += d;
                            a2 += d * d;
                        }
                    }
                } else {
                    float x, y;
                    boxmuller_f32(w.a, w.b, x, y);
                    if (VAR == MODEL_REF) {
                        const float sa = (float)mp.stau, sb = (float)mp.s1tau;
                        const float p = sa * Yf + sb * x, q = sa * y + sb * x;
                        const float d = p * p - q * q;
                        b1.add(d);
                        b2.add(d * d);
                        bab.add(p * p + q * q);
                    } else {
                        if (s >= i0) {
                            const float d = Yf + x;
                            b1.add(d);
                            b2.add(d * d);
                            bab.add(fabsf(Yf) + fabsf(x));
                        }
                        if (s + 1 < i1) {
                            const float d = Yf + y;
                            b1.add(d);
                            b2.add(d * d);
                            bab.add(fabsf(Yf) + fabsf(y));
                        }
                    }
                }
            }
            it++;
            continue;
        }
        // ---- end of a draw group: reduce and advance the state machine ----
        double r1 = 0., r2 = 0., rab = 0.;
        if (phase != PH_NEXT && phase != PH_EXIT) {
            if (!MIXED || fp64_pass) {
                r1 = group_sum<G>(a1);
                r2 = group_sum<G>(a2);
            } else {
                r1 = group_sum<G>(b1.value());
                r2 = group_sum<G>(b2.value());
                rab = group_sum<G>(bab.value());
            }
            a1 = a2 = aab = 0.;
            b1 = Kahan();
            b2 = Kahan();
            bab = Kahan();
        }
        // Transitions that need no draws run back to back until a draw group is
        // set up or the queue is empty.
        bool ready = false;
        while (!ready) {
            switch (phase) {
            case PH_NEXT: {
                unsigned long long nm = 0;
                if (gl == 0) nm = atomicAdd(A.queue, 1ull);
                nm = __shfl(nm, leader, 64);
                if (nm >= A.M) {
                    phase = PH_EXIT;
                    ready = true;
                    break;
                }
                m = nm;
                k1 = ((uint64_t)A.ell << 46) | (A.offset + m);
                // Outer draw: first normal of the stream {0, 0} (model_sampler.cpp:57-59).
                const u64x2 w = threefry2x64_20(0, 0, A.seed, k1);
                double x, y;
                boxmuller_f64(w.a, w.b, x, y);
                Y = x;
                Yf = (float)x;
                st[5] += 1;
                cur = A.cf.base;
                round = 0;
                fp64_pass = false;
                phase = PH_FT_BEGIN;
                break;
            }
            case PH_FT_BEGIN:
            case PH_CT_BEGIN: {
                const bool fine = phase == PH_FT_BEGIN;
                const LevelCounts& lc = fine ? A.cf : A.cc;
                if (lc.max <= 2 * cur) {
                    cur = lc.max;
                    if (fine) {
                        nf = cur;
                        st[1] += 1;
                        st[2] += (cur == lc.max);
                        st[3] += (cur == lc.base);
                        if (fine_only) phase = PH_FIN_BEGIN;
                        else {
                            cur = A.cc.base;
                            round = 0;
                            phase = PH_CT_BEGIN;
                        }
                    } else {
                        nc = cur;
                        phase = PH_FIN_BEGIN;
                    }
                    break;
                }
                st[0] += cur;
                st[5] += cur * nps;
                tag = (fine ? TAG_FINE_TEST : TAG_COARSE_TEST) + round;
                i0 = 0;
                i1 = cur;
                phase = fine ? PH_FT : PH_CT;
                ready = true;
                break;
            }
            case PH_FT:
            case PH_CT: {
                // sample_inner_delta2 + the refinement test (nested_sampler.hpp:120-129, 252-263).
                const bool fine = phase == PH_FT;
                const LevelCounts& lc = fine ? A.cf : A.cc;
                const double out_var = r2 / (double)(cur - 1) - (r1 * r1) / (double)(cur * (cur - 1));
                const double dist = start + r1 / (double)cur;
                const double delta2 = dist * dist / out_var;
                const double new_count =
                    (lc.rule == 1) ? lc.tmp / delta2 : pow(delta2, -A.ep.adaptive_r / 2) * lc.tmp;
                round++;
                if (new_count <= (double)cur) {
                    if (fine) {
                        nf = cur;
                        st[1] += 1;
                        st[2] += (cur == lc.max);
                        st[3] += (cur == lc.base);
                        if (fine_only) phase = PH_FIN_BEGIN;
                        else {
                            cur = A.cc.base;
                            round = 0;
                            phase = PH_CT_BEGIN;
                        }
                    } else {
                        nc = cur;
                        phase = PH_FIN_BEGIN;
                    }
                } else {
                    cur *= 2;
                    phase = fine ? PH_FT_BEGIN : PH_CT_BEGIN;
                }
                break;
            }
            case PH_FIN_BEGIN: {
                if (fine_only || anti) {
                    if (fine_only) {
                        big = small = nf;
                        f_big = true;
                    } else {
                        f_big = nf > nc;
                        big = f_big ? nf : nc;
                        small = f_big ? nc : nf;
                    }
                    groups = big / small;
                    coeff = (double)(1.f / (float)groups);
                    gi = 0;
                    s1_big = 0.;
                    sab_big = 0.;
                    phi_small = 0.;
                    near = false;
                    if (!fp64_pass) {
                        st[0] += big;
                        st[5] += big * nps;
                    }
                    tag = TAG_FINAL;
                    i0 = 0;
                    i1 = small;
                    phase = PH_FIN;
                } else {
                    if (!fp64_pass) {
                        st[0] += nf + nc;
                        st[5] += (nf + nc) * nps;
                    }
                    near = false;
                    tag = TAG_FINAL;
                    i0 = 0;
                    i1 = nf;
                    phase = PH_NA_F;
                }
                ready = true;
                break;
            }
            case PH_FIN: {
                // One antithetic group done (sample_phi_antithetic, nested_sampler.hpp:204-210).
                const double mean = start + r1 / (double)small;
                if (MIXED && !fp64_pass) near |= fabs(mean) <= kMixedBound * (rab / (double)small + fabs(start));
                phi_small = phi_scalar(mp.phi_type, mean, coeff, phi_small);
                s1_big += r1;
                sab_big += rab;
                gi++;
                if (gi < groups) {
                    i0 = gi * small;
                    i1 = i0 + small;
                    ready = true;
                    break;
                }
                const double mean_big = start + s1_big / (double)big;
                if (MIXED && !fp64_pass)
                    near |= fabs(mean_big) <= kMixedBound * (sab_big / (double)big + fabs(start));
                const double phi_big = phi_scalar(mp.phi_type, mean_big, 1., 0.);
                if (fine_only) {
                    phi_f = phi_big;
                    phi_c = 0.;
                } else {
                    phi_f = f_big ? phi_big : phi_small;
                    phi_c = f_big ? phi_small : phi_big;
                }
                phase = PH_EXIT + 1;  // done with this outer sample
                break;
            }
            case PH_NA_F: {
                const double mean = start + r1 / (double)nf;
                if (MIXED && !fp64_pass) near |= fabs(mean) <= kMixedBound * (rab / (double)nf + fabs(start));
                phi_f = phi_scalar(mp.phi_type, mean, 1., 0.);
                tag = TAG_FINAL_COARSE;
                i0 = 0;
                i1 = nc;
                phase = PH_NA_C;
                ready = true;
                break;
            }
            case PH_NA_C: {
                const double mean = start + r1 / (double)nc;
                if (MIXED && !fp64_pass) near |= fabs(mean) <= kMixedBound * (rab / (double)nc + fabs(start));
                phi_c = phi_scalar(mp.phi_type, mean, 1., 0.);
                phase = PH_EXIT + 1;
                break;
            }
            case PH_EXIT + 1: {
                // Outer sample complete.
                if (MIXED && near && !fp64_pass && mp.phi_type == PHI_INDICATOR) {
                    // Re-evaluate the final draws in FP64 from the same words.
                    phi32_f = phi_f;
                    phi32_c = phi_c;
                    fp64_pass = true;
                    phase = PH_FIN_BEGIN;
                    break;
                }
                if (MIXED && fp64_pass) {
                    st[6] += 1;
                    st[7] += (phi32_f != phi_f) || (phi32_c != phi_c);
                }
                add_moments(&phi_f, 1, f_sums, A.ep.fine_moments);
                if (!fine_only) {
                    const double dphi = phi_f - phi_c;
                    add_moments(&dphi, 1, d_sums, A.ep.diff_moments);
                }
                if (gl == 0) {
                    if (A.tr_phi_f) A.tr_phi_f[m] = phi_f;
                    if (A.tr_phi_c) A.tr_phi_c[m] = phi_c;
                    if (A.tr_nf) A.tr_nf[m] = nf;
                    if (A.tr_nc) A.tr_nc[m] = fine_only ? 0 : nc;
                }
                phase = PH_NEXT;
                break;
            }
            default:
                ready = true;
                break;
            }
        }
        if (phase == PH_EXIT) break;
        // Set up the iterations of the new draw group.
        u0 = i0 / spu;
        const uint64_t u1 = (i1 + spu - 1) / spu;
        nit = (u1 - u0 + G - 1) / G;
        it = 0;
    }
    const double work = (double)st[0];
    block_write_partial(gl == 0, f_sums, A.ep.fine_moments, d_sums, A.ep.diff_moments, work, st,
                        &A.partials[blockIdx.x]);
}

template <int VAR, bool MIXED>
static hipError_t launch_b(const ModelBatchedArgs& A, uint32_t G, uint32_t blocks, hipStream_t s) {
    const dim3 grid(blocks), block(256);
    switch (G) {
    case 1: hipLaunchKernelGGL((k_model_batched<VAR, 1, MIXED>), grid, block, 0, s, A); break;
    case 2: hipLaunchKernelGGL((k_model_batched<VAR, 2, MIXED>), grid, block, 0, s, A); break;
    case 4: hipLaunchKernelGGL((k_model_batched<VAR, 4, MIXED>), grid, block, 0, s, A); break;
    case 8: hipLaunchKernelGGL((k_model_batched<VAR, 8, MIXED>), grid, block, 0, s, A); break;
    case 16: hipLaunchKernelGGL((k_model_batched<VAR, 16, MIXED>), grid, block, 0, s, A); break;
    case 32: hipLaunchKernelGGL((k_model_batched<VAR, 32, MIXED>), grid, block, 0, s, A); break;
    default: hipLaunchKernelGGL((k_model_batched<VAR, 64, MIXED>), grid, block, 0, s, A); break;
    }
    return hipGetLastError();
}

__global__ void k_zero_u64(unsigned long long* p, int n) {
    if ((int)threadIdx.x < n) p[threadIdx.x] = 0ull;
}

hipError_t launch_model_batched(const ModelBatchedArgs& A, uint32_t G, uint32_t blocks, bool mixed,
                                hipStream_t s) {
    // Queue reset by a kernel (not a memset) so the sequence can be captured in a graph.
    hipLaunchKernelGGL(k_zero_u64, dim3(1), dim3(64), 0, s, A.queue, 1);
    if (A.mp.variant == MODEL_REF)
        return mixed ? launch_b<MODEL_REF, true>(A, G, blocks, s) : launch_b<MODEL_REF, false>(A, G, blocks, s);
    return mixed ? launch_b<MODEL_SHIFT, true>(A, G, blocks, s) : launch_b<MODEL_SHIFT, false>(A, G, blocks, s);
}

}  // namespace nmlmc
