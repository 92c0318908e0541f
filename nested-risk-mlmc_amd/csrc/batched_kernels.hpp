// Kernels of the batched engine (see batched.hpp), templated on the problem
// policy P and on the precision of the inner arithmetic.
//
// Launch order of one level call (launch_batched_level):
//   [k_b_zero unless the control block is clear already], [P's outer-draw kernel], k_b_init and k_b_test(r) for r = 0..R-1
//   (or k_b_fused and k_b_test(r) for r = 1..R-1), k_b_final_count (its
//   last block lays the final pool out), k_b_final_fill, k_b_final, [k_b_redo_split and
//   k_b_redo_seq when MIXED]
#pragma once
#include "batched.hpp"

namespace nmlmc {

// Waves per SIMD the test and final kernels are compiled for.  Mixed form:
// NMLMC_MIXED_WAVES (the inner loops wait on gathers, so occupancy matters
// more than a few spills); FP64 forms: no cap.  A policy with P::kMixedWaves
// false (the general portfolio loop, whose FP64 paths would spill heavily
// under the mixed cap) takes NMLMC_GENERAL_WAVES in both.
#ifndef NMLMC_MIXED_WAVES
#define NMLMC_MIXED_WAVES 4
#endif
#ifndef NMLMC_GENERAL_WAVES
#define NMLMC_GENERAL_WAVES 2
#endif
// FP64 forms (the batched FP64 form, the mixed form's rechecks): no cap; their
// kernels hold 140-220 VGPRs, two or three waves a SIMD.  Capped at 3 the
// model's batched FP64 sweep, the batch's FP64 sweep and the large book's
// mixed sweep ran within 1% of uncapped; at 4 the batch ran 1.18x slower.
#ifndef NMLMC_FP64_WAVES
#define NMLMC_FP64_WAVES 1
#endif
template <class P, bool MIXED>
constexpr int mixed_waves() {
    if (!P::kMixedWaves) return NMLMC_GENERAL_WAVES;
    return MIXED ? NMLMC_MIXED_WAVES : NMLMC_FP64_WAVES;
}

__device__ __forceinline__ uint32_t class_of(uint64_t n, uint64_t base_min) {
    uint32_t c = 0;
    while ((base_min << c) < n && c + 1 < (uint32_t)kMaxClasses) c++;
    return c;
}

// sample_inner_delta2 + the refinement test (nested_sampler.hpp:120-129, 252-263):
// true when the outer sample needs more inner samples than `cur`.
// RULE 0 is the reference's rule (a power of delta^2), RULE 1 the capped
// 3-sigma doubling; a compile-time choice keeps pow out of the rule-1 kernels.
template <int RULE>
__device__ __forceinline__ bool refine(const BatchedCommon& c, const LevelCounts& lc, double start,
                                       uint64_t cur, double s1, double s2) {
    const double out_var = s2 / (double)(cur - 1) - (s1 * s1) / (double)(cur * (cur - 1));
    const double dist = start + s1 / (double)cur;
    const double delta2 = dist * dist / out_var;
    const double new_count = (RULE == 1) ? lc.tmp / delta2 : pow(delta2, -c.ep.adaptive_r / 2) * lc.tmp;
    return !(new_count <= (double)cur);
}

// The count an outer sample moves to after a finished test group at `cur`
// (get_samples_count's loop body, nested_sampler.hpp:252-270): 2 cur with
// more = true while it refines, else the decided count (lc.max once the next
// doubling would reach it).
template <int RULE>
__device__ __forceinline__ uint64_t next_count(const BatchedCommon& c, const LevelCounts& lc, double start,
                                               uint64_t cur, double s1, double s2, bool& more) {
    uint64_t n = cur;
    more = refine<RULE>(c, lc, start, cur, s1, s2);
    if (more) {
        n = cur * 2;
        if (lc.max <= 2 * n) {
            n = lc.max;
            more = false;
        }
    }
    return n;
}

// Outer draws.  Round 0 tests every outer sample, so its queue is implicit
// (entry e is outer sample e); its entries are written zeroed when k_b_test
// may split a round-0 group over several segments: a group of more units than
// c.S, or than c.s_test_min when sparse rounds may shrink their segments to it.
template <class P>
__global__ __launch_bounds__(256) void k_b_init(typename P::Args A) {
    const BatchedCommon& c = A.c;
    BlockAcc<P::kMulti> acc;
    acc.template init<false>(c);
    const bool test_f = c.cf.max > 2 * c.cf.base;
    const bool test_c = c.ell > 0 && c.cc.max > 2 * c.cc.base;
    const uint64_t split_above = (c.s_test_min && c.s_test_min < c.S) ? c.s_test_min : c.S;
    const bool acc_f = test_f && (c.cf.base + P::spu - 1) / P::spu > split_above;
    const bool acc_c = test_c && (c.cc.base + P::spu - 1) / P::spu > split_above;
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        c.ctl->cnt_f[0] = test_f ? (uint32_t)c.M : 0;
        c.ctl->cnt_c[0] = test_c ? (uint32_t)c.M : 0;
    }
    const XcdRange xr = xcd_range(c.M, c.xcd_sliced);
    for (uint64_t m = xr.k0; m < xr.end; m += xr.stride) {
        ItemState& it = c.items[m];
        P::init_item(A, m, it, acc);
        const uint32_t q = it.inst;
        uint64_t nf = 0, nc = 0;
        if (!test_f) {
            nf = c.cf.max;
            acc.st(q, 1, 1);
            acc.st(q, 2, 1);
            acc.st(q, 3, c.cf.max == c.cf.base);
        } else {
            acc.st(q, 0, c.cf.base);
        }
        if (c.ell > 0) {
            if (!test_c) nc = c.cc.max;
            else acc.st(q, 0, c.cc.base);
        }
        it.nf = nf;
        it.nc = nc;
        if (acc_f) c.tf[0][m] = TestEntry{0., 0., 0., 0., (uint32_t)m, 0};
        if (acc_c) c.tc[0][m] = TestEntry{0., 0., 0., 0., (uint32_t)m, 0};
    }
    acc.template flush<false>(c);
}

// Mixed form: true when the refinement decision taken from FP32-accumulated
// sums may differ from the one the FP64 sums of the same draws give (the
// batched FP64 form's).  With |s1 - s1_64| <= e1 and |s2 - s2_64| <= e2
// (P::decision_bounds, plus 2^-21 s2 for the compensated sums' own rounding),
// |dist - dist_64| <= ed = 2 e1 / n and |var - var_64| <= ev (both doubled to
// leave room for the FP64 side's rounding), so delta^2_64 lies in
// [qlo, qhi] = [max(|dist| - ed, 0)^2 / (var + ev), (|dist| + ed)^2 / (var - ev)]
// and the FP64 count new_64 (decreasing in delta^2, as its power -1 or -r/2)
// in [nlo, nhi]: the decision is uncertain when that interval straddles cur.
// An outer sample whose mean sits within ed of the threshold has qlo = 0 and
// nhi infinite: it is certain to refine once nlo > cur, which is the common
// case (its delta^2 is near 0).  With ev at half of var or more (or a bound
// that is NaN or infinite) the decision goes to the recheck.  Shortcut: when
// both bounds are under a tenth of dist and var, delta^2_64 is within 40% of
// delta^2 and new_64 within a factor 1.7 of new (r < 2), so a decision with
// new at least a factor 2 from cur is certain without the interval.  The
// uncertain ones go to the FP64 recheck (k_b_test with RECHECK).
template <class P, int RULE>
__device__ __forceinline__ bool decision_uncertain(const BatchedCommon& c, const LevelCounts& lc, double start,
                                                   uint64_t cur, double s1, double s2, double sab, double smm) {
    const double n = (double)cur;
    double e1, e2;
    P::decision_bounds(n, start, sab, smm, e1, e2);
    e2 += 0x1p-21 * fabs(s2);
    const double var = s2 / (n - 1.) - (s1 * s1) / (n * (n - 1.));
    const double dist = start + s1 / n;
    const double ad = fabs(dist);
    const double ed = 2. * e1 / n;
    const double ev = 2. * (e2 + (2. * fabs(s1) + e1) * e1 / n) / (n - 1.);
    if (!(ed < INFINITY) || !(ev < 0.5 * var)) return true;  // also NaN bounds and var <= 0
    if (ed < 0.1 * ad && ev < 0.1 * var) {
        const double q = dist * dist / var;
        const double nc = (RULE == 1) ? lc.tmp / q : pow(q, -c.ep.adaptive_r / 2) * lc.tmp;
        if (nc >= 2. * n || nc <= 0.5 * n) return false;
    }
    const double lo = fmax(ad - ed, 0.);
    const double qlo = lo * lo / (var + ev), qhi = (ad + ed) * (ad + ed) / (var - ev);
    double nlo, nhi;
    if (RULE == 1) {
        nlo = lc.tmp / qhi;
        nhi = lc.tmp / qlo;
    } else {
        nlo = pow(qhi, -c.ep.adaptive_r / 2) * lc.tmp;
        nhi = pow(qlo, -c.ep.adaptive_r / 2) * lc.tmp;
    }
    return nlo <= n && n < nhi;
}

// Test round r of every outer sample still refining its fine or coarse count.
// MIXED: inner sums in FP32; a decision within the FP32 error of its threshold
// is not taken here but listed for the recheck.  RECHECK (after round r, mixed
// form only): the listed groups' sums again in FP64 from the same draws, and
// their decisions; their work and draw counts were taken in the FP32 pass.
template <class P, bool MIXED, int RULE, bool RECHECK = false>
__global__ __launch_bounds__(256, (mixed_waves<P, MIXED>())) void k_b_test(typename P::Args A, uint32_t r) {
    static_assert(!(MIXED && RECHECK), "the recheck runs in FP64");
    const BatchedCommon& c = A.c;
    const uint32_t nfe = aload_u32(RECHECK ? &c.ctl->rc_f[r] : &c.ctl->cnt_f[r]);
    const uint32_t nce = aload_u32(RECHECK ? &c.ctl->rc_c[r] : &c.ctl->cnt_c[r]);
    const uint64_t cur_f = c.cf.base << r, cur_c = c.cc.base << r;
    const uint64_t uf = (cur_f + P::spu - 1) / P::spu, uc = (cur_c + P::spu - 1) / P::spu;
    // Units per segment.  The recheck lists hold few groups: their units are
    // spread over the whole grid (one unit a lane while lanes last, at most
    // c.S), so a group's FP64 pass takes about one sample's latency instead of
    // c.S samples'.
    // A round whose units come to less than two segments of c.S per lane (the
    // later rounds test only the refining outer samples; round 0 of a level
    // with few outer samples) takes smaller segments, at least c.s_test_min
    // units (0: no smaller segments), so that every lane has work; per-sample
    // results do not depend on the segment size.  Every round's entries start
    // with zeroed sums (appended so, or written by k_b_init for round 0), so
    // any split is safe.
    uint64_t S = c.S;
    {
        const uint64_t lanes = (uint64_t)gridDim.x * blockDim.x;
        const uint64_t units_all = (uint64_t)nfe * uf + (uint64_t)nce * uc;
        if (RECHECK) {
            S = (units_all + lanes - 1) / lanes;
            S = S < 1 ? 1 : S;
        } else if (c.s_test_min && units_all < 2 * lanes * c.S) {
            S = (units_all + 2 * lanes - 1) / (2 * lanes);
            S = S < c.s_test_min ? c.s_test_min : S;
        }
        S = S > c.S ? c.S : S;
    }
    const uint64_t spe_f = (uf + S - 1) / S, spe_c = (uc + S - 1) / S;
    const uint64_t segs_f = (uint64_t)nfe * spe_f, total = segs_f + (uint64_t)nce * spe_c;
    if (total == 0) return;  // nothing refined into this round: nothing to flush
    BlockAcc<P::kMulti> acc;
    acc.template init<false>(c);
    TestEntry* const lf = RECHECK ? c.rq[0] : c.tf[r & 1];
    TestEntry* const lc_ = RECHECK ? c.rq[1] : c.tc[r & 1];
    const XcdRange xr = xcd_range(total, c.xcd_sliced);
    for (uint64_t k = xr.k0; k < xr.end; k += xr.stride) {
        const bool fine = k < segs_f;
        const uint64_t kk = fine ? k : k - segs_f;
        const uint64_t spe = fine ? spe_f : spe_c;
        const uint64_t e = kk / spe, j = kk % spe;
        TestEntry& te = fine ? lf[e] : lc_[e];
        const uint32_t m = (r == 0 && !RECHECK) ? (uint32_t)e : te.item;
        const ItemState& it = c.items[m];
        const uint64_t cur = fine ? cur_f : cur_c, units = fine ? uf : uc;
        const uint64_t u0 = j * S, u1 = (u0 + S < units) ? u0 + S : units;
        SegAcc s;
        P::template units<MIXED>(A, m, it, (fine ? TAG_FINE_TEST : TAG_COARSE_TEST) + r, u0, u1, 0, cur, s);
        const uint32_t q = it.inst;
        if (!RECHECK) {
            acc.work(q, s.work);
            acc.st(q, 4, s.nu);
            acc.st(q, 5, s.nn);
        }
        bool complete = spe == 1;
        double s1 = s.s1, s2 = s.s2, sab = s.sab, smm = s.smm;
        if (__ballot(spe > 1)) {  // wave-uniform
            const RunSums rs = run_sums(fine ? e : (e | (1ull << 63)), s.s1, s.s2);
            double rab = 0., rmm = 0.;
            if (MIXED) {
                const RunSums rm = run_sums(fine ? e : (e | (1ull << 63)), s.sab, s.smm);
                rab = rm.a;
                rmm = rm.b;
            }
            if (!complete && rs.tail) {
                aadd(&te.s1, rs.a);
                aadd(&te.s2, rs.b);
                if (MIXED) {
                    aadd(&te.sab, rab);
                    aadd(&te.smm, rmm);
                }
                wait_acks();
                if (acount(&te.done, rs.n) == spe - rs.n) {
                    complete = true;
                    s1 = aread(&te.s1);
                    s2 = aread(&te.s2);
                    if (MIXED) {
                        sab = aread(&te.sab);
                        smm = aread(&te.smm);
                    }
                }
            }
        }
        bool append = false, recheck = false;
        if (complete) {
            if (!RECHECK) acc.work(q, P::group_work(A, it));
            const LevelCounts& lc = fine ? c.cf : c.cc;
            if (MIXED) recheck = decision_uncertain<P, RULE>(c, lc, it.start, cur, s1, s2, sab, smm);
            if (RECHECK) acc.st(q, 8, 1);
            if (!recheck) {
                bool more;
                const uint64_t n = next_count<RULE>(c, lc, it.start, cur, s1, s2, more);
                if (more) {
                    append = true;
                    acc.st(q, 0, n);
                } else {
                    ItemState& w = c.items[m];
                    if (fine) {
                        w.nf = n;
                        acc.st(q, 1, 1);
                        acc.st(q, 2, n == lc.max);
                        acc.st(q, 3, n == lc.base);
                    } else {
                        w.nc = n;
                    }
                }
            }
        }
        // Refined outer samples go to round r + 1 at doubled count.
        const uint32_t sf = wave_append(&c.ctl->cnt_f[r + 1], append && fine);
        const uint32_t sc = wave_append(&c.ctl->cnt_c[r + 1], append && !fine);
        if (append) {
            if (fine) c.tf[(r + 1) & 1][sf] = TestEntry{0., 0., 0., 0., m, 0};
            else c.tc[(r + 1) & 1][sc] = TestEntry{0., 0., 0., 0., m, 0};
        }
        if (MIXED) {
            const uint32_t rf = wave_append(&c.ctl->rc_f[r], recheck && fine);
            const uint32_t rc = wave_append(&c.ctl->rc_c[r], recheck && !fine);
            if (recheck) c.rq[fine ? 0 : 1][fine ? rf : rc] = TestEntry{0., 0., 0., 0., m, 0};
        }
    }
    acc.template flush<false>(c);
}

// Final draw groups of an outer sample: count and shape.
struct FinalShape {
    uint32_t groups;  // entries
    uint64_t n;       // samples per antithetic group (kind 0)
    uint64_t big;
    bool f_big;
};

__device__ __forceinline__ FinalShape final_shape(const BatchedCommon& c, const ItemState& it) {
    FinalShape s;
    if (c.ell == 0) {
        s.groups = 1;
        s.n = s.big = it.nf;
        s.f_big = true;
    } else if (c.ep.antithetic) {
        s.f_big = it.nf > it.nc;
        s.big = s.f_big ? it.nf : it.nc;
        s.n = s.f_big ? it.nc : it.nf;
        s.groups = (uint32_t)(s.big / s.n);
    } else {
        s.groups = 2;
        s.n = 0;
        s.big = 0;
        s.f_big = true;
    }
    return s;
}

// Units of an outer sample's whole final (all its final draw groups).
template <class P>
__device__ __forceinline__ uint64_t final_units(const BatchedCommon& c, const ItemState& it, const FinalShape& s) {
    if (c.ell > 0 && !c.ep.antithetic) return (it.nf + P::spu - 1) / P::spu + (it.nc + P::spu - 1) / P::spu;
    return (uint64_t)s.groups * ((s.n + P::spu - 1) / P::spu);
}

// The sums of an antithetic final group that its two estimators take: the
// big side's (all draws of the larger count) and the small side's (groups of
// the smaller count).  A two-valued policy's group gives its fine values in
// s1 and its coarse values in s1c; the fine side is the big one when
// N_f > N_c (and at level 0).  Other policies have one value per sample.
template <class P>
__device__ __forceinline__ void side_sums(const BatchedCommon& c, const FinalShape& sh, double s1, double s1c,
                                          double& s_big, double& s_small) {
    if (!P::kTwoValued || c.ell == 0) {
        s_big = s_small = s1;
        return;
    }
    s_big = sh.f_big ? s1 : s1c;
    s_small = sh.f_big ? s1c : s1;
}

template <class P>
__global__ __launch_bounds__(256) void k_b_final_count(typename P::Args A) {
    const BatchedCommon& c = A.c;
    __shared__ uint32_t hist[kMaxClasses + 1];
    for (int i = threadIdx.x; i <= kMaxClasses; i += blockDim.x) hist[i] = 0;
    __syncthreads();
    BlockAcc<P::kMulti> acc;
    acc.template init<false>(c);
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    const uint64_t n_items = c.fused ? aload_u32(&c.ctl->n_pend) : c.M;
    for (uint64_t k = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; k < n_items; k += stride) {
        const uint64_t m = c.fused ? c.pend[k] : k;
        const ItemState& it = c.items[m];
        const FinalShape s = final_shape(c, it);
        const bool plain2 = c.ell > 0 && !c.ep.antithetic;
        acc.st(it.inst, 0, plain2 ? it.nf + it.nc : s.big);
        if (final_units<P>(c, it, s) <= c.S) {
            atomicAdd(&hist[kMaxClasses], 1u);  // one lane finishes it alone
        } else if (plain2) {
            atomicAdd(&hist[class_of(it.nf, c.base_min)], 1u);
            atomicAdd(&hist[class_of(it.nc, c.base_min)], 1u);
        } else {
            atomicAdd(&hist[class_of(s.n, c.base_min)], s.groups);
        }
    }
    __syncthreads();
    for (int i = threadIdx.x; i < kMaxClasses; i += blockDim.x)
        if (hist[i]) acount(&c.ctl->class_cnt[i], hist[i]);
    if (threadIdx.x == 0 && hist[kMaxClasses]) acount(&c.ctl->n_whole, hist[kMaxClasses]);
    acc.template flush<false>(c);
    // The block that counts last lays the pool out (one launch fewer than a
    // kernel of its own); its counts are acknowledged before it counts itself.
    __shared__ uint32_t last;
    wait_acks();
    __syncthreads();
    if (threadIdx.x == 0) last = acount(&c.ctl->count_done, 1u) == gridDim.x - 1;
    __syncthreads();
    if (last) final_alloc(c.ctl, c.fcap);
}

// Writes each outer sample's final entries: whole samples after the class
// regions, split samples into their size class's region; block ranks come
// from LDS counters, one global atomic per class and block.
template <class P>
__global__ __launch_bounds__(256) void k_b_final_fill(typename P::Args A) {
    const BatchedCommon& c = A.c;
    if (c.ctl->overflow) return;
    __shared__ uint32_t cnt[kMaxClasses + 1], base[kMaxClasses + 1];
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    const uint32_t whole_off = c.ctl->class_off[kMaxClasses];
    const uint64_t n_items = c.fused ? aload_u32(&c.ctl->n_pend) : c.M;
    for (uint64_t b0 = (uint64_t)blockIdx.x * blockDim.x; b0 < n_items; b0 += stride) {
        for (int i = threadIdx.x; i <= kMaxClasses; i += blockDim.x) cnt[i] = 0;
        __syncthreads();
        const uint64_t k = b0 + threadIdx.x;
        FinalShape s{};
        uint32_t cls_a = 0, cls_b = 0, rank_a = 0, rank_b = 0;
        const bool valid = k < n_items;
        const uint64_t m = valid ? (c.fused ? c.pend[k] : k) : 0;
        const bool plain2 = c.ell > 0 && !c.ep.antithetic;
        bool whole = false;
        if (valid) {
            ItemState& it = c.items[m];
            s = final_shape(c, it);
            whole = final_units<P>(c, it, s) <= c.S;
            if (whole) {
                cls_a = kMaxClasses;
                rank_a = atomicAdd(&cnt[kMaxClasses], 1u);
            } else {
                it.left = plain2 ? 2u : s.groups;
                it.near = 0;
                it.phi_acc = it.s1_big = it.sab_big = 0.;
                if (plain2) {
                    cls_a = class_of(it.nf, c.base_min);
                    cls_b = class_of(it.nc, c.base_min);
                    rank_a = atomicAdd(&cnt[cls_a], 1u);
                    rank_b = atomicAdd(&cnt[cls_b], 1u);
                } else {
                    cls_a = class_of(s.n, c.base_min);
                    rank_a = atomicAdd(&cnt[cls_a], s.groups);
                }
            }
        }
        __syncthreads();
        for (int i = threadIdx.x; i < kMaxClasses; i += blockDim.x)
            base[i] = cnt[i] ? atomicAdd(&c.ctl->class_cur[i], cnt[i]) : 0;
        if (threadIdx.x == 0) base[kMaxClasses] = cnt[kMaxClasses] ? atomicAdd(&c.ctl->n_whole, cnt[kMaxClasses]) : 0;
        __syncthreads();
        if (valid) {
            if (whole) {
                c.fpool[whole_off + base[kMaxClasses] + rank_a] = FinalEntry{0., 0., 0., (uint32_t)m, 0, 0, 3};
            } else if (plain2) {
                c.fpool[c.ctl->class_off[cls_a] + base[cls_a] + rank_a] = FinalEntry{0., 0., 0., (uint32_t)m, 0, 0, 1};
                c.fpool[c.ctl->class_off[cls_b] + base[cls_b] + rank_b] = FinalEntry{0., 0., 0., (uint32_t)m, 0, 0, 2};
            } else {
                FinalEntry* dst = c.fpool + c.ctl->class_off[cls_a] + base[cls_a] + rank_a;
                for (uint32_t g = 0; g < s.groups; g++) dst[g] = FinalEntry{0., 0., 0., (uint32_t)m, g, 0, 0};
            }
        }
        __syncthreads();
    }
}

// Records a finished outer sample: moments, trace (the reference's lines 446-451).
template <bool MULTI>
__device__ __forceinline__ void finish_outer(const BatchedCommon& c, uint64_t m, const ItemState& it,
                                             double phi_f, double phi_c, BlockAcc<MULTI>& acc) {
    acc.moments(it.inst, phi_f, c.ep.fine_moments, c.ell > 0, phi_f - phi_c, c.ep.diff_moments);
    if (c.tr_phi_f) c.tr_phi_f[m] = phi_f;
    if (c.tr_phi_c) c.tr_phi_c[m] = c.ell == 0 ? 0. : phi_c;
    if (c.tr_nf) c.tr_nf[m] = it.nf;
    if (c.tr_nc) c.tr_nc[m] = c.ell == 0 ? 0 : it.nc;
}

// phi of an outer sample from its finished final groups (sample_phi /
// sample_phi_antithetic, nested_sampler.hpp:164-216).
template <class P>
__device__ __forceinline__ void outer_phi(const BatchedCommon& c, const ItemState& it, double s1_big,
                                          double phi_acc, double phi_f_na, double phi_c_na,
                                          double& phi_f, double& phi_c) {
    if (c.ell == 0) {
        phi_f = P::phi(c, it.start + s1_big / (double)it.nf, 1., 0.);
        phi_c = 0.;
    } else if (c.ep.antithetic) {
        const FinalShape s = final_shape(c, it);
        const double phi_big = P::phi(c, it.start + s1_big / (double)s.big, 1., 0.);
        phi_f = s.f_big ? phi_big : phi_acc;
        phi_c = s.f_big ? phi_acc : phi_big;
    } else {
        phi_f = phi_f_na;
        phi_c = phi_c_na;
    }
}

// Sum of inner samples [i0, i1) of draw group `tag`, by one lane.
template <class P, bool F32>
__device__ __forceinline__ SegAcc lane_group(const typename P::Args& A, uint32_t m, const ItemState& it,
                                             uint32_t tag, uint64_t i0, uint64_t i1) {
    SegAcc s;
    P::template units<F32>(A, m, it, tag, i0 / P::spu, (i1 + P::spu - 1) / P::spu, i0, i1, s);
    return s;
}

// Finishing an outer sample of the FP64 redo (rare): its moments, trace and the
// redo counts (inner_stats[6], and [7] when the FP32 phi differed) go straight
// to the level accumulator with atomics.
__device__ __forceinline__ void redo_finish(const BatchedCommon& c, uint64_t m, const ItemState& it, double phi_f,
                                         double phi_c) {
    Partial* out = c.inst_out ? &c.inst_out[it.inst] : &c.ctl->acc;
    double p = 1.;
    for (uint32_t j = 0; j < c.ep.fine_moments; j++) {
        p *= phi_f;
        atomicAdd(&out->f[j], p);
    }
    if (c.ell > 0) {
        const double dphi = phi_f - phi_c;
        p = 1.;
        for (uint32_t j = 0; j < c.ep.diff_moments; j++) {
            p *= dphi;
            atomicAdd(&out->d[j], p);
        }
    }
    atomicAdd((unsigned long long*)&out->stats[6], 1ull);
    if (phi_f != it.phi32_f || phi_c != it.phi32_c) atomicAdd((unsigned long long*)&out->stats[7], 1ull);
    if (c.tr_phi_f) c.tr_phi_f[m] = phi_f;
    if (c.tr_phi_c) c.tr_phi_c[m] = c.ell == 0 ? 0. : phi_c;
    if (c.tr_nf) c.tr_nf[m] = it.nf;
    if (c.tr_nc) c.tr_nc[m] = c.ell == 0 ? 0 : it.nc;
}

// FP64 redo of the outer samples listed by the mixed final phase, from the same
// draws (their work and draw counts were taken in the FP32 pass).  An outer
// sample with a redo slot (the first kRedoSlots listed, at most kRedoGroups
// final groups, antithetic or level 0) is split by k_b_redo_split over up to
// kRedoParts blocks: lanes add their groups' FP64 sums into the slot, and the
// block completing the slot's count forms phi.  The others
// are summed by one block each in k_b_redo_seq, group after group.
// Two-valued policies' groups have two sums: they take the sequential redo.
template <class P>
__device__ __forceinline__ bool redo_split_shape(const BatchedCommon& c, const FinalShape& sh) {
    return !P::kTwoValued && !(c.ell > 0 && !c.ep.antithetic) && sh.groups <= kRedoGroups && sh.n % P::spu == 0;
}

template <class P>
__global__ __launch_bounds__(256) void k_b_redo_split(typename P::Args A) {
    const BatchedCommon& c = A.c;
    __shared__ uint32_t last;
    // The block's per-group sums: lanes add into LDS, then one global add per
    // group the block touched (global adds of every lane into one slot would
    // serialise on its L2 channel).
    __shared__ double gsum[kRedoGroups];
    const uint32_t n = aload_u32(&c.ctl->nreeval);
    // Blocks per listed outer sample: one per blockDim.x units of the level's
    // largest final (its fine or coarse count cap), at most kRedoParts.  The
    // grid's blocks are dealt out that many to a sample, so that the low
    // levels' small finals (thousands of listed samples) run side by side and
    // the high levels' large ones (a few samples of up to 10^6 inner samples)
    // split wide.  Every block computes the same count; the completing block
    // waits for exactly these.
    const uint64_t cmax = c.cf.max > c.cc.max ? c.cf.max : c.cc.max;
    const uint64_t want = ((cmax + P::spu - 1) / P::spu + blockDim.x - 1) / blockDim.x;
    const uint64_t cap = (uint64_t)gridDim.x * gridDim.y < kRedoParts ? (uint64_t)gridDim.x * gridDim.y : kRedoParts;
    const uint32_t parts = (uint32_t)(want < cap ? (want ? want : 1) : cap);
    const uint32_t b = blockIdx.y * gridDim.x + blockIdx.x;
    const uint32_t cols = (gridDim.x * gridDim.y) / parts;
    if (b >= cols * parts) return;
    const uint32_t col = b / parts, part = b % parts;
    const uint32_t n_split = n < kRedoSlots ? n : kRedoSlots;
    for (uint32_t w = col; w < n_split; w += cols) {
        const uint32_t m = c.reeval[w];
        const ItemState& it = c.items[m];
        const FinalShape sh = final_shape(c, it);
        const uint64_t gu = sh.n / P::spu;  // units per group
        const uint64_t U = (uint64_t)sh.groups * gu;
        const uint64_t p0 = U * part / parts, p1 = U * (part + 1) / parts;
        double* slot = c.redo_acc + (uint64_t)w * kRedoGroups;
        for (uint32_t g = threadIdx.x; g < kRedoGroups; g += blockDim.x) gsum[g] = 0.;
        __syncthreads();
        double cur = 0.;
        uint32_t cur_g = ~0u;
        for (uint64_t k = p0 + threadIdx.x; k < p1; k += blockDim.x) {
            const uint32_t g = (uint32_t)(k / gu);
            if (g != cur_g) {
                if (cur_g != ~0u) atomicAdd(&gsum[cur_g], cur);
                cur = 0.;
                cur_g = g;
            }
            SegAcc s;
            const uint64_t i0 = (uint64_t)g * sh.n;
            P::template units<false>(A, m, it, TAG_FINAL, k, k + 1, i0, i0 + sh.n, s);
            cur += s.s1;
        }
        if (cur_g != ~0u) atomicAdd(&gsum[cur_g], cur);
        __syncthreads();
        if (p1 > p0) {
            const uint32_t g_lo = (uint32_t)(p0 / gu), g_hi = (uint32_t)((p1 - 1) / gu);
            for (uint32_t g = g_lo + threadIdx.x; g <= g_hi; g += blockDim.x) aadd(&slot[g], gsum[g]);
        }
        wait_acks();
        __syncthreads();
        if (threadIdx.x == 0) last = acount(&c.redo_cnt[w], 1u) == parts - 1;
        __syncthreads();
        if (!last || threadIdx.x != 0) continue;
        const double coeff = (double)(1.f / (float)sh.groups);
        double s1_big = 0., phi_small = 0.;
        for (uint32_t g = 0; g < sh.groups; g++) {
            const double s1 = aread(&slot[g]);
            phi_small = P::phi(c, it.start + s1 / (double)sh.n, coeff, phi_small);
            s1_big += s1;
        }
        double phi_f, phi_c;
        outer_phi<P>(c, it, s1_big, phi_small, 0., 0., phi_f, phi_c);
        redo_finish(c, m, it, phi_f, phi_c);
    }
}

template <class P>
__global__ __launch_bounds__(256) void k_b_redo_seq(typename P::Args A) {
    const BatchedCommon& c = A.c;
    __shared__ double red[8];
    // Its list: the split list's entries past the slots, then the sequential
    // list (filled from the buffer's end); usually both are empty.  The draw
    // groups are summed at one call site of P::units: a helper taking A by
    // reference that the compiler does not inline makes it copy the kernel's
    // argument block into scratch memory on every lane at entry, which costs
    // more than the usual empty launch.
    const uint32_t n = aload_u32(&c.ctl->nreeval), ns = aload_u32(&c.ctl->nreeval_seq);
    const uint32_t extra = n > kRedoSlots ? n - kRedoSlots : 0;
    for (uint32_t q = blockIdx.x; q < extra + ns; q += gridDim.x) {
        const uint32_t m = q < extra ? c.reeval[kRedoSlots + q] : c.reeval[c.M - 1 - (q - extra)];
        const ItemState& it = c.items[m];
        const FinalShape sh = final_shape(c, it);
        // Unpaired fine and coarse means (two groups: the fine count, then the
        // coarse count) or the antithetic draw groups of sh.n.
        const bool unpaired = c.ell > 0 && !c.ep.antithetic;
        const uint32_t groups = unpaired ? 2u : sh.groups;
        const double coeff = (double)(1.f / (float)sh.groups);
        double phi_f = 0., phi_c = 0., s1_big = 0., phi_small = 0.;
        for (uint32_t g = 0; g < groups; g++) {
            const uint32_t tag = (unpaired && g == 1) ? TAG_FINAL_COARSE : TAG_FINAL;
            const uint64_t i0 = unpaired ? 0 : (uint64_t)g * sh.n;
            const uint64_t i1 = unpaired ? (g == 0 ? it.nf : it.nc) : i0 + sh.n;
            // FP64 sums of the group, block-wide: s1 and (two-valued) s1c.
            SegAcc sa;
            const uint64_t ub = i0 / P::spu, ue = (i1 + P::spu - 1) / P::spu;
            for (uint64_t u = ub + threadIdx.x; u < ue; u += blockDim.x)
                P::template units<false>(A, m, it, tag, u, u + 1, i0, i1, sa);
            const double v = wave_sum(sa.s1);
            const double vc = P::kTwoValued ? wave_sum(sa.s1c) : 0.;
            if ((threadIdx.x & 63) == 0) {
                red[threadIdx.x >> 6] = v;
                red[4 + (threadIdx.x >> 6)] = vc;
            }
            __syncthreads();
            const double s1 = ((red[0] + red[1]) + red[2]) + red[3];
            const double s1c = P::kTwoValued ? ((red[4] + red[5]) + red[6]) + red[7] : s1;
            __syncthreads();
            if (unpaired) {
                const double ph = P::phi(c, it.start + s1 / (double)(i1 - i0), 1., 0.);
                if (g == 0) phi_f = ph;
                else phi_c = ph;
            } else {
                double sb, ss;
                side_sums<P>(c, sh, s1, s1c, sb, ss);
                phi_small = P::phi(c, it.start + ss / (double)sh.n, coeff, phi_small);
                s1_big += sb;
            }
        }
        if (!unpaired) outer_phi<P>(c, it, s1_big, phi_small, 0., 0., phi_f, phi_c);
        if (threadIdx.x == 0) redo_finish(c, m, it, phi_f, phi_c);
    }
}

// The final draw groups of an outer sample small enough for one lane (at most
// S units in all): phi_f, phi_c in FP64 or (MIXED) FP32 arithmetic, with the
// work and draw counts.  Returns true, in the mixed form, when a mean came
// within the FP32 bound of the Heaviside threshold (the sample is then redone
// in FP64 instead of being finished).
template <class P, bool MIXED>
__device__ __forceinline__ bool whole_final(const typename P::Args& A, uint32_t m, const ItemState& it,
                                            BlockAcc<P::kMulti>& acc, double& phi_f, double& phi_c) {
    const BatchedCommon& c = A.c;
    const FinalShape sh = final_shape(c, it);
    bool near = false;
    if (c.ell > 0 && !c.ep.antithetic) {
        const SegAcc a = lane_group<P, MIXED>(A, m, it, TAG_FINAL, 0, it.nf);
        const SegAcc b = lane_group<P, MIXED>(A, m, it, TAG_FINAL_COARSE, 0, it.nc);
        acc.work(it.inst, a.work + b.work + 2 * P::group_work(A, it));
        acc.st(it.inst, 4, a.nu + b.nu);
        acc.st(it.inst, 5, a.nn + b.nn);
        const double mf = it.start + a.s1 / (double)it.nf, mc = it.start + b.s1 / (double)it.nc;
        near = MIXED && (fabs(mf) <= P::mixed_bound(a.sab / (double)it.nf, it.start) ||
                         fabs(mc) <= P::mixed_bound(b.sab / (double)it.nc, it.start));
        outer_phi<P>(c, it, 0., 0., P::phi(c, mf, 1., 0.), P::phi(c, mc, 1., 0.), phi_f,
                  phi_c);
    } else {
        const double coeff = (double)(1.f / (float)sh.groups);
        double s1_big = 0., sab_big = 0., phi_small = 0.;
        for (uint32_t g = 0; g < sh.groups; g++) {
            const SegAcc a = lane_group<P, MIXED>(A, m, it, TAG_FINAL, (uint64_t)g * sh.n, (uint64_t)(g + 1) * sh.n);
            acc.work(it.inst, a.work + P::group_work(A, it));
            acc.st(it.inst, 4, a.nu);
            acc.st(it.inst, 5, a.nn);
            double sb, ss;
            side_sums<P>(c, sh, a.s1, a.s1c, sb, ss);
            const double mean = it.start + ss / (double)sh.n;
            if (MIXED) near |= fabs(mean) <= P::mixed_bound(a.sab / (double)sh.n, it.start);
            phi_small = P::phi(c, mean, coeff, phi_small);
            s1_big += sb;
            sab_big += a.sab;
        }
        if (MIXED)
            near |= fabs(it.start + s1_big / (double)sh.big) <= P::mixed_bound(sab_big / (double)sh.big, it.start);
        outer_phi<P>(c, it, s1_big, phi_small, 0., 0., phi_f, phi_c);
    }
    return MIXED && (near || c.redo_all) && c.phi_type == PHI_INDICATOR;
}

// Lists an outer sample for the FP64 redo (mixed form): a final the split redo
// can take goes to the split list (from the buffer's start) and its slot is
// cleared; any other goes to the sequential list (from the buffer's end; the
// two lists hold at most the chunk's M outer samples together).  Every lane of
// the wave that reaches it must call it.
template <class P>
__device__ __forceinline__ void list_reeval(const BatchedCommon& c, bool reeval, uint32_t m) {
    const bool split = reeval && redo_split_shape<P>(c, final_shape(c, c.items[m]));
    const uint32_t slot = wave_append(&c.ctl->nreeval, split);
    const uint32_t sslot = wave_append(&c.ctl->nreeval_seq, reeval && !split);
    if (split) {
        c.reeval[slot] = m;
        if (slot < kRedoSlots) {
            for (uint32_t g = 0; g < kRedoGroups; g++) astore(&c.redo_acc[(uint64_t)slot * kRedoGroups + g], 0.);
            __hip_atomic_exchange(&c.redo_cnt[slot], 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
    } else if (reeval) {
        c.reeval[c.M - 1 - sslot] = m;
    }
}

// Final phase: every outer sample's final draw groups, in FP64 or (MIXED) FP32
// arithmetic.  In the mixed form an outer sample whose mean came within the
// FP32 bound of the Heaviside threshold (or met a payoff discontinuity) is not
// finished here but listed for the FP64 redo.
template <class P, bool MIXED>
__global__ __launch_bounds__(256, (mixed_waves<P, MIXED>())) void k_b_final(typename P::Args A) {
    const BatchedCommon& c = A.c;
    BlockAcc<P::kMulti> acc;
    acc.template init<true>(c);
    if (c.ctl->overflow) {
        acc.template flush<true>(c);
        return;
    }
    // Split entries (per size class, largest first in memory order of classes)
    // then whole outer samples.
    uint64_t total_split = 0;
    for (int i = 0; i < kMaxClasses; i++) {
        const uint64_t units = ((c.base_min << i) + P::spu - 1) / P::spu;
        total_split += (uint64_t)c.ctl->class_cnt[i] * ((units + c.S - 1) / c.S);
    }
    const uint32_t n_whole = c.ctl->n_whole;
    const uint32_t whole_off = c.ctl->class_off[kMaxClasses];
    const uint64_t total = total_split + n_whole;
    const XcdRange xr = xcd_range(total, c.xcd_sliced);
    for (uint64_t k = xr.k0; k < xr.end; k += xr.stride) {
        bool reeval = false;
        uint32_t m_done = 0;
        const bool split = k < total_split;
        uint64_t spe = 1, n_cls = 0, fidx = 0;
        SegAcc s;
        if (!split) {
            // ---- a whole outer sample on this lane: no atomics ----
            const FinalEntry& fe = c.fpool[whole_off + (k - total_split)];
            const uint32_t m = fe.item;
            ItemState& it = c.items[m];
            double phi_f, phi_c;
            if (whole_final<P, MIXED>(A, m, it, acc, phi_f, phi_c)) {
                it.phi32_f = phi_f;
                it.phi32_c = phi_c;
                reeval = true;
                m_done = m;
            } else {
                finish_outer(c, m, it, phi_f, phi_c, acc);
            }
        } else {
            // ---- one segment of a split final group ----
            uint64_t kk = k;
            int cls = 0;
            for (; cls < kMaxClasses; cls++) {
                n_cls = c.base_min << cls;
                const uint64_t units = (n_cls + P::spu - 1) / P::spu;
                spe = (units + c.S - 1) / c.S;
                const uint64_t segs = (uint64_t)c.ctl->class_cnt[cls] * spe;
                if (kk < segs) break;
                kk -= segs;
            }
            const uint64_t e = kk / spe, j = kk % spe;
            fidx = c.ctl->class_off[cls] + e;
            const FinalEntry& fe = c.fpool[fidx];
            const uint32_t m = fe.item;
            const uint64_t i0 = (fe.kind == 0) ? (uint64_t)fe.group * n_cls : 0;
            const uint64_t i1 = i0 + n_cls;
            const uint64_t ub = i0 / P::spu, ue = (i1 + P::spu - 1) / P::spu;
            const uint64_t u0 = ub + j * c.S, u1 = (u0 + c.S < ue) ? u0 + c.S : ue;
            const ItemState& it = c.items[m];
            P::template units<MIXED>(A, m, it, fe.kind == 2 ? TAG_FINAL_COARSE : TAG_FINAL, u0, u1, i0, i1, s);
            acc.work(it.inst, s.work);
            acc.st(it.inst, 4, s.nu);
            acc.st(it.inst, 5, s.nn);
        }
        bool complete = split && spe == 1;
        double s1 = s.s1, sab = s.sab, s1c = s.s1c;
        if (__ballot(split && spe > 1)) {  // wave-uniform
            FinalEntry& fe = c.fpool[fidx];
            const unsigned long long key = split ? fidx : ((1ull << 63) | k);
            const RunSums rs = run_sums(key, s.s1, s.sab);
            const double rc = P::kTwoValued ? run_sums(key, s.s1c, 0.).a : 0.;
            if (split && spe > 1 && rs.tail) {
                aadd(&fe.s1, rs.a);
                if (MIXED) aadd(&fe.sab, rs.b);
                if (P::kTwoValued) aadd(&fe.s1c, rc);
                wait_acks();
                if (acount(&fe.done, rs.n) == spe - rs.n) {
                    complete = true;
                    s1 = aread(&fe.s1);
                    sab = MIXED ? aread(&fe.sab) : 0.;
                    s1c = P::kTwoValued ? aread(&fe.s1c) : 0.;
                }
            }
        }
        if (complete) {
            const FinalEntry& fe = c.fpool[fidx];
            const uint32_t m = fe.item;
            ItemState& it = c.items[m];
            acc.work(it.inst, P::group_work(A, it));
            const FinalShape sh = final_shape(c, it);
            double s_big = s1, s_small = s1;
            if (fe.kind == 0) side_sums<P>(c, sh, s1, s1c, s_big, s_small);
            const double mean = it.start + s_small / (double)n_cls;
            const bool near = MIXED && fabs(mean) <= P::mixed_bound(sab / (double)n_cls, it.start);
            if (near) __hip_atomic_fetch_or(&it.near, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            if (fe.kind == 0) {
                const double coeff = (double)(1.f / (float)sh.groups);
                if (c.ell > 0) aadd(&it.phi_acc, P::phi(c, mean, coeff, 0.));
                aadd(&it.s1_big, s_big);
                if (MIXED) aadd(&it.sab_big, sab);
            } else if (fe.kind == 1) {
                astore(&it.phi_f, P::phi(c, mean, 1., 0.));
            } else {
                astore(&it.phi_c, P::phi(c, mean, 1., 0.));
            }
            wait_acks();
            if (__hip_atomic_fetch_sub(&it.left, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == 1u) {
                const double s1b = aread(&it.s1_big), pacc = aread(&it.phi_acc);
                double phi_f, phi_c;
                outer_phi<P>(c, it, s1b, pacc, aread(&it.phi_f), aread(&it.phi_c), phi_f, phi_c);
                bool is_near = MIXED && aread_u32(&it.near) != 0;
                if (MIXED && fe.kind == 0) {
                    const double mb = it.start + s1b / (double)sh.big;
                    is_near |= fabs(mb) <= P::mixed_bound(aread(&it.sab_big) / (double)sh.big, it.start);
                }
                if (MIXED && (is_near || c.redo_all) && c.phi_type == PHI_INDICATOR) {
                    it.phi32_f = phi_f;
                    it.phi32_c = phi_c;
                    reeval = true;
                    m_done = m;
                } else {
                    finish_outer(c, m, it, phi_f, phi_c, acc);
                }
            }
        }
        if (MIXED) list_reeval<P>(c, reeval, m_done);
    }
    acc.template flush<true>(c);
}

// Fused first pass, for level calls whose round-0 test groups fit one segment:
// the outer draw (k_b_init), the round-0 fine and coarse tests (k_b_test at
// r = 0) and, for an outer sample whose counts are both decided there and
// whose final fits one segment, its final draws and moments (k_b_final), all
// in the lane that draws the outer sample, with the item state in registers.
// Only outer samples that refine (queued for round 1), that have a larger
// final (both listed in pend for the final phase) or that need the FP64 redo
// (MIXED) are written out.  Draws, decisions and counts are those of the
// unfused sequence.
template <class P, bool MIXED, int RULE>
__global__ __launch_bounds__(256, (mixed_waves<P, MIXED>())) void k_b_fused(typename P::Args A) {
    const BatchedCommon& c = A.c;
    BlockAcc<P::kMulti> acc;
    acc.template init<true>(c);
    const bool test_f = c.cf.max > 2 * c.cf.base;
    const bool test_c = c.ell > 0 && c.cc.max > 2 * c.cc.base;
    const XcdRange xr = xcd_range(c.M, c.xcd_sliced);
    for (uint64_t m = xr.k0; m < xr.end; m += xr.stride) {
        ItemState it;
        P::init_item(A, m, it, acc);
        const uint32_t q = it.inst;
        it.nf = it.nc = 0;
        bool app_f = false, app_c = false, rck_f = false, rck_c = false;
        if (!test_f) {
            it.nf = c.cf.max;
            acc.st(q, 1, 1);
            acc.st(q, 2, 1);
            acc.st(q, 3, (c.cf.max == c.cf.base));
        } else {
            acc.st(q, 0, c.cf.base);
            const SegAcc t = lane_group<P, MIXED>(A, (uint32_t)m, it, TAG_FINE_TEST, 0, c.cf.base);
            acc.work(q, t.work + P::group_work(A, it));
            acc.st(q, 4, t.nu);
            acc.st(q, 5, t.nn);
            if (MIXED) rck_f = decision_uncertain<P, RULE>(c, c.cf, it.start, c.cf.base, t.s1, t.s2, t.sab, t.smm);
            if (!rck_f) {
                const uint64_t n = next_count<RULE>(c, c.cf, it.start, c.cf.base, t.s1, t.s2, app_f);
                if (app_f) {
                    acc.st(q, 0, n);
                } else {
                    it.nf = n;
                    acc.st(q, 1, 1);
                    acc.st(q, 2, (n == c.cf.max));
                    acc.st(q, 3, (n == c.cf.base));
                }
            }
        }
        if (c.ell > 0) {
            if (!test_c) {
                it.nc = c.cc.max;
            } else {
                acc.st(q, 0, c.cc.base);
                const SegAcc t = lane_group<P, MIXED>(A, (uint32_t)m, it, TAG_COARSE_TEST, 0, c.cc.base);
                acc.work(q, t.work + P::group_work(A, it));
                acc.st(q, 4, t.nu);
                acc.st(q, 5, t.nn);
                if (MIXED) rck_c = decision_uncertain<P, RULE>(c, c.cc, it.start, c.cc.base, t.s1, t.s2, t.sab, t.smm);
                if (!rck_c) {
                    const uint64_t n = next_count<RULE>(c, c.cc, it.start, c.cc.base, t.s1, t.s2, app_c);
                    if (app_c) acc.st(q, 0, n);
                    else it.nc = n;
                }
            }
        }
        // A decision left to the FP64 recheck (k_b_test<..., RECHECK> of round
        // 0) sends the outer sample to the final phase like a refined one.
        bool pending = app_f || app_c || rck_f || rck_c, reeval = false;
        if (!pending) {
            const FinalShape sh = final_shape(c, it);
            if (final_units<P>(c, it, sh) <= c.S) {
                // The final's inner samples, as k_b_final_count counts them.
                acc.st(q, 0, (c.ell > 0 && !c.ep.antithetic) ? it.nf + it.nc : sh.big);
                double phi_f, phi_c;
                if (whole_final<P, MIXED>(A, (uint32_t)m, it, acc, phi_f, phi_c)) {
                    it.phi32_f = phi_f;
                    it.phi32_c = phi_c;
                    reeval = true;
                } else {
                    finish_outer(c, m, it, phi_f, phi_c, acc);
                }
            } else {
                pending = true;
            }
        }
        if (pending || reeval) {
            ItemState& w = c.items[m];
            w.Y = it.Y;
            w.start = it.start;
            w.nf = it.nf;
            w.nc = it.nc;
            w.inst = it.inst;
            if (reeval) {
                w.phi32_f = it.phi32_f;
                w.phi32_c = it.phi32_c;
            }
        }
        const uint32_t sf = wave_append(&c.ctl->cnt_f[1], app_f);
        const uint32_t sc = wave_append(&c.ctl->cnt_c[1], app_c);
        const uint32_t sp = wave_append(&c.ctl->n_pend, pending);
        if (app_f) c.tf[1][sf] = TestEntry{0., 0., 0., 0., (uint32_t)m, 0};
        if (app_c) c.tc[1][sc] = TestEntry{0., 0., 0., 0., (uint32_t)m, 0};
        if (pending) c.pend[sp] = (uint32_t)m;
        if (MIXED) {
            list_reeval<P>(c, reeval, (uint32_t)m);
            const uint32_t rf = wave_append(&c.ctl->rc_f[0], rck_f);
            const uint32_t rc = wave_append(&c.ctl->rc_c[0], rck_c);
            if (rck_f) c.rq[0][rf] = TestEntry{0., 0., 0., 0., (uint32_t)m, 0};
            if (rck_c) c.rq[1][rc] = TestEntry{0., 0., 0., 0., (uint32_t)m, 0};
        }
    }
    acc.template flush<true>(c);
}

// True when the fused first pass applies: both round-0 test groups (those
// that run) fit one segment.
template <class P>
bool fused_applies(const BatchedCommon& c) {
    if (!P::kFuse) return false;
    const bool test_f = c.cf.max > 2 * c.cf.base;
    const bool test_c = c.ell > 0 && c.cc.max > 2 * c.cc.base;
    const bool fit_f = !test_f || (c.cf.base + P::spu - 1) / P::spu <= c.S;
    const bool fit_c = !test_c || (c.cc.base + P::spu - 1) / P::spu <= c.S;
    return c.pend != nullptr && fit_f && fit_c;
}

// The kernel sequence of one level call on stream s.
template <class P, bool MIXED>
hipError_t launch_batched_level(const typename P::Args& A0, uint32_t blocks, hipStream_t s) {
    typename P::Args A = A0;
    A.c.fused = fused_applies<P>(A.c) ? 1u : 0u;
    const bool rule1 = A.c.ep.adapt_rule == 1;
    if (A.c.zero_ctl) hipLaunchKernelGGL(k_b_zero, dim3(1), dim3(256), 0, s, A.c.ctl, A.c.inst_out, A.c.n_inst);
    hipError_t e = P::launch_outer(A, s);
    if (e != hipSuccess) return e;
    // Mixed form: the FP64 recheck of round r's uncertain decisions, before
    // round r + 1 reads its queue.  Its grid is the level's: a larger one
    // (measured at four times) cut the high levels' rechecked groups of
    // 10^5..10^6 samples into so many segments that their hand-off atomics on
    // the group's sums cost more than the shorter segments saved.
    auto recheck = [&](uint32_t r) {
        if (!MIXED) return;
        if (rule1) hipLaunchKernelGGL((k_b_test<P, false, 1, true>), dim3(blocks), dim3(256), 0, s, A, r);
        else hipLaunchKernelGGL((k_b_test<P, false, 0, true>), dim3(blocks), dim3(256), 0, s, A, r);
    };
    uint32_t r0 = 0;
    if constexpr (P::kFuse) {
        if (A.c.fused) {
            if (rule1) hipLaunchKernelGGL((k_b_fused<P, MIXED, 1>), dim3(blocks), dim3(256), 0, s, A);
            else hipLaunchKernelGGL((k_b_fused<P, MIXED, 0>), dim3(blocks), dim3(256), 0, s, A);
            recheck(0);
            r0 = 1;
        }
    }
    if (!A.c.fused) {
        hipLaunchKernelGGL((k_b_init<P>), dim3(blocks), dim3(256), 0, s, A);
    }
    for (uint32_t r = r0; r < A.c.rounds; r++) {
        if (rule1) hipLaunchKernelGGL((k_b_test<P, MIXED, 1>), dim3(blocks), dim3(256), 0, s, A, r);
        else hipLaunchKernelGGL((k_b_test<P, MIXED, 0>), dim3(blocks), dim3(256), 0, s, A, r);
        recheck(r);
    }
    // At most 256 blocks: each takes a ticket on one counter when done, and
    // the tickets of a 1024-block grid queue on that address for ~10 us.
    hipLaunchKernelGGL((k_b_final_count<P>), dim3(blocks < 256 ? blocks : 256), dim3(256), 0, s, A);
    hipLaunchKernelGGL((k_b_final_fill<P>), dim3(blocks), dim3(256), 0, s, A);
    hipLaunchKernelGGL((k_b_final<P, MIXED>), dim3(blocks), dim3(256), 0, s, A);
    if (MIXED) {
        hipLaunchKernelGGL((k_b_redo_split<P>), dim3(kRedoGridX, kRedoGridY), dim3(256), 0, s, A);
        hipLaunchKernelGGL((k_b_redo_seq<P>), dim3(256), dim3(256), 0, s, A);
    }
    return hipGetLastError();
}

}  // namespace nmlmc
