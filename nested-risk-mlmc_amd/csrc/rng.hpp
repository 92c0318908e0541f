// Counter-based RNG for gfx950: Threefry2x64-20 with the reference's stream
// semantics (Random123 as used by par_omp.hpp:72-121).
//
// A stream is a key {k0, k1} and a position P = number of u64 words already
// handed out.  Word q of a stream is element 1 - (q & 1) of the block
// threefry({q >> 1, 0}, key): random_state::rand() returns cur[1] before cur[0]
// (par_omp.hpp:92-99).  randn() takes two consecutive words (first, second) and
// evaluates boxmuller(u0 = second, u1 = first) - the argument order the
// reference build produces - returning x and caching y.  Because every position
// is a pure function of (key, P), any lane can compute any draw: the reference's
// sequential stream becomes a parallel one.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace nmlmc {

struct u64x2 {
    uint64_t a, b;
};

// 64-bit rotate by a compile-time amount as two v_alignbit_b32 funnel shifts
// (one per 32-bit half); a rotate by 32 is a register swap.
template <unsigned N>
__device__ __forceinline__ uint64_t rotl64(uint64_t x) {
    const uint32_t lo = (uint32_t)x, hi = (uint32_t)(x >> 32);
    uint32_t rlo, rhi;
    if constexpr (N == 32) {
        rlo = hi;
        rhi = lo;
    } else if constexpr (N < 32) {
        rhi = __builtin_amdgcn_alignbit(hi, lo, 32 - N);
        rlo = __builtin_amdgcn_alignbit(lo, hi, 32 - N);
    } else {
        rhi = __builtin_amdgcn_alignbit(lo, hi, 64 - N);
        rlo = __builtin_amdgcn_alignbit(hi, lo, 64 - N);
    }
    return ((uint64_t)rhi << 32) | rlo;
}

// Threefry2x64 with 20 rounds (Random123/threefry.h:190-293): Skein rotation
// constants {16,42,12,31,16,32,24,21}, key-schedule parity 0x1BD11BDAA9FC1A22,
// key injection every 4 rounds with the injection count added to word 1.
__device__ __forceinline__ u64x2 threefry2x64_20(uint64_t c0, uint64_t c1, uint64_t k0, uint64_t k1) {
    const uint64_t k2 = 0x1BD11BDAA9FC1A22ULL ^ k0 ^ k1;
    uint64_t x0 = c0 + k0, x1 = c1 + k1;
#define NMLMC_TF_R(R)     \
    x0 += x1;             \
    x1 = rotl64<R>(x1);   \
    x1 ^= x0;
    NMLMC_TF_R(16) NMLMC_TF_R(42) NMLMC_TF_R(12) NMLMC_TF_R(31)
    x0 += k1; x1 += k2 + 1;
    NMLMC_TF_R(16) NMLMC_TF_R(32) NMLMC_TF_R(24) NMLMC_TF_R(21)
    x0 += k2; x1 += k0 + 2;
    NMLMC_TF_R(16) NMLMC_TF_R(42) NMLMC_TF_R(12) NMLMC_TF_R(31)
    x0 += k0; x1 += k1 + 3;
    NMLMC_TF_R(16) NMLMC_TF_R(32) NMLMC_TF_R(24) NMLMC_TF_R(21)
    x0 += k1; x1 += k2 + 4;
    NMLMC_TF_R(16) NMLMC_TF_R(42) NMLMC_TF_R(12) NMLMC_TF_R(31)
    x0 += k2; x1 += k0 + 5;
#undef NMLMC_TF_R
    return {x0, x1};
}

// r123::u01 / uneg11 for W = 64 (Random123/uniform.hpp:144-180).  The scale is
// a power of two, so the product is exact and fused or unfused evaluation
// rounds identically.
__device__ __forceinline__ double u01_f64(uint64_t u) { return (double)u * 0x1p-64 + 0x1p-65; }
__device__ __forceinline__ double uneg11_f64(uint64_t u) {
    return (double)(int64_t)u * 0x1p-63 + 0x1p-64;
}
__device__ __forceinline__ float u01_f32(uint64_t u) { return (float)u * 0x1p-64f + 0x1p-65f; }

// boxmuller(uint64, uint64) (Random123/boxmuller.hpp:128-137) with the
// sincospi fallback sincos(PI * x) (boxmuller.hpp:104-107).
__device__ __forceinline__ void boxmuller_f64(uint64_t u0, uint64_t u1, double& x, double& y) {
    double s, c;
    sincos(3.1415926535897932 * uneg11_f64(u0), &s, &c);
    const double r = sqrt(-2. * log(u01_f64(u1)));
    x = s * r;
    y = c * r;
}

// FP32 Box-Muller for the mixed form: the same words, u1 -> (0,1] in float and
// u0 -> [-1,1], sin/cos(pi x) and log in hardware-accelerated single precision.
// The float of u1 carries a relative error up to 2^-24, i.e. an absolute error
// of about 2^-24 in ln u1; the radius r = sqrt(-2 ln u1) then has an error of
// about 2^-24 / r, which for u1 near 1 (small r) would reach 100% (a word of
// u1 >= 2^64 - 2^39 rounds to 1: radius 0 against an FP64 radius up to 2^-12).
// Words with u1 > 1 - 2^-8 (top byte all ones, 2^-8 of them) therefore take
// the radius from q = 1 - u01(u1) = ~u1 2^-64 + 2^-65, whose float is accurate
// to 2^-24 relative: -ln(1 - q) = q (1 + q/2 + q^2/3 + q^3/4), truncation below
// 2^-33 relative for q <= 2^-8.  Elsewhere r >= 2^-3.5 and the error stays
// below about 2^-20.5.
__device__ __forceinline__ void boxmuller_f32(uint64_t u0, uint64_t u1, float& x, float& y) {
    const float a = (float)(int64_t)u0 * 0x1p-63f + 0x1p-64f;
    float s, c;
    __sincosf(3.14159265358979f * a, &s, &c);
    float r2;
    if ((uint32_t)(u1 >> 32) < 0xFF000000u) {
        r2 = -2.f * __logf((float)u1 * 0x1p-64f + 0x1p-65f);
    } else {
        const float q = (float)(~u1) * 0x1p-64f + 0x1p-65f;
        r2 = 2.f * q * (1.f + q * (0.5f + q * (0x1.555556p-2f + q * 0.25f)));
    }
    const float r = __fsqrt_rn(r2);
    x = s * r;
    y = c * r;
}

// Word q of the stream keyed {k0, k1}.
__device__ __forceinline__ uint64_t stream_word(uint64_t k0, uint64_t k1, uint64_t q) {
    const u64x2 b = threefry2x64_20(q >> 1, 0, k0, k1);
    return (q & 1) ? b.a : b.b;
}

// The Box-Muller pair built from words (P, P+1) of a stream.
__device__ __forceinline__ void stream_bm_f64(uint64_t k0, uint64_t k1, uint64_t P, double& x,
                                              double& y) {
    uint64_t first, second;
    if ((P & 1) == 0) {
        const u64x2 b = threefry2x64_20(P >> 1, 0, k0, k1);
        first = b.b;
        second = b.a;
    } else {
        first = stream_word(k0, k1, P);
        second = stream_word(k0, k1, P + 1);
    }
    boxmuller_f64(second, first, x, y);
}

// A random_state (par_omp.hpp:72-121) positioned on a keyed counter sequence:
// block b of the stream is threefry({b, c1}, {k0, k1}); word q is element
// 1 - (q & 1) of block q >> 1.  c1 = 0 for the reference keying's per-thread
// streams; c1 = (tag << 40) | unit for a batched inner unit.
struct DStream {
    uint64_t k0, k1, c1, P;
    double saved;
    uint32_t is_saved;
    uint64_t blk;
    u64x2 cur;
    uint64_t nu, nn;  // randu / randn calls (inner_stats[4], [5])

    __device__ void init(uint64_t a, uint64_t b, uint64_t c, uint64_t p, double sv, uint32_t iss) {
        k0 = a;
        k1 = b;
        c1 = c;
        P = p;
        saved = sv;
        is_saved = iss;
        blk = ~0ull;
        nu = nn = 0;
    }
    __device__ uint64_t rand() {
        const uint64_t q = P++;
        const uint64_t b = q >> 1;
        if (b != blk) {
            cur = threefry2x64_20(b, c1, k0, k1);
            blk = b;
        }
        return (q & 1) ? cur.a : cur.b;
    }
    __device__ double randn() {
        nn++;
        if (is_saved) {
            is_saved = 0;
            return saved;
        }
        const uint64_t first = rand();
        const uint64_t second = rand();
        double x, y;
        boxmuller_f64(second, first, x, y);
        is_saved = 1;
        saved = y;
        return x;
    }
    __device__ double randu() {
        nu++;
        return u01_f64(rand());
    }
    // Move past the next k normals of a normal-only run of draws (the cached
    // one first, then Box-Muller pairs): the stream is then where k randn()
    // calls leave it, its last pair computed when k ends inside one.
    __device__ void skip_normals(uint64_t k) {
        if (k == 0) return;
        if (is_saved) {
            is_saved = 0;
            k--;
        }
        P += 2 * (k >> 1);
        if (k & 1) (void)randn();
    }
};

// Persistent state of one reference stream between calls: P words handed out,
// plus the cached second normal of the last Box-Muller pair.
struct StreamState {
    uint64_t P;
    double saved;
    uint32_t is_saved;
    uint32_t pad;
};

// Uniform (wave-uniform in practice) sequential draws on a stream state.
__device__ __forceinline__ double st_randn_f64(uint64_t k0, uint64_t k1, StreamState& s) {
    if (s.is_saved) {
        s.is_saved = 0;
        return s.saved;
    }
    double x, y;
    stream_bm_f64(k0, k1, s.P, x, y);
    s.P += 2;
    s.saved = y;
    s.is_saved = 1;
    return x;
}

__device__ __forceinline__ double st_randu_f64(uint64_t k0, uint64_t k1, StreamState& s) {
    return u01_f64(stream_word(k0, k1, s.P++));
}

// Normal number n (0-based) of the normal-only draw sequence that starts at
// state s: normal 0 is the cached one when s.is_saved, the rest come in
// Box-Muller pairs from words P, P+2, ...
__device__ __forceinline__ double normal_at_f64(uint64_t k0, uint64_t k1, const StreamState& s,
                                                uint64_t n) {
    if (s.is_saved) {
        if (n == 0) return s.saved;
        n -= 1;
    }
    double x, y;
    stream_bm_f64(k0, k1, s.P + 2 * (n >> 1), x, y);
    return (n & 1) ? y : x;
}

// State after consuming k normals from s (normal-only consumption).
__device__ __forceinline__ StreamState advance_normals_f64(uint64_t k0, uint64_t k1, StreamState s,
                                                           uint64_t k) {
    if (k == 0) return s;
    if (s.is_saved) {
        s.is_saved = 0;
        k -= 1;
    }
    const uint64_t pairs = (k + 1) >> 1;
    if (k & 1) {
        double x, y;
        stream_bm_f64(k0, k1, s.P + 2 * (pairs - 1), x, y);
        s.saved = y;
        s.is_saved = 1;
    }
    s.P += 2 * pairs;
    return s;
}

}  // namespace nmlmc
