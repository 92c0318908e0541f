// Model problem (model_sampler.cpp:15-92) on the device, and the launch
// interface of its kernels.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "batched.hpp"
#include "engine.hpp"

namespace nmlmc {

enum { MODEL_REF = 0, MODEL_SHIFT = 1 };

// model::data (model_sampler.cpp:21-39) plus the variant selector.
struct ModelParams {
    double stau, s1tau, max_loss;
    int phi_type;
    int variant;
};

// model::sampler::sample_inner without MODEL_ANTITHETIC (model_sampler.cpp:66-79):
// (sqrt(tau) Y + sqrt(1-tau) X)^2 - (sqrt(tau) Y2 + sqrt(1-tau) X)^2.
__device__ __forceinline__ double model_ref_inner(const ModelParams& mp, double Y, double X, double Y2) {
    const double a = mp.stau * Y + mp.s1tau * X;
    double r = a * a;
    const double b = mp.stau * Y2 + mp.s1tau * X;
    r -= b * b;
    return r;
}

struct ModelParityArgsHost {
    ModelParams mp;
    EngineParams ep;
    LevelCounts cf, cc;
    uint64_t seed;
    uint64_t M;
    uint32_t T;
    uint32_t ell;
    uint32_t G;
    uint32_t blocks;  // launched: the blocks of the virtual threads that hold outer samples
    StreamState* states;
    Partial* partials;
    double* tr_phi_f;
    double* tr_phi_c;
    uint64_t* tr_nf;
    uint64_t* tr_nc;
};

struct ModelBatchedHost {
    BatchedCommon c;
    ModelParams mp;
    uint32_t blocks;
    bool mixed;
};
hipError_t launch_model_batched(const ModelBatchedHost& H, hipStream_t s);

// Blocks of 256 threads covering T groups of G lanes; one Partial per block.
inline uint64_t parity_blocks(uint64_t T, uint32_t G) { return (T * G + 255) / 256; }
// Virtual threads that hold outer samples in a level call of M: libgomp's static
// schedule gives thread t a share only when t < M.  The others draw nothing and
// keep their streams, so their blocks are not launched (at least one block is).
inline uint64_t plain_active(uint64_t M, uint64_t T) { return M == 0 ? 1 : (M < T ? M : T); }

hipError_t launch_model_parity(const ModelParityArgsHost& H, hipStream_t s);
// Reduces n Partial records into out[0] in a fixed order; out has room for
// 1 + kReduceBlocks records (the rest are the reduction's staging records).
hipError_t launch_reduce_partials(const Partial* P, uint32_t n, Partial* out, hipStream_t s);

}  // namespace nmlmc
