// Fixed-order reduction of per-block Partial records into one record, in two
// passes: kReduceBlocks blocks each reduce a contiguous range of the records
// into a staging record, then one block reduces the staging records.  The
// ranges depend only on the record count, so the result does not depend on the
// order in which the producing blocks finished.
#include "batched.hpp"
#include "engine.hpp"

namespace nmlmc {

namespace {

constexpr int kRedF = kPartialFields;           // 8-byte words of a record
constexpr int kRedR = 256 / kRedF;              // records read side by side per block step
constexpr int kRedSum = 2 * kMaxPhi * kMaxMoments + 1;  // leading FP64 words; then the u64 stats
constexpr uint32_t kRedMinChunk = 16;

static_assert(sizeof(Partial) == 8 * kPartialFields, "a Partial is kPartialFields 8-byte words");

// Block b reduces records [b chunk, min(n, (b+1) chunk)) into out[b].  Thread
// (word f, row j) sums word f of the range's records j, j + kRedR, ... in
// that order (loads issued four steps ahead of the adds; the order of the adds
// is the plain loop's); the rows are then added in row order.  A record's
// words are contiguous, so each step reads kRedR whole records.
__global__ __launch_bounds__(256) void k_reduce_partials(const Partial* __restrict__ P, uint32_t n,
                                                         uint32_t chunk, Partial* __restrict__ out) {
    __shared__ uint64_t acc_s[kRedR][kRedF];
    const uint32_t tid = threadIdx.x;
    const uint32_t lo = blockIdx.x * chunk;
    const uint32_t hi = min(n, lo + chunk);
    const uint64_t* W = reinterpret_cast<const uint64_t*>(P);
    if (tid < (uint32_t)(kRedR * kRedF)) {
        const uint32_t f = tid % kRedF, j = tid / kRedF;
        const bool fp = f < (uint32_t)kRedSum;
        double ad = 0.;
        uint64_t au = 0;
        for (uint32_t r = lo + j; r < hi; r += 4 * kRedR) {
            uint64_t v[4];
#pragma unroll
            for (int k = 0; k < 4; k++) {
                const uint32_t rk = r + k * kRedR;
                v[k] = rk < hi ? W[(size_t)rk * kRedF + f] : 0ull;
            }
#pragma unroll
            for (int k = 0; k < 4; k++) {
                if (r + k * kRedR < hi) {
                    if (fp) ad += __longlong_as_double((long long)v[k]);
                    else au += v[k];
                }
            }
        }
        acc_s[j][f] = fp ? (uint64_t)__double_as_longlong(ad) : au;
    }
    __syncthreads();
    if (tid < (uint32_t)kRedF) {
        uint64_t* o = reinterpret_cast<uint64_t*>(out + blockIdx.x);
        if (tid < (uint32_t)kRedSum) {
            double a = 0.;
            for (int j = 0; j < kRedR; j++) a += __longlong_as_double((long long)acc_s[j][tid]);
            o[tid] = (uint64_t)__double_as_longlong(a);
        } else {
            uint64_t a = 0;
            for (int j = 0; j < kRedR; j++) a += acc_s[j][tid];
            o[tid] = a;
        }
    }
}

}  // namespace

hipError_t launch_reduce_partials(const Partial* P, uint32_t n, Partial* out, hipStream_t s) {
    const uint32_t chunk = max(kRedMinChunk, (n + kReduceBlocks - 1) / kReduceBlocks);
    const uint32_t nb = n ? (n + chunk - 1) / chunk : 1;
    if (nb == 1) {
        hipLaunchKernelGGL(k_reduce_partials, dim3(1), dim3(256), 0, s, P, n, chunk, out);
        return hipGetLastError();
    }
    Partial* stage = out + 1;  // nb <= kReduceBlocks staging records after the result
    hipLaunchKernelGGL(k_reduce_partials, dim3(nb), dim3(256), 0, s, P, n, chunk, stage);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_reduce_partials, dim3(1), dim3(256), 0, s, (const Partial*)stage, nb, nb, out);
    return hipGetLastError();
}

hipError_t launch_clear_ctl(BatchedCtl* ctl, hipStream_t s) {
    hipLaunchKernelGGL(k_b_zero, dim3(1), dim3(256), 0, s, ctl, (Partial*)nullptr, 0u);
    return hipGetLastError();
}

__global__ void k_b_zero(BatchedCtl* ctl, Partial* inst_out, uint32_t n_inst) {
    uint32_t* p = (uint32_t*)ctl;
    const uint32_t n = sizeof(BatchedCtl) / 4;
    for (uint32_t i = threadIdx.x; i < n; i += blockDim.x) p[i] = 0;
    uint32_t* q = (uint32_t*)inst_out;
    const uint32_t nq = inst_out ? n_inst * (uint32_t)(sizeof(Partial) / 4) : 0;
    for (uint32_t i = threadIdx.x; i < nq; i += blockDim.x) q[i] = 0;
}

}  // namespace nmlmc
