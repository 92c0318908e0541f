// Fixed-order reduction of per-block Partial records into one record: one
// 256-thread block per field, strided partial sums then an LDS tree, so the
// result does not depend on the order in which blocks finished.
#include "batched.hpp"
#include "engine.hpp"

namespace nmlmc {

__global__ __launch_bounds__(256) void k_reduce_partials(const Partial* __restrict__ P, uint32_t n,
                                                         Partial* out) {
    __shared__ double sd[256];
    __shared__ uint64_t su[256];
    constexpr int nsum = 2 * kMaxPhi * kMaxMoments;
    const int field = blockIdx.x;  // sums, then work, then the stats
    const int tid = threadIdx.x;
    if (field <= nsum) {
        double acc = 0.;
        for (uint32_t t = tid; t < n; t += 256) {
            const Partial& p = P[t];
            acc += field < kMaxPhi * kMaxMoments ? p.f[field]
                   : (field < nsum ? p.d[field - kMaxPhi * kMaxMoments] : p.work);
        }
        sd[tid] = acc;
        __syncthreads();
        for (int o = 128; o > 0; o >>= 1) {
            if (tid < o) sd[tid] += sd[tid + o];
            __syncthreads();
        }
        if (tid == 0) {
            if (field < kMaxPhi * kMaxMoments) out->f[field] = sd[0];
            else if (field < nsum) out->d[field - kMaxPhi * kMaxMoments] = sd[0];
            else out->work = sd[0];
        }
    } else {
        const int k = field - nsum - 1;
        uint64_t acc = 0;
        for (uint32_t t = tid; t < n; t += 256) acc += P[t].stats[k];
        su[tid] = acc;
        __syncthreads();
        for (int o = 128; o > 0; o >>= 1) {
            if (tid < o) su[tid] += su[tid + o];
            __syncthreads();
        }
        if (tid == 0) out->stats[k] = su[0];
    }
}

hipError_t launch_reduce_partials(const Partial* P, uint32_t n, Partial* out, hipStream_t s) {
    hipLaunchKernelGGL(k_reduce_partials, dim3(kPartialFields), dim3(256), 0, s, P, n, out);
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

__global__ void k_b_final_alloc(BatchedCtl* ctl, uint32_t fcap) {
    if (threadIdx.x != 0) return;
    uint32_t off = 0;
    for (int i = 0; i < kMaxClasses; i++) {
        ctl->class_off[i] = off;
        ctl->class_cur[i] = 0;
        off += ctl->class_cnt[i];
    }
    ctl->class_off[kMaxClasses] = off;
    ctl->final_total = off + ctl->n_whole;
    ctl->overflow = ctl->final_total > fcap;
    ctl->n_whole = 0;  // refilled as a cursor by k_b_final_fill
}

}  // namespace nmlmc
