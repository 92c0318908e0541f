// model::sampler (model_sampler.cpp:15-92) written as a class of the
// reference's sampler concept and run through nmlmc::init_samplers<SAMPLER> /
// nmlmc::sample<SAMPLER> (include/nmlmc_sampler.hpp): the proof that a
// reference-style sampler drops into the batched engine unchanged in shape.
// Its data lives in a __constant__ copy updated from the host, as the
// reference's gc_data (PAR_DEVICE_CONST / PAR_SET_CONST).
#include "../../include/nmlmc_sampler.hpp"

namespace model_class {

using nmlmc::real_t;
using nmlmc::work_t;

// model::data (model_sampler.cpp:21-39); phi_type 0 Indicator, 1 Max.
struct data {
    real_t stau, s1tau, max_loss;
    int phi_type;
};
__constant__ data gc_data;

class sampler : public nmlmc::inner_sampler_base {
    real_t Y;

public:
    static uint32_t GetWorkspaceSize(const data&) { return 0; }
    __device__ sampler(void*) : Y(NAN) {}
    __device__ void sample_outer(uint64_t, work_t&) { Y = randn_sync(); }
    __device__ void copy_outer(const sampler& rhs) { Y = rhs.Y; }
    __device__ real_t init_inner_sampler(uint64_t&, work_t&) { return -gc_data.max_loss; }
    __device__ real_t sample_inner(uint64_t, uint64_t, work_t& total_work) {
        total_work++;
        const real_t X = randn();
        const real_t Y2 = randn();
        const real_t a = gc_data.stau * Y + gc_data.s1tau * X;
        real_t r = a * a;
        const real_t b = gc_data.stau * Y2 + gc_data.s1tau * X;
        r -= b * b;
        return r;
    }
    __device__ static void add_phi(real_t avg, real_t coeff, real_t* phi) {
        if (gc_data.phi_type == 1) *phi += coeff * (avg > 0 ? avg : 0.);
        else *phi += coeff * (real_t)(avg >= 0.);
    }
};

}  // namespace model_class

extern "C" {

// model_init (model_sampler.cpp:100-104) for the sampler class: the data to the
// device constant, then init_samplers<sampler>.
int nmlmc_sampler_model_init(const model_data* d, uint64_t seed) {
    if (!d) return nmlmc::set_error(NMLMC_ERR_ARG, "model data is NULL");
    const model_class::data c{d->stau, d->s1tau, d->max_loss, d->phi_type};
    const hipError_t e = hipMemcpyToSymbol(HIP_SYMBOL(model_class::gc_data), &c, sizeof(c));
    if (e != hipSuccess) return nmlmc::hip_fail(e, "hipMemcpyToSymbol model data");
    return nmlmc::init_samplers<model_class::sampler>(seed, model_class::sampler::GetWorkspaceSize(c));
}

// model_sample (model_sampler.cpp:106-114) for the sampler class.
int nmlmc_sampler_model_sample(uint32_t ell, uint64_t M, sample_output* out) {
    if (!out) return nmlmc::set_error(NMLMC_ERR_ARG, "sample_output is NULL");
    out->M = M;
    return nmlmc::sample<model_class::sampler>(ell, 0, &out->M, out->fsums, out->diffsums, &out->total_work,
                                               out->inner_stats);
}

}  // extern "C"
