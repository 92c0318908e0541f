// Portfolio entry points (portfolio_sampler.cpp:618-756): host staging of the
// option tables and the level call.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <vector>

#include "../../include/nmlmc.h"
#include "host.hpp"

namespace nmlmc {

void Engine::free_portfolio() {
    for (void* p : pf_allocs) hipFree(p);
    pf_allocs.clear();
    std::memset(&pf, 0, sizeof(pf));
}

int Engine::sample_portfolio_parity(uint32_t, uint64_t, uint32_t, sample_output*, double*, double*,
                                    uint64_t*, uint64_t*) {
    return set_error(NMLMC_ERR_STATE, "portfolio plain form not available");
}

int Engine::sample_portfolio_batched(uint32_t, uint64_t, uint64_t, uint32_t, sample_output*, double*,
                                     double*, uint64_t*, uint64_t*) {
    return set_error(NMLMC_ERR_STATE, "portfolio batched form not available");
}

}  // namespace nmlmc

using nmlmc::g_engine;
using nmlmc::set_error;

extern "C" {

void portfolio_cleanup(void) {
    g_engine.free_portfolio();
    if (g_engine.problem == nmlmc::PROBLEM_PORTFOLIO) g_engine.problem = -1;
    set_error(NMLMC_OK, "");
}

void portfolio_update_data(const portfolio_data*) { set_error(NMLMC_ERR_STATE, "portfolio not available"); }
void portfolio_init(const portfolio_data*, uint64_t) { set_error(NMLMC_ERR_STATE, "portfolio not available"); }
void portfolio_comp_params(portfolio_data*) { set_error(NMLMC_ERR_STATE, "portfolio not available"); }
void portfolio_sample(uint32_t, uint64_t, uint64_t, sample_output*) {
    set_error(NMLMC_ERR_STATE, "portfolio not available");
}
int nmlmc_portfolio_set_correlation(const double*, uint32_t) {
    return set_error(NMLMC_ERR_STATE, "portfolio not available");
}

}  // extern "C"
