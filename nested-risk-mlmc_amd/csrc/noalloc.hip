// A level call of the plain FP64 form that allocates nothing: for callers that
// cannot allocate device memory (code under HIP graph capture, serving loops,
// harnesses that forbid hipMalloc).  No counterpart in the reference.
//
// The plain form's level call needs scratch beyond its inputs and outputs: one
// partial record per block of virtual threads (k_model_parity,
// k_portfolio_parity), the record the fixed-order reduction writes and its
// kReduceBlocks staging records.  Here they live in a file-scope device array
// of kNoAllocRecords + 1 + kReduceBlocks records (5.6 MB; each GPU the library
// runs on holds its own copy until the process ends).  Each block writes its
// whole record and each pass of the reduction writes whole records, so nothing
// needs clearing between calls.  The array's device address is taken once per
// device at set-up (noalloc_setup, from the library's init) and kept; host code
// uses that address, never the array's name.  Calls that overlap on one GPU
// share the array and race.
#include <hip/hip_runtime.h>

#include "host.hpp"

namespace nmlmc {

constexpr uint32_t kNoAllocRecords = 16384;  // 65536 virtual threads x 64 lanes / 256
constexpr int kNoAllocDevices = 64;

static __device__ __align__(256) Partial scratch[kNoAllocRecords + 1 + kReduceBlocks];

// The array's device address per device, from hipGetSymbolAddress.
static Partial* g_scratch_addr[kNoAllocDevices] = {};

int noalloc_setup(int device) {
    if (device < 0 || device >= kNoAllocDevices) return NMLMC_OK;  // the entry point then reports it
    if (g_scratch_addr[device]) return NMLMC_OK;
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
        (void)hipGetLastError();  // clear it: later launches check the last error
        return NMLMC_ERR_HIP;
    }
    g_scratch_addr[device] = (Partial*)p;
    return NMLMC_OK;
}

}  // namespace nmlmc

using nmlmc::g_engine;

extern "C" int nmlmc_sample_noalloc(uint32_t ell, uint64_t M, sample_output* out) {
    using namespace nmlmc;
    if (!g_engine.inited || !g_engine.stream)
        return set_error(NMLMC_ERR_STATE, "init_nestedmc_sampler has not been called");
    if (g_engine.form != NMLMC_FORM_PLAIN)
        return set_error(NMLMC_ERR_STATE, "nmlmc_sample_noalloc runs the plain form only (nmlmc_set_form)");
    const int dev = g_engine.device;
    Partial* base = (dev >= 0 && dev < kNoAllocDevices) ? g_scratch_addr[dev] : nullptr;
    if (!base) return set_error(NMLMC_ERR_STATE, "no-allocation scratch not set up on device %d", dev);
    g_engine.scr_partials = base;
    g_engine.scr_cap = kNoAllocRecords;
    g_engine.scr_result = base + kNoAllocRecords;
    const int rc = g_engine.sample(ell, M, UINT64_MAX, 0, out, nullptr, nullptr, nullptr, nullptr);
    g_engine.scr_partials = nullptr;
    g_engine.scr_cap = 0;
    g_engine.scr_result = nullptr;
    return rc;
}
