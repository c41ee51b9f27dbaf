// computeNextStep without device allocation (smc_compute_next_step_noalloc).
//
// smc_compute_next_step stages the weights, log-likelihoods, normalised weights and
// the (T, ESS) pair in engine buffers that grow with hipMalloc on the first call of a
// larger population.  This variant stages them in one file-scope device array of
// fixed size instead, so it can be called where allocation is not allowed (tight
// sampler loops, harnesses that forbid hipMalloc).  Each GPU holds its own copy of
// the array, at its own address, from module load to process end.
//
// Host code never uses the array's name as a device address: smc_noalloc_init,
// called by smc_engine_create with the engine's device current, looks its address
// up once per device with hipGetSymbolAddress and the engine keeps that pointer.
#include <mutex>

#include "engine_state.hpp"

namespace {

constexpr int kNoAllocMaxParticles = 65536;
// [w | lnL | normalised] x kNoAllocMaxParticles, then T and ESS
constexpr size_t kScratchDoubles = 3 * (size_t)kNoAllocMaxParticles + 2;

static __device__ __align__(256) double scratch[kScratchDoubles];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double *g_scratch_ptr[kMaxDevices] = {};

}  // namespace

int smc_noalloc_init(smc_engine *e)
{
    int dev = -1;
    HIP_TRY(e, hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) return smc_fail(e, SMC_ERR_ARG, "device ordinal above 63");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void *p = nullptr;
        hipError_t st = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (st != hipSuccess || !p)
            return smc_fail(e, SMC_ERR_HIP,
                            std::string("hipGetSymbolAddress(scratch) failed: ") + hipGetErrorString(st));
        g_scratch_ptr[dev] = static_cast<double *>(p);
    }
    e->d_noalloc_scratch = g_scratch_ptr[dev];
    return SMC_OK;
}

extern "C" int smc_compute_next_step_noalloc(smc_engine *e, int n, double *raw_weights, const double *raw_lnl,
                                             double *normalized, double *ess, int target_ess,
                                             double *current_temp, double last_temp)
{
    if (!e) return SMC_ERR_ARG;
    if (n < 1 || !raw_weights || !raw_lnl || !normalized || !ess || !current_temp)
        return smc_fail(e, SMC_ERR_ARG, "invalid computeNextStep arguments");
    if (n > kNoAllocMaxParticles)
        return smc_fail(e, SMC_ERR_ARG,
                        "smc_compute_next_step_noalloc: " + std::to_string(n) +
                            " particles exceed its fixed scratch of " + std::to_string(kNoAllocMaxParticles));
    if (!e->d_noalloc_scratch) return smc_fail(e, SMC_ERR_STATE, "no-allocation scratch not initialised");
    double *const w = e->d_noalloc_scratch;
    double *const l = w + kNoAllocMaxParticles;
    double *const norm = l + kNoAllocMaxParticles;
    double *const out2 = norm + kNoAllocMaxParticles;
    hipStream_t s = e->stream;
    // every element the kernel reads ([0, n) of w and lnL) is written here first, and it
    // writes all of [0, n) of the normalised weights and both outputs: nothing carries
    // over from an earlier call
    HIP_TRY(e, hipMemcpyAsync(w, raw_weights, (size_t)n * sizeof(double), hipMemcpyHostToDevice, s));
    HIP_TRY(e, hipMemcpyAsync(l, raw_lnl, (size_t)n * sizeof(double), hipMemcpyHostToDevice, s));
    launch_next_step(n, w, l, norm, out2, target_ess, last_temp, s);
    HIP_TRY(e, hipGetLastError());
    double res[2];
    HIP_TRY(e, hipMemcpyAsync(raw_weights, w, (size_t)n * sizeof(double), hipMemcpyDeviceToHost, s));
    HIP_TRY(e, hipMemcpyAsync(normalized, norm, (size_t)n * sizeof(double), hipMemcpyDeviceToHost, s));
    HIP_TRY(e, hipMemcpyAsync(res, out2, 2 * sizeof(double), hipMemcpyDeviceToHost, s));
    HIP_TRY(e, hipStreamSynchronize(s));
    *current_temp = res[0];
    *ess = res[1];
    return SMC_OK;
}
