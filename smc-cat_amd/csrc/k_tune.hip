// Proposal tuning from the resampled population on the GPU (main.cpp:213-289).
//
// Over slots [0, n): the variance of log shape and of log branch length (non-root
// nodes), the mean pInvar, and for GTR particles the rate-matrix move's factor: the
// uncentred second moment of the 190 log-exchangeabilities (Q21) scaled by
// 2.38^2 / 190 * 0.21, shrunk by 0.2 towards its mean diagonal, plus 1e-6 on the
// diagonal, then its lower Cholesky factor.  Variances are the mean of squares minus
// the squared mean (boost accumulators' variance).  Every sum runs in a fixed order,
// so results do not depend on timing.
#include <cmath>

#include "engine_state.hpp"

namespace {

constexpr int kTuneThreads = 256;
constexpr int kRates = 190;

__device__ double block_sum256(double v, double *sh)
{
    const int t = threadIdx.x;
    sh[t] = v;
    __syncthreads();
    for (int off = kTuneThreads / 2; off > 0; off >>= 1) {
        if (t < off) sh[t] += sh[t + off];
        __syncthreads();
    }
    const double r = sh[0];
    __syncthreads();
    return r;
}

// out[0..4] = sum log shape, sum log^2 shape, sum pinv, sum log blen, sum log^2 blen;
// out[5] = branches counted.
__global__ __launch_bounds__(kTuneThreads) void k_tune_scalars(const double *__restrict__ shape,
                                                               const double *__restrict__ pinv,
                                                               const double *__restrict__ blen,
                                                               const int *__restrict__ root, int n, int nn,
                                                               double *__restrict__ out)
{
    __shared__ double sh[kTuneThreads];
    double a = 0.0, b = 0.0, c = 0.0, d = 0.0, f = 0.0, g = 0.0;
    for (int p = threadIdx.x; p < n; p += kTuneThreads) {
        const double ls = log(shape[p]);
        a += ls;
        b += ls * ls;
        c += pinv[p];
        const double *bl = blen + (size_t)p * nn;
        for (int v = 0; v < nn; ++v) {
            if (v == root[p]) continue;
            const double l = log(bl[v]);
            d += l;
            f += l * l;
            g += 1.0;
        }
    }
    a = block_sum256(a, sh);
    b = block_sum256(b, sh);
    c = block_sum256(c, sh);
    d = block_sum256(d, sh);
    f = block_sum256(f, sh);
    g = block_sum256(g, sh);
    if (threadIdx.x == 0) {
        out[0] = a;
        out[1] = b;
        out[2] = c;
        out[3] = d;
        out[4] = f;
        out[5] = g;
    }
}

// m[i][j] = sum_p base[p][i] base[p][j] for j <= i (particle order), row-major 190 x 190.
__global__ __launch_bounds__(kTuneThreads) void k_tune_moment(const double *__restrict__ base, int n,
                                                              double *__restrict__ m)
{
    const int q = blockIdx.x * kTuneThreads + threadIdx.x;
    if (q >= kRates * kRates) return;
    const int i = q / kRates, j = q % kRates;
    if (j > i) {
        m[q] = 0.0;
        return;
    }
    double acc = 0.0;
    for (int p = 0; p < n; ++p) acc += base[(size_t)p * kRates + i] * base[(size_t)p * kRates + j];
    m[q] = acc;
}

// Scale and shrink the moment, then factor it in place (right-looking, one workgroup);
// *status = 1 when it is not positive definite.  diag_out = the shrinkage diagonal.
__global__ __launch_bounds__(kTuneThreads) void k_tune_factor(double *__restrict__ a, double *__restrict__ diag_out,
                                                              int *__restrict__ status)
{
    __shared__ double sh[kTuneThreads];
    __shared__ int bad;
    const int t = threadIdx.x;
    const double f = (2.38 * 2.38) / 190.0 * 0.21, shrink = 0.2;
    double dsum = 0.0;
    for (int i = t; i < kRates; i += kTuneThreads) dsum += a[i * kRates + i] * f;
    const double diag = block_sum256(dsum, sh) / 190.0 * shrink;
    for (int q = t; q < kRates * kRates; q += kTuneThreads) {
        const int i = q / kRates, j = q % kRates;
        if (j <= i) a[q] = a[q] * (f * (1.0 - shrink)) + (i == j ? diag + 1e-6 : 0.0);
    }
    if (t == 0) bad = 0;
    __syncthreads();
    for (int j = 0; j < kRates; ++j) {
        const double d2 = a[j * kRates + j];
        if (!(d2 > 0.0)) {
            if (t == 0) bad = 1;
            break;
        }
        const double d = sqrt(d2);
        __syncthreads();
        for (int i = j + 1 + t; i < kRates; i += kTuneThreads) a[i * kRates + j] /= d;
        if (t == 0) a[j * kRates + j] = d;
        __syncthreads();
        // trailing update of the lower part: a[i][k] -= a[i][j] a[k][j], j < k <= i
        const int rem = kRates - j - 1;
        for (int q = t; q < rem * rem; q += kTuneThreads) {
            const int i = j + 1 + q / rem, k = j + 1 + q % rem;
            if (k <= i) a[i * kRates + k] -= a[i * kRates + j] * a[k * kRates + j];
        }
        __syncthreads();
    }
    __syncthreads();
    for (int q = t; q < kRates * kRates; q += kTuneThreads) {
        const int i = q / kRates, j = q % kRates;
        if (j > i) a[q] = 0.0;
    }
    if (t == 0) {
        *status = bad;
        *diag_out = diag;
    }
}

}  // namespace

extern "C" int smc_proposal_tuning(smc_engine *e, int n, int using_lg, double *out, double *rate_cholesky)
{
    if (!e) return SMC_ERR_ARG;
    if (n < 1 || n > e->Pmax || !out) return smc_fail(e, SMC_ERR_ARG, "invalid tuning arguments");
    for (int p = 0; p < n; ++p)
        if (!e->h_valid[p]) return smc_fail(e, SMC_ERR_STATE, "tuning: slot not uploaded");
    hipStream_t s = e->stream;
    // workspace: shape[n], scalars[8], status, diag, moment[190 x 190]
    const size_t need = (size_t)n * 8 + 8 * 8 + 16 + (size_t)kRates * kRates * 8 + 1024;
    if (need > e->split_ws_bytes) {
        HIP_TRY(e, hipStreamSynchronize(s));
        if (e->d_split_ws) (void)hipFree(e->d_split_ws);
        e->d_split_ws = nullptr;
        e->split_ws_bytes = 0;
        HIP_TRY(e, hipMalloc((void **)&e->d_split_ws, need));
        e->split_ws_bytes = need;
    }
    uint8_t *ws = e->d_split_ws;
    double *d_moment = reinterpret_cast<double *>(ws);
    double *d_shape = d_moment + (size_t)kRates * kRates;
    double *d_scal = d_shape + n;
    double *d_diag = d_scal + 8;
    int *d_status = reinterpret_cast<int *>(d_diag + 1);
    int rc = h2d(e, d_shape, e->h_shape.data(), (size_t)n * sizeof(double));
    if (rc != SMC_OK) return rc;
    hipLaunchKernelGGL(k_tune_scalars, dim3(1), dim3(kTuneThreads), 0, s, d_shape, e->d_pinv, e->d_blen, e->d_root, n,
                       e->nn, d_scal);
    const bool factor = !using_lg && rate_cholesky;
    if (factor) {
        hipLaunchKernelGGL(k_tune_moment, dim3((kRates * kRates + kTuneThreads - 1) / kTuneThreads), dim3(kTuneThreads),
                           0, s, e->d_base, n, d_moment);
        hipLaunchKernelGGL(k_tune_factor, dim3(1), dim3(kTuneThreads), 0, s, d_moment, d_diag, d_status);
    }
    HIP_TRY(e, hipGetLastError());
    double sc[6];
    double diag = 0.0;
    int status = 0;
    HIP_TRY(e, hipMemcpyAsync(sc, d_scal, sizeof(sc), hipMemcpyDeviceToHost, s));
    if (factor) {
        HIP_TRY(e, hipMemcpyAsync(&diag, d_diag, sizeof(double), hipMemcpyDeviceToHost, s));
        HIP_TRY(e, hipMemcpyAsync(&status, d_status, sizeof(int), hipMemcpyDeviceToHost, s));
        HIP_TRY(e, hipMemcpyAsync(rate_cholesky, d_moment, (size_t)kRates * kRates * sizeof(double),
                                  hipMemcpyDeviceToHost, s));
    }
    HIP_TRY(e, hipStreamSynchronize(s));
    if (status) return smc_fail(e, SMC_ERR_STATE, "rate-matrix proposal covariance is not positive definite");
    const double mls = sc[0] / n, mbl = sc[5] > 0.0 ? sc[3] / sc[5] : 0.0;
    out[0] = sc[1] / n - mls * mls;
    out[1] = sc[5] > 0.0 ? sc[4] / sc[5] - mbl * mbl : 0.0;
    out[2] = sc[2] / n;
    out[3] = diag;
    return SMC_OK;
}
