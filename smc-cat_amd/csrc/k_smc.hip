// SMC reweighting kernels.
//
// k_next_step restates computeNextStep (Miscellaneous.hpp:42-109): bisection on
// the next temperature T in [T_last, 1] while high - low > 1e-5, where each probe
// is ESS(T) = 1 / sum(w_hat^2) of the max-shifted softmax of
// w_n + (T - T_last) lnL_n; then the weights are committed and normalised.  One
// workgroup of 1024 lanes; every reduction is a fixed-order tree, so results
// are deterministic.
//
// k_resample restates the systematic resampling walk of main.cpp:179-195.  It is
// a sequential walk with accumulated rounding (ruler += 1/N, cumulative +=
// w_k), and the assignments must be bit-identical to the reference's for the
// same normalized weights and u.  Both running sums are therefore accumulated
// in order, each by one lane over LDS; the walk itself is then a search: the
// rulers rise and the cumulative weights never fall, so ruler j goes to the
// first k with ruler_j <= cumulative_k (n - 1 when there is none, where the
// walk clamps).  Populations beyond the LDS arrays, or with a NaN weight (where
// that order breaks), take the single-lane walk.
#include "internal.hpp"

namespace smccat {

namespace {

constexpr int kSmcThreads = 1024;

__device__ double block_sum(double v, double *sh)
{
    const int t = threadIdx.x;
    sh[t] = v;
    __syncthreads();
    for (int off = kSmcThreads / 2; off > 0; off >>= 1) {
        if (t < off) sh[t] += sh[t + off];
        __syncthreads();
    }
    double r = sh[0];
    __syncthreads();
    return r;
}

__device__ double block_max(double v, double *sh)
{
    const int t = threadIdx.x;
    sh[t] = v;
    __syncthreads();
    for (int off = kSmcThreads / 2; off > 0; off >>= 1) {
        if (t < off) sh[t] = fmax(sh[t], sh[t + off]);
        __syncthreads();
    }
    double r = sh[0];
    __syncthreads();
    return r;
}

// ESS of the weights w + (temp - last) l (Miscellaneous.hpp:49-74).
__device__ double ess_at(int n, const double *w, const double *l, double temp, double last,
                         double *sh)
{
    double mx = -INFINITY;
    for (int i = threadIdx.x; i < n; i += kSmcThreads) mx = fmax(mx, w[i] + (temp - last) * l[i]);
    mx = block_max(mx, sh);
    double tot = 0.0;
    for (int i = threadIdx.x; i < n; i += kSmcThreads) tot += exp(w[i] + (temp - last) * l[i] - mx);
    tot = block_sum(tot, sh);
    double sq = 0.0;
    for (int i = threadIdx.x; i < n; i += kSmcThreads) {
        double v = exp(w[i] + (temp - last) * l[i] - mx) / tot;
        sq += v * v;
    }
    sq = block_sum(sq, sh);
    return 1.0 / sq;
}

__global__ __launch_bounds__(kSmcThreads) void k_next_step(int n, double *__restrict__ w,
                                                           const double *__restrict__ l,
                                                           double *__restrict__ norm,
                                                           double *__restrict__ out2,
                                                           int target_ess, double last_temp)
{
    __shared__ double sh[kSmcThreads];
    double low = last_temp, high = 1.0;
    while (high - low > 1e-5) {
        double mid = 0.5 * (low + high);
        double e = ess_at(n, w, l, mid, last_temp, sh);
        if (e > (double)target_ess)
            low = mid;
        else
            high = mid;
    }
    const double temp = high;
    double mx = -INFINITY;
    for (int i = threadIdx.x; i < n; i += kSmcThreads) {
        double v = w[i] + (temp - last_temp) * l[i];
        w[i] = v;
        mx = fmax(mx, v);
    }
    mx = block_max(mx, sh);
    double tot = 0.0;
    for (int i = threadIdx.x; i < n; i += kSmcThreads) {
        double v = exp(w[i] - mx);
        norm[i] = v;
        tot += v;
    }
    tot = block_sum(tot, sh);
    double sq = 0.0;
    for (int i = threadIdx.x; i < n; i += kSmcThreads) {
        double v = norm[i] / tot;
        norm[i] = v;
        sq += v * v;
    }
    sq = block_sum(sq, sh);
    if (threadIdx.x == 0) {
        out2[0] = temp;
        out2[1] = 1.0 / sq;
    }
}

constexpr int kResampleLds = 4096;  // particles of the LDS form (two arrays of doubles: 64 KB)

__device__ void resample_walk(int n, const double *__restrict__ norm, double u, int *__restrict__ assign)
{
    double cumulative = norm[0];
    int k = 0, cnt = 0;
    double ruler = u / (double)n;
    const double step = 1.0 / (double)n;
    // The walk ends after at most 2n steps unless rounding pushes the ruler
    // past 1.0 (where the reference would not terminate); bound it anyway.
    long guard = 4L * n + 64;
    while (cnt < n && guard-- > 0) {
        if (ruler <= cumulative) {
            assign[cnt++] = k;
            ruler += step;
        } else {
            if (++k >= n) {
                k = n - 1;
                cumulative = 1.0;
            } else {
                cumulative += norm[k];
            }
        }
    }
    while (cnt < n) assign[cnt++] = n - 1;
}

__global__ __launch_bounds__(kSmcThreads) void k_resample(int n, const double *__restrict__ norm, double u,
                                                          int *__restrict__ assign)
{
    __shared__ double cum[kResampleLds], ruler[kResampleLds];
    __shared__ int bad;
    const int t = threadIdx.x;
    if (n > kResampleLds) {
        if (t == 0) resample_walk(n, norm, u, assign);
        return;
    }
    for (int i = t; i < n; i += kSmcThreads) cum[i] = norm[i];
    __syncthreads();
    if (t == 0) {  // cumulative weights, in the walk's order
        double c = cum[0];
        int nan = c != c;
        for (int k = 1; k < n; k++) {
            c += cum[k];
            cum[k] = c;
            nan |= c != c;
        }
        bad = nan;
    } else if (t == 64) {  // the rulers, in the walk's order
        double r = u / (double)n;
        const double step = 1.0 / (double)n;
        for (int j = 0; j < n; j++) {
            ruler[j] = r;
            r += step;
        }
    }
    __syncthreads();
    if (bad) {
        if (t == 0) resample_walk(n, norm, u, assign);
        return;
    }
    for (int j = t; j < n; j += kSmcThreads) {
        const double r = ruler[j];
        int lo = 0, hi = n;  // first k with r <= cum[k]
        while (lo < hi) {
            const int mid = (lo + hi) >> 1;
            if (r <= cum[mid])
                hi = mid;
            else
                lo = mid + 1;
        }
        assign[j] = lo < n ? lo : n - 1;
    }
}

// Row gather: row dst_rows[r] (r when null) of dst is row src_rows[r] (r when null) of
// src, rows of row_elems elements of W (16-, 4- or 1-byte words).  One workgroup per row.
template <typename W>
__global__ __launch_bounds__(256) void k_gather_rows(const W *__restrict__ src, W *__restrict__ dst, size_t row_elems,
                                                     const int *__restrict__ src_rows, const int *__restrict__ dst_rows)
{
    const size_t r = blockIdx.x;
    const W *s = src + (size_t)(src_rows ? src_rows[r] : (int)r) * row_elems;
    W *d = dst + (size_t)(dst_rows ? dst_rows[r] : (int)r) * row_elems;
    for (size_t i = threadIdx.x; i < row_elems; i += blockDim.x) d[i] = s[i];
}

// Scatter of packed rows: segment k copies segs[k].words 32-bit words from
// src + segs[k].src_word to the device address segs[k].dst.  One workgroup per
// segment.
__global__ __launch_bounds__(256) void k_scatter_segments(const uint32_t *__restrict__ src,
                                                          const ScatterSeg *__restrict__ segs)
{
    const ScatterSeg sg = segs[blockIdx.x];
    uint32_t *d = reinterpret_cast<uint32_t *>(sg.dst);
    const uint32_t *sp = src + sg.src_word;
    for (uint32_t i = threadIdx.x; i < sg.words; i += blockDim.x) d[i] = sp[i];
}

// Tip codes in each particle's class-sorted slot order: pcodes[p][t][slot] =
// min(codes[t][perm[p][slot]], 20) for particles [first, first + count).  One
// workgroup per (particle, tip) row.
__global__ __launch_bounds__(256) void k_permute_codes(const uint8_t *__restrict__ codes, const int *__restrict__ perm,
                                                       uint8_t *__restrict__ pcodes, int T, int S_pad, int first)
{
    const int p = first + (int)blockIdx.x, t = (int)blockIdx.y;
    const int *pm = perm + (size_t)p * S_pad;
    const uint8_t *src = codes + (size_t)t * S_pad;
    uint8_t *dst = pcodes + ((size_t)p * T + t) * S_pad;
    for (int sl = threadIdx.x; sl < S_pad; sl += blockDim.x) {
        const int c = src[pm[sl]];
        dst[sl] = (uint8_t)(c < 20 ? c : 20);
    }
}

}  // namespace

void launch_permute_codes(const uint8_t *codes, const int *perm, uint8_t *pcodes, int T, int S_pad, int first,
                          int count, hipStream_t s)
{
    if (count <= 0) return;
    hipLaunchKernelGGL(k_permute_codes, dim3(count, T), dim3(256), 0, s, codes, perm, pcodes, T, S_pad, first);
}

void launch_gather_rows(const void *src, void *dst, size_t row_bytes, const int *src_rows, const int *dst_rows,
                        int n, hipStream_t s)
{
    if (n <= 0 || row_bytes == 0) return;
    if (row_bytes % 16 == 0)
        hipLaunchKernelGGL(k_gather_rows<uint4>, dim3(n), dim3(256), 0, s, (const uint4 *)src, (uint4 *)dst,
                           row_bytes / 16, src_rows, dst_rows);
    else if (row_bytes % 4 == 0)
        hipLaunchKernelGGL(k_gather_rows<uint32_t>, dim3(n), dim3(256), 0, s, (const uint32_t *)src,
                           (uint32_t *)dst, row_bytes / 4, src_rows, dst_rows);
    else
        hipLaunchKernelGGL(k_gather_rows<uint8_t>, dim3(n), dim3(256), 0, s, (const uint8_t *)src, (uint8_t *)dst,
                           row_bytes, src_rows, dst_rows);
}

void launch_scatter_segments(const uint32_t *src, const ScatterSeg *segs, int n, hipStream_t s)
{
    if (n <= 0) return;
    hipLaunchKernelGGL(k_scatter_segments, dim3(n), dim3(256), 0, s, src, segs);
}

void launch_next_step(int n, double *w, const double *l, double *norm, double *out2,
                      int target_ess, double last_temp, hipStream_t s)
{
    hipLaunchKernelGGL(k_next_step, dim3(1), dim3(kSmcThreads), 0, s, n, w, l, norm, out2,
                       target_ess, last_temp);
}

void launch_resample(int n, const double *norm, double u, int *assign, hipStream_t s)
{
    hipLaunchKernelGGL(k_resample, dim3(1), dim3(kSmcThreads), 0, s, n, norm, u, assign);
}

}  // namespace smccat
