// Single-site likelihood of one alignment column under many candidate CAT
// profiles: the per-site, per-class conditional that the Gibbs partition move
// (Particle.cpp:857-941) and the learned allocation proposal both need.
//
// One wavefront scores one (entry, candidate, rate): entry = (particle, site).
// The tree is walked in a Strahler-ordered post-order prepared on the host, so
// only O(log T) internal partials are live; they sit on a small LDS stack.
// P(t) is never formed: a child's contribution is A (e^{lam t r} (B x)),
// 2 x 400 FMA per child instead of the 8,000 of a back-transform.  Lane i
// (< 20) owns state i and keeps row i of A and B and lam_i in registers.
// Partials are scaled by 2^-e (frexp of the max) with cumulative exponents, as
// in the sweep kernels.  A second kernel folds the rates: rate weights 1, the
// invariant term, the reference or log-space mixture, times the tempering.
#include "internal.hpp"
#include "root_mixture.hpp"

namespace smccat {

namespace {

constexpr double kLn2s = 0.69314718055994530942;

__global__ __launch_bounds__(64) void k_site_class(
    const int *__restrict__ ent_particle, const int *__restrict__ ent_site,
    const int *__restrict__ ent_ncand, int n_cand, const double *__restrict__ cand_eig,
    const int *__restrict__ children, const int *__restrict__ prog, const uint8_t *__restrict__ codes,
    const double *__restrict__ blen, const double *__restrict__ rates,
    double *__restrict__ rate_lnl, int T, int R, int S_pad)
{
    __shared__ double stack[kSiteStack][20];
    __shared__ int stack_e[kSiteStack];     // cumulative exponent of the entry
    __shared__ int stack_node[kSiteStack];  // node id of the entry
    __shared__ double tmp[20];

    const int id = blockIdx.x;  // (entry * n_cand + cand) * R + r
    const int r = id % R;
    const int ec = id / R;
    const int cand = ec % n_cand, entry = ec / n_cand;
    if (cand >= ent_ncand[entry]) return;  // whole block exits; no barrier crossed
    const int p = ent_particle[entry], site = ent_site[entry];
    const int lane = threadIdx.x;
    const bool act = lane < 20;
    const int i = act ? lane : 0;
    const double *eg = cand_eig + (size_t)ec * kEigStride;
    double arow[20], brow[20];
#pragma unroll
    for (int k = 0; k < 20; k++) {
        arow[k] = eg[i * 20 + k];
        brow[k] = eg[kEigB + i * 20 + k];
    }
    const double lam = eg[kEigLam + i];
    const double rate = rates[(size_t)p * R + r];
    const int nn = 2 * T - 1;
    const int *ch = children + (size_t)p * (T - 1) * 2;
    const int *pg = prog + (size_t)p * (T - 1);
    const double *bl = blen + (size_t)p * nn;

    int top = 0;  // stack size (wave-uniform)
    double z = 0.0;
    int ez = 0;
    for (int step = 0; step < T - 1; step++) {
        const int n = pg[step];
        const int c0 = ch[(n - T) * 2], c1 = ch[(n - T) * 2 + 1];
        // internal children are the top stack entries; match them by node id
        int slot0 = -1, slot1 = -1;
        if (c0 >= T && c1 >= T) {
            const bool c1_on_top = stack_node[top - 1] == c1;
            slot1 = c1_on_top ? top - 1 : top - 2;
            slot0 = c1_on_top ? top - 2 : top - 1;
            top -= 2;
        } else if (c0 >= T) {
            slot0 = top - 1;
            top -= 1;
        } else if (c1 >= T) {
            slot1 = top - 1;
            top -= 1;
        }
        double y = 1.0;
        int e = 0;
#pragma unroll
        for (int c = 0; c < 2; c++) {
            const int child = c ? c1 : c0;
            const int slot = c ? slot1 : slot0;
            double u = 0.0;
            if (child < T) {
                const int code = codes[(size_t)child * S_pad + site];
                if (code < 20) {
                    u = eg[kEigB + i * 20 + code];
                } else {
#pragma unroll
                    for (int k = 0; k < 20; k++) u += brow[k];
                }
            } else {
#pragma unroll
                for (int k = 0; k < 20; k++) u = fma(brow[k], stack[slot][k], u);
                e += stack_e[slot];
            }
            const double v = exp(lam * bl[child] * rate) * u;
            __syncthreads();
            if (act) tmp[i] = v;
            __syncthreads();
            double yc = 0.0;
#pragma unroll
            for (int k = 0; k < 20; k++) yc = fma(arow[k], tmp[k], yc);
            y *= yc;
        }
        // binary-exponent scaling of the 20 states (lanes >= 20 hold 0)
        double m = act ? y : 0.0;
#pragma unroll
        for (int off = 16; off >= 1; off >>= 1) m = fmax(m, __shfl_xor(m, off, 64));
        m = fmax(m, __shfl_xor(m, 32, 64));
        int ex;
        (void)frexp(m, &ex);
        const bool ok = m > 0.0 && isfinite(m);
        z = ok ? ldexp(y, -ex) : NAN;
        ez = e + (ok ? ex : 0);
        if (step < T - 2) {  // not the root: push
            __syncthreads();
            if (act) stack[top][i] = z;
            if (lane == 0) {
                stack_e[top] = ez;
                stack_node[top] = n;
            }
            __syncthreads();
            top++;
        }
    }
    // root: l_r = log(pi . z) + e ln 2
    double d = act ? eg[kEigPi + i] * z : 0.0;
#pragma unroll
    for (int off = 16; off >= 1; off >>= 1) d += __shfl_xor(d, off, 64);
    d += __shfl_xor(d, 32, 64);
    if (lane == 0) rate_lnl[id] = log(d) + (double)ez * kLn2s;
}

__global__ void k_site_class_fold(const int *__restrict__ ent_particle, const int *__restrict__ ent_site,
                                  const int *__restrict__ ent_ncand, int n_cand,
                                  const double *__restrict__ cand_eig,
                                  const double *__restrict__ rate_lnl,
                                  const double *__restrict__ pinv, const double *__restrict__ isinv,
                                  const int *__restrict__ invchar, int R, double tempering,
                                  int mix_mode, int count, double *__restrict__ out)
{
    const int ec = blockIdx.x * blockDim.x + threadIdx.x;
    if (ec >= count * n_cand) return;
    const int entry = ec / n_cand, cand = ec % n_cand;
    if (cand >= ent_ncand[entry]) {
        out[ec] = NAN;
        return;
    }
    const int p = ent_particle[entry], site = ent_site[entry];
    double m = -INFINITY, s = 0.0;
    for (int r = 0; r < R; r++) lse_push(m, s, rate_lnl[(size_t)ec * R + r], mix_mode);
    const double *pi = cand_eig + (size_t)ec * kEigStride + kEigPi;
    out[ec] = tempering * site_log_likelihood(m, s, pinv[p], pi[invchar[site]], isinv[site], mix_mode);
}

}  // namespace

void launch_site_class(const SiteClassArgs &a, hipStream_t s)
{
    if (a.count <= 0) return;
    const unsigned blocks = (unsigned)((size_t)a.count * a.n_cand * a.R);
    hipLaunchKernelGGL(k_site_class, dim3(blocks), dim3(64), 0, s, a.ent_particle, a.ent_site,
                       a.ent_ncand, a.n_cand, a.cand_eig, a.children, a.prog, a.codes, a.blen,
                       a.rates, a.rate_lnl, a.T, a.R, a.S_pad);
    const int n = a.count * a.n_cand;
    hipLaunchKernelGGL(k_site_class_fold, dim3((n + 127) / 128), dim3(128), 0, s, a.ent_particle,
                       a.ent_site, a.ent_ncand, a.n_cand, a.cand_eig, a.rate_lnl, a.pinv, a.isinv,
                       a.invchar, a.R, a.tempering, a.mix_mode, a.count, a.out);
}

}  // namespace smccat
