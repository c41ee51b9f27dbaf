// Single-site likelihood of one alignment column under many candidate CAT
// profiles: the per-site, per-class conditional that the Gibbs partition move
// (Particle.cpp:857-941) and the learned allocation proposal both need.
//
// k_site_class_lv (trees up to kSiteLevelMaxTaxa taxa) walks the tree by height
// levels with twelve nodes in flight; k_site_class, for larger trees, is one
// wavefront per (entry, candidate, rate): entry = (particle, site), walking a
// Strahler-ordered post-order prepared on the host, so only O(log T) internal
// partials are live; they sit on a small LDS stack.
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


// LDS written by lanes of one wavefront is visible to its other lanes after this
__device__ __forceinline__ void wave_sync_lds()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Where candidate ec of an entry of particle p keeps its eigensystem (see SiteClassArgs).
struct CandEig {
    const double *cand_eig;
    const int *cls;
    const double *cache;
    int max_classes;
    const double *pool;
};
__device__ __forceinline__ const double *cand_eigensystem(const CandEig &c, int ec, int p)
{
    const int k = c.cls ? c.cls[ec] : -1;
    if (k >= 0) return c.cache + ((size_t)p * c.max_classes + k) * kEigStride;
    if (k <= -2) return c.pool + (size_t)(-2 - k) * kEigStride;
    return c.cand_eig + (size_t)ec * kEigStride;
}

__global__ __launch_bounds__(64) void k_site_class(
    const int *__restrict__ ent_particle, const int *__restrict__ ent_site,
    const int *__restrict__ ent_ncand, int n_cand, CandEig ce,
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
    const double *eg = cand_eigensystem(ce, ec, p);
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
    const int *pg = prog + (size_t)p * site_prog_stride(T);
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
    if (lane == 0) rate_lnl[id] = rate_log_term(d, ez);
}

// Level-parallel form: one 256-lane workgroup per (entry, candidate, rate).  Each
// wavefront holds three 20-lane groups (lanes 60-63 idle), twelve per workgroup;
// the groups share the nodes of one height level, so a level costs one node's
// latency and a C4-sized tree (200 taxa, height about 17) takes about 30 steps
// instead of 199.  Every internal partial stays in LDS; a group exchanges the
// vector between its two matrix-vector products inside its own wavefront.  The
// arithmetic of every node is the same as in k_site_class, in the same order.
constexpr int kLvGroups = 12;

__global__ __launch_bounds__(256) void k_site_class_lv(
    const int *__restrict__ ent_particle, const int *__restrict__ ent_site,
    const int *__restrict__ ent_ncand, int n_cand, CandEig ce,
    const int *__restrict__ children, const int *__restrict__ prog, const uint8_t *__restrict__ codes,
    const double *__restrict__ blen, const double *__restrict__ rates,
    double *__restrict__ rate_lnl, int T, int R, int S_pad)
{
    extern __shared__ double lv_smem[];
    const int Ti = T - 1;
    double *zs = lv_smem;                    // [T-1][20] scaled partials by node - T
    double *tmp = zs + (size_t)Ti * 20;      // [12][20] per-group exchange
    int *ezs = (int *)(tmp + kLvGroups * 20);  // [T-1] cumulative exponents

    const int id = blockIdx.x;  // (entry * n_cand + cand) * R + r
    const int r = id % R;
    const int ec = id / R;
    const int cand = ec % n_cand, entry = ec / n_cand;
    if (cand >= ent_ncand[entry]) return;  // whole workgroup exits; no barrier crossed
    const int p = ent_particle[entry], site = ent_site[entry];
    const int tid = threadIdx.x, wl = tid & 63;
    const int gw = wl / 20;                // group within the wavefront (3: idle lanes)
    const bool act = gw < 3;
    const int i = act ? wl - 20 * gw : 0;  // state owned by this lane
    const int g = (tid >> 6) * 3 + gw;     // group within the workgroup
    double *tg = tmp + (act ? g : 0) * 20;
    const double *eg = cand_eigensystem(ce, ec, p);
    double arow[20], brow[20];
#pragma unroll
    for (int k = 0; k < 20; k++) {
        arow[k] = eg[i * 20 + k];
        brow[k] = eg[kEigB + i * 20 + k];
    }
    const double lam = eg[kEigLam + i];
    const double rate = rates[(size_t)p * R + r];
    const int nn = 2 * T - 1;
    const int *ch = children + (size_t)p * Ti * 2;
    const int *lvn = prog + (size_t)p * site_prog_stride(T) + Ti;  // nodes by height
    const int H = lvn[Ti];
    const int *off = lvn + Ti + 1;
    const double *bl = blen + (size_t)p * nn;

    for (int h = 0; h < H; h++) {
        if (act) {
            for (int k = off[h] + g; k < off[h + 1]; k += kLvGroups) {
                const int n = lvn[k];  // uniform over the group's 20 lanes
                const int c0 = ch[(n - T) * 2], c1 = ch[(n - T) * 2 + 1];
                double y = 1.0;
                int e = 0;
#pragma unroll
                for (int c = 0; c < 2; c++) {
                    const int child = c ? c1 : c0;
                    double u = 0.0;
                    if (child < T) {
                        const int code = codes[(size_t)child * S_pad + site];
                        if (code < 20) {
                            u = eg[kEigB + i * 20 + code];
                        } else {
#pragma unroll
                            for (int q = 0; q < 20; q++) u += brow[q];
                        }
                    } else {
                        const double *zc = zs + (size_t)(child - T) * 20;
#pragma unroll
                        for (int q = 0; q < 20; q++) u = fma(brow[q], zc[q], u);
                        e += ezs[child - T];
                    }
                    const double v = exp(lam * bl[child] * rate) * u;
                    tg[i] = v;
                    wave_sync_lds();
                    double yc = 0.0;
#pragma unroll
                    for (int q = 0; q < 20; q++) yc = fma(arow[q], tg[q], yc);
                    wave_sync_lds();  // the group's reads are done before tg is rewritten
                    y *= yc;
                }
                // binary-exponent scaling of the 20 states
                tg[i] = y;
                wave_sync_lds();
                double m = 0.0;
#pragma unroll
                for (int q = 0; q < 20; q++) m = fmax(m, tg[q]);
                wave_sync_lds();
                int ex;
                (void)frexp(m, &ex);
                const bool ok = m > 0.0 && isfinite(m);
                zs[(size_t)(n - T) * 20 + i] = ok ? ldexp(y, -ex) : NAN;
                if (i == 0) ezs[n - T] = e + (ok ? ex : 0);
            }
        }
        __syncthreads();  // this level's partials are visible to the next
    }
    // root (the one node of the top level): l_r = log(pi . z) + e ln 2
    if (tid < 64) {
        const int root = lvn[Ti - 1];
        const double d0 = tid < 20 ? eg[kEigPi + tid] * zs[(size_t)(root - T) * 20 + tid] : 0.0;
        double d = d0;
#pragma unroll
        for (int o = 16; o >= 1; o >>= 1) d += __shfl_xor(d, o, 64);
        d += __shfl_xor(d, 32, 64);
        if (tid == 0) rate_lnl[id] = rate_log_term(d, ezs[root - T]);
    }
}

__global__ void k_site_class_fold(const int *__restrict__ ent_particle, const int *__restrict__ ent_site,
                                  const int *__restrict__ ent_ncand, int n_cand, CandEig ce,
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
    const double *pi = cand_eigensystem(ce, ec, p) + kEigPi;
    out[ec] = tempering * site_log_likelihood(m, s, pinv[p], pi[invchar[site]], isinv[site], mix_mode);
}

}  // namespace

void launch_site_class(const SiteClassArgs &a, hipStream_t s)
{
    if (a.count <= 0) return;
    const unsigned blocks = (unsigned)((size_t)a.count * a.n_cand * a.R);
    const CandEig ce{a.cand_eig, a.cand_cls, a.eig_cache, a.max_classes, a.pool};
    if (a.T <= kSiteLevelMaxTaxa) {
        const size_t lds = ((size_t)(a.T - 1) * 20 + kLvGroups * 20) * sizeof(double) + (size_t)(a.T - 1) * sizeof(int);
        hipLaunchKernelGGL(k_site_class_lv, dim3(blocks), dim3(256), lds, s, a.ent_particle, a.ent_site,
                           a.ent_ncand, a.n_cand, ce, a.children, a.prog, a.codes, a.blen,
                           a.rates, a.rate_lnl, a.T, a.R, a.S_pad);
    } else {
        hipLaunchKernelGGL(k_site_class, dim3(blocks), dim3(64), 0, s, a.ent_particle, a.ent_site,
                           a.ent_ncand, a.n_cand, ce, a.children, a.prog, a.codes, a.blen,
                           a.rates, a.rate_lnl, a.T, a.R, a.S_pad);
    }
    const int n = a.count * a.n_cand;
    hipLaunchKernelGGL(k_site_class_fold, dim3((n + 127) / 128), dim3(128), 0, s, a.ent_particle,
                       a.ent_site, a.ent_ncand, a.n_cand, ce, a.rate_lnl, a.pinv, a.isinv,
                       a.invchar, a.R, a.tempering, a.mix_mode, a.count, a.out);
}

}  // namespace smccat
