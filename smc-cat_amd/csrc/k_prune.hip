// Felsenstein pruning over site patterns with per-node binary log-scaling and
// the root mixture fused in.
//
// Reference: Particle::refreshLikelihood (Particle.cpp:577-675).  The reference
// walks node -> child -> class -> rate -> member site doing a 20x20 matvec per
// (site, rate), then divides each partial by its max and stores log(max)
// (:614-624), and at the root sums log(pi . x) and the log-scales of all nodes,
// log-sum-exps over rates with weight 1 (Q4) and adds the invariant-site term
// (:634-674).
//
// Here one launch processes one level of the trees (all (particle, internal node)
// pairs whose children are done).  One wavefront handles one tile of <= 64
// class-sorted site slots of one (particle, node): lane = site slot, loop over
// rates.  The tile's P matrices are wave-uniform (same particle, class, child,
// rate), so they come through scalar loads; partials are stored state-major
// [node][rate][state][site] so every load/store of a state is a coalesced
// 64-lane row.  Instead of dividing by the max, each partial is multiplied by
// 2^-e with e the binary exponent of its max (exact), and the cumulative
// exponent of the subtree is kept per (node, rate, site), so the root needs no
// sum over all nodes.  At the root the tile's per-site log-likelihoods are
// summed (wavefront butterfly, fixed order) into one value per tile.
#include "internal.hpp"
#include "root_mixture.hpp"

namespace smccat {

namespace {


template <typename TX>
struct PruneTypes;

template <>
struct PruneTypes<double> {
    using acc = double;
};
template <>
struct PruneTypes<float> {
    using acc = float;
};

// y = P x for the first child.  Tips: x is one-hot (code < 20) or all ones, so
// P x is column `code` of P or its row sums (the stored 21st column).
template <typename TX>
__device__ __forceinline__ void child_vector(int child, int cslot, int T, int r, int s, int orig,
                                             const TX *__restrict__ Pc,
                                             const uint8_t *__restrict__ codes, int S_pad,
                                             const TX *__restrict__ Xp, const int *__restrict__ Ep,
                                             int R, TX y[20], int &e)
{
    if (child < T) {
        int code = codes[(size_t)child * S_pad + orig];
        int col = code < 20 ? code : 20;
#pragma unroll
        for (int i = 0; i < 20; i++) y[i] = Pc[i * 21 + col];
        e = 0;
    } else {
        const size_t cl = (size_t)cslot;
        const TX *x = Xp + ((cl * R + r) * 20) * (size_t)S_pad + s;
        TX xv[20];
#pragma unroll
        for (int j = 0; j < 20; j++) xv[j] = x[(size_t)j * S_pad];
        e = Ep[(cl * R + r) * (size_t)S_pad + s];
#pragma unroll
        for (int i = 0; i < 20; i++) {
            TX acc = 0;
#pragma unroll
            for (int j = 0; j < 20; j++) acc = fma(Pc[i * 21 + j], xv[j], acc);
            y[i] = acc;
        }
    }
}

// y *= P x for the second child, row by row, so only y and x are live.
template <typename TX>
__device__ __forceinline__ void child_multiply(int child, int cslot, int T, int r, int s, int orig,
                                               const TX *__restrict__ Pc,
                                               const uint8_t *__restrict__ codes, int S_pad,
                                               const TX *__restrict__ Xp,
                                               const int *__restrict__ Ep, int R, TX y[20],
                                               int &e)
{
    if (child < T) {
        int code = codes[(size_t)child * S_pad + orig];
        int col = code < 20 ? code : 20;
#pragma unroll
        for (int i = 0; i < 20; i++) y[i] *= Pc[i * 21 + col];
        e = 0;
    } else {
        const size_t cl = (size_t)cslot;
        const TX *x = Xp + ((cl * R + r) * 20) * (size_t)S_pad + s;
        TX xv[20];
#pragma unroll
        for (int j = 0; j < 20; j++) xv[j] = x[(size_t)j * S_pad];
        e = Ep[(cl * R + r) * (size_t)S_pad + s];
#pragma unroll
        for (int i = 0; i < 20; i++) {
            TX acc = 0;
#pragma unroll
            for (int j = 0; j < 20; j++) acc = fma(Pc[i * 21 + j], xv[j], acc);
            y[i] *= acc;
        }
    }
}

template <typename TX, bool IS_ROOT>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3))) void k_prune(const Item *__restrict__ items,
                                               const Tile *__restrict__ tiles,
                                               const int *__restrict__ ntiles, int tile_stride,
                                               const int *__restrict__ children,
                                               const int *__restrict__ root_of,
                                               const int *__restrict__ perm,
                                               const uint8_t *__restrict__ codes,
                                               const TX *__restrict__ Pm,
                                               const double *__restrict__ eig,
                                               const double *__restrict__ pinv,
                                               const double *__restrict__ isinv,
                                               const int *__restrict__ invchar,
                                               TX *__restrict__ X, int *__restrict__ E,
                                               double *__restrict__ tile_lnl, int T, int R,
                                               int S_pad, int max_classes, int blocks_per_item,
                                               const int *__restrict__ pcls_off, int p_base, int mix_mode,
                                               const uint8_t *__restrict__ xflag, int WS,
                                               const double *__restrict__ pweight)
{
    // 1-D grid: consecutive blocks walk the tiles of one item, 4 tiles (one per
    // wavefront) per block, so concurrently running waves stream contiguous
    // stretches of each partial row.
    const int item = blockIdx.x / blocks_per_item;
    const int wave = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    const int tile = (blockIdx.x % blocks_per_item) * 4 + wave;
    const Item it = items[item];
    const int p = it.particle, n = it.node;
    if (tile >= ntiles[p]) return;  // whole wavefront exits; no block barrier follows
    const Tile tl = tiles[(size_t)p * tile_stride + tile];
    const int lane = threadIdx.x & 63;
    const bool active = lane >= tl.lo && lane < tl.len;
    const int s = tl.start + (active ? lane : tl.lo);
    const int orig = perm[(size_t)p * S_pad + s];
    const size_t g = (size_t)p * max_classes + tl.cls;
    const int nn = 2 * T - 1;
    const int nl = n - T;
    const int c0 = children[((size_t)p * (T - 1) + nl) * 2];
    const int c1 = children[((size_t)p * (T - 1) + nl) * 2 + 1];
    constexpr bool is_root = IS_ROOT;
    (void)root_of;

    const size_t node_elems = (size_t)R * 20 * S_pad;
    const size_t lp = (size_t)(p - p_base);  // slot of the particle in the wave's partials
    TX *Xp = X + lp * WS * (T - 1) * node_elems;
    int *Ep = E + lp * WS * (T - 1) * R * S_pad;
    const uint8_t *xf = xflag ? xflag + (size_t)p * (T - 1) : nullptr;
    auto nslot = [&](int node) { return (xf ? (int)xf[node - T] * (T - 1) : 0) + node - T; };
    const int cs0 = c0 >= T ? nslot(c0) : 0, cs1 = c1 >= T ? nslot(c1) : 0;
    const TX *Pg = Pm + ((size_t)pcls_off[p] + tl.cls) * nn * R * kPStride;
    const double *pi = eig + g * kEigStride + kEigPi;

    double lse_m = -INFINITY, lse_s = 0.0;
    for (int r = 0; r < R; r++) {
        TX y0[20];
        int e0, e1;
        child_vector<TX>(c0, cs0, T, r, s, orig, Pg + ((size_t)c0 * R + r) * kPStride, codes, S_pad,
                         Xp, Ep, R, y0, e0);
        child_multiply<TX>(c1, cs1, T, r, s, orig, Pg + ((size_t)c1 * R + r) * kPStride, codes, S_pad,
                           Xp, Ep, R, y0, e1);
        TX mx = 0;
#pragma unroll
        for (int i = 0; i < 20; i++) mx = fmax(mx, y0[i]);
        int ex;
        (void)frexp(mx, &ex);
        const bool ok = (mx > (TX)0) && isfinite(mx);
        // the reference divides by the max (0/0 -> NaN for an all-zero partial)
        const TX nanv = (TX)NAN;
#pragma unroll
        for (int i = 0; i < 20; i++) y0[i] = ok ? (TX)ldexp(y0[i], -ex) : nanv;
        const int etot = e0 + e1 + (ok ? ex : 0);
        if constexpr (!is_root) {
            if (active) {
                const int ns = nslot(n);
                TX *xo = Xp + (((size_t)ns * R + r) * 20) * S_pad + s;
#pragma unroll
                for (int i = 0; i < 20; i++) xo[(size_t)i * S_pad] = y0[i];
                Ep[((size_t)ns * R + r) * S_pad + s] = etot;
            }
        } else {
            double dot = 0.0;
#pragma unroll
            for (int i = 0; i < 20; i++) dot = fma(pi[i], (double)y0[i], dot);
            double l = rate_log_term(dot, etot);
            // running log-sum-exp over rates (rate weights are 1, Particle.cpp:661-665)
            lse_push(lse_m, lse_s, l, mix_mode);
        }
    }
    if constexpr (is_root) {
        double lf = 0.0;
        if (active) {
            lf = pweight[(size_t)p * S_pad + s] *
                 site_log_likelihood(lse_m, lse_s, pinv[p], pi[invchar[orig]], isinv[orig], mix_mode);
        }
#pragma unroll
        for (int off = 32; off >= 1; off >>= 1) lf += __shfl_xor(lf, off, 64);
        if (lane == 0) tile_lnl[(size_t)p * tile_stride + tile] = lf;
    }
}

// One wavefront per particle: lane l sums tiles l, l + 64, ... in order, then a
// butterfly over the lanes; the order is fixed, so the sum does not depend on timing.
__global__ __launch_bounds__(256) void k_reduce_lnl(const double *__restrict__ tile_lnl,
                                                     const int *__restrict__ ntiles, int tile_stride, int first,
                                                     int count, double *__restrict__ lnl,
                                                     uint8_t *__restrict__ nonfinite)
{
    const int i = (int)((blockIdx.x * blockDim.x + threadIdx.x) >> 6);
    const int lane = threadIdx.x & 63;
    if (i >= count) return;  // whole wavefront
    const int p = first + i;
    const double *t = tile_lnl + (size_t)p * tile_stride;
    const int nt = ntiles[p];
    double acc = 0.0;
    for (int k = lane; k < nt; k += 64) acc += t[k];
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) acc += __shfl_xor(acc, off, 64);
    if (lane == 0) {
        lnl[p] = acc;
        nonfinite[p] = isfinite(acc) ? 0 : 1;
    }
}

}  // namespace

void launch_prune(const PruneArgs &a, hipStream_t s)
{
    if (a.n_items <= 0) return;
    const int bpi = (a.tile_stride + 3) / 4;
    dim3 grid((unsigned)((size_t)a.n_items * bpi));
#define SMC_PRUNE_ARGS                                                                          \
    a.items, a.tiles, a.ntiles, a.tile_stride, a.children, a.root_of, a.perm, a.codes,         \
        (const TXa *)a.P, a.eig, a.pinv, a.isinv, a.invchar, (TXa *)a.X, a.E, a.tile_lnl, a.T, \
        a.R, a.S_pad, a.max_classes, bpi, a.pcls_off, a.p_base, a.mix_mode, a.xflag, a.WS, a.pweight
    if (a.fp32) {
        using TXa = float;
        if (a.roots)
            hipLaunchKernelGGL((k_prune<float, true>), grid, dim3(256), 0, s, SMC_PRUNE_ARGS);
        else
            hipLaunchKernelGGL((k_prune<float, false>), grid, dim3(256), 0, s, SMC_PRUNE_ARGS);
    } else {
        using TXa = double;
        if (a.roots)
            hipLaunchKernelGGL((k_prune<double, true>), grid, dim3(256), 0, s, SMC_PRUNE_ARGS);
        else
            hipLaunchKernelGGL((k_prune<double, false>), grid, dim3(256), 0, s, SMC_PRUNE_ARGS);
    }
#undef SMC_PRUNE_ARGS
}

void launch_reduce_lnl(const double *tile_lnl, const int *ntiles, int tile_stride, int first,
                       int count, double *lnl, uint8_t *nonfinite, hipStream_t s)
{
    if (count <= 0) return;
    hipLaunchKernelGGL(k_reduce_lnl, dim3((count + 3) / 4), dim3(256), 0, s, tile_lnl, ntiles, tile_stride,
                       first, count, lnl, nonfinite);
}

}  // namespace smccat
