// Felsenstein pruning on the matrix cores (v_mfma_f64_16x16x4f64 on the plain
// FP64 path, v_mfma_f32_16x16x4f32 on the FP32 fast path).
//
// Reference: Particle::refreshLikelihood (Particle.cpp:577-675); semantics as in
// k_prune.hip (binary exponent scaling per node, cumulative exponents, root
// mixture with rate weights 1 and the invariant term fused at the root).
//
// Per (particle, node, rate, child) the contraction is Y[20 x n] = P[20 x 20] X[20 x n]
// over the n sites of a tile.  One wavefront owns one tile of <= 64 class-sorted
// site slots and walks it in 16-site column blocks:
//   A operand: P rows 16*rb + (lane & 15), columns 4*kb + (lane >> 4)   (rb < 2, kb < 5)
//              rows 20..31 are zero padding, so 62.5% of the MFMA work is useful;
//   B operand: partial rows 4*kb + (lane >> 4) at site 16*sb + (lane & 15);
//   C/D:       at site 16*sb + (lane & 15), register v < 4 holds row
//              16*rb + 4*(lane >> 4) + v for f32 and 16*rb + (lane >> 4) + 4*v for f64
//              (the two builtins place result rows differently; A and B agree).
// P stays in VGPRs for all column blocks of the tile; partial rows are read and written
// as 16-lane contiguous runs of the state-major [state][site] layout.  Tip
// children are not multiplied: their P x is a column (or the row sums) of P,
// gathered straight into the C layout.
#include "internal.hpp"

namespace smccat {

namespace {

constexpr double kLn2m = 0.69314718055994530942;

template <typename TX>
struct Mfma;

template <>
struct Mfma<double> {
    using v4 = __attribute__((ext_vector_type(4))) double;
    static __device__ __forceinline__ v4 mma(double a, double b, v4 c)
    {
        return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
    }
    // result row held by lane group g4, register v, of row block rb
    static __device__ __forceinline__ int crow(int rb, int g4, int v) { return 16 * rb + g4 + 4 * v; }
};

template <>
struct Mfma<float> {
    using v4 = __attribute__((ext_vector_type(4))) float;
    static __device__ __forceinline__ v4 mma(float a, float b, v4 c)
    {
        return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
    }
    static __device__ __forceinline__ int crow(int rb, int g4, int v) { return 16 * rb + 4 * g4 + v; }
};

// Y (C layout, [rb]) of one child over one 16-site column block.
template <typename TX>
__device__ __forceinline__ void child_block(bool tip, const TX (&a)[2][5], const TX *__restrict__ Pc,
                                            const TX *__restrict__ xcol, int S_pad, int col,
                                            bool valid, int g4, typename Mfma<TX>::v4 (&y)[2])
{
    using V4 = typename Mfma<TX>::v4;
    if (tip) {
#pragma unroll
        for (int rb = 0; rb < 2; rb++)
#pragma unroll
            for (int v = 0; v < 4; v++) {
                const int row = Mfma<TX>::crow(rb, g4, v);
                y[rb][v] = (row < 20) ? Pc[row * 21 + col] : (TX)0;
            }
    } else {
        TX b[5];
#pragma unroll
        for (int kb = 0; kb < 5; kb++) b[kb] = valid ? xcol[(size_t)(4 * kb + g4) * S_pad] : (TX)0;
#pragma unroll
        for (int rb = 0; rb < 2; rb++) {
            V4 acc = {0, 0, 0, 0};
#pragma unroll
            for (int kb = 0; kb < 5; kb++) acc = Mfma<TX>::mma(a[rb][kb], b[kb], acc);
            y[rb] = acc;
        }
    }
}

template <typename TX, bool IS_ROOT>
__global__ __launch_bounds__(256) void k_prune_mfma(
    const Item *__restrict__ items, const Tile *__restrict__ tiles, const int *__restrict__ ntiles,
    int tile_stride, const int *__restrict__ children, const int *__restrict__ root_of,
    const int *__restrict__ perm, const uint8_t *__restrict__ codes, const TX *__restrict__ Pm,
    const double *__restrict__ eig, const double *__restrict__ pinv,
    const double *__restrict__ isinv, const int *__restrict__ invchar, TX *__restrict__ X,
    int *__restrict__ E, double *__restrict__ tile_lnl, int T, int R, int S_pad, int max_classes,
    int blocks_per_item)
{
    using V4 = typename Mfma<TX>::v4;
    const int item = blockIdx.x / blocks_per_item;
    const int wave = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    const int tile = (blockIdx.x % blocks_per_item) * 4 + wave;
    const Item it = items[item];
    const int p = it.particle, n = it.node;
    if (tile >= ntiles[p]) return;  // whole wavefront exits; no block barrier follows
    const Tile tl = tiles[(size_t)p * tile_stride + tile];
    const int lane = threadIdx.x & 63;
    const int c16 = lane & 15, g4 = lane >> 4;
    const size_t g = (size_t)p * max_classes + tl.cls;
    const int nn = 2 * T - 1, nl = n - T;
    const int c0 = children[((size_t)p * (T - 1) + nl) * 2];
    const int c1 = children[((size_t)p * (T - 1) + nl) * 2 + 1];
    const bool tip0 = c0 < T, tip1 = c1 < T;
    constexpr bool is_root = IS_ROOT;
    (void)root_of;
    const size_t node_elems = (size_t)R * 20 * S_pad;
    TX *Xp = X + (size_t)p * (T - 1) * node_elems;
    int *Ep = E + (size_t)p * (T - 1) * R * S_pad;
    const TX *Pg = Pm + g * nn * R * kPStride;
    const double *pi = eig + g * kEigStride + kEigPi;
    const int nsb = (tl.len + 15) >> 4;  // 16-site column blocks in this tile (wave-uniform)

    // running log-sum-exp over rates of the root's 4 column blocks
    double lse_m[4], lse_s[4];
#pragma unroll
    for (int sb = 0; sb < 4; sb++) {
        lse_m[sb] = -INFINITY;
        lse_s[sb] = 0.0;
    }

    for (int r = 0; r < R; r++) {
        const TX *P0 = Pg + ((size_t)c0 * R + r) * kPStride;
        const TX *P1 = Pg + ((size_t)c1 * R + r) * kPStride;
        TX a0[2][5], a1[2][5];
#pragma unroll
        for (int rb = 0; rb < 2; rb++)
#pragma unroll
            for (int kb = 0; kb < 5; kb++) {
                const int row = 16 * rb + c16;
                a0[rb][kb] = (!tip0 && row < 20) ? P0[row * 21 + 4 * kb + g4] : (TX)0;
                a1[rb][kb] = (!tip1 && row < 20) ? P1[row * 21 + 4 * kb + g4] : (TX)0;
            }
        const TX *X0 = tip0 ? Xp : Xp + ((size_t)(c0 - T) * R + r) * 20 * S_pad;
        const TX *X1 = tip1 ? Xp : Xp + ((size_t)(c1 - T) * R + r) * 20 * S_pad;
        const int *E0 = Ep + ((size_t)(tip0 ? 0 : c0 - T) * R + r) * S_pad;
        const int *E1 = Ep + ((size_t)(tip1 ? 0 : c1 - T) * R + r) * S_pad;

#pragma unroll 1
        for (int sb = 0; sb < nsb; sb++) {
            const int o = 16 * sb + c16;
            const bool valid = o < tl.len;
            const int slot = tl.start + (valid ? o : 0);
            const int orig = perm[(size_t)p * S_pad + slot];
            int col0 = 20, col1 = 20;
            if (tip0) {
                const int code = codes[(size_t)c0 * S_pad + orig];
                col0 = code < 20 ? code : 20;
            }
            if (tip1) {
                const int code = codes[(size_t)c1 * S_pad + orig];
                col1 = code < 20 ? code : 20;
            }
            V4 y[2], w[2];
            child_block<TX>(tip0, a0, P0, X0 + slot, S_pad, col0, valid, g4, y);
            child_block<TX>(tip1, a1, P1, X1 + slot, S_pad, col1, valid, g4, w);
            const int e = (tip0 ? 0 : E0[slot]) + (tip1 ? 0 : E1[slot]);
            // z = y * w and the per-site max over the 20 real rows (4 lane groups x 8 values)
            TX m = 0;
#pragma unroll
            for (int rb = 0; rb < 2; rb++)
#pragma unroll
                for (int v = 0; v < 4; v++) {
                    y[rb][v] *= w[rb][v];
                    m = fmax(m, y[rb][v]);
                }
            m = fmax(m, __shfl_xor(m, 16, 64));
            m = fmax(m, __shfl_xor(m, 32, 64));
            int ex;
            (void)frexp(m, &ex);
            const bool ok = (m > (TX)0) && isfinite(m);
            const TX nanv = (TX)NAN;
#pragma unroll
            for (int rb = 0; rb < 2; rb++)
#pragma unroll
                for (int v = 0; v < 4; v++) y[rb][v] = ok ? (TX)ldexp(y[rb][v], -ex) : nanv;
            const int etot = e + (ok ? ex : 0);
            if constexpr (!is_root) {
                if (valid) {
                    TX *xo = Xp + ((size_t)nl * R + r) * 20 * S_pad + slot;
#pragma unroll
                    for (int rb = 0; rb < 2; rb++)
#pragma unroll
                        for (int v = 0; v < 4; v++) {
                            const int row = Mfma<TX>::crow(rb, g4, v);
                            if (row < 20) xo[(size_t)row * S_pad] = y[rb][v];
                        }
                    if (g4 == 0) Ep[((size_t)nl * R + r) * S_pad + slot] = etot;
                }
            } else {
                double dot = 0.0;
#pragma unroll
                for (int rb = 0; rb < 2; rb++)
#pragma unroll
                    for (int v = 0; v < 4; v++) {
                        const int row = Mfma<TX>::crow(rb, g4, v);
                        if (row < 20) dot = fma(pi[row], (double)y[rb][v], dot);
                    }
                dot += __shfl_xor(dot, 16, 64);
                dot += __shfl_xor(dot, 32, 64);
                const double l = log(dot) + (double)etot * kLn2m;
                // sb is wave-uniform and < 4: select the accumulator without dynamic indexing
#pragma unroll
                for (int k = 0; k < 4; k++)
                    if (k == sb) {
                        if (l > lse_m[k]) {
                            lse_s[k] = lse_s[k] * exp(lse_m[k] - l) + 1.0;
                            lse_m[k] = l;
                        } else {
                            lse_s[k] += exp(l - lse_m[k]);
                        }
                    }
            }
        }
    }
    if constexpr (is_root) {
        double lf = 0.0;
        const double pv = pinv[p];
        const double *pic = pi;
#pragma unroll
        for (int sb = 0; sb < 4; sb++) {
            const int o = 16 * sb + c16;
            if (g4 == 0 && o < tl.len) {
                const int orig = perm[(size_t)p * S_pad + tl.start + o];
                const double lsite = lse_m[sb] + log(lse_s[sb]);
                const double fin = exp(lsite) * (1.0 - pv) + pv * pic[invchar[orig]] * isinv[orig];
                lf += log(fin);
            }
        }
#pragma unroll
        for (int off = 32; off >= 1; off >>= 1) lf += __shfl_xor(lf, off, 64);
        if (lane == 0) tile_lnl[(size_t)p * tile_stride + tile] = lf;
    }
}

}  // namespace

void launch_prune_mfma(const PruneArgs &a, hipStream_t s)
{
    if (a.n_items <= 0) return;
    const int bpi = (a.tile_stride + 3) / 4;
    dim3 grid((unsigned)((size_t)a.n_items * bpi));
#define SMC_PRUNE_ARGS                                                                          \
    a.items, a.tiles, a.ntiles, a.tile_stride, a.children, a.root_of, a.perm, a.codes,         \
        (const TXa *)a.P, a.eig, a.pinv, a.isinv, a.invchar, (TXa *)a.X, a.E, a.tile_lnl, a.T, \
        a.R, a.S_pad, a.max_classes, bpi
    if (a.fp32) {
        using TXa = float;
        if (a.roots)
            hipLaunchKernelGGL((k_prune_mfma<float, true>), grid, dim3(256), 0, s, SMC_PRUNE_ARGS);
        else
            hipLaunchKernelGGL((k_prune_mfma<float, false>), grid, dim3(256), 0, s, SMC_PRUNE_ARGS);
    } else {
        using TXa = double;
        if (a.roots)
            hipLaunchKernelGGL((k_prune_mfma<double, true>), grid, dim3(256), 0, s, SMC_PRUNE_ARGS);
        else
            hipLaunchKernelGGL((k_prune_mfma<double, false>), grid, dim3(256), 0, s, SMC_PRUNE_ARGS);
    }
#undef SMC_PRUNE_ARGS
}

}  // namespace smccat
