// Felsenstein pruning on the matrix cores (v_mfma_f64_16x16x4f64 on the plain
// FP64 path, v_mfma_f32_16x16x4f32 on the FP32 fast path).
//
// Reference: Particle::refreshLikelihood (Particle.cpp:577-675); semantics as in
// k_prune.hip (binary exponent scaling per node, cumulative exponents, root
// mixture with rate weights 1 and the invariant term fused at the root).
//
// Per (particle, node, rate, child) the contraction is Y[20 x n] = P[20 x 20] X[20 x n]
// over the n sites of a tile.  One wavefront owns one tile -- the slots of one
// class inside a 64-slot aligned window of the class-sorted slot order, so that
// every column block starts on a cache-line boundary -- and walks it in 16-site
// column blocks:
//   A operand: P row Mfma::arow(rb, lane & 15), columns 4*kb + (lane >> 4)   (rb < 2, kb < 5)
//              rows 20..31 are zero padding, so 62.5% of the MFMA work is useful;
//   B operand: partial rows 4*kb + (lane >> 4) at site 16*sb + (lane & 15);
//   C/D:       at site 16*sb + (lane & 15), register v < 4 holds row
//              16*rb + (lane >> 4) + 4*v (for f32 through a row permutation of A: the
//              two builtins place result rows differently; see Mfma<float>), so the
//              rb = 0 registers and register 0 of rb = 1 are the B operand of a parent.
// (With unroll >= 2 the sites of a lane are paired instead; see the column mapping in
// prune_unit.)  P stays in VGPRs for all column blocks of the tile; partial rows are
// read and written as 16-lane contiguous runs of the state-major [state][site] layout.
// Tip children are not multiplied: their P x is a column (or the row sums) of P,
// gathered straight into the C layout.  Each step issues every load of both children
// before its MFMAs, and a tile's tip codes are gathered once for all rates.
//
// Cherries (internal nodes whose children are both tips, a third of the nodes
// of a random tree) are never stored: the parent builds the cherry's partial
// -- the product of two tip columns of P -- straight into its B operand, which
// saves the cherry's write and the parent's read (about a third of a sweep's
// HBM bytes).  Each child is therefore a tip, a cherry or a stored partial; items are
// launched per (child mode, child mode) pair, roots decide at run time.
#include "internal.hpp"
#include "root_mixture.hpp"

#include <algorithm>
#include <cstdint>
#include <type_traits>

namespace smccat {

namespace {


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
    // P row that lane c16 supplies as A-operand row c16 of row block rb
    static __device__ __forceinline__ int arow(int rb, int c16) { return 16 * rb + c16; }
};

template <>
struct Mfma<float> {
    using v4 = __attribute__((ext_vector_type(4))) float;
    static __device__ __forceinline__ v4 mma(float a, float b, v4 c)
    {
        return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
    }
    // The f32 builtin leaves result row i of a block in lane group i / 4, register
    // i % 4.  Its A operand is row-permuted (lane c16 of block rb supplies P row
    // 16 rb + c16 / 4 + 4 (c16 % 4)) so that results land as they do for f64: lane
    // group g4, register v holds row 16 rb + g4 + 4 v, which for rb = 0 and v = kb
    // (and rb = 1, v = 0 for kb = 4) is B-operand row 4 kb + g4 of the same lane.
    // A node's result is then the B operand of its parent without any lane
    // exchange (the chain kernel carries it in registers).  Each result element
    // is the same dot product in the same order, so only its position changes.
    static __device__ __forceinline__ int crow(int rb, int g4, int v) { return 16 * rb + g4 + 4 * v; }
    static __device__ __forceinline__ int arow(int rb, int c16) { return 16 * rb + (c16 >> 2) + 4 * (c16 & 3); }
};

// Child modes; MODE_RT decides per item at run time (roots).
constexpr int MODE_TIP = 0, MODE_CHERRY = 1, MODE_STORED = 2, MODE_RT = 3;

// Y (C layout) of a tip child: column `col` of P (row sums for col 20).
template <typename TX>
__device__ __forceinline__ void tip_block(const TX *__restrict__ Pc, int col, int g4,
                                          typename Mfma<TX>::v4 (&y)[2])
{
#pragma unroll
    for (int rb = 0; rb < 2; rb++)
#pragma unroll
        for (int v = 0; v < 4; v++) {
            const int row = Mfma<TX>::crow(rb, g4, v);
            y[rb][v] = (row < 20) ? Pc[row * 21 + col] : (TX)0;
        }
}

// B operand (states 4 kb + g4 of this lane's site) of a cherry: the product of
// its tips' P columns.  It is not rescaled: the products of two probabilities
// stay in range, and the parent's own rescale by a power of two makes the
// result exactly the one a rescaled, stored cherry gives (the exponents only
// split differently).
template <typename TX>
__device__ __forceinline__ int cherry_b(const TX *__restrict__ Pa, const TX *__restrict__ Pb, int ca, int cb,
                                        bool valid, int g4, TX (&b)[5])
{
#pragma unroll
    for (int kb = 0; kb < 5; kb++) {
        const int row = 4 * kb + g4;
        const TX x = Pa[row * 21 + ca] * Pb[row * 21 + cb];
        b[kb] = valid ? x : (TX)0;
    }
    return 0;
}

template <typename TX>
struct Vec2;
template <>
struct Vec2<float> {
    using t = __attribute__((ext_vector_type(2))) float;
};
template <>
struct Vec2<double> {
    using t = __attribute__((ext_vector_type(2))) double;
};
using int2v = __attribute__((ext_vector_type(2))) int;

// B operands of two adjacent slots (xpair: the even one) with one vector load per
// row; a slot outside its tile contributes zeros
template <typename TX>
__device__ __forceinline__ void stored_b2(const TX *__restrict__ xpair, int S_pad, bool v0, bool v1, int g4,
                                          TX (&b0)[5], TX (&b1)[5])
{
    using T2 = typename Vec2<TX>::t;
#pragma unroll
    for (int kb = 0; kb < 5; kb++) {
        const T2 x = *(const T2 *)(xpair + (size_t)(4 * kb + g4) * S_pad);
        b0[kb] = v0 ? x.x : (TX)0;
        b1[kb] = v1 ? x.y : (TX)0;
    }
}

// (xcol must be a readable slot even when `valid` is false: the load is unconditional)
template <typename TX>
__device__ __forceinline__ void stored_b(const TX *__restrict__ xcol, int S_pad, bool valid, int g4, TX (&b)[5])
{
#pragma unroll
    for (int kb = 0; kb < 5; kb++) {
        const TX x = xcol[(size_t)(4 * kb + g4) * S_pad];
        b[kb] = valid ? x : (TX)0;
    }
}

template <typename TX>
__device__ __forceinline__ void mma_block(const TX (&a)[2][5], const TX (&b)[5], typename Mfma<TX>::v4 (&y)[2])
{
    using V4 = typename Mfma<TX>::v4;
#pragma unroll
    for (int rb = 0; rb < 2; rb++) {
        V4 acc = {0, 0, 0, 0};
#pragma unroll
        for (int kb = 0; kb < 5; kb++) acc = Mfma<TX>::mma(a[rb][kb], b[kb], acc);
        y[rb] = acc;
    }
}

// One child of a run-time-mode item over one column block: its loads, then its
// products into Y (C layout); returns the exponent the child adds.
template <typename TX>
__device__ __forceinline__ int child_block_rt(int mode, const TX (&a)[2][5], const TX *__restrict__ Pc,
                                              const TX *__restrict__ Pa, const TX *__restrict__ Pb,
                                              const TX *__restrict__ xcol, const int *__restrict__ ecol,
                                              int S_pad, int col, int ca, int cb, bool valid, int g4,
                                              typename Mfma<TX>::v4 (&y)[2])
{
    if (mode == MODE_TIP) {
        tip_block<TX>(Pc, col, g4, y);
        return 0;
    }
    TX b[5];
    int e = 0;
    if (mode == MODE_CHERRY) {
        e = cherry_b<TX>(Pa, Pb, ca, cb, valid, g4, b);
    } else {
        stored_b<TX>(xcol, S_pad, valid, g4, b);
        e = *ecol;
    }
    mma_block<TX>(a, b, y);
    return e;
}

// Kernel arguments (one struct, passed by value).
template <typename TX>
struct KArgs {
    const Item *items;
    const Tile *tiles;
    const int *ntiles;
    int tile_stride;
    const int *children, *perm;
    const uint8_t *codes;
    const uint8_t *pcodes;  // [P][T][S_pad] tip codes in the particle's slot order, clamped to 20
    const double *pweight;  // [P][S_pad] columns each slot stands for
    const TX *Pm;
    const double *eig, *pinv, *isinv;
    const int *invchar;
    TX *X;
    int *E;
    double *tile_lnl;
    int T, R, S_pad, max_classes, blocks_per_item;
    const int *pcls_off;
    int p_base, mix_mode;
    const uint8_t *xflag;
    int WS;
    int xcds;  // workgroup i runs on XCD i mod xcds: the units of an item share one XCD's L2
};

// One unit of work: tile group `tile_group` (4 tiles, one per wavefront) of
// item `item`.  U is the number of 16-site column blocks processed together
// (independent MFMA chains and loads in flight); a tile has at most 4 blocks.
// LDS per wavefront for the tips' P matrices of cherry children (2 per cherry).  (A
// tip child's own P gathered from LDS as well was measured: no gain in FP32, slower in
// FP64.)
__host__ __device__ constexpr int lds_mats(int mode) { return mode == MODE_CHERRY || mode == MODE_RT ? 2 : 0; }
template <int M0, int M1>
struct CherryLds {
    static constexpr int mats = lds_mats(M0) + lds_mats(M1);
    static constexpr int elems = mats > 0 ? mats * kPStride : 1;
};

// Copy of one P matrix (kPStride values, 16-byte aligned at both ends) into LDS.
template <typename TX>
__device__ __forceinline__ void stage_p(TX *__restrict__ dst, const TX *__restrict__ src, int lane)
{
    using Q = __attribute__((ext_vector_type(4))) unsigned int;
    constexpr int n = kPStride * (int)sizeof(TX) / 16;
    static_assert(kPStride * sizeof(TX) % 16 == 0, "P matrices must be whole 16-byte blocks");
    for (int i = lane; i < n; i += 64) ((Q *)dst)[i] = ((const Q *)src)[i];
}

__device__ __forceinline__ void wave_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

template <typename TX, int M0, int M1, bool IS_ROOT, int U, int W>
__device__ __forceinline__ void prune_unit(const KArgs<TX> &ka, int item, int tile_group,
                                           TX (&lds)[W][CherryLds<M0, M1>::elems])
{
    using V4 = typename Mfma<TX>::v4;
    const Item *__restrict__ items = ka.items;
    const Tile *__restrict__ tiles = ka.tiles;
    const int *__restrict__ ntiles = ka.ntiles;
    const int tile_stride = ka.tile_stride;
    const int *__restrict__ children = ka.children;
    const int *__restrict__ perm = ka.perm;
    const TX *__restrict__ Pm = ka.Pm;
    const double *__restrict__ eig = ka.eig;
    const double *__restrict__ pinv = ka.pinv;
    const double *__restrict__ isinv = ka.isinv;
    const int *__restrict__ invchar = ka.invchar;
    TX *__restrict__ X = ka.X;
    int *__restrict__ E = ka.E;
    double *__restrict__ tile_lnl = ka.tile_lnl;
    const int T = ka.T, R = ka.R, S_pad = ka.S_pad, max_classes = ka.max_classes;
    const int *__restrict__ pcls_off = ka.pcls_off;
    const int p_base = ka.p_base, mix_mode = ka.mix_mode, WS = ka.WS;
    const uint8_t *__restrict__ xflag = ka.xflag;
    // W wavefronts per workgroup, one tile each
    const int wave = W == 1 ? 0 : __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    const int tile = tile_group * W + wave;
    const Item it = items[item];
    const int p = it.particle, n = it.node;
    if (tile >= ntiles[p]) return;  // this wavefront has no tile here; no block barrier follows
    const Tile tl = tiles[(size_t)p * tile_stride + tile];
    const int lane = threadIdx.x & 63;
    const int c16 = lane & 15, g4 = lane >> 4;
    const size_t g = (size_t)p * max_classes + tl.cls;
    const int nn = 2 * T - 1, nl = n - T;
    const int *__restrict__ chp = children + (size_t)p * (T - 1) * 2;
    int c0 = chp[nl * 2], c1 = chp[nl * 2 + 1];
    auto mode_of = [&](int c) {
        if (c < T) return MODE_TIP;
        return (chp[(c - T) * 2] < T && chp[(c - T) * 2 + 1] < T) ? MODE_CHERRY : MODE_STORED;
    };
    int m0 = mode_of(c0), m1 = mode_of(c1);
    if (m0 > m1) {  // canonical order: the launch's (M0, M1) pair is sorted
        const int tc = c0, tm = m0;
        c0 = c1;
        m0 = m1;
        c1 = tc;
        m1 = tm;
    }
    const int md0 = M0 == MODE_RT ? m0 : M0, md1 = M1 == MODE_RT ? m1 : M1;
    // tips of cherry children
    const int ga0 = md0 == MODE_CHERRY ? chp[(c0 - T) * 2] : 0;
    const int gb0 = md0 == MODE_CHERRY ? chp[(c0 - T) * 2 + 1] : 0;
    const int ga1 = md1 == MODE_CHERRY ? chp[(c1 - T) * 2] : 0;
    const int gb1 = md1 == MODE_CHERRY ? chp[(c1 - T) * 2 + 1] : 0;
    const size_t node_elems = (size_t)R * 20 * S_pad;
    const size_t lp = (size_t)(p - p_base);  // slot of the particle in the wave's partials
    // partial slots of a particle: WS workspaces x (T-1) internal nodes; a node's
    // current workspace is its flag (accept/reject double buffering)
    TX *Xp = X + lp * WS * (T - 1) * node_elems;
    int *Ep = E + lp * WS * (T - 1) * R * S_pad;
    const uint8_t *xf = xflag ? xflag + (size_t)p * (T - 1) : nullptr;
    auto nslot = [&](int node) { return (xf ? (int)xf[node - T] * (T - 1) : 0) + node - T; };
    const TX *Pg = Pm + ((size_t)pcls_off[p] + tl.cls) * nn * R * kPStride;
    const double *pi = eig + g * kEigStride + kEigPi;
    // Column mapping.  Unpaired (U = 1): lane c16 of block sb holds window offset
    // 16 sb + c16.  Paired (U >= 2): blocks 2j and 2j + 1 of a step hold the even and
    // odd offsets of one 32-slot run, so a lane owns two adjacent slots and reads and
    // writes them with one 8-byte (FP32) or 16-byte (FP64) access per row: 16 lanes
    // then cover a whole 128-byte line (FP32) where single slots cover half of one.
    constexpr bool PAIRED = U >= 2;
    using T2 = typename Vec2<TX>::t;
    auto site_of = [](int sb0, int u, int c) {
        return PAIRED ? 16 * sb0 + 32 * (u >> 1) + 2 * c + (u & 1) : 16 * (sb0 + u) + c;
    };
    const int nsb = (tl.len + 15) >> 4;               // 16-slot column blocks of the window in use
    const int sb_begin = ((tl.lo >> 4) / U) * U;      // the first step holding a valid slot

    // Root mixture: after the column blocks of a rate, lane group g4 holds the
    // column (site c16) of block g4, so each (site, rate) log and log-sum-exp step is
    // evaluated once per wavefront rather than once per lane group (the FP64 log
    // and exp dominate the root's instruction count otherwise).  The running
    // log-sum-exp over rates of that column:
    double lse_m = -INFINITY, lse_s = 0.0;

    // workspaces of the stored children and of this node (read once: the flags do
    // not change during a launch)
    const int cs0 = md0 == MODE_STORED ? nslot(c0) : 0, cs1 = md1 == MODE_STORED ? nslot(c1) : 0;
    const int ns = IS_ROOT ? 0 : nslot(n);
    // Tip codes of the tile's (at most 4) column positions, gathered once for all
    // rates: byte 0 / 1 the tip of child 0 or its cherry's two tips, bytes 2 / 3 the
    // same for child 1 (20 stands for a gap or unknown state).  Position k is the
    // one a step's block u takes when k = sb0 + u.
    // (read in the particle's slot order: contiguous across the lanes of a block)
    const uint8_t *__restrict__ pcp = ka.pcodes + (size_t)p * T * S_pad;
    auto tip_codes = [&](int k) -> uint32_t {
        if (md0 == MODE_STORED && md1 == MODE_STORED) return 0;
        const int o = site_of(k / U * U, k % U, c16);
        const int sl = tl.start + ((o >= tl.lo && o < tl.len) ? o : tl.lo);
        auto cc = [&](int tip) { return (uint32_t)pcp[(size_t)tip * S_pad + sl]; };
        uint32_t wd = 0;
        if (md0 == MODE_TIP) wd |= cc(c0);
        if (md0 == MODE_CHERRY) wd |= cc(ga0) | cc(gb0) << 8;
        if (md1 == MODE_TIP) wd |= cc(c1) << 16;
        if (md1 == MODE_CHERRY) wd |= cc(ga1) << 16 | cc(gb1) << 24;
        return wd;
    };
    // Four scalars, not an array (a run-time index into an array sends it to scratch).
    // Each rate's column loop walks a working copy that is shifted by U positions
    // per step, so step sb0's block u always reads working register u: indices known
    // at compile time, no per-step selects or branches.
    const uint32_t tc0 = tip_codes(0), tc1 = tip_codes(1), tc2 = tip_codes(2), tc3 = tip_codes(3);
    uint32_t w0, w1, w2, w3;
    auto code_word = [&](int u) { return u == 0 ? w0 : u == 1 ? w1 : u == 2 ? w2 : w3; };

    for (int r = 0; r < R; r++) {
        const TX *P0 = Pg + ((size_t)c0 * R + r) * kPStride;
        const TX *P1 = Pg + ((size_t)c1 * R + r) * kPStride;
        // the tips' P matrices of cherry children, staged in this wavefront's LDS by
        // 16-byte loads (the gathers by tip code then stay out of the vector L1)
        TX *lw = lds[wave];
        TX *l0 = lw, *l1 = lw + lds_mats(md0) * kPStride;
        const TX *Pa0 = l0, *Pb0 = l0 + kPStride, *Pa1 = l1, *Pb1 = l1 + kPStride;

        if constexpr (CherryLds<M0, M1>::mats > 0) {
            if (lds_mats(md0) + lds_mats(md1) > 0) {
                wave_sync();  // the previous rate's reads are done
                if (md0 == MODE_CHERRY) {
                    stage_p<TX>(l0, Pg + ((size_t)ga0 * R + r) * kPStride, lane);
                    stage_p<TX>(l0 + kPStride, Pg + ((size_t)gb0 * R + r) * kPStride, lane);
                }
                if (md1 == MODE_CHERRY) {
                    stage_p<TX>(l1, Pg + ((size_t)ga1 * R + r) * kPStride, lane);
                    stage_p<TX>(l1 + kPStride, Pg + ((size_t)gb1 * R + r) * kPStride, lane);
                }
                wave_sync();
            }
        }
        TX a0[2][5], a1[2][5];
#pragma unroll
        for (int rb = 0; rb < 2; rb++)
#pragma unroll
            for (int kb = 0; kb < 5; kb++) {
                const int row = Mfma<TX>::arow(rb, c16);
                a0[rb][kb] = (md0 != MODE_TIP && row < 20) ? P0[row * 21 + 4 * kb + g4] : (TX)0;
                a1[rb][kb] = (md1 != MODE_TIP && row < 20) ? P1[row * 21 + 4 * kb + g4] : (TX)0;
            }
        const TX *X0 = Xp + ((size_t)cs0 * R + r) * 20 * S_pad;
        const TX *X1 = Xp + ((size_t)cs1 * R + r) * 20 * S_pad;
        const int *E0 = Ep + ((size_t)cs0 * R + r) * S_pad;
        const int *E1 = Ep + ((size_t)cs1 * R + r) * S_pad;

        w0 = tc0;
        w1 = tc1;
        w2 = tc2;
        w3 = tc3;
        for (int q = 0; q < sb_begin; q++) {  // the first step's positions (wave-uniform)
            w0 = w1;
            w1 = w2;
            w2 = w3;
        }
        double rdot = 0.0;  // root: pi . x and exponent of this lane group's column
        int rexp = 0;
#pragma unroll 1
        for (int sb0 = sb_begin; sb0 < nsb; sb0 += U) {
            V4 y[U][2], w[U][2];
            int slot[U], etot[U];
            bool valid[U];
#pragma unroll
            for (int u = 0; u < U; u++) {
                const int o = site_of(sb0, u, c16);
                valid[u] = o >= tl.lo && o < tl.len;
                slot[u] = tl.start + (valid[u] ? o : tl.lo);  // a lane outside the tile reads a slot inside it
            }
            // Phase 1: every load of the step for both children -- stored partial rows
            // and exponents, tip codes and their P columns, cherry products -- issued
            // before any MFMA, so the two children's memory latencies overlap.  The
            // loads are unconditional (every lane's slot is inside the tile's window)
            // and lanes outside the tile are zeroed afterwards.
            auto fetch = [&](auto mode_c, int md, const TX *Pc, const TX *Pa, const TX *Pb, const TX *Xc,
                             const int *Ec, int shift, TX(&bv)[U][5], int(&ex)[U], V4(&acc)[U][2]) {
                constexpr int MC = decltype(mode_c)::value;
                if (MC == MODE_STORED || (MC == MODE_RT && md == MODE_STORED)) {
                    if constexpr (PAIRED) {
#pragma unroll
                        for (int j = 0; j < U / 2; j++) {
                            const int pb = tl.start + site_of(sb0, 2 * j, c16);  // even, in the window
                            stored_b2<TX>(Xc + pb, S_pad, valid[2 * j], valid[2 * j + 1], g4, bv[2 * j],
                                          bv[2 * j + 1]);
                            const int2v ev = *(const int2v *)(Ec + pb);
                            ex[2 * j] = ev.x;
                            ex[2 * j + 1] = ev.y;
                        }
                    } else {
#pragma unroll
                        for (int u = 0; u < U; u++) {
                            stored_b<TX>(Xc + slot[u], S_pad, valid[u], g4, bv[u]);
                            ex[u] = Ec[slot[u]];
                        }
                    }
                    return;
                }
#pragma unroll
                for (int u = 0; u < U; u++) {
                    const uint32_t wd = code_word(u) >> shift;
                    const int ca = (int)(wd & 0xff), cb = (int)((wd >> 8) & 0xff);
                    ex[u] = 0;
                    if (MC == MODE_TIP || (MC == MODE_RT && md == MODE_TIP)) {
                        tip_block<TX>(Pc, ca, g4, acc[u]);
                    } else {
                        cherry_b<TX>(Pa, Pb, ca, cb, valid[u], g4, bv[u]);
                    }
                }
            };
            int e1[U];
            if constexpr (M0 == MODE_RT) {
                // child modes decided per item: per child, its loads then its products,
                // which keeps this general kernel's register count (and occupancy)
                auto child_rt = [&](int md, const TX(&a)[2][5], const TX *Pc, const TX *Pa, const TX *Pb,
                                    const TX *Xc, const int *Ec, int shift, V4(&acc)[U][2], int(&ex)[U]) {
                    if constexpr (PAIRED) {
                        if (md == MODE_STORED) {
#pragma unroll
                            for (int j = 0; j < U / 2; j++) {
                                const int pb = tl.start + site_of(sb0, 2 * j, c16);
                                TX b0[5], b1[5];
                                stored_b2<TX>(Xc + pb, S_pad, valid[2 * j], valid[2 * j + 1], g4, b0, b1);
                                const int2v ev = *(const int2v *)(Ec + pb);
                                mma_block<TX>(a, b0, acc[2 * j]);
                                mma_block<TX>(a, b1, acc[2 * j + 1]);
                                ex[2 * j] = ev.x;
                                ex[2 * j + 1] = ev.y;
                            }
                            return;
                        }
                    }
#pragma unroll
                    for (int u = 0; u < U; u++) {
                        const uint32_t wd = code_word(u) >> shift;
                        const int ca = (int)(wd & 0xff), cb = (int)((wd >> 8) & 0xff);
                        ex[u] = child_block_rt<TX>(md, a, Pc, Pa, Pb, Xc + slot[u], Ec + slot[u], S_pad, ca, ca,
                                                   cb, valid[u], g4, acc[u]);
                    }
                };
                child_rt(md0, a0, P0, Pa0, Pb0, X0, E0, 0, y, etot);
                child_rt(md1, a1, P1, Pa1, Pb1, X1, E1, 16, w, e1);
            } else {
                TX bv0[U][5], bv1[U][5];
                fetch(std::integral_constant<int, M0>{}, md0, P0, Pa0, Pb0, X0, E0, 0, bv0, etot, y);
                fetch(std::integral_constant<int, M1>{}, md1, P1, Pa1, Pb1, X1, E1, 16, bv1, e1, w);
                // Phase 2: the products of the children that are not tips.  The
                // scheduling groups ask the compiler to issue the step's vector loads
                // before its MFMAs (it otherwise interleaves them to shorten register
                // live ranges, which serialises the two children's memory latencies);
                // only in kinds with a stored child: elsewhere the longer live ranges
                // cost more occupancy than the overlap gains.
                // vector-memory loads of a step: per stored child 5 rows + the exponent
                // per block (half as many instructions when paired), per tip child 8
                // gathers per block; cherry products gather from LDS
                constexpr int n_st = (M0 == MODE_STORED) + (M1 == MODE_STORED);
                constexpr int n_tip = (M0 == MODE_TIP) + (M1 == MODE_TIP);
                constexpr int n_ch = (M0 == MODE_CHERRY) + (M1 == MODE_CHERRY);
                constexpr int n_vmem = n_st * (PAIRED ? 3 * U : 6 * U) + n_tip * 8 * U;
                if constexpr (n_st > 0) {
                    __builtin_amdgcn_sched_group_barrier(0x020, n_vmem, 0);
                    __builtin_amdgcn_sched_group_barrier(0x008, 20 * U, 0);
                } else if constexpr (n_ch > 0) {
                    // cherry products: their LDS gathers first, then the MFMAs
                    __builtin_amdgcn_sched_group_barrier(0x100, 10 * U * n_ch, 0);
                    __builtin_amdgcn_sched_group_barrier(0x008, 10 * U * n_ch, 0);
                }
                if (M0 != MODE_TIP && md0 != MODE_TIP) {
#pragma unroll
                    for (int u = 0; u < U; u++) mma_block<TX>(a0, bv0[u], y[u]);
                }
                if (M1 != MODE_TIP && md1 != MODE_TIP) {
#pragma unroll
                    for (int u = 0; u < U; u++) mma_block<TX>(a1, bv1[u], w[u]);
                }
            }
#pragma unroll
            for (int q = 0; q < U; q++) {  // the next step's positions
                w0 = w1;
                w1 = w2;
                w2 = w3;
            }
#pragma unroll
            for (int u = 0; u < U; u++) {
                etot[u] += e1[u];
                // z = y * w and the per-site max over the 20 real rows (4 lane groups x 8 values)
                TX m = 0;
#pragma unroll
                for (int rb = 0; rb < 2; rb++)
#pragma unroll
                    for (int v = 0; v < 4; v++) {
                        y[u][rb][v] *= w[u][rb][v];
                        m = fmax(m, y[u][rb][v]);
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
                    for (int v = 0; v < 4; v++)
                        y[u][rb][v] = ok ? (TX)ldexp(y[u][rb][v], -ex) : nanv;
                etot[u] += ok ? ex : 0;
                if constexpr (IS_ROOT) {
                    double dot = 0.0;
#pragma unroll
                    for (int rb = 0; rb < 2; rb++)
#pragma unroll
                        for (int v = 0; v < 4; v++) {
                            const int row = Mfma<TX>::crow(rb, g4, v);
                            if (row < 20) dot = fma(pi[row], (double)y[u][rb][v], dot);
                        }
                    dot += __shfl_xor(dot, 16, 64);
                    dot += __shfl_xor(dot, 32, 64);
                    // every lane group now holds the column's pi . x; group g4 keeps
                    // block g4's (the block index sb0 + u is wave-uniform and < 4)
                    if (g4 == sb0 + u) {
                        rdot = dot;
                        rexp = etot[u];
                    }
                }
            }
            if constexpr (!IS_ROOT) {
                TX *xo = Xp + ((size_t)ns * R + r) * 20 * S_pad;
                int *eo = Ep + ((size_t)ns * R + r) * S_pad;
                if constexpr (PAIRED) {
#pragma unroll
                    for (int j = 0; j < U / 2; j++) {
                        const int pb = tl.start + site_of(sb0, 2 * j, c16);
                        const bool v0 = valid[2 * j], v1 = valid[2 * j + 1];
#pragma unroll
                        for (int rb = 0; rb < 2; rb++)
#pragma unroll
                            for (int v = 0; v < 4; v++) {
                                const int row = Mfma<TX>::crow(rb, g4, v);
                                if (row >= 20) continue;
                                TX *q = xo + (size_t)row * S_pad + pb;
                                if (v0 && v1) {
                                    *(T2 *)q = T2{y[2 * j][rb][v], y[2 * j + 1][rb][v]};
                                } else if (v0) {
                                    q[0] = y[2 * j][rb][v];
                                } else if (v1) {
                                    q[1] = y[2 * j + 1][rb][v];
                                }
                            }
                        if (g4 == 0) {
                            if (v0 && v1)
                                *(int2v *)(eo + pb) = int2v{etot[2 * j], etot[2 * j + 1]};
                            else if (v0)
                                eo[pb] = etot[2 * j];
                            else if (v1)
                                eo[pb + 1] = etot[2 * j + 1];
                        }
                    }
                } else {
#pragma unroll
                    for (int u = 0; u < U; u++) {
                        if (!valid[u]) continue;
#pragma unroll
                        for (int rb = 0; rb < 2; rb++)
#pragma unroll
                            for (int v = 0; v < 4; v++) {
                                const int row = Mfma<TX>::crow(rb, g4, v);
                                if (row < 20) xo[(size_t)row * S_pad + slot[u]] = y[u][rb][v];
                            }
                        if (g4 == 0) eo[slot[u]] = etot[u];
                    }
                }
            }
        }
        // (a block outside the tile leaves rdot = 0: its lanes hold no valid site)
        if constexpr (IS_ROOT) lse_push(lse_m, lse_s, rate_log_term(rdot, rexp), mix_mode);
    }
    if constexpr (IS_ROOT) {
        double lf = 0.0;
        const int o = site_of(g4 / U * U, g4 % U, c16);  // this lane's column: block g4
        if (o >= tl.lo && o < tl.len) {
            const size_t sl = (size_t)p * S_pad + tl.start + o;
            const int orig = perm[sl];
            // a slot stands for the identical columns of its class (weight 1 uncompressed)
            lf = ka.pweight[sl] *
                 site_log_likelihood(lse_m, lse_s, pinv[p], pi[invchar[orig]], isinv[orig], mix_mode);
        }
#pragma unroll
        for (int off = 32; off >= 1; off >>= 1) lf += __shfl_xor(lf, off, 64);
        if (lane == 0) tile_lnl[(size_t)p * tile_stride + tile] = lf;
    }
}

// Dirty-path chains (incremental refreshes).  A chain is a run of internal nodes
// n_0, n_1 = parent(n_0), ..., in which every node after the first has exactly
// one child that is dirty and stored: the node below it.  One wavefront walks one
// tile of a chain from the bottom up, rate by rate, and keeps each node's scaled
// partial in registers (rows 4 kb + g4 of the lane's two slots, which is the
// parent's B operand, see Mfma) together with its exponent, so a node after the
// first reads only its other child -- a tip, a cherry built in registers or a
// stored partial -- and writes its own partial and exponent.  The first node reads
// both children.  Every node runs the products, scaling and exponent sums of
// prune_unit in the same order, so a chain gives the per-node launches' partials
// bit for bit.  Column mapping: paired (unroll 2) -- block k of the 64-slot window
// holds, at lane c16, window offset 32 (k / 2) + 2 c16 + (k % 2).  A workgroup
// takes one tile and its wavefronts take its rates (wave w: rates w, w + 4, ...):
// a chain node depends on the one below, so splitting the rates shortens each
// wavefront's dependent sequence fourfold where the per-node kernels split tiles.
template <typename TX>
__device__ __forceinline__ void chain_unit(const KArgs<TX> &ka, const int *__restrict__ chain_nodes,
                                           const int *__restrict__ chain_len, int item, int tile_group,
                                           TX (&lds)[4][4 * kPStride])
{
    using V4 = typename Mfma<TX>::v4;
    const int *__restrict__ ntiles = ka.ntiles;
    const int *__restrict__ children = ka.children;

    const TX *__restrict__ Pm = ka.Pm;
    const int T = ka.T, R = ka.R, S_pad = ka.S_pad;
    const int wave = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    const int tile = tile_group;
    const Item it = ka.items[item];
    const int p = it.particle, start = it.node, len = chain_len[item];
    if (tile >= ntiles[p]) return;  // no block barrier follows
    const Tile tl = ka.tiles[(size_t)p * ka.tile_stride + tile];
    const int lane = threadIdx.x & 63;
    const int c16 = lane & 15, g4 = lane >> 4;
    const int nn = 2 * T - 1;
    const int *__restrict__ chp = children + (size_t)p * (T - 1) * 2;
    auto mode_of = [&](int c) {
        if (c < T) return MODE_TIP;
        return (chp[(c - T) * 2] < T && chp[(c - T) * 2 + 1] < T) ? MODE_CHERRY : MODE_STORED;
    };
    const size_t node_elems = (size_t)R * 20 * S_pad;
    const size_t lp = (size_t)(p - ka.p_base);
    TX *Xp = ka.X + lp * ka.WS * (T - 1) * node_elems;
    int *Ep = ka.E + lp * ka.WS * (T - 1) * R * S_pad;
    const uint8_t *xf = ka.xflag ? ka.xflag + (size_t)p * (T - 1) : nullptr;
    auto nslot = [&](int node) { return (xf ? (int)xf[node - T] * (T - 1) : 0) + node - T; };
    const TX *Pg = Pm + ((size_t)ka.pcls_off[p] + tl.cls) * nn * R * kPStride;
    using T2 = typename Vec2<TX>::t;
    const int nsb = (tl.len + 15) >> 4;
    const int sb_begin = ((tl.lo >> 4) / 2) * 2;
    // window offsets and slots of the lane's 4 positions (block k)
    int off[4], slk[4];
    bool valid[4];
#pragma unroll
    for (int k = 0; k < 4; k++) {
        off[k] = 32 * (k >> 1) + 2 * c16 + (k & 1);
        valid[k] = off[k] >= tl.lo && off[k] < tl.len;
        slk[k] = tl.start + (valid[k] ? off[k] : tl.lo);  // a lane outside the tile reads a slot inside it
    }
    const uint8_t *__restrict__ pcp = ka.pcodes + (size_t)p * T * S_pad;
    auto code_at = [&](int tip, int k) { return (int)pcp[(size_t)tip * S_pad + slk[k]]; };

    for (int r = wave; r < R; r += 4) {
        TX cy[4][5];  // the previous chain node's partial (B layout) and exponent, per block
        int ce[4];
#pragma unroll
        for (int k = 0; k < 4; k++) {
            ce[k] = 0;
#pragma unroll
            for (int kb = 0; kb < 5; kb++) cy[k][kb] = (TX)0;
        }
        for (int j = 0; j < len; j++) {
            const int n = chain_nodes[start + j];
            const int c0 = chp[(n - T) * 2], c1 = chp[(n - T) * 2 + 1];
            // child A: the chain node below (carried), or for the first node its
            // first child; child B: the other child
            const int below = j > 0 ? chain_nodes[start + j - 1] : -1;
            const int cA = j > 0 ? below : c0;
            const int cB = j > 0 ? (c0 == below ? c1 : c0) : c1;
            const int mA = j > 0 ? MODE_RT : mode_of(cA);  // MODE_RT here: carried in registers
            const int mB = mode_of(cB);
            const TX *PA = Pg + ((size_t)cA * R + r) * kPStride;
            const TX *PB = Pg + ((size_t)cB * R + r) * kPStride;
            TX *lA = lds[wave], *lB = lds[wave] + 2 * kPStride;
            if (mA == MODE_CHERRY || mB == MODE_CHERRY) {
                wave_sync();  // the previous node's reads are done
                if (mA == MODE_CHERRY) {
                    stage_p<TX>(lA, Pg + ((size_t)chp[(cA - T) * 2] * R + r) * kPStride, lane);
                    stage_p<TX>(lA + kPStride, Pg + ((size_t)chp[(cA - T) * 2 + 1] * R + r) * kPStride, lane);
                }
                if (mB == MODE_CHERRY) {
                    stage_p<TX>(lB, Pg + ((size_t)chp[(cB - T) * 2] * R + r) * kPStride, lane);
                    stage_p<TX>(lB + kPStride, Pg + ((size_t)chp[(cB - T) * 2 + 1] * R + r) * kPStride, lane);
                }
                wave_sync();
            }
            TX aA[2][5], aB[2][5];
#pragma unroll
            for (int rb = 0; rb < 2; rb++)
#pragma unroll
                for (int kb = 0; kb < 5; kb++) {
                    const int row = Mfma<TX>::arow(rb, c16);
                    aA[rb][kb] = (mA != MODE_TIP && row < 20) ? PA[row * 21 + 4 * kb + g4] : (TX)0;
                    aB[rb][kb] = (mB != MODE_TIP && row < 20) ? PB[row * 21 + 4 * kb + g4] : (TX)0;
                }
            const int sA = mA == MODE_STORED ? nslot(cA) : 0, sB = mB == MODE_STORED ? nslot(cB) : 0;
            const TX *XA = Xp + ((size_t)sA * R + r) * 20 * S_pad;
            const TX *XB = Xp + ((size_t)sB * R + r) * 20 * S_pad;
            const int *EA = Ep + ((size_t)sA * R + r) * S_pad;
            const int *EB = Ep + ((size_t)sB * R + r) * S_pad;
            TX *xo = Xp + ((size_t)nslot(n) * R + r) * 20 * S_pad;
            int *eo = Ep + ((size_t)nslot(n) * R + r) * S_pad;
            // one child over the two blocks k0, k0 + 1 of a step: its products into
            // acc (C layout) and its exponents
            auto child = [&](int mode, int c, const TX(&a)[2][5], const TX *Pc, const TX *lc, const TX *Xc,
                             const int *Ec, int k0, V4(&acc)[2][2], int(&ex)[2]) {
                const int pb = tl.start + off[k0];  // the even slot of the lane's pair, in the window
                if (mode == MODE_STORED) {
                    TX b0[5], b1[5];
                    stored_b2<TX>(Xc + pb, S_pad, valid[k0], valid[k0 + 1], g4, b0, b1);
                    const int2v ev = *(const int2v *)(Ec + pb);
                    mma_block<TX>(a, b0, acc[0]);
                    mma_block<TX>(a, b1, acc[1]);
                    ex[0] = ev.x;
                    ex[1] = ev.y;
                    return;
                }
#pragma unroll
                for (int u = 0; u < 2; u++) {
                    const int k = k0 + u;
                    ex[u] = 0;
                    if (mode == MODE_TIP) {
                        tip_block<TX>(Pc, code_at(c, k), g4, acc[u]);
                    } else if (mode == MODE_CHERRY) {
                        TX b[5];
                        cherry_b<TX>(lc, lc + kPStride, code_at(chp[(c - T) * 2], k),
                                     code_at(chp[(c - T) * 2 + 1], k), valid[k], g4, b);
                        mma_block<TX>(a, b, acc[u]);
                    } else {  // carried from the chain node below
                        TX b[5];
#pragma unroll
                        for (int kb = 0; kb < 5; kb++) b[kb] = cy[k][kb];
                        mma_block<TX>(a, b, acc[u]);
                        ex[u] = ce[k];
                    }
                }
            };
#pragma unroll
            for (int k0 = 0; k0 < 4; k0 += 2) {
                // one step's products at a time (the steps' accumulators never live together)
                __builtin_amdgcn_sched_barrier(0);
                if (k0 < sb_begin || k0 >= nsb) continue;  // no valid slot in this step (wave-uniform)
                V4 y[2][2], w[2][2];
                int ey[2], ew[2];
                child(mA, cA, aA, PA, lA, XA, EA, k0, y, ey);
                child(mB, cB, aB, PB, lB, XB, EB, k0, w, ew);
#pragma unroll
                for (int u = 0; u < 2; u++) {
                    const int k = k0 + u;
                    int etot = ey[u] + ew[u];
                    TX m = 0;
#pragma unroll
                    for (int rb = 0; rb < 2; rb++)
#pragma unroll
                        for (int v = 0; v < 4; v++) {
                            y[u][rb][v] *= w[u][rb][v];
                            m = fmax(m, y[u][rb][v]);
                        }
                    m = fmax(m, __shfl_xor(m, 16, 64));
                    m = fmax(m, __shfl_xor(m, 32, 64));
                    int e;
                    (void)frexp(m, &e);
                    const bool ok = (m > (TX)0) && isfinite(m);
                    const TX nanv = (TX)NAN;
#pragma unroll
                    for (int rb = 0; rb < 2; rb++)
#pragma unroll
                        for (int v = 0; v < 4; v++) y[u][rb][v] = ok ? (TX)ldexp(y[u][rb][v], -e) : nanv;
                    etot += ok ? e : 0;
#pragma unroll
                    for (int kb = 0; kb < 4; kb++) cy[k][kb] = y[u][0][kb];
                    cy[k][4] = y[u][1][0];
                    ce[k] = etot;
                }
                // store the step's two slots: one paired access per row where both are valid
                const int pb = tl.start + off[k0];
                const bool v0 = valid[k0], v1 = valid[k0 + 1];
#pragma unroll
                for (int rb = 0; rb < 2; rb++)
#pragma unroll
                    for (int v = 0; v < 4; v++) {
                        const int row = Mfma<TX>::crow(rb, g4, v);
                        if (row >= 20) continue;
                        TX *q = xo + (size_t)row * S_pad + pb;
                        if (v0 && v1) {
                            *(T2 *)q = T2{y[0][rb][v], y[1][rb][v]};
                        } else if (v0) {
                            q[0] = y[0][rb][v];
                        } else if (v1) {
                            q[1] = y[1][rb][v];
                        }
                    }
                if (g4 == 0) {
                    if (v0 && v1)
                        *(int2v *)(eo + pb) = int2v{ce[k0], ce[k0 + 1]};
                    else if (v0)
                        eo[pb] = ce[k0];
                    else if (v1)
                        eo[pb + 1] = ce[k0 + 1];
                }
            }
        }
    }
}

// Walks (incremental refreshes: the changed particles of a call in one launch per
// walk level).  A walk is a piece of a particle's dirty stored nodes in post-order
// (short dirty sets are one piece with the root last; whole trees are cut into
// pieces of at most kWalkPiece nodes, and a piece runs in the launch after the
// pieces cut below it, which it reads back).  One wavefront, a workgroup of its
// own, takes one (tile, half, rate) of one piece: the half h of the tile's 64-slot
// window (blocks 2 h, 2 h + 1) and the rates r, r + 4, ..., and runs the whole
// piece, so every dependency between its nodes stays inside one wavefront and no
// two wavefronts of a launch depend on each other (nothing is shared through LDS;
// single-wavefront workgroups measured faster than the tile's eight wavefronts in
// one workgroup, 16.6 against 18.8 ms per particle of the C4 batched MCMC: they
// start and end one by one, beside the other stream's level launches).  It
// generalises the chains below: a node's result stays in registers as the B
// operand of the next entry when that entry is its parent (child A carried); any
// other stored child -- a committed partial, a node of an earlier piece, or a dirty
// node this wavefront wrote earlier in the piece -- is read from memory (a
// wavefront's own earlier stores are visible to its later loads).  A node issues
// the loads of both children before its products; tip columns and cherry products
// are gathered from the P matrices by code (no LDS, so occupancy is set by the
// registers alone).  Per node the products, scaling and exponent sums are those of
// prune_unit, so the partials equal the per-level launches' bit for bit.  The root
// entry writes each rate's pi . x and exponent per site to a buffer, and
// k_walk_mixture forms the tiles' sums from it (a separate launch: the mixture's
// FP64 transcendentals would otherwise set the walk's register count).  Entries
// carry their children's modes and partial slots (see WalkItem), so a node needs
// no lookups before its loads.  Column mapping: block k of the window holds, at
// lane c16, window offset 32 (k / 2) + 2 c16 + (k % 2).
template <typename TX>
__device__ __forceinline__ void walk_unit(const KArgs<TX> &ka, const int4 *__restrict__ entries, const WalkItem wi,
                                          int tile, int half, int rate0, double *__restrict__ rdot,
                                          int *__restrict__ rexp)
{
    using V4 = typename Mfma<TX>::v4;
    using T2 = typename Vec2<TX>::t;
    const int T = ka.T, R = ka.R, S_pad = ka.S_pad;
    const int p = wi.particle;
    const Tile tl = ka.tiles[(size_t)p * ka.tile_stride + tile];
    const int lane = threadIdx.x & 63;
    const int c16 = lane & 15, g4 = lane >> 4;
    const int nn = 2 * T - 1;
    const size_t node_elems = (size_t)R * 20 * S_pad;
    const size_t lp = (size_t)(p - ka.p_base);
    TX *Xp = ka.X + lp * ka.WS * (T - 1) * node_elems;
    int *Ep = ka.E + lp * ka.WS * (T - 1) * R * S_pad;
    const size_t g = (size_t)p * ka.max_classes + tl.cls;
    const TX *Pg = ka.Pm + ((size_t)ka.pcls_off[p] + tl.cls) * nn * R * kPStride;
    const int nsb = (tl.len + 15) >> 4;
    const int sb_begin = ((tl.lo >> 4) / 2) * 2;
    const int k0 = 2 * half;  // this wavefront's blocks k0, k0 + 1
    const bool active = k0 >= sb_begin && k0 < nsb;  // wave-uniform: a valid slot in the blocks
    const int pb = tl.start + 2 * c16 + 32 * half;   // the lane's even slot (block k0), in the window
    const bool v0 = pb - tl.start >= tl.lo && pb - tl.start < tl.len;
    const bool v1 = pb + 1 - tl.start >= tl.lo && pb + 1 - tl.start < tl.len;
    const bool valid[2] = {v0, v1};
    const int slk0 = v0 ? pb : tl.start + tl.lo, slk1 = v1 ? pb + 1 : tl.start + tl.lo;
    const uint8_t *__restrict__ pcp = ka.pcodes + (size_t)p * T * S_pad;
    // The tip codes an entry needs at the lane's two slots, packed in one word per
    // child: a tip's codes in bytes 0 and 1 (blocks k0, k0 + 1), a cherry's tip a in
    // bytes 0-1 and tip b in bytes 2-3.  (Loading the next entry's codes during this
    // one was measured: 26 more registers, 3 wavefronts per SIMD, slower.)
    auto codes_of = [&](int mode, int c, int aux) -> uint32_t {
        auto two = [&](int tip) {
            const uint8_t *q = pcp + (size_t)tip * S_pad;
            return (uint32_t)q[slk0] | (uint32_t)q[slk1] << 8;
        };
        if (mode == WALK_TIP) return two(c);
        if (mode == WALK_CHERRY) return two(aux & 0xffff) | two(aux >> 16) << 16;
        return 0u;
    };
    // the root's per-site pi . x and exponents: [wave slot][tile][rate][64 lanes]
    const size_t root_base = ((size_t)lp * ka.tile_stride + tile) * R;
    const int4 *ent = entries + 2 * (size_t)wi.first;

    for (int r = rate0; r < R && active; r += 4) {
        TX cy[2][5];  // the previous entry's partial (B layout) and exponent, per block
        int ce[2] = {0, 0};
#pragma unroll
        for (int u = 0; u < 2; u++)
#pragma unroll
            for (int kb = 0; kb < 5; kb++) cy[u][kb] = (TX)0;
        for (int j = 0; j < wi.count; j++) {
            const int4 en = ent[2 * j], ex = ent[2 * j + 1];
            const int cA = en.y, cB = en.z;
            const int mA = en.w & 7, mB = (en.w >> 4) & 7;
            const bool is_root = (en.w >> 8) & 1;
            const int taA = ex.y & 0xffff, tbA = ex.y >> 16, taB = ex.z & 0xffff, tbB = ex.z >> 16;
            const int oslot = ex.x, sA = mA == WALK_STORED ? ex.y : 0, sB = mB == WALK_STORED ? ex.z : 0;
            const uint32_t kA = codes_of(mA, cA, ex.y), kB = codes_of(mB, cB, ex.z);
            const TX *PA = Pg + ((size_t)cA * R + r) * kPStride;
            const TX *PB = Pg + ((size_t)cB * R + r) * kPStride;
            TX aA[2][5], aB[2][5];
#pragma unroll
            for (int rb = 0; rb < 2; rb++)
#pragma unroll
                for (int kb = 0; kb < 5; kb++) {
                    const int row = Mfma<TX>::arow(rb, c16);
                    aA[rb][kb] = (mA != WALK_TIP && row < 20) ? PA[row * 21 + 4 * kb + g4] : (TX)0;
                    aB[rb][kb] = (mB != WALK_TIP && row < 20) ? PB[row * 21 + 4 * kb + g4] : (TX)0;
                }
            // the children's loads before any product: B operands of stored children
            // (child A's into cy), exponents
            TX xb[2][5];
            int eb[2] = {0, 0};
            if (mA == WALK_STORED) {
                const TX *Xc = Xp + ((size_t)sA * R + r) * 20 * S_pad;
                stored_b2<TX>(Xc + pb, S_pad, v0, v1, g4, cy[0], cy[1]);
                const int2v ev = *(const int2v *)(Ep + ((size_t)sA * R + r) * S_pad + pb);
                ce[0] = ev.x;
                ce[1] = ev.y;
            }
            if (mB == WALK_STORED) {
                const TX *Xc = Xp + ((size_t)sB * R + r) * 20 * S_pad;
                stored_b2<TX>(Xc + pb, S_pad, v0, v1, g4, xb[0], xb[1]);
                const int2v ev = *(const int2v *)(Ep + ((size_t)sB * R + r) * S_pad + pb);
                eb[0] = ev.x;
                eb[1] = ev.y;
            }
            // one child's products over the two blocks into acc (C layout); a tip's
            // P x is a column of its P, a cherry's B operand the product of its
            // tips' P columns (both gathered by code from the P matrices)
            auto child = [&](int mode, uint32_t kw, int ta, int tb, const TX(&a)[2][5], const TX *Pc,
                             const TX(&bs)[2][5], V4(&acc)[2][2]) {
#pragma unroll
                for (int u = 0; u < 2; u++) {
                    const int ca = (int)((kw >> (8 * u)) & 0xff), cb = (int)((kw >> (8 * u + 16)) & 0xff);
                    if (mode == WALK_TIP) {
                        tip_block<TX>(Pc, ca, g4, acc[u]);
                    } else if (mode == WALK_CHERRY) {
                        TX b[5];
                        cherry_b<TX>(Pg + ((size_t)ta * R + r) * kPStride, Pg + ((size_t)tb * R + r) * kPStride,
                                     ca, cb, valid[u], g4, b);
                        mma_block<TX>(a, b, acc[u]);
                    } else {  // stored or carried: the B operand is in registers
                        mma_block<TX>(a, bs[u], acc[u]);
                    }
                }
            };
            V4 y[2][2], w[2][2];
            child(mA, kA, taA, tbA, aA, PA, cy, y);
            child(mB, kB, taB, tbB, aB, PB, xb, w);
            const bool stA = mA == WALK_STORED || mA == WALK_CARRIED, stB = mB == WALK_STORED;
#pragma unroll
            for (int u = 0; u < 2; u++) {
                int etot = (stA ? ce[u] : 0) + (stB ? eb[u] : 0);
                TX m = 0;
#pragma unroll
                for (int rb = 0; rb < 2; rb++)
#pragma unroll
                    for (int v = 0; v < 4; v++) {
                        y[u][rb][v] *= w[u][rb][v];
                        m = fmax(m, y[u][rb][v]);
                    }
                m = fmax(m, __shfl_xor(m, 16, 64));
                m = fmax(m, __shfl_xor(m, 32, 64));
                int e;
                (void)frexp(m, &e);
                const bool ok = (m > (TX)0) && isfinite(m);
                const TX nanv = (TX)NAN;
#pragma unroll
                for (int rb = 0; rb < 2; rb++)
#pragma unroll
                    for (int v = 0; v < 4; v++) y[u][rb][v] = ok ? (TX)ldexp(y[u][rb][v], -e) : nanv;
                etot += ok ? e : 0;
#pragma unroll
                for (int kb = 0; kb < 4; kb++) cy[u][kb] = y[u][0][kb];
                cy[u][4] = y[u][1][0];
                ce[u] = etot;
            }
            if (is_root) {
                // (the stationary vector is read here, not kept in registers for the walk)
                const double *pi = ka.eig + g * kEigStride + kEigPi;
                asm volatile("" : "+v"(pi));
#pragma unroll
                for (int u = 0; u < 2; u++) {
                    // pi . x of the block's column -- the same sum as prune_unit's root
                    double dot = 0.0;
#pragma unroll
                    for (int rb = 0; rb < 2; rb++)
#pragma unroll
                        for (int v = 0; v < 4; v++) {
                            const int row = Mfma<TX>::crow(rb, g4, v);
                            if (row < 20) dot = fma(pi[row], (double)y[u][rb][v], dot);
                        }
                    dot += __shfl_xor(dot, 16, 64);
                    dot += __shfl_xor(dot, 32, 64);
                    if (g4 == 0) {
                        rdot[(root_base + r) * 64 + 16 * (k0 + u) + c16] = dot;
                        rexp[(root_base + r) * 64 + 16 * (k0 + u) + c16] = ce[u];
                    }
                }
                continue;  // the root's partial is not stored
            }
            // store both slots: one paired access per row where both are valid
            TX *xo = Xp + ((size_t)oslot * R + r) * 20 * S_pad;
            int *eo = Ep + ((size_t)oslot * R + r) * S_pad;
#pragma unroll
            for (int rb = 0; rb < 2; rb++)
#pragma unroll
                for (int v = 0; v < 4; v++) {
                    const int row = Mfma<TX>::crow(rb, g4, v);
                    if (row >= 20) continue;
                    TX *q = xo + (size_t)row * S_pad + pb;
                    if (v0 && v1) {
                        *(T2 *)q = T2{y[0][rb][v], y[1][rb][v]};
                    } else if (v0) {
                        q[0] = y[0][rb][v];
                    } else if (v1) {
                        q[1] = y[1][rb][v];
                    }
                }
            if (g4 == 0) {
                if (v0 && v1)
                    *(int2v *)(eo + pb) = int2v{ce[0], ce[1]};
                else if (v0)
                    eo[pb] = ce[0];
                else if (v1)
                    eo[pb + 1] = ce[1];
            }
        }
    }
}

// The root mixture of walked tiles (one wavefront per (walk item, tile)): lane
// 16 k + c16 takes block k's column c16 of the tile (window offset
// 32 (k / 2) + 2 c16 + (k % 2)), sums the rates' terms in rate order and the
// lanes' site log-likelihoods with prune_unit's root butterfly, so tile_lnl equals
// the root launch's bit for bit.  Blocks with no valid slot were never written by
// the walk: their lanes are outside the tile.
__global__ __launch_bounds__(256) void k_walk_mixture(const WalkItem *__restrict__ witems, int n,
                                                      const Tile *__restrict__ tiles, const int *__restrict__ ntiles,
                                                      int tile_stride, const double *__restrict__ rdot,
                                                      const int *__restrict__ rexp, int R, int p_base, int S_pad,
                                                      int max_classes, const int *__restrict__ perm,
                                                      const double *__restrict__ pweight,
                                                      const double *__restrict__ pinv, const double *__restrict__ eig,
                                                      const int *__restrict__ invchar,
                                                      const double *__restrict__ isinv, int mix_mode,
                                                      double *__restrict__ tile_lnl)
{
    const int64_t w = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const int lane = threadIdx.x & 63;
    if (w >= (int64_t)n * tile_stride) return;  // whole wavefront
    const int q = (int)(w / tile_stride), tile = (int)(w % tile_stride);
    const int p = witems[q].particle;
    if (tile >= ntiles[p]) return;  // whole wavefront
    const Tile tl = tiles[(size_t)p * tile_stride + tile];
    const int g4 = lane >> 4, c16 = lane & 15;
    const int o = 32 * (g4 >> 1) + 2 * c16 + (g4 & 1);
    const bool in = o >= tl.lo && o < tl.len;
    const size_t base = ((size_t)(p - p_base) * tile_stride + tile) * R;
    double lse_m = -INFINITY, lse_s = 0.0;
    for (int r = 0; r < R; r++) {
        const double d = in ? rdot[(base + r) * 64 + lane] : 0.0;
        const int ez = in ? rexp[(base + r) * 64 + lane] : 0;
        lse_push(lse_m, lse_s, rate_log_term(d, ez), mix_mode);
    }
    double lf = 0.0;
    if (in) {
        const size_t sl = (size_t)p * S_pad + tl.start + o;
        const int orig = perm[sl];
        const double *pi = eig + ((size_t)p * max_classes + tl.cls) * kEigStride + kEigPi;
        lf = pweight[sl] * site_log_likelihood(lse_m, lse_s, pinv[p], pi[invchar[orig]], isinv[orig], mix_mode);
    }
#pragma unroll
    for (int o2 = 32; o2 >= 1; o2 >>= 1) lf += __shfl_xor(lf, o2, 64);
    if (lane == 0) tile_lnl[(size_t)p * tile_stride + tile] = lf;
}


// Persistent grid-stride form: the launch walks the units of items
// [base, base + n), read from the device (bucket_off[bucket], [bucket + 1]) when
// bucket_off is given, so one launch configuration serves every topology and a
// captured graph can be replayed after the schedule changes.
template <typename TX, int M0, int M1, bool IS_ROOT, int U, int W>
__global__ __launch_bounds__(64 * W) void k_prune_mfma(KArgs<TX> ka, const int *__restrict__ bucket_off,
                                                    int bucket, int base, int n)
{
    if (bucket_off) {
        base = bucket_off[bucket];
        n = bucket_off[bucket + 1] - base;
    }
    // unit w runs on XCD w mod X (the grid is a multiple of X); XCD x takes the
    // items x, x + X, ..., each with all its tile groups, so the P matrices and
    // tip codes an item reads are fetched into one L2 only
    __shared__ __attribute__((aligned(16))) TX lds[W][CherryLds<M0, M1>::elems];
    const int bpi = ka.blocks_per_item, X = ka.xcds;
    const int n_work = (n + X - 1) / X * X * bpi;
    for (int w = blockIdx.x; w < n_work; w += gridDim.x) {
        const int l = w / X;
        const int q = (l / bpi) * X + (w - l * X);
        if (q < n) prune_unit<TX, M0, M1, IS_ROOT, U, W>(ka, base + q, l % bpi, lds);
    }
}

// Chains, persistent grid-stride like k_prune_mfma with one tile per workgroup
// (ka.blocks_per_item = tile stride): items [0, n) of `items`, chain i's nodes at
// chain_nodes[items[i].node ...], chain_len[i] of them.
template <typename TX>
__global__ __launch_bounds__(256) void k_prune_chain(KArgs<TX> ka, const int *__restrict__ chain_nodes,
                                                     const int *__restrict__ chain_len, int n)
{
    __shared__ __attribute__((aligned(16))) TX lds[4][4 * kPStride];
    const int bpi = ka.blocks_per_item, X = ka.xcds;
    const int n_work = (n + X - 1) / X * X * bpi;
    for (int w = blockIdx.x; w < n_work; w += gridDim.x) {
        const int l = w / X;
        const int q = (l / bpi) * X + (w - l * X);
        if (q < n) chain_unit<TX>(ka, chain_nodes, chain_len, q, l % bpi, lds);
    }
}

// Walks, one (tile, half, rate) per single-wavefront workgroup (ka.blocks_per_item =
// 8 x tile stride; grid-stride, so any grid works): walk items [0, n) of `witems`.
template <typename TX>
__global__ __launch_bounds__(64) void k_prune_walk(KArgs<TX> ka, const WalkItem *__restrict__ witems,
                                                    const int4 *__restrict__ entries, int n,
                                                    double *__restrict__ rdot, int *__restrict__ rexp)
{
    const int bpi = ka.blocks_per_item, X = ka.xcds;
    const int n_work = (n + X - 1) / X * X * bpi;
    for (int w = blockIdx.x; w < n_work; w += gridDim.x) {
        const int l = w / X;
        const int q = (l / bpi) * X + (w - l * X);
        if (q >= n) continue;
        const WalkItem wi = witems[q];
        const int u = l % bpi, tile = u >> 3, half = (u >> 2) & 1, rate = u & 3;
        if (tile >= ka.ntiles[wi.particle]) continue;  // workgroup-uniform
        walk_unit<TX>(ka, entries, wi, tile, half, rate, rdot, rexp);
    }
}

}  // namespace

namespace {

template <typename TX, int M0, int M1, bool IS_ROOT, int U>
void launch_one(const PruneArgs &a, const KArgs<TX> &ka, unsigned grid, hipStream_t s)
{
    if constexpr (U == 2) {
        if (a.single_wave) {
            hipLaunchKernelGGL((k_prune_mfma<TX, M0, M1, IS_ROOT, U, 1>), dim3(grid), dim3(64), 0, s, ka,
                               a.bucket_off, a.bucket, a.item_base, a.n_items);
            return;
        }
    }
    hipLaunchKernelGGL((k_prune_mfma<TX, M0, M1, IS_ROOT, U, 4>), dim3(grid), dim3(256), 0, s, ka, a.bucket_off,
                       a.bucket, a.item_base, a.n_items);
}

template <typename TX, int M0, int M1>
void launch_pair(const PruneArgs &a, const KArgs<TX> &ka, unsigned grid, hipStream_t s)
{
    if (a.unroll >= 4)
        launch_one<TX, M0, M1, false, 4>(a, ka, grid, s);
    else if (a.unroll == 2)
        launch_one<TX, M0, M1, false, 2>(a, ka, grid, s);
    else
        launch_one<TX, M0, M1, false, 1>(a, ka, grid, s);
}

// launches whose child modes are decided per item (roots, merged small buckets)
template <typename TX, bool IS_ROOT>
void launch_rt(const PruneArgs &a, const KArgs<TX> &ka, unsigned grid, hipStream_t s)
{
    if (a.unroll >= 4)
        launch_one<TX, MODE_RT, MODE_RT, IS_ROOT, 4>(a, ka, grid, s);
    else if (a.unroll == 2)
        launch_one<TX, MODE_RT, MODE_RT, IS_ROOT, 2>(a, ka, grid, s);
    else
        launch_one<TX, MODE_RT, MODE_RT, IS_ROOT, 1>(a, ka, grid, s);
}

template <typename TX>
KArgs<TX> kargs(const PruneArgs &a)
{
    const int bpi = (a.tile_stride + 3) / 4;
    return KArgs<TX>{a.items, a.tiles, a.ntiles, a.tile_stride, a.children, a.perm, a.codes, a.pcodes, a.pweight,
                     (const TX *)a.P, a.eig, a.pinv, a.isinv, a.invchar, (TX *)a.X, a.E, a.tile_lnl,
                     a.T, a.R, a.S_pad, a.max_classes, bpi, a.pcls_off, a.p_base, a.mix_mode, a.xflag, a.WS,
                     std::max(1, a.xcds)};
}

template <typename TX>
void launch_chain_typed(const PruneArgs &a, const int *chain_nodes, const int *chain_len, hipStream_t s)
{
    KArgs<TX> ka = kargs<TX>(a);
    ka.blocks_per_item = a.tile_stride;  // one workgroup per tile, its wavefronts per rate
    const int X = ka.xcds, bpi = ka.blocks_per_item;
    const int64_t units = (int64_t)(a.n_items + X - 1) / X * X * bpi;
    const unsigned grid = (unsigned)std::max<int64_t>(1, std::min<int64_t>(units, a.grid > 0 ? a.grid : units));
    hipLaunchKernelGGL((k_prune_chain<TX>), dim3(grid), dim3(256), 0, s, ka, chain_nodes, chain_len, a.n_items);
}

template <typename TX>
void launch_walk_typed(const PruneArgs &a, const WalkItem *witems, const int4 *entries, double *rdot, int *rexp,
                       hipStream_t s)
{
    KArgs<TX> ka = kargs<TX>(a);
    ka.blocks_per_item = 8 * a.tile_stride;  // one single-wavefront workgroup per (tile, half, rate)
    const int X = ka.xcds, bpi = ka.blocks_per_item;
    const int64_t units = (int64_t)(a.n_items + X - 1) / X * X * bpi;
    // one workgroup per unit: short-lived workgroups let the dispatcher interleave
    // the second stream's level launches as they finish (a persistent grid of half
    // the pruning grid, or of the resident capacity, measured slower)
    const unsigned grid = (unsigned)std::max<int64_t>(1, std::min<int64_t>(units, INT32_MAX));
    hipLaunchKernelGGL((k_prune_walk<TX>), dim3(grid), dim3(64), 0, s, ka, witems, entries, a.n_items, rdot,
                       rexp);
}

template <typename TX>
void launch_typed(const PruneArgs &a, hipStream_t s)
{
    // single-wavefront workgroups (full sweeps at unroll 2: measured 3.5% faster there,
    // 1% slower in MCMC rounds) take one tile each, four-wavefront ones four tiles;
    // the persistent grid keeps the same number of wavefronts
    const bool one = a.single_wave && a.unroll == 2;
    const int bpi = one ? a.tile_stride : (a.tile_stride + 3) / 4;
    KArgs<TX> ka = kargs<TX>(a);
    ka.blocks_per_item = bpi;
    // without device bucket offsets the work is known: launch exactly what it needs
    // (items rounded up to whole XCD groups), capped at the persistent grid
    const int X = ka.xcds;
    int64_t units = a.bucket_off ? INT64_MAX : (int64_t)(a.n_items + X - 1) / X * X * bpi;
    const int64_t cap = a.grid > 0 ? (one ? 4 * (int64_t)a.grid : a.grid) : units;
    unsigned grid = (unsigned)std::max<int64_t>(1, std::min<int64_t>(units, cap));
    if (a.roots) {
        launch_rt<TX, true>(a, ka, grid, s);
        return;
    }
    switch (a.kind) {
        case PRUNE_TT: launch_pair<TX, MODE_TIP, MODE_TIP>(a, ka, grid, s); break;
        case PRUNE_TC: launch_pair<TX, MODE_TIP, MODE_CHERRY>(a, ka, grid, s); break;
        case PRUNE_TS: launch_pair<TX, MODE_TIP, MODE_STORED>(a, ka, grid, s); break;
        case PRUNE_CC: launch_pair<TX, MODE_CHERRY, MODE_CHERRY>(a, ka, grid, s); break;
        case PRUNE_CS: launch_pair<TX, MODE_CHERRY, MODE_STORED>(a, ka, grid, s); break;
        case PRUNE_SS: launch_pair<TX, MODE_STORED, MODE_STORED>(a, ka, grid, s); break;
        default: launch_rt<TX, false>(a, ka, grid, s); break;
    }
}

}  // namespace

void launch_prune_chain(const PruneArgs &a, const int *chain_nodes, const int *chain_len, hipStream_t s)
{
    if (a.n_items <= 0) return;
    if (a.fp32)
        launch_chain_typed<float>(a, chain_nodes, chain_len, s);
    else
        launch_chain_typed<double>(a, chain_nodes, chain_len, s);
}

void launch_prune_walk(const PruneArgs &a, const WalkItem *witems, const int4 *entries, double *rdot, int *rexp,
                       hipStream_t s)
{
    if (a.n_items <= 0) return;
    if (a.fp32)
        launch_walk_typed<float>(a, witems, entries, rdot, rexp, s);
    else
        launch_walk_typed<double>(a, witems, entries, rdot, rexp, s);
}

void launch_walk_mixture(const PruneArgs &a, const WalkItem *witems, const double *rdot, const int *rexp,
                         hipStream_t s)
{
    if (a.n_items <= 0) return;
    const int64_t waves = (int64_t)a.n_items * a.tile_stride;
    hipLaunchKernelGGL(k_walk_mixture, dim3((unsigned)((waves + 3) / 4)), dim3(256), 0, s, witems, a.n_items,
                       a.tiles, a.ntiles, a.tile_stride, rdot, rexp, a.R, a.p_base, a.S_pad, a.max_classes, a.perm,
                       a.pweight, a.pinv, a.eig, a.invchar, a.isinv, a.mix_mode, a.tile_lnl);
}

void launch_prune_mfma(const PruneArgs &a, hipStream_t s)
{
    if (!a.bucket_off && a.n_items <= 0) return;
    if (a.fp32)
        launch_typed<float>(a, s);
    else
        launch_typed<double>(a, s);
}

}  // namespace smccat
