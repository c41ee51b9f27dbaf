// Internal declarations shared by the engine's translation units.
#pragma once

#include <utility>

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace smccat {

constexpr int kStates = 20;        // amino-acid states
constexpr int kTile = 64;          // sites per pruning tile = one wavefront
constexpr int kEigStride = 840;    // A[400] | lam[20] | B[400] | pi[20]
constexpr int kEigLam = 400;
constexpr int kEigB = 420;
constexpr int kEigPi = 820;
constexpr int kPStride = 420;      // P row-major with a 21st column: P[i][20] = sum_j P[i][j]

extern const double kLG190[190];

// A pruning tile: the slots of class `cls` inside the 64-slot window that starts at
// slot `start` (a multiple of 64) of the class-sorted slot order, i.e. window
// offsets [lo, len).  Windows are aligned so that every 16-slot column block starts
// on a cache-line boundary; a window holding the end of one class and the start of
// the next is two tiles.
struct Tile {
    int start;
    int len;
    int cls;
    int lo;
};

// One (particle, internal node) pair of a pruning level.
struct Item {
    int particle;
    int node;
};

// Walks of incremental refreshes: a particle's dirty stored nodes in post-order
// (the root last), entries [first, first + count) of the entry list.  An entry is
// two int4: {node, child A, child B, modes} -- bits 0-2 the mode of child A, bits
// 4-6 the mode of child B (WALK_*), bit 8 set for the root -- and {the node's
// partial slot, aux A, aux B, 0}, where a partial slot is workspace * (T - 1) +
// node - T and a child's aux is its partial slot when stored, tip a | tip b << 16
// when a cherry.  A carried child (the entry just before) is always child A.
struct WalkItem {
    int particle;
    int first;
    int count;
};
enum { WALK_TIP = 0, WALK_CHERRY = 1, WALK_STORED = 2, WALK_CARRIED = 3 };
constexpr int kWalkMaxRates = 8;
constexpr int kWalkPiece = 24;  // longest walk piece (stored nodes); longer walks are cut

// Host-side discretizeGamma (Miscellaneous.hpp:22-37).
void discretize_gamma(double shape, int num, double *out);

// ---- kernel launchers (stream-ordered, no synchronisation) ----
struct EigenArgs {
    const double *base;      // [P][190]
    const double *logits;    // [P][Kmax][20]
    const int *num_classes;  // [P]
    double *eig;             // [P*Kmax][840]
    int max_classes;
    int first;               // first particle
    int count;               // particles
    const int *plist;        // optional [count] particle list (else first..first+count-1)
};
void launch_eigen(const EigenArgs &a, hipStream_t s);

struct PmatArgs {
    const double *eig;
    const double *blen;      // [P][nn]
    const double *rates;     // [P][R]
    const int *num_classes;
    const int *pcls_off;     // [P] first compact class slot of each particle in the wave
    void *P;                 // [wave classes][nn][R][420] (double or float)
    int nn, R, max_classes, first, count;
    bool fp32;
    const int *plist;        // optional [count] particle list (else first..first+count-1)
};
void launch_pmat(const PmatArgs &a, hipStream_t s);
// P of listed nodes only: particle plist[i] recomputes nodes[noff[i] .. noff[i+1])
// for every class and rate.
void launch_pmat_nodes(const PmatArgs &a, const int *plist, const int *noff, const int *nodes, int n_particles,
                       int max_nodes, hipStream_t s);

// Child modes of a pruning item (sorted pair): a child is a tip (T), a cherry
// (C: an internal node whose children are both tips) or a stored partial (S).
// The matrix-core kernel never stores cherries (their parents build them);
// the VALU kernel stores them like any internal node.
enum { PRUNE_TT = 0, PRUNE_TC = 1, PRUNE_TS = 2, PRUNE_CC = 3, PRUNE_CS = 4, PRUNE_SS = 5, PRUNE_KINDS = 6,
       PRUNE_ANY = 6 };

// PRUNE_* kind of internal node `node` of a children array [T-1][2].
inline int prune_kind(const int *ch, int T, int node)
{
    auto mode = [&](int c) {
        if (c < T) return 0;
        return (ch[2 * (c - T)] < T && ch[2 * (c - T) + 1] < T) ? 1 : 2;
    };
    int a = mode(ch[2 * (node - T)]), b = mode(ch[2 * (node - T) + 1]);
    if (a > b) std::swap(a, b);
    static const int table[3][3] = {{PRUNE_TT, PRUNE_TC, PRUNE_TS}, {PRUNE_TC, PRUNE_CC, PRUNE_CS},
                                    {PRUNE_TS, PRUNE_CS, PRUNE_SS}};
    return table[a][b];
}

// Number of children of a PRUNE_* kind read as stored partials (0, 1, 2).
inline int prune_stored(int kind)
{
    static const int n[PRUNE_KINDS] = {0, 0, 1, 0, 1, 2};
    return kind < PRUNE_KINDS ? n[kind] : 1;
}

struct PruneArgs {
    const Item *items;       // this level's items
    int n_items;
    const Tile *tiles;       // [P][tile_stride]
    const int *ntiles;       // [P]
    int tile_stride;
    const int *children;     // [P][T-1][2]
    const int *root_of;      // [P]
    const int *perm;         // [P][S_pad] slot -> alignment column
    const uint8_t *codes;    // [T][S_pad]
    const uint8_t *pcodes;   // [P][T][S_pad] tip codes in each particle's slot order (20: gap or unknown)
    const double *pweight;   // [P][S_pad] columns each slot stands for (roots weight its log-likelihood)
    const void *P;           // [wave classes][nn][R][420]
    const int *pcls_off;     // [P] compact class slot of (particle, class 0) in P
    const double *eig;       // [P*Kmax][840]
    const double *pinv;      // [P]
    const double *isinv;     // [S_pad]
    const int *invchar;      // [S_pad]
    void *X;                 // partials [wave slot][T-1][R][20][S_pad]
    int *E;                  // cumulative binary exponents [wave slot][T-1][R][S_pad]
    int p_base;              // first particle of the wave (wave slot = particle - p_base)
    const uint8_t *xflag;    // [P][T-1] current workspace of each internal node (null: 0)
    int WS;                  // workspaces per particle (1, or 2 for accept/reject)
    double *tile_lnl;        // [P][tile_stride]
    int T, R, S_pad, max_classes;
    bool fp32;
    bool roots;              // items are tree roots: fuse the root mixture, store nothing
    int mix_mode;            // SMC_MIX_REFERENCE or SMC_MIX_LOGSPACE (roots only)
    int kind;                // PRUNE_* child modes of every item (PRUNE_ANY: decided per item)
    // matrix-core kernel only: items [item_base, item_base + n_items) of `items`,
    // or, when bucket_off is set, the range bucket_off[bucket] .. [bucket + 1]
    // read on the device (a launch that does not depend on the topology)
    int item_base = 0;
    const int *bucket_off = nullptr;
    int bucket = 0;
    int grid = 0;            // persistent grid (workgroups); 0: one per unit of work
    int unroll = 1;          // 16-site column blocks per step (1, 2 or 4)
    int xcds = 1;            // keep an item's workgroups on one of this many XCDs
    bool single_wave = false;  // matrix-core levels of unroll 2: one wavefront (one tile) per workgroup
};
void launch_prune(const PruneArgs &a, hipStream_t s);       // VALU, scalar-operand P
void launch_prune_mfma(const PruneArgs &a, hipStream_t s);  // matrix cores
// Dirty-path chains on the matrix cores: items [0, n_items) of a.items are chains,
// item i = {particle, index of its first node in chain_nodes}, chain_len[i] nodes
// from the bottom of the path upwards (every node after the first has the node
// before it as its only dirty stored child).
void launch_prune_chain(const PruneArgs &a, const int *chain_nodes, const int *chain_len, hipStream_t s);
// walks: items [0, a.n_items) of `witems`, their entries in `entries` (R <= kWalkMaxRates);
// rdot / rexp: the roots' per-site pi . x and exponents, [wave slots][tile_stride][R][64]
void launch_prune_walk(const PruneArgs &a, const WalkItem *witems, const int4 *entries, double *rdot, int *rexp,
                       hipStream_t s);
// the root mixture of the tiles of walked roots (items [0, a.n_items)) into tile_lnl
void launch_walk_mixture(const PruneArgs &a, const WalkItem *witems, const double *rdot, const int *rexp,
                         hipStream_t s);

// Single-site x candidate-profile scoring (Gibbs / learned allocation proposal).
constexpr int kSiteStack = 32;  // LDS stack entries for live internal partials
// Per-particle scoring program ([P][3T] ints): the Strahler-ordered post-order of
// the internal nodes (T-1), the internal nodes sorted by height (T-1), the level
// count H, then H+1 offsets of the height levels in the sorted list.
__host__ __device__ inline int site_prog_stride(int T) { return 3 * T; }
constexpr int kSiteLevelMaxTaxa = 320;  // the level kernel keeps every internal partial in LDS
struct SiteClassArgs {
    const int *ent_particle;  // [count] particle of each entry
    const int *ent_site;      // [count] alignment column of each entry
    const int *ent_ncand;     // [count] candidates of each entry (<= n_cand)
    int n_cand;
    const double *cand_eig;   // [count][n_cand][840]
    // Optional [count][n_cand] class codes: a candidate coded >= 0 (a class of the
    // entry's particle, in eig_cache [particles][max_classes]) or <= -2 (pool[-2 - v])
    // is read where it lies; -1 (or no codes) reads cand_eig.
    const int *cand_cls = nullptr;
    const double *eig_cache = nullptr;
    int max_classes = 0;
    const double *pool = nullptr;
    const int *children;      // [P][T-1][2]
    const int *prog;          // [P][3T] scoring programs (site_prog_stride)
    const uint8_t *codes;     // [T][S_pad]
    const double *blen;       // [P][nn]
    const double *rates;      // [P][R]
    const double *pinv;
    const double *isinv;
    const int *invchar;
    double *rate_lnl;         // [count][n_cand][R] scratch
    double *out;              // [count][n_cand] tempering * log L
    int count, T, R, S_pad, mix_mode;
    double tempering;
};
void launch_site_class(const SiteClassArgs &a, hipStream_t s);

// Eigen-decomposition of candidate profiles: entry e uses the base matrix of
// particle ent_particle[e] and logits[e][c][20] for c < ent_ncand[e].
// cand_cls (nullable) [count][n_cand]: >= 0 copies the class's eigensystem from
// eig_cache ([particles][max_classes] layout), <= -2 copies pool[-2 - v] (neither is
// copied with copy_known false: the site-class kernels then read them where they
// lie), -1 decomposes the logits.
void launch_eigen_cand(const double *base, const int *ent_particle, const int *ent_ncand,
                       const double *logits, int count, int n_cand, double *eig, const int *cand_cls,
                       const double *eig_cache, int max_classes, const double *pool, hipStream_t s,
                       bool copy_known = true);

void launch_reduce_lnl(const double *tile_lnl, const int *ntiles, int tile_stride, int first,
                       int count, double *lnl, uint8_t *nonfinite, hipStream_t s);

void launch_next_step(int n, double *w, const double *l, double *norm, double *out2,
                      int target_ess, double last_temp, hipStream_t s);
void launch_resample(int n, const double *norm, double u, int *assign, hipStream_t s);
// pcodes[p][t][slot] = min(codes[t][perm[p][slot]], 20) for particles [first, first + count).
void launch_permute_codes(const uint8_t *codes, const int *perm, uint8_t *pcodes, int T, int S_pad, int first,
                          int count, hipStream_t s);
// dst row dst_rows[i] (i when null) <- src row src_rows[i] (i when null), rows of
// row_bytes bytes, i < n (buffers aligned to the largest of 16, 4, 1 dividing row_bytes).
// A row of a packed upload: `words` 32-bit words from word `src_word` of the
// packed buffer to the device address `dst`.
struct ScatterSeg {
    uint64_t dst;
    uint32_t src_word;
    uint32_t words;
};
void launch_scatter_segments(const uint32_t *src, const ScatterSeg *segs, int n, hipStream_t s);
void launch_gather_rows(const void *src, void *dst, size_t row_bytes, const int *src_rows, const int *dst_rows,
                        int n, hipStream_t s);

// Proposal-network batch (mlp_layout.hpp): B examples, row r uses example
// idx[r] (idx may be null: row r is example r).  Activations are per row.
struct MlpArgs {
    int B = 0;
    const int *idx = nullptr;
    const float *hdr = nullptr;     // [examples][kHdr]
    const int *K = nullptr;         // [examples] classes
    const int *label = nullptr;     // [examples] chosen slot (training)
    const int64_t *coff = nullptr;  // [examples] first class entry
    const float *cls = nullptr;     // [entries][kCls]
    const float *theta = nullptr;
    float *Z1 = nullptr, *H1 = nullptr, *Z2 = nullptr, *H2 = nullptr, *dH2 = nullptr;  // [B][kH]
    float *loss = nullptr;          // [B]
    float *part = nullptr;          // [ceil(B / kRows)][kNParam]
    float *logq = nullptr;          // [B][out_stride] (inference)
    int out_stride = 0;
    float inv_B = 0.f;
};

// The learned allocation move's proposal draw and the site-class entries of its
// MH step (see k_mlp_draw).
struct MoveDrawArgs {
    int count = 0, stride = 0, kmax = 0;
    const float *logq = nullptr;   // [count][stride] proposal log-weights
    const int *K = nullptr;        // [count] classes without the site
    const int *aux = nullptr;      // [count][4]: particle slot, site, current class, its class code
    const int *codes = nullptr;    // [count][kmax + 1]: class code of each slot (engine class,
                                   // or -2 - pool index for a fresh profile)
    const double *u = nullptr;     // [count] uniforms of the draws
    int *pick = nullptr;           // [count] drawn slot
    int *ent = nullptr;            // [3][count] entries: particle, site, candidate count (2 or 0)
    int *cand_cls = nullptr;       // [count][2] candidate class codes: current, drawn
};
void launch_mlp_draw(const MoveDrawArgs &a, hipStream_t s);
void launch_mlp_trunk(const MlpArgs &a, bool train, hipStream_t s);
void launch_mlp_head(const MlpArgs &a, bool train, hipStream_t s);
void launch_mlp_backward(const MlpArgs &a, hipStream_t s);
void launch_mlp_reduce(int rows, const float *part, float *grad, const float *loss, int B,
                       double *loss_acc, hipStream_t s);
void launch_mlp_adam(float *theta, float *m, float *v, const float *g, float lr, float b1, float b2,
                     float eps, float c1, float c2, hipStream_t s);

}  // namespace smccat
