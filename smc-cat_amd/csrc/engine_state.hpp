// Engine state shared by the translation units of the C-ABI (not installed).
#pragma once

#include <hip/hip_runtime.h>

#include <array>
#include <chrono>
#include <functional>
#include <cstdint>
#include <string>
#include <vector>

#include "internal.hpp"
#include "smccat_hip.h"

using namespace smccat;

struct PendingTiming {
    int family;
    hipEvent_t a, b;
};

// Fresh candidate profiles drawn ahead and decomposed in one launch: q per slot
// of a move's batch, consumed in draw order (logits [count][q][20], device
// eigensystems in the same order).
struct ProfilePool {
    int count = 0, q = 0;
    std::vector<int> next;
    std::vector<double> logits;
    double *d_eig = nullptr, *d_logits = nullptr;
    int *d_ent = nullptr;
    int64_t cap = 0, lcap = 0, ecap = 0;
};

struct smc_engine {
    smc_engine_config cfg{};
    int T = 0, S = 0, S_pad = 0, R = 0, nn = 0, Pmax = 0, Kmax = 0, tile_stride = 0;
    int Wmax = 0;          // partial slots (particles per wave)
    int WS = 1;            // workspaces per particle: 2 enables accept/reject
    int64_t Cmax = 0;      // classes per wave in Pmat
    bool fp32 = false;
    size_t tx_size = 8;
    hipStream_t stream = nullptr;
    // a second stream (default priority) for work that overlaps the first within
    // one call: an incremental refresh's whole-tree levels beside its walks; every
    // use joins the first stream before the call returns
    hipStream_t stream2 = nullptr;
    hipEvent_t ev_fork = nullptr, ev_join = nullptr;
    std::string err;
    int64_t device_bytes = 0;
    bool have_alignment = false;
    int prune_variant = SMC_PRUNE_MFMA;
    int mix_mode = SMC_MIX_REFERENCE;

    // device buffers
    uint8_t *d_codes = nullptr;
    uint8_t *d_pcodes = nullptr;  // [P][T][S_pad] tip codes in each particle's class-sorted slot order
    double *d_pweight = nullptr;  // [P][S_pad] columns each slot stands for (site patterns within a class)
    bool compress = true;         // merge identical columns of a class into one weighted slot
    std::vector<int> h_pattern;   // [S] pattern id of each column (identical columns share one)
    int n_patterns = 0;
    std::vector<int> h_nslots;    // [P] slots (patterns) in use per particle
    double *d_isinv = nullptr;
    int *d_invchar = nullptr;
    int *d_children = nullptr;
    double *d_blen = nullptr;
    double *d_base = nullptr;
    double *d_logits = nullptr;
    int *d_nclass = nullptr;
    double *d_pinv = nullptr;
    double *d_rates = nullptr;
    int *d_root = nullptr;
    int *d_perm = nullptr;
    Tile *d_tiles = nullptr;
    int *d_ntiles = nullptr;
    double *d_eig = nullptr;
    void *d_P = nullptr;
    // a second P buffer for multi-wave sweeps, allocated when memory allows: the next
    // wave's eigensystems and P are computed into it on stream2 beside this wave's
    // pruning (pipe_ev: [0..1] models of a buffer done, [2..3] pruning of a buffer done)
    void *d_P2 = nullptr;
    bool p2_tried = false;
    hipEvent_t pipe_ev[4] = {nullptr, nullptr, nullptr, nullptr};
    void *d_X = nullptr;
    int *d_E = nullptr;
    double *d_tile_lnl = nullptr;
    double *d_lnl = nullptr;
    uint8_t *d_nonfinite = nullptr;
    Item *d_items = nullptr;
    // staging of smc_compute_next_step_noalloc: this device's copy of a file-scope
    // device array (never allocated or freed by the engine)
    double *d_noalloc_scratch = nullptr;
    // SMC scratch (grown on demand)
    int smc_cap = 0;
    double *d_w = nullptr, *d_l = nullptr, *d_norm = nullptr, *d_out2 = nullptr;
    int *d_assign = nullptr;

    // host mirrors
    std::vector<int> h_children;  // [P][T-1][2]
    std::vector<int> h_root;      // [P]
    std::vector<int> h_height;    // [P][T-1]
    std::vector<int> h_prog;      // [P][3T] host copy of the scoring programs
    std::vector<uint8_t> h_prog_stale;  // [P] topology changed since its program was derived
    int64_t n_prog_stale = 0;
    std::vector<int> h_valid;     // [P] particle uploaded
    std::vector<int> h_nclass;    // [P] classes per particle
    // host copies of the uploaded particle state (for moves that edit it)
    std::vector<double> h_blen;   // [P][nn]
    std::vector<double> h_base;   // [P][190]
    std::vector<double> h_logits; // [P][Kmax][20]
    std::vector<int> h_assign;    // [P][S]
    std::vector<double> h_pinv, h_shape;  // [P]
    std::vector<float> h_sitefreq;        // [S][20] amino-acid frequencies of each column
    std::vector<double> h_rates;          // [P][R] discrete-gamma rates of h_shape
    int *d_pcls_off = nullptr;    // [P] compact class slot of (particle, class 0) in its wave
    int *d_prog = nullptr;        // [P][3T] scoring programs: Strahler post-order, height levels
    uint8_t *d_xflag = nullptr;   // [P][T-1] current workspace of each internal node
    int *d_plist = nullptr;       // [P] particle list scratch for partial recomputes
    // accept/reject state (Particle.cpp:343-478): committed flags, lnL and parameters
    std::vector<uint8_t> h_xflag, h_xflag_old;  // [P][T-1]
    std::vector<double> h_lnl, h_lnl_old;       // [P]
    std::vector<uint8_t> h_dirty;               // [P] bit0 P changed, bit1 eigen changed
    std::vector<std::vector<int>> h_pnodes;     // [P] nodes whose P changed since the commit
    std::vector<uint8_t> h_touched;             // [P] pInvar changed by the last parameter upload
    // Rejects restore P lazily: the recomputes are merged into the next incremental
    // refresh of the particle (one launch with that refresh's own, and none where
    // the particle recomputes every P anyway), or dropped by a full refresh.
    std::vector<int> pend_full;                 // particles whose every P is recomputed
    std::vector<Item> pend_nodes;               // (particle, node) P recomputes
    // Eigensystems of a particle before its first uncommitted recompute, copied
    // back on a reject instead of decomposed again: [Wmax * Kmax][840] by slot
    double *d_eig_save = nullptr;
    std::vector<int> h_eig_save;                // [P] save slot holding the committed eigensystems, or -1
    std::vector<int> eig_save_free;             // free save slots
    // [Wmax] the particle whose committed partials fill wave slot w (-1: none): a
    // particle p refreshed with p_base b is resident while h_owner[p - b] == p
    std::vector<int> h_owner;
    uint8_t *d_gather = nullptr;                 // row-gather staging of smc_copy_particles
    uint8_t *d_scatter = nullptr;                // packed parameter rows of upload_params (grown on demand)
    int64_t scatter_cap = 0;
    int64_t gather_cap = 0;
    uint8_t *d_split_ws = nullptr;               // split-posterior workspace (grown on demand)
    size_t split_ws_bytes = 0;
    Item *d_pnodes = nullptr;                   // [P * nn] node lists for P updates
    int *d_chain_nodes = nullptr;               // [P * (T-1)] dirty-path chains' nodes, bottom up
    int *d_chain_len = nullptr;                 // [P * (T-1)] chain lengths, aligned with d_items
    int4 *d_walk = nullptr;                     // [P * T * 2] walk entries (two int4 each)
    WalkItem *d_witems = nullptr;               // [P * (T + 1)] walk pieces, then the roots
    double *d_walk_rdot = nullptr;              // [Wmax][tile_stride][R][64] walked roots' pi . x
    int *d_walk_rexp = nullptr;                 // the same, exponents (both allocated on first use)
    struct Committed {
        std::vector<int32_t> children;
        std::vector<double> blen, base;
        double pinv = 0.0, shape = 1.0;
        int root = -1;             // derived from children: a restore needs no re-derivation
        std::vector<int> heights;  // [T-1]
    };
    std::vector<Committed> committed;           // [P] parameters at the last commit
    // site x candidate scoring scratch (grown on demand)
    int64_t sc_cap = 0;           // entries * candidates
    int sc_ent_cap = 0;
    int *d_sc_ent = nullptr;      // [3][entries]: particle, site, candidate count
    int *d_sc_cls = nullptr;      // [entries][candidates] engine class of a candidate or -1
    ProfilePool pool;             // fresh profiles of the CAT moves
    double *d_sc_logits = nullptr, *d_sc_eig = nullptr, *d_sc_rate = nullptr, *d_sc_out = nullptr;

    // pruning schedule cache: waves of particles, each with its level/kind buckets
    struct Wave {
        int first, count;
        int root_off;                 // roots of the wave's particles
        int max_level;
        int graph_levels;             // max_level rounded up: the levels a captured sweep walks
        int boff;                     // first entry of this wave's offsets in d_bucket_off
        // items of (level L, kind k): [off[K L + k], off[K L + k + 1]) for L <= graph_levels
        // (K = PRUNE_KINDS), then the roots: [off[nb], off[nb + 1]) with nb = K (graph_levels + 1)
        std::vector<int> bucket_off;
    };
    int *d_bucket_off = nullptr;      // the waves' bucket offsets on the device
    int64_t bucket_cap = 0;

    // launch tuning (smc_engine_autotune) and the captured full sweep
    struct Tune {
        int blocks_per_cu = 32;       // persistent grid of the pruning launches: CUs x this
        int unroll[3] = {2, 1, 1};    // 16-site column blocks per step, by stored children (0, 1, 2);
                                      // set per precision at creation
        int xcd_groups = 8;           // an item's workgroups share one of this many XCDs
        int unroll_rt = 2;            // the same for launches that decide child modes per item
        int chain_max = 48;           // incremental refreshes: dirty sets of at most this many
                                      // stored nodes run as chains (tree moves touch one or two
                                      // paths to the root)
    };
    Tune tune;
    int num_cus = 256;
    bool use_graph = true;
    // captured sweeps keyed by their launch sequence (particle range, waves, level
    // counts, settings); the least recently used one is dropped beyond kMaxGraphs,
    // so a one-particle-at-a-time caller replays one graph per slot instead of
    // capturing anew on every call
    struct CachedGraph {
        std::vector<int64_t> key;
        hipGraphExec_t exec;
        uint64_t used;
    };
    static constexpr size_t kMaxGraphs = 256;
    std::vector<CachedGraph> graphs;
    uint64_t graph_clock = 0;
    int64_t graph_launches = 0, graph_captures = 0;
    int sched_first = -1, sched_count = -1;
    // single-wave class layout last written to d_pcls_off (incremental path)
    int pcls_first = -1, pcls_count = -1;
    uint64_t pcls_version = 0, class_version = 1;
    uint64_t sched_version = 0, topo_version = 1;
    std::vector<Wave> waves;

    // host time per entry point (smc_engine_phase_stats)
    double phase_ms[SMC_PHASE_COUNT] = {};
    int64_t phase_calls[SMC_PHASE_COUNT] = {};

    // pinned staging of the uploads of the per-step calls (h2d): bump-allocated,
    // recycled at sync_stream
    uint8_t *h_stage = nullptr;
    size_t stage_cap = 0, stage_off = 0;
    // pinned readback of the per-step calls' results (d2h): bump-allocated, copied
    // out to their destinations and recycled at sync_stream
    struct Readback {
        void *dst;
        size_t off, bytes;
    };
    uint8_t *h_read = nullptr;
    size_t read_cap = 0, read_off = 0;
    std::vector<Readback> readbacks;

    // The lane: an incremental refresh that runs beside the caller's next calls (the
    // whole-tree proposals of a lockstep batch, while the other particles go on with
    // their rounds).  It has its own stream, staging and readback buffers, item and
    // model lists and lnL output, swapped into the fields above while it is issued
    // (swap_lane); its particles are busy (h_busy) until lane_finish.
    struct Lane {
        hipStream_t stream = nullptr;
        hipEvent_t fork = nullptr, done = nullptr;
        uint8_t *h_stage = nullptr;
        size_t stage_cap = 0, stage_off = 0;
        uint8_t *h_read = nullptr;
        size_t read_cap = 0, read_off = 0;
        std::vector<Readback> readbacks;
        double *d_lnl = nullptr;
        uint8_t *d_nonfinite = nullptr;
        Item *d_items = nullptr;
        int *d_chain_nodes = nullptr, *d_chain_len = nullptr, *d_plist = nullptr;
        Item *d_pnodes = nullptr;
        std::vector<double> lnl;     // [P] lnL read back by the lane's refresh
        std::vector<int> particles;  // the particles it recomputes
        bool in_flight = false;
    } lane;
    std::vector<uint8_t> h_busy;  // [P] particle recomputed by the lane (no commit, no reject)
    bool lane_enabled = true;     // rejuvenation runs whole-tree steps on the lane

    // timing
    bool timing = false;
    std::vector<hipEvent_t> ev_pool;
    std::vector<PendingTiming> pending;
    int64_t launches[SMC_K_COUNT] = {};
    double total_ms[SMC_K_COUNT] = {};
};

int smc_fail(smc_engine *e, int code, const std::string &msg);

// Looks up this device's address of the no-allocation scratch array (once per
// device, with the engine's device current) into e->d_noalloc_scratch.
int smc_noalloc_init(smc_engine *e);

// Queues a copy of `bytes` host bytes to device `dst` on the engine's stream
// through the pinned staging buffer: it returns without waiting for the device,
// and `src` may be changed at once.  (Copies from pageable memory wait for the
// stream to drain first, which leaves the GPU idle while the host catches up.)
int h2d(smc_engine *e, void *dst, const void *src, size_t bytes);
// Queues a copy of `bytes` device bytes at `src` to host `dst` on the engine's
// stream through the pinned readback buffer; `dst` holds them after the next
// sync_stream.  (A copy to pageable memory is staged by the runtime with a wait of
// its own.)
int d2h(smc_engine *e, void *dst, const void *src, size_t bytes);
// Waits for the engine's stream and copies the queued readbacks out; the staging
// and readback buffers are free again afterwards.
int sync_stream(smc_engine *e);
// Incremental refresh of slots [first, first + count) (smc_refresh_incremental);
// with on_lane, issued on the lane and not waited for: the lnL of its changed
// particles arrive at lane_finish.
int refresh_incremental(smc_engine *e, int first, int count, const uint8_t *node_flags,
                        const uint8_t *eig_flags, double *lnl_out, uint8_t *nonfinite_out, bool on_lane);
// Completes the lane's refresh: with wait, waits for it; without, returns at once
// with *done false while it still runs.  On completion the lane's particles hold
// their new lnL (h_lnl) and are no longer busy.
int lane_finish(smc_engine *e, bool wait, bool *done);
// Site x candidate scoring buffers (d_sc_*) for `count` entries and `need` (entry,
// candidate) pairs, grown geometrically.
int ensure_site_class_buffers(smc_engine *e, int count, int64_t need);
// Queues the site-class scoring of the entries in d_sc_ent ([particle | site |
// candidate count] x count) with, when with_cls, the class codes in d_sc_cls
// (existing classes and pool profiles are read where they lie); `decompose` first
// decomposes the candidates coded -1 (every candidate without codes) from
// d_sc_logits.  The scores land in `out` ([count][n_cand]; null: d_sc_out).
int issue_site_class(smc_engine *e, int count, int n_cand, bool with_cls, bool decompose, double tempering,
                     double *out = nullptr);

// Adds the wall time of its scope to e->phase_ms[phase].
struct PhaseTimer {
    smc_engine *e;
    int phase;
    std::chrono::steady_clock::time_point t0;
    PhaseTimer(smc_engine *ee, int ph) : e(ee), phase(ph), t0(std::chrono::steady_clock::now()) {}
    ~PhaseTimer()
    {
        e->phase_ms[phase] +=
            std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
        e->phase_calls[phase]++;
    }
};

// Receives every draw of the Gibbs allocation move: the site, the classes
// (logits) and sizes without the site, and the chosen slot (K = a new class).
struct AllocationSink {
    virtual ~AllocationSink() = default;
    virtual void add(int site, const std::vector<std::array<double, 20>> &logits,
                     const std::vector<int> &sizes, int label, double tempering, double dpp_alpha) = 0;
};

// Replaces the classes (K, logits [Kmax][20]) and allocation of one uploaded
// slot, keeping its other parameters.
int smc_upload_partition(smc_engine *e, int slot, int K, const double *logits, const int32_t *assign);

// smc_rejuvenate that also adds the CAT move's site updates and site moves to
// cat_stats[0..1] (nullable).  With deferred_cat (nullable) the slots chosen for
// the CAT move are appended there instead of moved: the caller runs smc_cat_moves
// over them later (the move of a slot depends on nothing but that slot and the
// seed, so the result is the same) and refreshes their likelihoods.
int smc_rejuvenate_ex(smc_engine *e, int first, int count, double tempering,
                      const smc_rejuvenation_options *opt, double *lnl_out, int64_t *stats_out,
                      int64_t *cat_stats, std::vector<int32_t> *deferred_cat = nullptr);
// The CAT move of smc_rejuvenate (learned, recording Gibbs or Gibbs, as the options
// say) over any list of slots, all in lockstep.
int smc_cat_moves(smc_engine *e, const std::vector<int32_t> &slots, double tempering,
                  const smc_rejuvenation_options &o, int64_t *cat_stats);

// smc_site_class_lnl where cand_class[i][c] >= 0 (nullable) names a class of
// the entry's particle whose eigensystem the engine holds (its logits are then
// not read); -1 decomposes cand_logits[i][c].
extern "C" int smc_site_class_lnl_ex(smc_engine *e, int count, const int32_t *particles,
                                     const int32_t *sites, int n_cand, const int32_t *ncand,
                                     const double *cand_logits, const int32_t *cand_class, double tempering,
                                     double *out);

// Fills the profile pool: q profiles per listed slot, all decomposed in one
// launch.  With fresh = 0 (a refill within one move) each slot keeps its
// unconsumed profiles first and draw(i, logits20) tops it up; fresh = 1 starts
// a new pool.  A site-class candidate -2 - (i q + j) names profile j of entry i.
int smc_pool_refill(smc_engine *e, int count, const int32_t *slots, int q,
                    const std::function<void(int, double *)> &draw, int fresh);

// Copies pool profile pool_index's eigensystem into class slot `cls` of
// particle `slot` (stream-ordered), so a class a move creates is scored from
// the engine's eigensystems like the others.
int smc_pool_to_class(smc_engine *e, int slot, int pool_index, int cls);

// Recomputes the eigensystems of every class of the listed slots (from the
// uploaded base matrices and logits), so that cached candidates are current.
int smc_refresh_eigen(smc_engine *e, int count, const int32_t *slots);

// The Gibbs allocation move (smc_gibbs_partition_move) with an optional sink.
int smc_gibbs_move(smc_engine *e, int count, const int32_t *slots, double tempering, uint64_t seed,
                   int num_iter, double dpp_alpha, int64_t *stats_out, AllocationSink *sink);

// Commits the current state of slots [first, first+count): workspace flags,
// lnL and parameters become the ones a rejected proposal returns to.
void smc_commit_range(smc_engine *e, int first, int count);
void drop_pending(smc_engine *e, int first, int count);
int ensure_programs(smc_engine *e);
int scatter_upload(smc_engine *e, const std::vector<ScatterSeg> &segs, const std::vector<uint32_t> &data);

// refreshLikelihood(true) of slots [first, first+count) (one resident wave, two
// workspaces) that recomputes the partials only of slots whose committed partials
// are not resident (h_owner); the others keep theirs and their lnL.  Eigensystems
// and P are recomputed for the whole range.  Commits the range.
int smc_refresh_stale(smc_engine *e, int first, int count);

// refreshLikelihood(true) of the slots of [first, first+count) with which[i] != 0
// (one resident wave, two workspaces): eigensystems, P and every node of those
// slots, committed; the other slots keep their lnL and are not touched.
int smc_refresh_slots(smc_engine *e, int first, int count, const uint8_t *which);

// Validates the children array of particle p ([T-1][2]) and derives its root,
// the heights of its internal nodes and the Strahler-ordered scoring program.
int smc_topology(smc_engine *e, int p, const int32_t *ch, int *root_out, int *heights,
                 int *prog_out);

#define HIP_TRY(e, call)                                                                      \
    do {                                                                                      \
        hipError_t _st = (call);                                                              \
        if (_st != hipSuccess)                                                                \
            return smc_fail((e), SMC_ERR_HIP,                                                 \
                            std::string(#call) + " failed: " + hipGetErrorString(_st));      \
    } while (0)
