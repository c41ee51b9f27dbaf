// C-ABI of the particle-likelihood engine: device memory, level scheduling of
// the per-particle trees and the launch sequence of one likelihood sweep.
//
// Device layout (all resident in HBM, sized at create time):
//   codes    uint8  [T][S_pad]                 alignment, original column order
//   perm     int32  [P][S_pad]                 slot -> column, class-sorted per particle
//   tiles    Tile   [P][tile_stride]           one class's slots in an aligned 64-slot window
//   eig      f64    [P*Kmax][840]              A | lam | B | pi per class
//   Pmat     TX     [C][2T-1][K_r][20x21]      transition matrices (+ row sums) of the
//                                              C = pmat_classes classes of one wave
//   X        TX     [W][T-1][K_r][20][S_pad]   internal-node partials, state-major, for the
//                                              W = partial_slots particles of one wave
//   E        int32  [W][T-1][K_r][S_pad]       cumulative binary exponents
//   tile_lnl f64    [P][tile_stride]           per-tile site-lnL sums
// P = max_particles particle states stay resident; a sweep over more particles than W
// (or with more classes than C) runs in waves.  TX is double on the plain path and
// float on the fast path.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <functional>
#include <mutex>
#include <string>
#include <thread>
#include <utility>
#include <vector>

#include "engine_state.hpp"

int smc_fail(smc_engine *e, int code, const std::string &msg)
{
    if (e) e->err = msg;
    return code;
}

int sync_stream(smc_engine *e)
{
    HIP_TRY(e, hipStreamSynchronize(e->stream));
    if (e->stream2) HIP_TRY(e, hipStreamSynchronize(e->stream2));
    e->stage_off = 0;  // every staged copy has completed
    for (const smc_engine::Readback &r : e->readbacks) std::memcpy(r.dst, e->h_read + r.off, r.bytes);
    e->readbacks.clear();
    e->read_off = 0;
    return SMC_OK;
}

int d2h(smc_engine *e, void *dst, const void *src, size_t bytes)
{
    if (bytes == 0) return SMC_OK;
    const size_t need = (bytes + 255) & ~(size_t)255;
    if (need > e->read_cap) {  // larger than the readback buffer: copy and wait
        int rc = sync_stream(e);
        if (rc != SMC_OK) return rc;
        HIP_TRY(e, hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, e->stream));
        return sync_stream(e);
    }
    if (e->read_off + need > e->read_cap) {  // full: wait and copy the queued ones out
        int rc = sync_stream(e);
        if (rc != SMC_OK) return rc;
    }
    HIP_TRY(e, hipMemcpyAsync(e->h_read + e->read_off, src, bytes, hipMemcpyDeviceToHost, e->stream));
    e->readbacks.push_back(smc_engine::Readback{dst, e->read_off, bytes});
    e->read_off += need;
    return SMC_OK;
}

int h2d(smc_engine *e, void *dst, const void *src, size_t bytes)
{
    if (bytes == 0) return SMC_OK;
    const size_t need = (bytes + 255) & ~(size_t)255;
    if (need > e->stage_cap) {  // larger than the staging buffer: copy and wait
        HIP_TRY(e, hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, e->stream));
        return sync_stream(e);
    }
    if (e->stage_off + need > e->stage_cap) {  // full: wait for the queued copies
        int rc = sync_stream(e);
        if (rc != SMC_OK) return rc;
    }
    uint8_t *st = e->h_stage + e->stage_off;
    std::memcpy(st, src, bytes);
    e->stage_off += need;
    HIP_TRY(e, hipMemcpyAsync(dst, st, bytes, hipMemcpyHostToDevice, e->stream));
    return SMC_OK;
}

namespace {

int fail(smc_engine *e, int code, const std::string &msg) { return smc_fail(e, code, msg); }


template <typename T>
int dev_alloc(smc_engine *e, T **ptr, size_t count)
{
    size_t bytes = count * sizeof(T);
    if (bytes == 0) bytes = 16;
    hipError_t st = hipMalloc((void **)ptr, bytes);
    if (st != hipSuccess) {
        *ptr = nullptr;
        return fail(e, SMC_ERR_NOMEM,
                    "hipMalloc of " + std::to_string(bytes) + " bytes failed: " +
                        hipGetErrorString(st));
    }
    e->device_bytes += (int64_t)bytes;
    return SMC_OK;
}

hipEvent_t take_event(smc_engine *e)
{
    if (!e->ev_pool.empty()) {
        hipEvent_t ev = e->ev_pool.back();
        e->ev_pool.pop_back();
        return ev;
    }
    hipEvent_t ev = nullptr;
    if (hipEventCreate(&ev) != hipSuccess) return nullptr;
    return ev;
}

// Bracket a launch with events when timing is on.
struct TimedLaunch {
    smc_engine *e;
    int family;
    hipEvent_t a = nullptr, b = nullptr;
    TimedLaunch(smc_engine *eng, int fam) : e(eng), family(fam)
    {
        if (e->timing) {
            a = take_event(e);
            b = take_event(e);
            if (a) (void)hipEventRecord(a, e->stream);
        }
    }
    ~TimedLaunch()
    {
        if (e->timing && a && b) {
            (void)hipEventRecord(b, e->stream);
            e->pending.push_back({family, a, b});
        }
    }
};

// Call after a stream synchronisation.
void collect_timings(smc_engine *e)
{
    for (auto &pt : e->pending) {
        float ms = 0.f;
        if (hipEventElapsedTime(&ms, pt.a, pt.b) == hipSuccess) {
            e->launches[pt.family] += 1;
            e->total_ms[pt.family] += ms;
        }
        e->ev_pool.push_back(pt.a);
        e->ev_pool.push_back(pt.b);
    }
    e->pending.clear();
}

// Build the pruning schedule for slots [first, first+count).  Particles are
// grouped into waves of at most Wmax particles and Cmax classes (partials and P
// of one wave are resident at a time).  Within a wave, bucket (L, k) holds the
// non-root internal nodes of height L (tips have height 0) whose children are
// of PRUNE_* kind k; buckets are stored level-major, and the wave's roots follow
// in one final list, so the level kernels never carry the root mixture.
int build_schedule(smc_engine *e, int first, int count)
{
    if (e->sched_first == first && e->sched_count == count && e->sched_version == e->topo_version)
        return SMC_OK;
    const int T = e->T, Ti = T - 1;
    auto kind_of = [&](int p, int i) { return prune_kind(&e->h_children[(size_t)p * Ti * 2], T, T + i); };
    e->waves.clear();
    std::vector<Item> items((size_t)count * Ti);
    std::vector<int> pcls((size_t)count, 0);
    size_t item_pos = 0;
    int p = first;
    while (p < first + count) {
        smc_engine::Wave w;
        w.first = p;
        int64_t ncls = 0;
        while (p < first + count && p - w.first < e->Wmax && ncls + e->h_nclass[p] <= e->Cmax) {
            pcls[p - first] = (int)ncls;
            ncls += e->h_nclass[p];
            ++p;
        }
        w.count = p - w.first;
        if (w.count == 0) return fail(e, SMC_ERR_STATE, "a particle has more classes than Pmat holds");
        int maxh = 0;
        for (int q = w.first; q < w.first + w.count; ++q)
            for (int i = 0; i < Ti; ++i)
                if (T + i != e->h_root[q]) maxh = std::max(maxh, e->h_height[(size_t)q * Ti + i]);
        // a captured sweep walks graph_levels levels: rounding keeps it valid while the
        // population's tallest tree varies a little
        const int glev = std::min(T - 1, (maxh + 3) / 4 * 4);
        const int nb = (glev + 1) * PRUNE_KINDS;
        std::vector<int> cnt(nb + 1, 0);
        for (int q = w.first; q < w.first + w.count; ++q)
            for (int i = 0; i < Ti; ++i)
                if (T + i != e->h_root[q]) cnt[e->h_height[(size_t)q * Ti + i] * PRUNE_KINDS + kind_of(q, i)]++;
        w.bucket_off.assign(nb + 2, (int)item_pos);
        for (int b = 0; b < nb; ++b) w.bucket_off[b + 1] = w.bucket_off[b] + cnt[b];
        w.bucket_off[nb + 1] = w.bucket_off[nb] + w.count;
        w.max_level = maxh;
        w.graph_levels = glev;
        std::vector<int> fill(w.bucket_off);
        w.root_off = w.bucket_off[nb];
        for (int q = w.first; q < w.first + w.count; ++q)
            for (int i = 0; i < Ti; ++i) {
                if (T + i == e->h_root[q]) {
                    items[w.root_off + (q - w.first)] = Item{q, T + i};
                } else {
                    const int b = e->h_height[(size_t)q * Ti + i] * PRUNE_KINDS + kind_of(q, i);
                    items[fill[b]++] = Item{q, T + i};
                }
            }
        item_pos = (size_t)w.root_off + w.count;
        e->waves.push_back(std::move(w));
    }
    std::vector<int> boffs;
    for (smc_engine::Wave &w : e->waves) {
        w.boff = (int)boffs.size();
        boffs.insert(boffs.end(), w.bucket_off.begin(), w.bucket_off.end());
    }
    if ((int64_t)boffs.size() > e->bucket_cap) {
        // captured sweeps still in flight read the old buffer
        HIP_TRY(e, hipStreamSynchronize(e->stream));
        if (e->d_bucket_off) (void)hipFree(e->d_bucket_off);
    if (e->d_P2) (void)hipFree(e->d_P2);
    for (hipEvent_t ev : e->pipe_ev)
        if (ev) (void)hipEventDestroy(ev);
        e->d_bucket_off = nullptr;
        e->bucket_cap = 0;
        const int64_t cap = std::max<int64_t>((int64_t)boffs.size(), 1024);
        HIP_TRY(e, hipMalloc((void **)&e->d_bucket_off, (size_t)cap * sizeof(int)));
        e->bucket_cap = cap;
        for (smc_engine::CachedGraph &c : e->graphs)  // captured sweeps read the old buffer
            (void)hipGraphExecDestroy(c.exec);
        e->graphs.clear();
    }
    HIP_TRY(e, hipMemcpyAsync(e->d_bucket_off, boffs.data(), boffs.size() * sizeof(int),
                              hipMemcpyHostToDevice, e->stream));
    HIP_TRY(e, hipMemcpyAsync(e->d_items, items.data(), items.size() * sizeof(Item),
                              hipMemcpyHostToDevice, e->stream));
    HIP_TRY(e, hipMemcpyAsync(e->d_pcls_off + first, pcls.data(), pcls.size() * sizeof(int),
                              hipMemcpyHostToDevice, e->stream));
    HIP_TRY(e, hipStreamSynchronize(e->stream));
    e->pcls_first = -1;  // the wave layout of the schedule replaced any single-wave layout
    e->sched_first = first;
    e->sched_count = count;
    e->sched_version = e->topo_version;
    return SMC_OK;
}

}  // namespace

void smc_commit_range(smc_engine *e, int first, int count)
{
    const int Ti = e->T - 1, nn = e->nn;
    for (int p = first; p < first + count; ++p) {
        std::copy(e->h_xflag.begin() + (size_t)p * Ti, e->h_xflag.begin() + (size_t)(p + 1) * Ti,
                  e->h_xflag_old.begin() + (size_t)p * Ti);
        e->h_lnl_old[p] = e->h_lnl[p];
        e->h_dirty[p] = 0;
        e->h_pnodes[p].clear();
        if (e->h_eig_save[p] >= 0) {
            e->eig_save_free.push_back(e->h_eig_save[p]);
            e->h_eig_save[p] = -1;
        }
        smc_engine::Committed &c = e->committed[p];
        c.children.assign(e->h_children.begin() + (size_t)p * Ti * 2,
                          e->h_children.begin() + (size_t)(p + 1) * Ti * 2);
        c.blen.assign(e->h_blen.begin() + (size_t)p * nn, e->h_blen.begin() + (size_t)(p + 1) * nn);
        c.base.assign(e->h_base.begin() + (size_t)p * 190, e->h_base.begin() + (size_t)(p + 1) * 190);
        c.pinv = e->h_pinv[p];
        c.shape = e->h_shape[p];
        c.root = e->h_root[p];
        c.heights.assign(e->h_height.begin() + (size_t)p * Ti, e->h_height.begin() + (size_t)(p + 1) * Ti);
    }
}

// Drops the deferred reject restores of slots [first, first + count) (their P is
// recomputed or replaced).
void drop_pending(smc_engine *e, int first, int count)
{
    auto in = [&](int p) { return p >= first && p < first + count; };
    e->pend_full.erase(std::remove_if(e->pend_full.begin(), e->pend_full.end(), in), e->pend_full.end());
    e->pend_nodes.erase(std::remove_if(e->pend_nodes.begin(), e->pend_nodes.end(),
                                       [&](const Item &it) { return in(it.particle); }),
                        e->pend_nodes.end());
}

int smc_topology(smc_engine *e, int p, const int32_t *ch, int *root_out, int *heights, int *prog_out)
{
    const int T = e->T, Ti = T - 1, nn = e->nn;
    // topology checks: every non-root node has exactly one parent (scratch
    // vectors reused across calls: this runs for every proposed topology)
    thread_local std::vector<int> parent, h, order, stv;
    parent.assign(nn, -1);
    for (int k = 0; k < Ti; ++k)
        for (int c = 0; c < 2; ++c) {
            int d = ch[2 * k + c];
            if (d < 0 || d >= nn || d == T + k || parent[d] != -1)
                return fail(e, SMC_ERR_ARG, "particle " + std::to_string(p) +
                                                ": invalid children array");
            parent[d] = T + k;
        }
    int root = -1, nroots = 0;
    for (int n = T; n < nn; ++n)
        if (parent[n] == -1) {
            root = n;
            nroots++;
        }
    for (int t = 0; t < T; ++t)
        if (parent[t] == -1) nroots = -1;
    if (nroots != 1)
        return fail(e, SMC_ERR_ARG, "particle " + std::to_string(p) + ": not a rooted binary tree");
    // heights (tips 0); iterate in an order where children come first
    h.assign(nn, 0);
    order.clear();
    {
        std::vector<int> &st = stv;
        st.assign(1, root);
        while (!st.empty()) {
            int n = st.back();
            st.pop_back();
            order.push_back(n);
            for (int c = 0; c < 2; ++c) {
                int d = ch[2 * (n - T) + c];
                if (d >= T) st.push_back(d);
            }
        }
        if ((int)order.size() != Ti)
            return fail(e, SMC_ERR_ARG, "particle " + std::to_string(p) + ": disconnected tree");
        for (int k = Ti - 1; k >= 0; --k) {
            int n = order[k];
            h[n] = 1 + std::max(h[ch[2 * (n - T)]], h[ch[2 * (n - T) + 1]]);
        }
    }
    // evaluation program for single-site scoring (prog_out null: not wanted now;
    // the caller marks the program stale): post-order that descends into the
    // child with the larger Strahler number first, so at most max-Strahler + 1
    // internal partials are live; checked against kSiteStack
    if (prog_out) {
        std::vector<int> st(nn, 0);
        for (int k = Ti - 1; k >= 0; --k) {
            const int n = order[k];
            const int a = st[ch[2 * (n - T)]], b = st[ch[2 * (n - T) + 1]];
            st[n] = a == b ? a + 1 : std::max(a, b);
        }
        std::vector<int> prog;
        prog.reserve(Ti);
        std::vector<std::pair<int, int>> stk{{root, 0}};
        while (!stk.empty()) {
            auto &[n, state] = stk.back();
            if (state == 0) {
                state = 1;
                int a = ch[2 * (n - T)], b = ch[2 * (n - T) + 1];
                if (st[b] > st[a]) std::swap(a, b);
                // push b first so a (the larger subtree) is evaluated first
                if (b >= T) stk.push_back({b, 0});
                if (a >= T) stk.push_back({a, 0});
            } else {
                prog.push_back(n);
                stk.pop_back();
            }
        }
        int depth = 0, maxd = 0;
        for (int n : prog) {
            depth -= (ch[2 * (n - T)] >= T) + (ch[2 * (n - T) + 1] >= T);
            if (depth < 0) return fail(e, SMC_ERR_ARG, "internal error: scoring program");
            if (n != root) maxd = std::max(maxd, ++depth);
        }
        if ((int)prog.size() != Ti || (maxd > kSiteStack && T > kSiteLevelMaxTaxa))
            return fail(e, SMC_ERR_ARG, "particle " + std::to_string(p) +
                                            ": tree too deep for single-site scoring");
        std::copy(prog.begin(), prog.end(), prog_out);
    }
    // the same nodes by height (counting sort, node order within a level), for the
    // level-parallel scoring kernel: level h holds the nodes whose children all
    // lie below it
    if (prog_out) {
        int H = 0;
        for (int n = T; n < nn; ++n) H = std::max(H, h[n]);
        int *lv = prog_out + Ti, *off = prog_out + 2 * Ti + 1;
        prog_out[2 * Ti] = H;
        std::fill(off, off + H + 1, 0);
        for (int n = T; n < nn; ++n) off[h[n]]++;  // off[h] counts height h (1..H)
        int run = 0;
        for (int l = 0; l <= H; ++l) {  // off[l] becomes the start of height l + 1
            const int c = l + 1 <= H ? off[l + 1] : 0;
            off[l] = run;
            run += c;
        }
        std::vector<int> fill(off, off + H);
        for (int n = T; n < nn; ++n) lv[fill[h[n] - 1]++] = n;
    }
    *root_out = root;
    for (int k = 0; k < Ti; ++k) heights[k] = h[T + k];
    return SMC_OK;
}

// One packed upload of rows (segments) and one scatter launch.
int scatter_upload(smc_engine *e, const std::vector<ScatterSeg> &segs, const std::vector<uint32_t> &data)
{
    if (segs.empty()) return SMC_OK;
    // [segments | words], the words 16-byte aligned
    const size_t seg_bytes = (segs.size() * sizeof(ScatterSeg) + 15) & ~(size_t)15;
    const size_t bytes = seg_bytes + data.size() * sizeof(uint32_t);
    if ((int64_t)bytes > e->scatter_cap) {
        if (e->d_scatter) {  // an earlier scatter may still read it
            int rc = sync_stream(e);
            if (rc != SMC_OK) return rc;
            (void)hipFree(e->d_scatter);
        }
        e->d_scatter = nullptr;
        e->scatter_cap = 0;
        HIP_TRY(e, hipMalloc((void **)&e->d_scatter, bytes * 2));
        e->scatter_cap = (int64_t)bytes * 2;
    }
    std::vector<uint8_t> packed(bytes, 0);
    std::memcpy(packed.data(), segs.data(), segs.size() * sizeof(ScatterSeg));
    std::memcpy(packed.data() + seg_bytes, data.data(), data.size() * sizeof(uint32_t));
    int rc = h2d(e, e->d_scatter, packed.data(), bytes);
    if (rc != SMC_OK) return rc;
    launch_scatter_segments(reinterpret_cast<const uint32_t *>(e->d_scatter + seg_bytes),
                            reinterpret_cast<const ScatterSeg *>(e->d_scatter), (int)segs.size(), e->stream);
    HIP_TRY(e, hipGetLastError());
    return SMC_OK;
}

// Scoring programs of the particles whose topology changed since theirs was
// derived (parameter uploads leave them stale: MCMC rounds do not score sites).
int ensure_programs(smc_engine *e)
{
    if (e->n_prog_stale == 0) return SMC_OK;
    const int T = e->T, Ti = T - 1, ps = site_prog_stride(T);
    std::vector<ScatterSeg> segs;
    std::vector<uint32_t> data;
    std::vector<int> hh(Ti);
    for (int p = 0; p < e->Pmax; ++p) {
        if (!e->h_prog_stale[p]) continue;
        int root = -1;
        int *prog = &e->h_prog[(size_t)p * ps];
        int rc = smc_topology(e, p, &e->h_children[(size_t)p * Ti * 2], &root, hh.data(), prog);
        if (rc != SMC_OK) return rc;
        const size_t o = data.size();
        segs.push_back(ScatterSeg{(uint64_t)(uintptr_t)(e->d_prog + (size_t)p * ps), (uint32_t)o, (uint32_t)ps});
        data.resize(o + ps);
        std::memcpy(data.data() + o, prog, (size_t)ps * sizeof(int));
        e->h_prog_stale[p] = 0;
    }
    e->n_prog_stale = 0;
    return scatter_upload(e, segs, data);
}


int ensure_site_class_buffers(smc_engine *e, int count, int64_t need)
{
    if (need <= e->sc_cap && count <= e->sc_ent_cap) return SMC_OK;
    void *bufs[] = {e->d_sc_ent, e->d_sc_logits, e->d_sc_eig, e->d_sc_rate, e->d_sc_out, e->d_sc_cls};
    for (void *b : bufs)
        if (b) (void)hipFree(b);
    e->d_sc_ent = e->d_sc_cls = nullptr;
    e->d_sc_logits = e->d_sc_eig = e->d_sc_rate = e->d_sc_out = nullptr;
    // grow geometrically: the moves call this with varying batch sizes
    const int64_t cap = std::max<int64_t>(need, e->sc_cap * 3 / 2);
    const int ecap = std::max(count, e->sc_ent_cap * 3 / 2);
    e->sc_cap = 0;
    e->sc_ent_cap = 0;
    int rc = dev_alloc(e, &e->d_sc_ent, (size_t)3 * ecap);
    if (rc == SMC_OK) rc = dev_alloc(e, &e->d_sc_logits, (size_t)cap * 20);
    if (rc == SMC_OK) rc = dev_alloc(e, &e->d_sc_eig, (size_t)cap * kEigStride);
    if (rc == SMC_OK) rc = dev_alloc(e, &e->d_sc_rate, (size_t)cap * e->R);
    if (rc == SMC_OK) rc = dev_alloc(e, &e->d_sc_out, (size_t)cap);
    if (rc == SMC_OK) rc = dev_alloc(e, &e->d_sc_cls, (size_t)cap);
    if (rc != SMC_OK) return rc;
    e->sc_cap = cap;
    e->sc_ent_cap = ecap;
    return SMC_OK;
}

int issue_site_class(smc_engine *e, int count, int n_cand, bool with_cls, bool decompose, double tempering,
                     double *out)
{
    {
        int rc = ensure_programs(e);  // the kernels walk the trees' scoring programs
        if (rc != SMC_OK) return rc;
    }
    hipStream_t s = e->stream;
    if (decompose) {  // candidates coded -1 (or every candidate without codes)
        TimedLaunch tl(e, SMC_K_EIGEN);
        launch_eigen_cand(e->d_base, e->d_sc_ent, e->d_sc_ent + 2 * count, e->d_sc_logits, count, n_cand,
                          e->d_sc_eig, with_cls ? e->d_sc_cls : nullptr, e->d_eig, e->Kmax, e->pool.d_eig, s,
                          false);
    }
    {
        TimedLaunch tl(e, SMC_K_SITE);
        SiteClassArgs a{};
        a.ent_particle = e->d_sc_ent;
        a.ent_site = e->d_sc_ent + count;
        a.ent_ncand = e->d_sc_ent + 2 * count;
        a.n_cand = n_cand;
        a.cand_eig = e->d_sc_eig;
        a.cand_cls = with_cls ? e->d_sc_cls : nullptr;  // existing classes and pool profiles read in place
        a.eig_cache = e->d_eig;
        a.max_classes = e->Kmax;
        a.pool = e->pool.d_eig;
        a.children = e->d_children;
        a.prog = e->d_prog;
        a.codes = e->d_codes;
        a.blen = e->d_blen;
        a.rates = e->d_rates;
        a.pinv = e->d_pinv;
        a.isinv = e->d_isinv;
        a.invchar = e->d_invchar;
        a.rate_lnl = e->d_sc_rate;
        a.out = out ? out : e->d_sc_out;
        a.count = count;
        a.T = e->T;
        a.R = e->R;
        a.S_pad = e->S_pad;
        a.mix_mode = e->mix_mode;
        a.tempering = tempering;
        launch_site_class(a, s);
    }
    HIP_TRY(e, hipGetLastError());
    return SMC_OK;
}

extern "C" {

int smc_engine_create(const smc_engine_config *cfg, smc_engine **out)
{
    if (!cfg || !out) return SMC_ERR_ARG;
    *out = nullptr;
    if (cfg->taxa < 3 || cfg->sites < 1 || cfg->num_rates < 1 || cfg->num_rates > 64 ||
        cfg->max_particles < 1 || cfg->max_classes < 1 || cfg->max_classes > 4096 ||
        (cfg->precision != SMC_PREC_FP64 && cfg->precision != SMC_PREC_FP32) ||
        cfg->partial_slots < 0 || cfg->pmat_classes < 0 || cfg->workspaces < 0 ||
        cfg->workspaces > 2 || cfg->reserved != 0)
        return SMC_ERR_ARG;
    smc_engine *e = new smc_engine();
    e->cfg = *cfg;
    e->T = cfg->taxa;
    e->S = cfg->sites;
    e->S_pad = (cfg->sites + kTile - 1) / kTile * kTile;
    e->R = cfg->num_rates;
    e->nn = 2 * cfg->taxa - 1;
    e->Pmax = cfg->max_particles;
    e->Kmax = cfg->max_classes;
    e->WS = cfg->workspaces == 2 ? 2 : 1;
    e->Wmax = cfg->partial_slots > 0 ? std::min(cfg->partial_slots, cfg->max_particles)
                                     : cfg->max_particles;
    e->Cmax = cfg->pmat_classes > 0 ? cfg->pmat_classes : (int64_t)e->Wmax * cfg->max_classes;
    if (e->Cmax < cfg->max_classes) e->Cmax = cfg->max_classes;
    e->tile_stride = (cfg->sites + kTile - 1) / kTile + cfg->max_classes;
    e->fp32 = cfg->precision == SMC_PREC_FP32;
    if (e->fp32) {  // measured defaults (the autotuner refines them for a workload)
        e->tune.unroll[0] = e->tune.unroll[1] = e->tune.unroll[2] = 2;
    } else {
        e->tune.unroll[0] = e->tune.unroll[1] = 1;
        e->tune.unroll[2] = 2;
    }
    e->tx_size = e->fp32 ? 4 : 8;
    if (e->tile_stride > 65535) {
        int rc = fail(e, SMC_ERR_ARG, "too many tiles per particle");
        smc_engine_destroy(e);
        return rc;
    }
    hipError_t st = hipSetDevice(cfg->device);
    if (st == hipSuccess) st = hipStreamCreateWithFlags(&e->stream, hipStreamNonBlocking);
    if (st == hipSuccess) st = hipStreamCreateWithFlags(&e->stream2, hipStreamNonBlocking);
    if (st == hipSuccess) st = hipEventCreateWithFlags(&e->ev_fork, hipEventDisableTiming);
    if (st == hipSuccess) st = hipEventCreateWithFlags(&e->ev_join, hipEventDisableTiming);
    if (st != hipSuccess) {
        if (e->stream) (void)hipStreamDestroy(e->stream);
        if (e->stream2) (void)hipStreamDestroy(e->stream2);
        if (e->ev_fork) (void)hipEventDestroy(e->ev_fork);
        delete e;
        return SMC_ERR_HIP;
    }
    {
        int ncu = 0;
        if (hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, cfg->device) == hipSuccess &&
            ncu > 0)
            e->num_cus = ncu;
    }
    const size_t P = e->Pmax, T = e->T, Sp = e->S_pad, R = e->R, K = e->Kmax, nn = e->nn;
    int rc = SMC_OK;
    auto A = [&](auto **ptr, size_t n) {
        if (rc == SMC_OK) rc = dev_alloc(e, ptr, n);
    };
    A(&e->d_codes, T * Sp);
    A(&e->d_pcodes, P * T * Sp);
    A(&e->d_pweight, P * Sp);
    A(&e->d_isinv, Sp);
    A(&e->d_invchar, Sp);
    A(&e->d_children, P * (T - 1) * 2);
    A(&e->d_blen, P * nn);
    A(&e->d_base, P * 190);
    A(&e->d_logits, P * K * 20);
    A(&e->d_nclass, P);
    A(&e->d_pinv, P);
    A(&e->d_rates, P * R);
    A(&e->d_root, P);
    A(&e->d_perm, P * Sp);
    A(&e->d_tiles, P * e->tile_stride);
    A(&e->d_ntiles, P);
    A(&e->d_eig, P * K * kEigStride);
    A(&e->d_tile_lnl, P * e->tile_stride);
    A(&e->d_lnl, P);
    A(&e->d_nonfinite, P);
    A(&e->d_items, P * (T - 1));
    A(&e->d_chain_nodes, P * (T - 1));
    A(&e->d_chain_len, P * (T - 1));
    A(&e->d_walk, P * T * 2);
    A(&e->d_witems, P * (T + 1));  // walk pieces (at most one per node) and the roots
    A(&e->d_pcls_off, P);
    A(&e->d_prog, P * site_prog_stride(T));
    A(&e->d_plist, 4 * P);      // two model jobs' lists (see upload_models)
    A(&e->d_pnodes, 2 * P * nn);  // and their node lists
    const size_t W = e->Wmax, C = (size_t)e->Cmax;
    const size_t WS = e->WS;
    A(&e->d_E, W * WS * (T - 1) * R * Sp);
    A(&e->d_xflag, P * (T - 1));
    if (e->fp32) {
        A((float **)&e->d_P, C * nn * R * kPStride);
        A((float **)&e->d_X, W * WS * (T - 1) * R * 20 * Sp);
    } else {
        A((double **)&e->d_P, C * nn * R * kPStride);
        A((double **)&e->d_X, W * WS * (T - 1) * R * 20 * Sp);
    }
    if (rc != SMC_OK) {
        std::string msg = e->err;
        smc_engine_destroy(e);
        fprintf(stderr, "smc_engine_create: %s\n", msg.c_str());
        return rc;
    }
    rc = smc_noalloc_init(e);  // this device's address of the no-allocation scratch
    if (rc != SMC_OK) {
        std::string msg = e->err;
        smc_engine_destroy(e);
        fprintf(stderr, "smc_engine_create: %s\n", msg.c_str());
        return rc;
    }
    (void)hipMemsetAsync(e->d_ntiles, 0, P * sizeof(int), e->stream);
    (void)hipMemsetAsync(e->d_nclass, 0, P * sizeof(int), e->stream);
    (void)hipMemsetAsync(e->d_xflag, 0, P * (T - 1), e->stream);
    (void)hipStreamSynchronize(e->stream);
    {
        constexpr size_t kStageBytes = size_t(16) << 20;
        void *hs = nullptr;
        if (hipHostMalloc(&hs, kStageBytes, hipHostMallocDefault) != hipSuccess) {
            smc_engine_destroy(e);
            fprintf(stderr, "smc_engine_create: pinned staging buffer of %zu bytes failed\n", kStageBytes);
            return SMC_ERR_HIP;
        }
        e->h_stage = (uint8_t *)hs;
        e->stage_cap = kStageBytes;
        constexpr size_t kReadBytes = size_t(4) << 20;
        void *hr = nullptr;
        if (hipHostMalloc(&hr, kReadBytes, hipHostMallocDefault) != hipSuccess) {
            smc_engine_destroy(e);
            fprintf(stderr, "smc_engine_create: pinned readback buffer of %zu bytes failed\n", kReadBytes);
            return SMC_ERR_HIP;
        }
        e->h_read = (uint8_t *)hr;
        e->read_cap = kReadBytes;
    }
    e->h_children.assign(P * (T - 1) * 2, 0);
    e->h_root.assign(P, -1);
    e->h_height.assign(P * (T - 1), 0);
    e->h_prog.assign(P * site_prog_stride(T), 0);
    e->h_valid.assign(P, 0);
    e->h_nclass.assign(P, 1);
    e->h_blen.assign(P * nn, 0.0);
    e->h_base.assign(P * 190, 0.0);
    e->h_logits.assign(P * K * 20, 0.0);
    e->h_assign.assign(P * (size_t)e->S, 0);
    e->h_pinv.assign(P, 0.0);
    e->h_shape.assign(P, 1.0);
    e->h_rates.assign((size_t)P * e->R, 1.0);
    e->h_xflag.assign(P * (T - 1), 0);
    e->h_xflag_old.assign(P * (T - 1), 0);
    e->h_lnl.assign(P, 0.0);
    e->h_lnl_old.assign(P, 0.0);
    e->h_dirty.assign(P, 0);
    e->h_eig_save.assign(P, -1);
    e->h_prog_stale.assign(P, 0);
    e->h_touched.assign(P, 0);
    e->h_busy.assign(P, 0);
    e->lane.lnl.assign(P, 0.0);
    e->h_pnodes.assign(P, {});
    e->committed.resize(P);
    e->h_owner.assign(e->Wmax, -1);
    e->h_nslots.assign(P, 0);
    *out = e;
    return SMC_OK;
}

int smc_engine_destroy(smc_engine *e)
{
    if (!e) return SMC_ERR_ARG;
    void *bufs[] = {e->d_codes,  e->d_pcodes, e->d_pweight, e->d_isinv, e->d_invchar, e->d_children, e->d_blen,
                    e->d_base,   e->d_logits, e->d_nclass, e->d_pinv,     e->d_rates,
                    e->d_root,   e->d_perm,  e->d_tiles,   e->d_ntiles,   e->d_eig,
                    e->d_P,      e->d_X,     e->d_E,       e->d_tile_lnl, e->d_lnl,
                    e->d_nonfinite, e->d_items, e->d_w,    e->d_l,        e->d_norm,
                    e->d_out2,   e->d_assign, e->d_pcls_off, e->d_prog, e->d_sc_ent, e->d_xflag, e->d_plist,
                    e->d_sc_logits, e->d_sc_eig, e->d_sc_rate, e->d_sc_out, e->d_sc_cls,
                    e->d_chain_nodes, e->d_chain_len, e->d_walk, e->d_witems, e->d_walk_rdot, e->d_walk_rexp,
                    e->d_eig_save};
    if (e->stream) (void)hipStreamSynchronize(e->stream);
    if (e->stream2) (void)hipStreamSynchronize(e->stream2);
    if (e->lane.stream) (void)hipStreamSynchronize(e->lane.stream);
    if (e->h_stage) (void)hipHostFree(e->h_stage);
    if (e->h_read) (void)hipHostFree(e->h_read);
    {
        smc_engine::Lane &l = e->lane;
        if (l.h_stage) (void)hipHostFree(l.h_stage);
        if (l.h_read) (void)hipHostFree(l.h_read);
        for (void *b : {(void *)l.d_lnl, (void *)l.d_nonfinite, (void *)l.d_items, (void *)l.d_chain_nodes,
                        (void *)l.d_chain_len, (void *)l.d_plist, (void *)l.d_pnodes})
            if (b) (void)hipFree(b);
        if (l.stream) (void)hipStreamDestroy(l.stream);
        if (l.fork) (void)hipEventDestroy(l.fork);
        if (l.done) (void)hipEventDestroy(l.done);
    }
    for (smc_engine::CachedGraph &c : e->graphs) (void)hipGraphExecDestroy(c.exec);
    if (e->d_bucket_off) (void)hipFree(e->d_bucket_off);
    if (e->d_pnodes) (void)hipFree(e->d_pnodes);
    if (e->d_gather) (void)hipFree(e->d_gather);
    if (e->d_scatter) (void)hipFree(e->d_scatter);
    if (e->d_split_ws) (void)hipFree(e->d_split_ws);
    for (void *b : {(void *)e->pool.d_eig, (void *)e->pool.d_logits, (void *)e->pool.d_ent})
        if (b) (void)hipFree(b);
    for (void *b : bufs)
        if (b) (void)hipFree(b);
    for (auto &pt : e->pending) {
        e->ev_pool.push_back(pt.a);
        e->ev_pool.push_back(pt.b);
    }
    for (hipEvent_t ev : e->ev_pool) (void)hipEventDestroy(ev);
    if (e->stream) (void)hipStreamDestroy(e->stream);
    if (e->stream2) (void)hipStreamDestroy(e->stream2);
    if (e->ev_fork) (void)hipEventDestroy(e->ev_fork);
    if (e->ev_join) (void)hipEventDestroy(e->ev_join);
    delete e;
    return SMC_OK;
}

const char *smc_engine_last_error(const smc_engine *e) { return e ? e->err.c_str() : "null engine"; }

int64_t smc_engine_device_bytes(const smc_engine *e) { return e ? e->device_bytes : 0; }

int smc_set_alignment(smc_engine *e, const int8_t *codes)
{
    if (!e) return SMC_ERR_ARG;
    if (!codes) return fail(e, SMC_ERR_ARG, "codes is null");
    const int T = e->T, S = e->S, Sp = e->S_pad;
    std::vector<uint8_t> hc((size_t)T * Sp, 22);
    for (int t = 0; t < T; ++t)
        for (int s = 0; s < S; ++s) {
            int c = codes[(size_t)t * S + s];
            if (c < 0 || c > 22)
                return fail(e, SMC_ERR_ARG, "alignment code out of range 0..22 at taxon " +
                                                std::to_string(t) + " site " + std::to_string(s));
            hc[(size_t)t * Sp + s] = (uint8_t)c;
        }
    // invariant detection, Particle.cpp:44-72 (exact code equality, gaps included: Q5)
    std::vector<double> inv(Sp, 0.0);
    std::vector<int> ich(Sp, 0);
    for (int s = 0; s < S; ++s) {
        int c1 = hc[s];
        bool same = true;
        for (int t = 1; t < T && same; ++t) same = hc[(size_t)t * Sp + s] == c1;
        inv[s] = same ? 1.0 : 0.0;
        ich[s] = (same && c1 < 20) ? c1 : 0;
    }
    // site patterns: identical columns (raw codes, so the invariant terms agree too)
    // share one id
    {
        std::vector<int> order(S);
        for (int s = 0; s < S; ++s) order[s] = s;
        auto col_less = [&](int a, int b) {
            for (int t = 0; t < T; ++t) {
                const uint8_t x = hc[(size_t)t * Sp + a], y = hc[(size_t)t * Sp + b];
                if (x != y) return x < y;
            }
            return false;
        };
        std::stable_sort(order.begin(), order.end(), col_less);
        e->h_pattern.assign(S, 0);
        int id = -1;
        for (int k = 0; k < S; ++k) {
            if (k == 0 || col_less(order[k - 1], order[k])) ++id;
            e->h_pattern[order[k]] = id;
        }
        e->n_patterns = id + 1;
    }
    // amino-acid frequencies of each column (codes >= 20 carry no state)
    e->h_sitefreq.assign((size_t)S * 20, 0.f);
    for (int s = 0; s < S; ++s) {
        int obs = 0;
        for (int t = 0; t < T; ++t) obs += hc[(size_t)t * Sp + s] < 20;
        for (int t = 0; t < T; ++t) {
            const int c = hc[(size_t)t * Sp + s];
            if (c < 20) e->h_sitefreq[(size_t)s * 20 + c] += 1.f / (float)obs;
        }
    }
    HIP_TRY(e, hipMemcpyAsync(e->d_codes, hc.data(), hc.size(), hipMemcpyHostToDevice, e->stream));
    HIP_TRY(e, hipMemcpyAsync(e->d_isinv, inv.data(), Sp * sizeof(double), hipMemcpyHostToDevice,
                              e->stream));
    HIP_TRY(e, hipMemcpyAsync(e->d_invchar, ich.data(), Sp * sizeof(int), hipMemcpyHostToDevice,
                              e->stream));
    HIP_TRY(e, hipStreamSynchronize(e->stream));
    e->have_alignment = true;
    // particles already uploaded: their slots (patterns) and tip codes follow the new
    // alignment (runs of valid slots, re-uploaded from the host copies)
    const int Ti = T - 1, nn = e->nn, K = e->Kmax, R = e->R;
    (void)R;
    for (int p = 0; p < e->Pmax;) {
        if (!e->h_valid[p]) {
            ++p;
            continue;
        }
        int q = p;
        while (q < e->Pmax && e->h_valid[q]) ++q;
        const size_t c = (size_t)(q - p);
        std::vector<int32_t> ch(e->h_children.begin() + (size_t)p * Ti * 2, e->h_children.begin() + (size_t)q * Ti * 2);
        std::vector<double> bl(e->h_blen.begin() + (size_t)p * nn, e->h_blen.begin() + (size_t)q * nn);
        std::vector<double> base(e->h_base.begin() + (size_t)p * 190, e->h_base.begin() + (size_t)q * 190);
        std::vector<int32_t> nk(e->h_nclass.begin() + p, e->h_nclass.begin() + q);
        std::vector<double> lg(e->h_logits.begin() + (size_t)p * K * 20, e->h_logits.begin() + (size_t)q * K * 20);
        std::vector<int32_t> asg(e->h_assign.begin() + (size_t)p * S, e->h_assign.begin() + (size_t)q * S);
        std::vector<double> pv(e->h_pinv.begin() + p, e->h_pinv.begin() + q);
        std::vector<double> sh(e->h_shape.begin() + p, e->h_shape.begin() + q);
        (void)c;
        const int rc = smc_set_particles(e, p, q - p, ch.data(), bl.data(), base.data(), nk.data(), lg.data(),
                                         asg.data(), pv.data(), sh.data());
        if (rc != SMC_OK) return rc;
        p = q;
    }
    return SMC_OK;
}

int smc_set_particles(smc_engine *e, int first, int count, const int32_t *children,
                      const double *branch_len, const double *base, const int32_t *num_classes,
                      const double *logits, const int32_t *assignment, const double *pinv,
                      const double *shape)
{
    if (!e) return SMC_ERR_ARG;
    PhaseTimer phase_timer(e, SMC_PHASE_SET_PARTICLES);
    if (first < 0 || count < 1 || first + count > e->Pmax)
        return fail(e, SMC_ERR_ARG, "particle range out of bounds");
    if (!children || !branch_len || !base || !num_classes || !logits || !assignment || !pinv ||
        !shape)
        return fail(e, SMC_ERR_ARG, "null particle array");
    const int T = e->T, Ti = T - 1, nn = e->nn, S = e->S, Sp = e->S_pad, R = e->R, K = e->Kmax;
    const int ts = e->tile_stride;
    for (int &o : e->h_owner)  // new states: resident partials of these slots are stale
        if (o >= first && o < first + count) o = -1;
    std::vector<double> blen((size_t)count * nn), rates((size_t)count * R), weight((size_t)count * Sp, 0.0);
    std::vector<int> perm((size_t)count * Sp, 0), ntiles(count), roots(count);
    // a column's pattern: its id (uploaded alignment), else the column itself
    auto pattern_of = [&](int s) { return e->h_pattern.empty() ? s : e->h_pattern[s]; };
    std::vector<Tile> tiles((size_t)count * ts, Tile{0, 0, 0, 0});
    std::vector<int> progs((size_t)count * site_prog_stride(T)), hts((size_t)count * Ti);
    // topologies and rates first, in order: their checks report through the engine's
    // error text, and the rates call lgamma, which sets a process-wide sign
    for (int i = 0; i < count; ++i) {
        const int p = first + i;
        const int32_t *ch = children + (size_t)i * Ti * 2;
        int rc = smc_topology(e, p, ch, &roots[i], hts.data() + (size_t)i * Ti,
                              progs.data() + (size_t)i * site_prog_stride(T));
        if (rc != SMC_OK) return rc;
        if (e->h_prog_stale[p]) {
            e->h_prog_stale[p] = 0;
            e->n_prog_stale--;
        }
        // rates: discretizeGamma when K_r > 1 (Particle.cpp:224-225), else 1.0
        if (R > 1) {
            if (!(shape[i] > 0.0))
                return fail(e, SMC_ERR_ARG, "particle " + std::to_string(p) + ": shape must be > 0");
            discretize_gamma(shape[i], R, &rates[(size_t)i * R]);
        } else {
            rates[(size_t)i] = 1.0;
        }
    }
    // the rest of a particle's setup touches only its own rows: particles split over
    // host threads, and an error names the lowest failing particle, as in order
    std::atomic<int> next{0};
    std::mutex err_mu;
    int err_i = count, err_code = SMC_OK;
    std::string err_msg;
    auto worker = [&]() {
        // pattern stamps are keyed by (particle, class), so each thread keeps its own
        std::vector<int> stamp(std::max(e->n_patterns, S), -1), pslot(std::max(e->n_patterns, S), 0);
        for (int i = next.fetch_add(1); i < count; i = next.fetch_add(1)) {
            const int p = first + i;
            const int32_t *ch = children + (size_t)i * Ti * 2;
            auto one = [&]() -> std::pair<int, std::string> {
                auto bad = [](int code, std::string msg) { return std::make_pair(code, std::move(msg)); };
                for (int k = 0; k < Ti; ++k) {
                    e->h_height[(size_t)p * Ti + k] = hts[(size_t)i * Ti + k];
                    e->h_children[((size_t)p * Ti + k) * 2] = ch[2 * k];
                    e->h_children[((size_t)p * Ti + k) * 2 + 1] = ch[2 * k + 1];
                }
                std::copy_n(progs.data() + (size_t)i * site_prog_stride(T), site_prog_stride(T),
                            e->h_prog.data() + (size_t)p * site_prog_stride(T));
                const int root = roots[i];
                e->h_root[p] = root;
                for (int n = 0; n < nn; ++n) {
                    double bl = branch_len[(size_t)i * nn + n];
                    blen[(size_t)i * nn + n] = (n == root) ? 0.0 : bl;
                }
                std::copy_n(&rates[(size_t)i * R], R, &e->h_rates[(size_t)p * R]);
                // class-sorted slot order and tiles (each tile is one class)
                const int Kp = num_classes[i];
                if (Kp < 1 || Kp > K)
                    return bad(SMC_ERR_ARG, "particle " + std::to_string(p) + ": num_classes out of range");
                const int32_t *as = assignment + (size_t)i * S;
                std::vector<int> cnt(Kp + 1, 0);
                for (int s = 0; s < S; ++s) {
                    if (as[s] < 0 || as[s] >= Kp)
                        return bad(SMC_ERR_ARG, "particle " + std::to_string(p) + ": assignment out of range");
                    cnt[as[s] + 1]++;
                }
                for (int k = 0; k < Kp; ++k) cnt[k + 1] += cnt[k];
                // class-sorted slots: one per column, or (compressed) one per distinct pattern of
                // a class, weighted by its columns; the slot holds its first column
                std::vector<int> fillp(cnt.begin(), cnt.end() - 1), bycls(S);
                for (int s = 0; s < S; ++s) bycls[fillp[as[s]]++] = s;  // sites grouped by class, in site order
                std::vector<int> used(Kp, 0);
                for (int k = 0; k < Kp; ++k)  // one class at a time: its patterns' stamps stay its own
                    for (int j = cnt[k]; j < cnt[k + 1]; ++j) {
                        const int s = bycls[j];
                        const int pid = pattern_of(s);
                        if (e->compress && stamp[pid] == i * K + k) {
                            weight[(size_t)i * Sp + pslot[pid]] += 1.0;
                            continue;
                        }
                        stamp[pid] = i * K + k;
                        const int sl = cnt[k] + used[k]++;
                        pslot[pid] = sl;
                        perm[(size_t)i * Sp + sl] = s;
                        weight[(size_t)i * Sp + sl] = 1.0;
                    }
                // slots of class k: [cnt[k], cnt[k] + used[k]); pack them to the front
                {
                    std::vector<int> pk(Sp, 0);
                    std::vector<double> pw(Sp, 0.0);
                    int o = 0;
                    for (int k = 0; k < Kp; ++k) {
                        const int c0 = cnt[k];
                        cnt[k] = o;
                        for (int j = 0; j < used[k]; ++j, ++o) {
                            pk[o] = perm[(size_t)i * Sp + c0 + j];
                            pw[o] = weight[(size_t)i * Sp + c0 + j];
                        }
                    }
                    cnt[Kp] = o;
                    std::copy(pk.begin(), pk.end(), perm.begin() + (size_t)i * Sp);
                    std::copy(pw.begin(), pw.end(), weight.begin() + (size_t)i * Sp);
                    e->h_nslots[p] = o;
                }
                // tiles: each class's slots cut at the 64-slot window boundaries (at most
                // ceil(S / 64) windows plus one split window per class boundary)
                int nt = 0;
                for (int k = 0; k < Kp; ++k)
                    for (int w = cnt[k] / kTile * kTile; w < cnt[k + 1]; w += kTile) {
                        const int lo = std::max(cnt[k], w) - w, hi = std::min(cnt[k + 1], w + kTile) - w;
                        if (nt >= ts) return bad(SMC_ERR_STATE, "tile table overflow");
                        tiles[(size_t)i * ts + nt++] = Tile{w, hi, k, lo};
                    }
                ntiles[i] = nt;
                e->h_valid[p] = 1;
                e->h_nclass[p] = Kp;
                std::copy(blen.begin() + (size_t)i * nn, blen.begin() + (size_t)(i + 1) * nn,
                          e->h_blen.begin() + (size_t)p * nn);
                std::copy(base + (size_t)i * 190, base + (size_t)(i + 1) * 190,
                          e->h_base.begin() + (size_t)p * 190);
                std::copy(logits + (size_t)i * K * 20, logits + (size_t)(i + 1) * K * 20,
                          e->h_logits.begin() + (size_t)p * K * 20);
                std::copy(as, as + S, e->h_assign.begin() + (size_t)p * S);
                e->h_pinv[p] = pinv[i];
                e->h_shape[p] = shape[i];
                return {SMC_OK, std::string()};
            };
            std::pair<int, std::string> r = one();
            if (r.first != SMC_OK) {
                std::lock_guard<std::mutex> lk(err_mu);
                if (i < err_i) {
                    err_i = i;
                    err_code = r.first;
                    err_msg = std::move(r.second);
                }
            }
        }
    };
    const int nthr = count >= 64 ? (int)std::max(1u, std::min(16u, std::thread::hardware_concurrency())) : 1;
    if (nthr > 1) {
        std::vector<std::thread> pool;
        for (int t = 0; t < nthr; ++t) pool.emplace_back(worker);
        for (std::thread &t : pool) t.join();
    } else {
        worker();
    }
    if (err_code != SMC_OK) return fail(e, err_code, err_msg);
    std::vector<int> hch((size_t)count * Ti * 2);
    std::memcpy(hch.data(), children, hch.size() * sizeof(int));
    hipStream_t s = e->stream;
    const size_t P0 = first;
    HIP_TRY(e, hipMemcpyAsync(e->d_children + P0 * Ti * 2, hch.data(), hch.size() * sizeof(int),
                              hipMemcpyHostToDevice, s));
    HIP_TRY(e, hipMemcpyAsync(e->d_blen + P0 * nn, blen.data(), blen.size() * sizeof(double),
                              hipMemcpyHostToDevice, s));
    HIP_TRY(e, hipMemcpyAsync(e->d_base + P0 * 190, base, (size_t)count * 190 * sizeof(double),
                              hipMemcpyHostToDevice, s));
    HIP_TRY(e, hipMemcpyAsync(e->d_logits + P0 * K * 20, logits,
                              (size_t)count * K * 20 * sizeof(double), hipMemcpyHostToDevice, s));
    HIP_TRY(e, hipMemcpyAsync(e->d_nclass + P0, num_classes, (size_t)count * sizeof(int),
                              hipMemcpyHostToDevice, s));
    HIP_TRY(e, hipMemcpyAsync(e->d_pinv + P0, pinv, (size_t)count * sizeof(double),
                              hipMemcpyHostToDevice, s));
    HIP_TRY(e, hipMemcpyAsync(e->d_rates + P0 * R, rates.data(), rates.size() * sizeof(double),
                              hipMemcpyHostToDevice, s));
    HIP_TRY(e, hipMemcpyAsync(e->d_root + P0, roots.data(), roots.size() * sizeof(int),
                              hipMemcpyHostToDevice, s));
    HIP_TRY(e, hipMemcpyAsync(e->d_perm + P0 * Sp, perm.data(), perm.size() * sizeof(int),
                              hipMemcpyHostToDevice, s));
    HIP_TRY(e, hipMemcpyAsync(e->d_pweight + P0 * Sp, weight.data(), weight.size() * sizeof(double),
                              hipMemcpyHostToDevice, s));
    HIP_TRY(e, hipMemcpyAsync(e->d_tiles + P0 * ts, tiles.data(), tiles.size() * sizeof(Tile),
                              hipMemcpyHostToDevice, s));
    HIP_TRY(e, hipMemcpyAsync(e->d_ntiles + P0, ntiles.data(), ntiles.size() * sizeof(int),
                              hipMemcpyHostToDevice, s));
    HIP_TRY(e, hipMemcpyAsync(e->d_prog + P0 * site_prog_stride(T), progs.data(), progs.size() * sizeof(int),
                              hipMemcpyHostToDevice, s));
    if (e->have_alignment) launch_permute_codes(e->d_codes, e->d_perm, e->d_pcodes, T, Sp, first, count, s);
    HIP_TRY(e, hipGetLastError());
    HIP_TRY(e, hipStreamSynchronize(s));
    drop_pending(e, first, count);  // new states: nothing of the old ones is restored
    for (int p = first; p < first + count; ++p)
        if (e->h_eig_save[p] >= 0) {
            e->eig_save_free.push_back(e->h_eig_save[p]);
            e->h_eig_save[p] = -1;
        }
    e->class_version++;
    e->topo_version++;
    return SMC_OK;
}

namespace {

PruneArgs prune_args(smc_engine *e, int p_base)
{
    PruneArgs a{};
    a.tiles = e->d_tiles;
    a.ntiles = e->d_ntiles;
    a.tile_stride = e->tile_stride;
    a.children = e->d_children;
    a.root_of = e->d_root;
    a.perm = e->d_perm;
    a.codes = e->d_codes;
    a.pcodes = e->d_pcodes;
    a.pweight = e->d_pweight;
    a.P = e->d_P;
    a.pcls_off = e->d_pcls_off;
    a.eig = e->d_eig;
    a.pinv = e->d_pinv;
    a.isinv = e->d_isinv;
    a.invchar = e->d_invchar;
    a.X = e->d_X;
    a.E = e->d_E;
    a.p_base = p_base;
    a.tile_lnl = e->d_tile_lnl;
    a.T = e->T;
    a.R = e->R;
    a.S_pad = e->S_pad;
    a.max_classes = e->Kmax;
    a.fp32 = e->fp32;
    a.mix_mode = e->mix_mode;
    a.xflag = e->WS == 2 ? e->d_xflag : nullptr;
    a.WS = e->WS;
    a.grid = e->tune.blocks_per_cu > 0 ? e->num_cus * e->tune.blocks_per_cu : 0;
    a.xcds = e->tune.xcd_groups;
    return a;
}

// The launches of a full sweep over the scheduled waves.  `device_buckets`: the
// pruning launches read their item ranges from d_bucket_off (persistent grid, no
// host-known sizes), as a captured graph needs; otherwise each launch covers its
// bucket exactly.
void issue_sweep(smc_engine *e, int first, int count, bool device_buckets)
{
    hipStream_t s = e->stream;
    const bool mfma = e->prune_variant != SMC_PRUNE_VALU;
    // With a second P buffer, wave w + 1's eigensystems and P are computed on stream2
    // into the buffer wave w - 1 used while wave w prunes (4% faster at C4): the
    // stream2 work of a buffer waits for the pruning that read it, and a wave's pruning
    // for its models.  (Error returns of the event calls are left to the launches'
    // hipGetLastError: issue_sweep also runs under graph capture.)
    const bool pipe = e->d_P2 && e->waves.size() > 1 && !e->timing;
    auto models = [&](const smc_engine::Wave &w, void *P, hipStream_t st) {
        {
            TimedLaunch tl(e, SMC_K_EIGEN);
            launch_eigen(EigenArgs{e->d_base, e->d_logits, e->d_nclass, e->d_eig, e->Kmax, w.first,
                                   w.count},
                         st);
        }
        {
            TimedLaunch tl(e, SMC_K_PMAT);
            launch_pmat(PmatArgs{e->d_eig, e->d_blen, e->d_rates, e->d_nclass, e->d_pcls_off,
                                 P, e->nn, e->R, e->Kmax, w.first, w.count, e->fp32},
                        st);
        }
    };
    // buffers alternate so that the last wave's P ends in d_P, where the incremental
    // refreshes of its resident particles read it
    const size_t nw = e->waves.size();
    auto bix = [&](size_t wi) { return pipe ? (int)((nw - 1 - wi) & 1) : 0; };
    auto buffer = [&](size_t wi) { return bix(wi) ? e->d_P2 : e->d_P; };
    if (pipe) {
        (void)hipEventRecord(e->ev_fork, s);
        (void)hipStreamWaitEvent(e->stream2, e->ev_fork, 0);
        models(e->waves[0], buffer(0), e->stream2);
        (void)hipEventRecord(e->pipe_ev[bix(0)], e->stream2);
    }
    for (size_t wi = 0; wi < nw; ++wi) {
        const smc_engine::Wave &w = e->waves[wi];
        if (pipe) {
            (void)hipStreamWaitEvent(s, e->pipe_ev[bix(wi)], 0);
            if (wi + 1 < nw) {
                if (wi >= 1) (void)hipStreamWaitEvent(e->stream2, e->pipe_ev[2 + bix(wi + 1)], 0);
                models(e->waves[wi + 1], buffer(wi + 1), e->stream2);
                (void)hipEventRecord(e->pipe_ev[bix(wi + 1)], e->stream2);
            }
        } else {
            models(w, e->d_P, s);
        }
        PruneArgs a = prune_args(e, w.first);
        a.P = buffer(wi);
        a.single_wave = true;  // full sweeps: one tile per workgroup
        const std::vector<int> &bo = w.bucket_off;
        const int nb = PRUNE_KINDS * (w.graph_levels + 1);
        auto run = [&](int b, int kind, bool roots) {
            const int off = bo[b], n = bo[b + 1] - bo[b];
            if (!device_buckets && n <= 0) return;
            TimedLaunch tl(e, SMC_K_PRUNE);
            a.kind = kind;
            a.roots = roots;
            a.unroll = roots ? e->tune.unroll_rt : e->tune.unroll[prune_stored(kind)];
            if (device_buckets) {
                a.items = e->d_items;
                a.bucket_off = e->d_bucket_off + w.boff;
                a.bucket = b;
            } else {
                a.items = e->d_items + off;
                a.item_base = 0;
                a.n_items = n;
                a.bucket_off = nullptr;
            }
            if (mfma)
                launch_prune_mfma(a, s);
            else
                launch_prune(a, s);
        };
        const int levels = device_buckets ? w.graph_levels : w.max_level;
        for (int L = 1; L <= levels; ++L) {
            if (!mfma) {
                a.items = e->d_items + bo[PRUNE_KINDS * L];
                a.n_items = bo[PRUNE_KINDS * (L + 1)] - bo[PRUNE_KINDS * L];
                if (a.n_items > 0) {
                    TimedLaunch tl(e, SMC_K_PRUNE);
                    a.kind = PRUNE_ANY;
                    a.roots = false;
                    launch_prune(a, s);
                }
                continue;
            }
            // cherries (tip-tip, all of height 1) are built by their parents; every
            // node above height 1 has an internal child
            if (L >= 2)
                for (int k = PRUNE_TC; k < PRUNE_KINDS; ++k) run(PRUNE_KINDS * L + k, k, false);
        }
        if (!mfma) {
            a.items = e->d_items + w.root_off;
            a.n_items = w.count;
            a.roots = true;
            a.kind = PRUNE_ANY;
            TimedLaunch tl(e, SMC_K_PRUNE);
            launch_prune(a, s);
        } else {
            run(nb, PRUNE_ANY, true);
        }
        if (pipe) (void)hipEventRecord(e->pipe_ev[2 + bix(wi)], s);
    }
    TimedLaunch tl(e, SMC_K_REDUCE);
    launch_reduce_lnl(e->d_tile_lnl, e->d_ntiles, e->tile_stride, first, count, e->d_lnl, e->d_nonfinite,
                      s);
}

// One full sweep: replayed from a captured graph when the launch sequence is
// the one captured (same waves, level counts and settings), else captured anew;
// eager when timing per kernel family or with the VALU variant.
// The second P buffer of pipelined multi-wave sweeps, once, when the device keeps
// kPipeMarginBytes free after it (later lazy allocations need room).
constexpr size_t kPipeMarginBytes = size_t(16) << 30;

int ensure_p2(smc_engine *e)
{
    if (e->p2_tried || e->waves.size() < 2) return SMC_OK;
    e->p2_tried = true;
    const size_t bytes = (size_t)e->Cmax * e->nn * e->R * kPStride * e->tx_size;
    size_t free_b = 0, total_b = 0;
    HIP_TRY(e, hipMemGetInfo(&free_b, &total_b));
    if (free_b < bytes + kPipeMarginBytes) return SMC_OK;  // sweeps stay unpipelined
    for (hipEvent_t &ev : e->pipe_ev) HIP_TRY(e, hipEventCreateWithFlags(&ev, hipEventDisableTiming));
    if (hipMalloc(&e->d_P2, bytes) != hipSuccess) {
        e->d_P2 = nullptr;
        (void)hipGetLastError();
        return SMC_OK;
    }
    e->device_bytes += (int64_t)bytes;
    return SMC_OK;
}

int run_sweep(smc_engine *e, int first, int count)
{
    hipStream_t s = e->stream;
    int rc0 = ensure_p2(e);
    if (rc0 != SMC_OK) return rc0;
    const bool graph = e->use_graph && !e->timing && e->prune_variant != SMC_PRUNE_VALU;
    if (!graph) {
        issue_sweep(e, first, count, false);
        return SMC_OK;
    }
    std::vector<int64_t> key = {first, count, e->fp32, e->mix_mode, e->WS, e->tune.blocks_per_cu,
                                e->tune.unroll[0], e->tune.unroll[1], e->tune.unroll[2], e->tune.xcd_groups,
                                e->tune.unroll_rt, (int64_t)(intptr_t)e->d_bucket_off, e->d_P2 != nullptr};
    for (const smc_engine::Wave &w : e->waves) {
        key.push_back(w.first);
        key.push_back(w.count);
        key.push_back(w.graph_levels);
        key.push_back(w.boff);
    }
    smc_engine::CachedGraph *hit = nullptr;
    for (smc_engine::CachedGraph &c : e->graphs)
        if (c.key == key) {
            hit = &c;
            break;
        }
    if (!hit) {
        if (e->graphs.size() >= smc_engine::kMaxGraphs) {  // drop the least recently used
            // the evicted graph may still be running on the stream: let it finish first
            HIP_TRY(e, hipStreamSynchronize(s));
            auto lru = std::min_element(e->graphs.begin(), e->graphs.end(),
                                        [](const smc_engine::CachedGraph &a, const smc_engine::CachedGraph &b) {
                                            return a.used < b.used;
                                        });
            HIP_TRY(e, hipGraphExecDestroy(lru->exec));
            e->graphs.erase(lru);
        }
        hipGraph_t g = nullptr;
        HIP_TRY(e, hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
        issue_sweep(e, first, count, true);
        HIP_TRY(e, hipStreamEndCapture(s, &g));
        hipGraphExec_t exec = nullptr;
        hipError_t st = hipGraphInstantiate(&exec, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        if (st != hipSuccess)
            return smc_fail(e, SMC_ERR_HIP, std::string("hipGraphInstantiate failed: ") + hipGetErrorString(st));
        e->graphs.push_back(smc_engine::CachedGraph{key, exec, 0});
        hit = &e->graphs.back();
        e->graph_captures++;
    }
    hit->used = ++e->graph_clock;
    HIP_TRY(e, hipGraphLaunch(hit->exec, s));
    e->graph_launches++;
    return SMC_OK;
}

}  // namespace

int smc_refresh_likelihood(smc_engine *e, int first, int count, double *lnl_out,
                           uint8_t *nonfinite_out)
{
    if (!e) return SMC_ERR_ARG;
    PhaseTimer phase_timer(e, SMC_PHASE_REFRESH);
    if (first < 0 || count < 1 || first + count > e->Pmax)
        return fail(e, SMC_ERR_ARG, "particle range out of bounds");
    if (!e->have_alignment) return fail(e, SMC_ERR_STATE, "no alignment uploaded");
    for (int p = first; p < first + count; ++p)
        if (!e->h_valid[p])
            return fail(e, SMC_ERR_STATE, "particle slot " + std::to_string(p) + " not uploaded");
    int rc = build_schedule(e, first, count);
    if (rc != SMC_OK) return rc;
    hipStream_t s = e->stream;
    const int Ti = e->T - 1;
    if (e->WS == 2) {
        // forced refresh: every internal node writes its other workspace
        // (Particle.cpp:566-571), so the committed partials stay intact until commit
        for (size_t k = (size_t)first * Ti; k < (size_t)(first + count) * Ti; ++k) e->h_xflag[k] ^= 1;
        rc = h2d(e, e->d_xflag + (size_t)first * Ti, e->h_xflag.data() + (size_t)first * Ti, (size_t)count * Ti);
        if (rc != SMC_OK) return rc;
    }
    rc = run_sweep(e, first, count);
    if (rc != SMC_OK) return rc;
    HIP_TRY(e, hipGetLastError());
    rc = d2h(e, e->h_lnl.data() + first, e->d_lnl + first, (size_t)count * sizeof(double));
    if (rc == SMC_OK && nonfinite_out) rc = d2h(e, nonfinite_out, e->d_nonfinite + first, (size_t)count);
    if (rc == SMC_OK) rc = sync_stream(e);
    if (rc != SMC_OK) return rc;
    collect_timings(e);
    for (const smc_engine::Wave &w : e->waves)  // the last wave of each slot stays resident
        for (int k = 0; k < w.count; ++k) e->h_owner[k] = w.first + k;
    if (lnl_out) std::copy(e->h_lnl.begin() + first, e->h_lnl.begin() + first + count, lnl_out);
    smc_commit_range(e, first, count);
    drop_pending(e, first, count);  // every P of the range was recomputed
    return SMC_OK;
}

// ------------------------------------------------------------ incremental path
namespace {

// Compact class slots of one resident wave [first, first+count); fails if the
// wave does not fit the partial slots or the P buffer.
// The lane's resources, allocated on its first use: the sizes of the engine's own.
int lane_alloc(smc_engine *e)
{
    smc_engine::Lane &l = e->lane;
    if (l.stream) return SMC_OK;
    const size_t P = e->Pmax, T = e->T, nn = e->nn;
    int rc = SMC_OK;
    auto A = [&](auto **ptr, size_t n) {
        if (rc == SMC_OK && !*ptr) rc = dev_alloc(e, ptr, n);
    };
    A(&l.d_lnl, P);
    A(&l.d_nonfinite, P);
    A(&l.d_items, P * (T - 1));
    A(&l.d_chain_nodes, P * (T - 1));
    A(&l.d_chain_len, P * (T - 1));
    A(&l.d_plist, 4 * P);
    A(&l.d_pnodes, 2 * P * nn);
    if (rc != SMC_OK) return rc;
    if (!l.h_stage) {
        void *hs = nullptr;
        HIP_TRY(e, hipHostMalloc(&hs, e->stage_cap, hipHostMallocDefault));
        l.h_stage = (uint8_t *)hs;
        l.stage_cap = e->stage_cap;
    }
    if (!l.h_read) {
        void *hr = nullptr;
        HIP_TRY(e, hipHostMalloc(&hr, e->read_cap, hipHostMallocDefault));
        l.h_read = (uint8_t *)hr;
        l.read_cap = e->read_cap;
    }
    if (!l.fork) HIP_TRY(e, hipEventCreateWithFlags(&l.fork, hipEventDisableTiming));
    if (!l.done) HIP_TRY(e, hipEventCreateWithFlags(&l.done, hipEventDisableTiming));
    HIP_TRY(e, hipStreamCreateWithFlags(&l.stream, hipStreamNonBlocking));  // last: marks the lane ready
    return SMC_OK;
}

// Exchanges the lane's stream, staging, readback and scratch buffers with the
// engine's own (twice: in, then out).
void swap_lane(smc_engine *e)
{
    smc_engine::Lane &l = e->lane;
    std::swap(e->stream, l.stream);
    std::swap(e->h_stage, l.h_stage);
    std::swap(e->stage_cap, l.stage_cap);
    std::swap(e->stage_off, l.stage_off);
    std::swap(e->h_read, l.h_read);
    std::swap(e->read_cap, l.read_cap);
    std::swap(e->read_off, l.read_off);
    std::swap(e->readbacks, l.readbacks);
    std::swap(e->d_lnl, l.d_lnl);
    std::swap(e->d_nonfinite, l.d_nonfinite);
    std::swap(e->d_items, l.d_items);
    std::swap(e->d_chain_nodes, l.d_chain_nodes);
    std::swap(e->d_chain_len, l.d_chain_len);
    std::swap(e->d_plist, l.d_plist);
    std::swap(e->d_pnodes, l.d_pnodes);
}

struct LaneGuard {
    smc_engine *e;
    bool on;
    ~LaneGuard()
    {
        if (on) swap_lane(e);
    }
};

int set_single_wave(smc_engine *e, int first, int count)
{
    if (count > e->Wmax) return fail(e, SMC_ERR_ARG, "range exceeds partial_slots");
    if (e->pcls_first == first && e->pcls_count == count && e->pcls_version == e->class_version)
        return SMC_OK;  // the device already holds this layout
    std::vector<int> pcls(count);
    int64_t ncls = 0;
    for (int i = 0; i < count; ++i) {
        pcls[i] = (int)ncls;
        ncls += e->h_nclass[first + i];
    }
    if (ncls > e->Cmax) return fail(e, SMC_ERR_ARG, "range exceeds pmat_classes");
    int rc = h2d(e, e->d_pcls_off + first, pcls.data(), count * sizeof(int));
    if (rc != SMC_OK) return rc;
    e->pcls_first = first;
    e->pcls_count = count;
    e->pcls_version = e->class_version;
    return SMC_OK;
}

// Eigensystem and P recomputes in two halves, so that their uploads can go ahead
// of a stream fork: upload_models stages the lists (eigen list at plist, full-P
// list at plist + Pmax, node lists in pnodes), launch_models launches on a stream.
struct ModelJob {
    int n_eig = 0, n_p = 0, np = 0, max_nodes = 0;
    int *d_eig = nullptr, *d_p = nullptr, *d_nodes = nullptr;
};

int upload_models(smc_engine *e, const std::vector<int> &eig_list, const std::vector<int> &p_list,
                  const std::vector<Item> &p_nodes, int *plist, Item *pnodes, ModelJob &job)
{
    if (!p_nodes.empty()) {  // P of single (particle, node) pairs: branch-length moves
        if ((int64_t)p_nodes.size() > (int64_t)e->Pmax * e->nn)
            return fail(e, SMC_ERR_STATE, "too many P nodes");
        // grouped by particle (the lists are built particle by particle):
        // [particles | offsets | nodes] in one upload
        std::vector<int> buf;
        std::vector<int> plist_n, noff{0}, nodes;
        for (size_t i = 0; i < p_nodes.size(); ++i) {
            if (i == 0 || p_nodes[i].particle != p_nodes[i - 1].particle) {
                if (i) noff.push_back((int)nodes.size());
                plist_n.push_back(p_nodes[i].particle);
            }
            nodes.push_back(p_nodes[i].node);
        }
        noff.push_back((int)nodes.size());
        job.np = (int)plist_n.size();
        for (int i = 0; i < job.np; ++i) job.max_nodes = std::max(job.max_nodes, noff[i + 1] - noff[i]);
        buf.insert(buf.end(), plist_n.begin(), plist_n.end());
        buf.insert(buf.end(), noff.begin(), noff.end());
        buf.insert(buf.end(), nodes.begin(), nodes.end());
        job.d_nodes = reinterpret_cast<int *>(pnodes);  // capacity: 2 ints per (particle, node)
        int rc = h2d(e, job.d_nodes, buf.data(), buf.size() * sizeof(int));
        if (rc != SMC_OK) return rc;
    }
    if (!eig_list.empty()) {
        job.n_eig = (int)eig_list.size();
        job.d_eig = plist;
        int rc = h2d(e, job.d_eig, eig_list.data(), eig_list.size() * sizeof(int));
        if (rc != SMC_OK) return rc;
    }
    if (!p_list.empty()) {
        job.n_p = (int)p_list.size();
        job.d_p = plist + e->Pmax;  // apart from the eigen list, which its launch may still read
        int rc = h2d(e, job.d_p, p_list.data(), p_list.size() * sizeof(int));
        if (rc != SMC_OK) return rc;
    }
    return SMC_OK;
}

void launch_models(smc_engine *e, int first, const ModelJob &job, hipStream_t s)
{
    if (job.np > 0) {
        TimedLaunch tl(e, SMC_K_PMAT);
        PmatArgs pa{e->d_eig, e->d_blen, e->d_rates, e->d_nclass, e->d_pcls_off, e->d_P, e->nn,
                    e->R, e->Kmax, first, 0, e->fp32, nullptr};
        const int *d = job.d_nodes;
        launch_pmat_nodes(pa, d, d + job.np, d + 2 * job.np + 1, job.np, job.max_nodes, s);
    }
    if (job.n_eig > 0) {
        TimedLaunch tl(e, SMC_K_EIGEN);
        EigenArgs ea{e->d_base, e->d_logits, e->d_nclass, e->d_eig, e->Kmax, first, job.n_eig, job.d_eig};
        launch_eigen(ea, s);
    }
    if (job.n_p > 0) {
        TimedLaunch tl(e, SMC_K_PMAT);
        PmatArgs pa{e->d_eig, e->d_blen, e->d_rates, e->d_nclass, e->d_pcls_off, e->d_P, e->nn,
                    e->R, e->Kmax, first, job.n_p, e->fp32, job.d_p};
        launch_pmat(pa, s);
    }
}

// Eigen-decompositions and/or P matrices of the particles in the lists.
int recompute_models(smc_engine *e, int first, const std::vector<int> &eig_list,
                     const std::vector<int> &p_list, const std::vector<Item> &p_nodes = {})
{
    ModelJob job;
    int rc = upload_models(e, eig_list, p_list, p_nodes, e->d_plist, e->d_pnodes, job);
    if (rc != SMC_OK) return rc;
    launch_models(e, first, job, e->stream);
    return SMC_OK;
}

// Uploads new topology/branch lengths/exchangeabilities/pInvar/shape for slots
// [first, first+count) without touching their CAT classes.  Only the rows that
// differ from the uploaded state go up: packed into one copy and scattered on the
// device by one launch (a move changes one or two of a particle's rows; finished
// chains of a lockstep batch change none).
// `slots` (optional): row i is slot slots[i] instead of first + i; `committed`:
// the rows are the slots' committed state, whose topologies need no derivation.
int upload_params(smc_engine *e, int first, int count, const int32_t *children,
                  const double *branch_len, const double *base, const double *pinv,
                  const double *shape, const int *slots = nullptr, bool committed = false)
{
    const int T = e->T, Ti = T - 1, nn = e->nn, R = e->R;
    std::vector<ScatterSeg> segs;
    std::vector<uint32_t> data;
    auto add = [&](const void *dst, const void *src, size_t bytes) {
        const size_t o = data.size();
        segs.push_back(ScatterSeg{(uint64_t)(uintptr_t)dst, (uint32_t)o, (uint32_t)(bytes / 4)});
        data.resize(o + bytes / 4);
        std::memcpy(data.data() + o, src, bytes);
    };
    std::vector<double> bl(nn);
    std::vector<int> hh(Ti);
    for (int i = 0; i < count; ++i) {
        const int p = slots ? slots[i] : first + i;
        if (!e->h_valid[p]) return fail(e, SMC_ERR_STATE, "slot " + std::to_string(p) + " not uploaded");
        const int32_t *ch = children + (size_t)i * Ti * 2;
        int *mch = &e->h_children[(size_t)p * Ti * 2];
        if (!std::equal(ch, ch + Ti * 2, mch) && committed && e->committed[p].root >= 0) {
            // back to the committed topology: its root and heights as committed
            const smc_engine::Committed &c = e->committed[p];
            std::copy(c.heights.begin(), c.heights.end(), &e->h_height[(size_t)p * Ti]);
            std::copy(ch, ch + Ti * 2, mch);
            e->h_root[p] = c.root;
            if (!e->h_prog_stale[p]) {
                e->h_prog_stale[p] = 1;
                e->n_prog_stale++;
            }
            add(e->d_children + (size_t)p * Ti * 2, mch, (size_t)Ti * 2 * sizeof(int));
            add(e->d_root + p, &e->h_root[p], sizeof(int));
        } else if (!std::equal(ch, ch + Ti * 2, mch)) {
            // a new topology: its root and heights (unchanged topologies --
            // branch-length, rate and invariant moves, restores of them -- keep the
            // ones already derived); its scoring program is derived when a site-class
            // kernel or a particle copy next needs it (ensure_programs)
            int root = -1;
            int rc = smc_topology(e, p, ch, &root, hh.data(), nullptr);
            if (rc != SMC_OK) return rc;
            std::copy(hh.begin(), hh.end(), &e->h_height[(size_t)p * Ti]);
            std::copy(ch, ch + Ti * 2, mch);
            e->h_root[p] = root;
            if (!e->h_prog_stale[p]) {
                e->h_prog_stale[p] = 1;
                e->n_prog_stale++;
            }
            add(e->d_children + (size_t)p * Ti * 2, mch, (size_t)Ti * 2 * sizeof(int));
            add(e->d_root + p, &e->h_root[p], sizeof(int));
        }
        const int root = e->h_root[p];
        for (int n = 0; n < nn; ++n) bl[n] = n == root ? 0.0 : branch_len[(size_t)i * nn + n];
        double *mbl = &e->h_blen[(size_t)p * nn];
        if (!std::equal(bl.begin(), bl.end(), mbl)) {
            std::copy(bl.begin(), bl.end(), mbl);
            add(e->d_blen + (size_t)p * nn, mbl, (size_t)nn * sizeof(double));
        }
        const double *bs = base + (size_t)i * 190;
        double *mbs = &e->h_base[(size_t)p * 190];
        if (!std::equal(bs, bs + 190, mbs)) {
            std::copy(bs, bs + 190, mbs);
            add(e->d_base + (size_t)p * 190, mbs, 190 * sizeof(double));
        }
        if (pinv[i] != e->h_pinv[p]) {
            e->h_touched[p] = 1;
            e->h_pinv[p] = pinv[i];
            add(e->d_pinv + p, &e->h_pinv[p], sizeof(double));
        }
        if (R > 1) {
            if (!(shape[i] > 0.0)) return fail(e, SMC_ERR_ARG, "shape must be > 0");
            if (shape[i] != e->h_shape[p]) {  // rates of an unchanged shape are kept
                discretize_gamma(shape[i], R, &e->h_rates[(size_t)p * R]);
                add(e->d_rates + (size_t)p * R, &e->h_rates[(size_t)p * R], (size_t)R * sizeof(double));
            }
        }
        e->h_shape[p] = shape[i];
    }
    int rc = scatter_upload(e, segs, data);
    if (rc != SMC_OK) return rc;
    e->topo_version++;
    return SMC_OK;
}

}  // namespace

int smc_update_params(smc_engine *e, int first, int count, const int32_t *children,
                      const double *branch_len, const double *base, const double *pinv,
                      const double *shape)
{
    if (!e) return SMC_ERR_ARG;
    PhaseTimer phase_timer(e, SMC_PHASE_UPDATE_PARAMS);
    if (first < 0 || count < 1 || first + count > e->Pmax || !children || !branch_len || !base ||
        !pinv || !shape)
        return fail(e, SMC_ERR_ARG, "invalid parameter update");
    return upload_params(e, first, count, children, branch_len, base, pinv, shape);
}

extern "C++" {
int refresh_incremental(smc_engine *e, int first, int count, const uint8_t *node_flags,
                        const uint8_t *eig_flags, double *lnl_out, uint8_t *nonfinite_out, bool on_lane)
{
    if (!e) return SMC_ERR_ARG;
    PhaseTimer phase_timer(e, SMC_PHASE_REFRESH_INCREMENTAL);
    if (e->WS != 2) return fail(e, SMC_ERR_STATE, "incremental refresh needs workspaces = 2");
    if (first < 0 || count < 1 || first + count > e->Pmax || !node_flags)
        return fail(e, SMC_ERR_ARG, "invalid incremental refresh arguments");
    for (int p = first; p < first + count; ++p)
        if (!e->h_valid[p]) return fail(e, SMC_ERR_STATE, "slot " + std::to_string(p) + " not uploaded");
    if (on_lane) {
        if (e->lane.in_flight || e->timing) return fail(e, SMC_ERR_STATE, "the lane is busy or timing is on");
        int rc0 = lane_alloc(e);
        if (rc0 != SMC_OK) return rc0;
        // the lane starts after what the caller queued on the engine's stream (the
        // proposals' parameter uploads)
        HIP_TRY(e, hipEventRecord(e->lane.fork, e->stream));
    }
    LaneGuard lane_guard{e, false};
    if (on_lane) {
        swap_lane(e);
        lane_guard.on = true;
        HIP_TRY(e, hipStreamWaitEvent(e->stream, e->lane.fork, 0));
    }
    int rc = set_single_wave(e, first, count);
    if (rc != SMC_OK) return rc;
    const int T = e->T, Ti = T - 1, nn = e->nn;
    std::vector<int> eig_list, p_list;
    std::vector<std::vector<Item>> buckets;  // (level * PRUNE_KINDS + kind)
    std::vector<Item> roots;
    std::vector<uint8_t> dirty(nn);
    std::vector<int> parent(nn), dlevel(nn, 0), dl_order, clevel(nn, 0);
    std::vector<std::vector<Item>> chains;              // per chain level: {particle, first node}
    std::vector<std::vector<int>> chain_lens_by_level;  // their lengths
    std::vector<int> chain_nodes;
    std::vector<Item> p_nodes;
    std::vector<int> keep;
    std::vector<double> keep_lnl;
    // Chains pay off where the per-level launches would be small: a one-particle-at-a-
    // time caller, or the last rounds of a lockstep batch.  With many particles moving,
    // the per-level launches fill the GPU and their kind-specialised kernels stream
    // faster than the chain kernel (whose wavefronts run ~20 dependent steps each):
    // chains are used when the particles that move in this call give per-level
    // launches of fewer units than the merged-small-launch bound below.
    int moving = 0;
    std::vector<uint8_t> mov(count, 0);  // particles with flagged nodes or eigensystems
    for (int i = 0; i < count; ++i) {
        bool any = eig_flags && eig_flags[i];
        const uint8_t *nf = node_flags + (size_t)i * nn;
        for (int n = 0; n < nn && !any; ++n) any = nf[n] != 0;
        if (e->h_busy[first + i]) any = false;  // on the lane: left alone
        mov[i] = any ? 1 : 0;
        moving += any ? 1 : 0;
    }
    const bool chains_fit = (int64_t)moving * ((e->tile_stride + 3) / 4) <= (int64_t)e->num_cus * 8;
    // walks: every changed particle in one launch (see launch_prune_walk)
    const bool walk_ok = e->R <= kWalkMaxRates;
    const bool use_walk = e->prune_variant == SMC_PRUNE_MFMA_WALK && walk_ok;
    std::vector<int4> walk;                     // two int4 per entry
    std::vector<std::vector<WalkItem>> wlevels;  // walk pieces by launch level
    std::vector<WalkItem> wroots;               // the pieces holding the roots
    std::vector<int> po, wst;
    std::vector<int> wsize(nn), wlev(nn), wtop(nn);
    std::vector<uint8_t> wcut(nn), walked(count, 0);
    for (int i = 0; i < count; ++i) {
        const int p = first + i;
        // the lane's refresh takes only its flagged particles; the others leave the
        // lane's particles alone (their lnL stays until lane_finish)
        if (on_lane && !mov[i]) continue;
        if (!on_lane && e->h_busy[p]) {
            keep.push_back(p);
            keep_lnl.push_back(e->h_lnl[p]);
            continue;
        }
        const uint8_t *nf = node_flags + (size_t)i * nn;
        const bool eig = eig_flags && eig_flags[i];
        // P: every node of a particle whose eigensystems change, else the nodes
        // whose branch length (or rate) changed; kept for a reject to restore
        bool any_tp = eig;
        for (int n = 0; n < nn; ++n)
            if (nf[n] & 2) {
                any_tp = true;
                if (!eig) {
                    p_nodes.push_back(Item{p, n});
                    e->h_pnodes[p].push_back(n);
                }
            }
        if (eig) {
            eig_list.push_back(p);
            p_list.push_back(p);
        }
        e->h_dirty[p] |= (any_tp ? 1 : 0) | (eig ? 2 : 0);
        // dirty internal nodes: flagged, parents of changed branches, and every
        // ancestor of a dirty node (Tree.cpp marks the path to the root the same way)
        const int *ch = &e->h_children[(size_t)p * Ti * 2];
        std::fill(parent.begin(), parent.end(), -1);
        for (int k = 0; k < Ti; ++k) {
            parent[ch[2 * k]] = T + k;
            parent[ch[2 * k + 1]] = T + k;
        }
        std::fill(dirty.begin(), dirty.end(), 0);
        for (int n = 0; n < nn; ++n) {
            if (eig || (nf[n] & 2)) {
                if (parent[n] >= 0) dirty[parent[n]] = 1;
                if (eig && n >= T) dirty[n] = 1;
            }
            if ((nf[n] & 1) && n >= T) dirty[n] = 1;
        }
        for (int n = T; n < nn; ++n)
            if (dirty[n])
                for (int q = parent[n]; q >= 0 && !dirty[q]; q = parent[q]) dirty[q] = 1;
        const int root = e->h_root[p];
        // launch level of a dirty node: one above the highest level among its dirty
        // children that are stored partials (a cherry is built by its parent, a tip
        // needs nothing), so the k-th node from the bottom of every particle's dirty
        // path shares one launch; a node's height in the tree would spread a path
        // over as many launches as the tree is tall.  The dirty nodes form a subtree
        // holding the root (every ancestor of a dirty node is dirty): walk it in
        // pre-order, then assign levels in reverse, children first.
        dl_order.clear();
        if (dirty[root]) dl_order.push_back(root);
        for (size_t i = 0; i < dl_order.size(); ++i) {
            const int v = dl_order[i];
            for (int j = 0; j < 2; ++j) {
                const int c = ch[2 * (v - T) + j];
                if (c >= T && dirty[c]) dl_order.push_back(c);
            }
        }
        for (size_t i = dl_order.size(); i-- > 0;) {
            const int v = dl_order[i];
            int L = 1;
            for (int j = 0; j < 2; ++j) {
                const int c = ch[2 * (v - T) + j];
                if (c >= T && dirty[c] && prune_kind(ch, T, c) != PRUNE_TT) L = std::max(L, dlevel[c] + 1);
            }
            dlevel[v] = L;
        }
        // Short dirty paths run as chains (see launch_prune_chain): a chain starts at
        // every dirty stored node that has no dirty stored child or two of them, and
        // climbs while the parent's only dirty stored child is the chain's top; a
        // chain runs after the chains that end at its first node's children.  Larger
        // dirty sets (rate-matrix and shape moves, big subtree scalings) keep the
        // per-level bucket launches and their kind-specialised kernels.
        int n_stored = 0;
        for (int n = T; n < nn; ++n)
            if (dirty[n] && n != root && prune_kind(ch, T, n) != PRUNE_TT) n_stored++;
        const bool use_chains = e->prune_variant == SMC_PRUNE_MFMA && chains_fit && n_stored <= e->tune.chain_max;
        // the variant's walks, or (SMC_PRUNE_MFMA) the small dirty sets when the call
        // holds many: dirty paths walked in one launch, whole trees level by level
        const bool walk_p = !on_lane && (use_walk || (walk_ok && e->prune_variant == SMC_PRUNE_MFMA &&
                                                      !chains_fit && n_stored <= e->tune.chain_max));
        walked[i] = walk_p ? 1 : 0;
        for (int n = T; n < nn; ++n) {
            if (!dirty[n] || n == root) continue;
            e->h_xflag[(size_t)p * Ti + (n - T)] ^= 1;
            if (use_chains || walk_p) continue;
            const int L = dlevel[n];
            const int kind = prune_kind(ch, T, n);
            if ((int)buckets.size() < PRUNE_KINDS * (L + 1)) buckets.resize(PRUNE_KINDS * (L + 1));
            buckets[PRUNE_KINDS * L + kind].push_back(Item{p, n});
        }
        if (use_chains) {
            auto stored_dirty = [&](int c) { return c >= T && dirty[c] && prune_kind(ch, T, c) != PRUNE_TT; };
            auto dsc = [&](int v) {
                return (int)stored_dirty(ch[2 * (v - T)]) + (int)stored_dirty(ch[2 * (v - T) + 1]);
            };
            // dl_order is a pre-order of the dirty subtree: walked backwards, a node's
            // dirty children come before it
            for (size_t k = dl_order.size(); k-- > 0;) {
                const int v = dl_order[k];
                if (v == root || !stored_dirty(v)) continue;
                const int d = dsc(v);
                if (d == 1) continue;  // continues the chain of its dirty stored child
                int L = 1;
                if (d == 2)
                    L = 1 + std::max(clevel[ch[2 * (v - T)]], clevel[ch[2 * (v - T) + 1]]);
                if ((int)chains.size() < L + 1) chains.resize(L + 1);
                chains[L].push_back(Item{p, (int)chain_nodes.size()});
                int q = v, len = 0;
                for (;;) {
                    chain_nodes.push_back(q);
                    clevel[q] = L;
                    len++;
                    const int up = parent[q];
                    if (up == root || dsc(up) != 1) break;
                    q = up;
                }
                chain_lens_by_level.resize(chains.size());
                chain_lens_by_level[L].push_back(len);
            }
        }
        bool changed = eig || e->h_touched[p];  // pInvar enters at the root only
        e->h_touched[p] = 0;
        for (int n = 0; n < nn && !changed; ++n) changed = nf[n] != 0;
        if (changed && e->h_owner[i] != p) e->h_owner[i] = -1;  // wave slot i held another slot's partials
        if (changed && walk_p) {
            // post-order of the dirty stored nodes (cherries are built by their parents;
            // every ancestor of a dirty node is dirty, so all are reachable from the
            // root), the root last
            po.clear();
            wst.assign(1, root);
            while (!wst.empty()) {
                const int v = wst.back();
                wst.pop_back();
                po.push_back(v);
                for (int j = 0; j < 2; ++j) {
                    const int c = ch[2 * (v - T) + j];
                    if (c >= T && dirty[c] && prune_kind(ch, T, c) != PRUNE_TT) wst.push_back(c);
                }
            }
            std::reverse(po.begin(), po.end());
            // Long walks (whole trees) are cut into pieces of at most kWalkPiece nodes,
            // so that no wavefront walks a long dependent sequence: in post-order, a
            // node whose uncut subtree grows beyond the bound cuts its largest child
            // subtrees off as pieces of their own.  A piece runs in the launch after
            // every piece cut below it (its level); its cut children are read back.
            auto stored_kid = [&](int c) { return c >= T && dirty[c] && prune_kind(ch, T, c) != PRUNE_TT; };
            for (int v : po) {
                int sz = 1, lev = 0;
                int kid[2], nk = 0;
                for (int j = 0; j < 2; ++j) {
                    const int c = ch[2 * (v - T) + j];
                    if (stored_kid(c)) {
                        wcut[c] = 0;
                        kid[nk++] = c;
                        sz += wsize[c];
                    }
                }
                while (sz > kWalkPiece && nk > 0) {
                    int b = 0;
                    for (int k = 1; k < nk; ++k)
                        if (wsize[kid[k]] > wsize[kid[b]]) b = k;
                    wcut[kid[b]] = 1;
                    sz -= wsize[kid[b]];
                    kid[b] = kid[--nk];
                }
                for (int j = 0; j < 2; ++j) {
                    const int c = ch[2 * (v - T) + j];
                    if (stored_kid(c)) lev = std::max(lev, wlev[c] + (wcut[c] ? 1 : 0));
                }
                wsize[v] = sz;
                wlev[v] = lev;
            }
            wcut[root] = 1;
            // the piece of each node: its nearest cut ancestor-or-self
            for (size_t k = po.size(); k-- > 0;) {
                const int v = po[k];
                wtop[v] = (wcut[v] || v == root) ? v : wtop[parent[v]];
            }
            const uint8_t *xfp = &e->h_xflag[(size_t)p * Ti];
            auto pslot = [&](int c) { return (int)xfp[c - T] * Ti + (c - T); };
            for (size_t k = 0; k < po.size(); ++k) {
                const int t = po[k];
                if (wtop[t] != t) continue;  // not a piece's top: pieces are emitted at their top
                // the piece's nodes in post-order: those of po whose top is t
                const int first = (int)(walk.size() / 2);
                int prev = -1, cnt = 0;
                for (size_t q = 0; q <= k; ++q) {
                    const int v = po[q];
                    if (wtop[v] != t) continue;
                    int c[2] = {ch[2 * (v - T)], ch[2 * (v - T) + 1]};
                    int m[2], aux[2];
                    for (int j = 0; j < 2; ++j) {
                        const int cc = c[j];
                        if (cc < T) {
                            m[j] = WALK_TIP;
                            aux[j] = 0;
                        } else if (prune_kind(ch, T, cc) == PRUNE_TT) {
                            m[j] = WALK_CHERRY;
                            aux[j] = ch[2 * (cc - T)] | ch[2 * (cc - T) + 1] << 16;
                        } else {
                            m[j] = cc == prev ? WALK_CARRIED : WALK_STORED;
                            aux[j] = pslot(cc);
                        }
                    }
                    const int a = m[1] == WALK_CARRIED ? 1 : 0, b = 1 - a;  // a carried child is child A
                    walk.push_back(make_int4(v, c[a], c[b], m[a] | m[b] << 4 | (v == root ? 1 << 8 : 0)));
                    walk.push_back(make_int4(v == root ? 0 : pslot(v), aux[a], aux[b], 0));
                    prev = v;
                    cnt++;
                }
                const int lev = wlev[t];
                if ((int)wlevels.size() <= lev) wlevels.resize(lev + 1);
                wlevels[lev].push_back(WalkItem{p, first, cnt});
                if (t == root) wroots.push_back(WalkItem{p, first, cnt});
            }
        } else if (changed) {
            roots.push_back(Item{p, root});
        }
        if (!changed) {  // nothing to recompute: its lnL stays (its tile sums may be stale)
            keep.push_back(p);
            keep_lnl.push_back(e->h_lnl[p]);
        }
    }
    // committed eigensystems of particles about to be decomposed again: kept for a
    // reject (a particle without a free save slot is decomposed again on a reject)
    if (!eig_list.empty() && !e->d_eig_save) {
        rc = dev_alloc(e, &e->d_eig_save, (size_t)e->Wmax * e->Kmax * kEigStride);
        if (rc != SMC_OK) return rc;
        for (int k = e->Wmax; k-- > 0;) e->eig_save_free.push_back(k);
    }
    for (int p : eig_list)
        if (e->h_eig_save[p] < 0 && !e->eig_save_free.empty()) {
            const int slot = e->eig_save_free.back();
            e->eig_save_free.pop_back();
            HIP_TRY(e, hipMemcpyAsync(e->d_eig_save + (size_t)slot * e->Kmax * kEigStride,
                                      e->d_eig + (size_t)p * e->Kmax * kEigStride,
                                      (size_t)e->h_nclass[p] * kEigStride * sizeof(double), hipMemcpyDeviceToDevice,
                                      e->stream));
            e->h_eig_save[p] = slot;
        }
    // deferred reject restores of this range join this refresh's recomputes (none
    // where the particle recomputes every P anyway)
    if (!e->pend_full.empty() || !e->pend_nodes.empty()) {
        std::vector<uint8_t> full(count, 0);
        for (int p : p_list) full[p - first] = 1;
        auto in = [&](int p) {
            return p >= first && p < first + count && (on_lane ? mov[p - first] != 0 : !e->h_busy[p]);
        };
        std::vector<int> keep_full;
        for (int p : e->pend_full) {
            if (!in(p)) {
                keep_full.push_back(p);
            } else if (!full[p - first]) {
                full[p - first] = 1;
                p_list.push_back(p);
            }
        }
        std::vector<Item> keep_nodes;
        std::vector<uint8_t> has_node;  // (particle, node) pairs already listed
        for (const Item &it : e->pend_nodes) {
            if (!in(it.particle)) {
                keep_nodes.push_back(it);
            } else if (!full[it.particle - first]) {
                if (has_node.empty()) {
                    has_node.assign((size_t)count * nn, 0);
                    for (const Item &q : p_nodes) has_node[(size_t)(q.particle - first) * nn + q.node] = 1;
                }
                uint8_t &h = has_node[(size_t)(it.particle - first) * nn + it.node];
                if (!h) p_nodes.push_back(it);
                h = 1;
            }
        }
        e->pend_full.swap(keep_full);
        e->pend_nodes.swap(keep_nodes);
        // the node lists go up grouped by particle
        std::stable_sort(p_nodes.begin(), p_nodes.end(),
                         [](const Item &x, const Item &y) { return x.particle < y.particle; });
    }
    // With walks and level launches in one call (dirty paths walked, whole trees by
    // level), the levels run on the second stream beside the walks: their narrow
    // deep launches then fill the gaps of the walk instead of following it; the
    // models of the levels' particles are recomputed on that stream too.
    const bool split = !on_lane && !e->timing && !wroots.empty() && !roots.empty();
    ModelJob job_w, job_t;  // (split) the levels' and the walks' models
    if (split) {
        std::vector<int> eig_w, eig_t, pl_w, pl_t;
        std::vector<Item> pn_w, pn_t;
        for (int p : eig_list) (walked[p - first] ? eig_t : eig_w).push_back(p);
        for (int p : p_list) (walked[p - first] ? pl_t : pl_w).push_back(p);
        for (const Item &it : p_nodes) (walked[it.particle - first] ? pn_t : pn_w).push_back(it);
        rc = upload_models(e, eig_w, pl_w, pn_w, e->d_plist, e->d_pnodes, job_w);
        if (rc == SMC_OK)
            rc = upload_models(e, eig_t, pl_t, pn_t, e->d_plist + 2 * (size_t)e->Pmax,
                               e->d_pnodes + (size_t)e->Pmax * nn, job_t);
    } else {
        rc = recompute_models(e, first, eig_list, p_list, p_nodes);
    }
    if (rc != SMC_OK) return rc;
    hipStream_t s = e->stream;
    if (on_lane) {  // only its particles' rows: the engine's stream may be writing the others
        for (int i = 0; i < count && rc == SMC_OK; ++i)
            if (mov[i])
                rc = h2d(e, e->d_xflag + (size_t)(first + i) * Ti, e->h_xflag.data() + (size_t)(first + i) * Ti,
                         (size_t)Ti);
    } else {
        rc = h2d(e, e->d_xflag + (size_t)first * Ti, e->h_xflag.data() + (size_t)first * Ti, (size_t)count * Ti);
    }
    if (rc != SMC_OK) return rc;
    std::vector<Item> items;
    std::vector<int> off;
    for (auto &b : buckets) {
        off.push_back((int)items.size());
        items.insert(items.end(), b.begin(), b.end());
    }
    off.push_back((int)items.size());
    std::vector<int> chain_off;  // chain level L: items [chain_off[L], chain_off[L + 1])
    std::vector<int> lens(items.size(), 0);
    for (size_t L = 0; L < chains.size(); ++L) {
        chain_off.push_back((int)items.size());
        items.insert(items.end(), chains[L].begin(), chains[L].end());
        lens.insert(lens.end(), chain_lens_by_level[L].begin(), chain_lens_by_level[L].end());
    }
    chain_off.push_back((int)items.size());
    const int root_off = (int)items.size();
    items.insert(items.end(), roots.begin(), roots.end());
    rc = h2d(e, e->d_items, items.data(), items.size() * sizeof(Item));
    if (rc == SMC_OK && !chain_nodes.empty()) {
        rc = h2d(e, e->d_chain_nodes, chain_nodes.data(), chain_nodes.size() * sizeof(int));
        if (rc == SMC_OK) rc = h2d(e, e->d_chain_len, lens.data(), lens.size() * sizeof(int));
    }
    if (rc != SMC_OK) return rc;
    e->sched_first = -1;  // the item buffer no longer holds the full schedule
    PruneArgs a = prune_args(e, first);
    a.xflag = e->d_xflag;
    a.WS = 2;
    hipStream_t sl = s;  // the level launches' stream
    if (split) {  // fork after every upload
        HIP_TRY(e, hipEventRecord(e->ev_fork, s));
        HIP_TRY(e, hipStreamWaitEvent(e->stream2, e->ev_fork, 0));
        sl = e->stream2;
        launch_models(e, first, job_w, sl);
        launch_models(e, first, job_t, s);
    }
    auto run = [&](int o, int n, int kind, bool is_roots) {
        if (n <= 0) return;
        TimedLaunch tl(e, SMC_K_PRUNE);
        a.items = e->d_items + o;
        a.n_items = n;
        a.kind = kind;
        a.roots = is_roots;
        a.unroll = (is_roots || kind == PRUNE_ANY) ? e->tune.unroll_rt : e->tune.unroll[prune_stored(kind)];
        if (e->prune_variant == SMC_PRUNE_VALU)
            launch_prune(a, sl);
        else
            launch_prune_mfma(a, sl);
    };
    const int bpi = (e->tile_stride + 3) / 4;
    const int64_t small = (int64_t)e->num_cus * 8;  // units that do not fill the GPU twice over
    for (size_t b = 0; b + 1 < off.size(); ++b) {
        const int kind = (int)(b % PRUNE_KINDS);
        if (e->prune_variant == SMC_PRUNE_VALU) {
            run(off[b], off[b + 1] - off[b], PRUNE_ANY, false);
        } else if (kind == PRUNE_TC) {
            // a level's non-cherry buckets are contiguous: when they are small, one
            // launch that decides the child modes per item saves four launches
            const size_t end = std::min(off.size() - 1, b + (PRUNE_KINDS - 1));
            const int n = off[end] - off[b];
            if ((int64_t)n * bpi <= small) {
                run(off[b], n, PRUNE_ANY, false);
                b = end - 1;
            } else {
                run(off[b], off[b + 1] - off[b], kind, false);
            }
        } else if (kind != PRUNE_TT) {  // cherries are built by their parents
            run(off[b], off[b + 1] - off[b], kind, false);
        }
    }
    for (size_t L = 0; L + 1 < chain_off.size(); ++L) {
        const int n = chain_off[L + 1] - chain_off[L];
        if (n <= 0) continue;
        TimedLaunch tl(e, SMC_K_PRUNE);
        a.items = e->d_items + chain_off[L];
        a.n_items = n;
        a.roots = false;
        launch_prune_chain(a, e->d_chain_nodes, e->d_chain_len + chain_off[L], sl);
    }
    run(root_off, (int)roots.size(), PRUNE_ANY, true);
    if (split) HIP_TRY(e, hipEventRecord(e->ev_join, e->stream2));
    if (!wroots.empty()) {
        if (!e->d_walk_rdot) {
            const size_t nr = (size_t)e->Wmax * e->tile_stride * e->R * 64;
            rc = dev_alloc(e, &e->d_walk_rdot, nr);
            if (rc == SMC_OK) rc = dev_alloc(e, &e->d_walk_rexp, nr);
            if (rc != SMC_OK) return rc;
        }
        // item lists: the levels' pieces (longest first, so a launch does not end on
        // one), then the roots' pieces for the mixture
        std::vector<WalkItem> wall;
        std::vector<int> woff{0};
        for (auto &lv : wlevels) {
            std::stable_sort(lv.begin(), lv.end(),
                             [](const WalkItem &x, const WalkItem &y) { return x.count > y.count; });
            wall.insert(wall.end(), lv.begin(), lv.end());
            woff.push_back((int)wall.size());
        }
        const int roots_at = (int)wall.size();
        wall.insert(wall.end(), wroots.begin(), wroots.end());
        rc = h2d(e, e->d_walk, walk.data(), walk.size() * sizeof(int4));
        if (rc == SMC_OK) rc = h2d(e, e->d_witems, wall.data(), wall.size() * sizeof(WalkItem));
        if (rc != SMC_OK) return rc;
        TimedLaunch tl(e, SMC_K_PRUNE);
        for (size_t L = 0; L + 1 < woff.size(); ++L) {
            a.n_items = woff[L + 1] - woff[L];
            launch_prune_walk(a, e->d_witems + woff[L], e->d_walk, e->d_walk_rdot, e->d_walk_rexp, s);
        }
        a.n_items = (int)wroots.size();
        launch_walk_mixture(a, e->d_witems + roots_at, e->d_walk_rdot, e->d_walk_rexp, s);
    }
    if (split) HIP_TRY(e, hipStreamWaitEvent(s, e->ev_join, 0));  // the levels are done
    {
        TimedLaunch tl(e, SMC_K_REDUCE);
        launch_reduce_lnl(e->d_tile_lnl, e->d_ntiles, e->tile_stride, first, count, e->d_lnl,
                          e->d_nonfinite, s);
    }
    HIP_TRY(e, hipGetLastError());
    if (on_lane) {  // no wait: lane_finish collects the lnL of its particles
        rc = d2h(e, e->lane.lnl.data() + first, e->d_lnl + first, (size_t)count * sizeof(double));
        if (rc != SMC_OK) return rc;
        HIP_TRY(e, hipEventRecord(e->lane.done, e->stream));
        e->lane.particles.clear();
        for (int i = 0; i < count; ++i)
            if (mov[i]) {
                e->lane.particles.push_back(first + i);
                e->h_busy[first + i] = 1;
            }
        e->lane.in_flight = true;
        return SMC_OK;
    }
    rc = d2h(e, e->h_lnl.data() + first, e->d_lnl + first, (size_t)count * sizeof(double));
    if (rc == SMC_OK && nonfinite_out) rc = d2h(e, nonfinite_out, e->d_nonfinite + first, (size_t)count);
    if (rc == SMC_OK) rc = sync_stream(e);
    if (rc != SMC_OK) return rc;
    collect_timings(e);
    for (size_t k = 0; k < keep.size(); ++k) {
        e->h_lnl[keep[k]] = keep_lnl[k];
        if (nonfinite_out) nonfinite_out[keep[k] - first] = std::isfinite(keep_lnl[k]) ? 0 : 1;
    }
    if (lnl_out) std::copy(e->h_lnl.begin() + first, e->h_lnl.begin() + first + count, lnl_out);
    return SMC_OK;
}

int lane_finish(smc_engine *e, bool wait, bool *done)
{
    smc_engine::Lane &l = e->lane;
    *done = true;
    if (!l.in_flight) return SMC_OK;
    if (!wait) {
        const hipError_t q = hipEventQuery(l.done);
        if (q == hipErrorNotReady) {
            (void)hipGetLastError();  // not an error: clear it
            *done = false;
            return SMC_OK;
        }
        HIP_TRY(e, q);
    }
    HIP_TRY(e, hipEventSynchronize(l.done));
    for (const smc_engine::Readback &r : l.readbacks) std::memcpy(r.dst, l.h_read + r.off, r.bytes);
    l.readbacks.clear();
    l.read_off = 0;
    l.stage_off = 0;
    for (int p : l.particles) {
        e->h_lnl[p] = l.lnl[p];
        e->h_busy[p] = 0;
    }
    l.particles.clear();
    l.in_flight = false;
    return SMC_OK;
}
}  // extern "C++"

int smc_refresh_incremental(smc_engine *e, int first, int count, const uint8_t *node_flags,
                            const uint8_t *eig_flags, double *lnl_out, uint8_t *nonfinite_out)
{
    return refresh_incremental(e, first, count, node_flags, eig_flags, lnl_out, nonfinite_out, false);
}

int smc_accept_reject(smc_engine *e, int first, int count, const uint8_t *accept)
{
    if (!e) return SMC_ERR_ARG;
    PhaseTimer phase_timer(e, SMC_PHASE_ACCEPT_REJECT);
    if (first < 0 || count < 1 || first + count > e->Pmax || !accept)
        return fail(e, SMC_ERR_ARG, "invalid accept/reject arguments");
    const int T = e->T, Ti = T - 1, nn = e->nn;
    std::vector<int> eig_list, p_list, rej;
    std::vector<Item> p_nodes;
    for (int i = 0; i < count; ++i) {
        const int p = first + i;
        if (e->h_busy[p]) continue;  // recomputed on the lane: decided after lane_finish
        if (accept[i]) {
            smc_commit_range(e, p, 1);
            continue;
        }
        rej.push_back(p);
        std::copy(e->h_xflag_old.begin() + (size_t)p * Ti, e->h_xflag_old.begin() + (size_t)(p + 1) * Ti,
                  e->h_xflag.begin() + (size_t)p * Ti);
        e->h_lnl[p] = e->h_lnl_old[p];
        if (e->h_dirty[p] & 2) {
            if (e->h_eig_save[p] >= 0) {  // the committed eigensystems come back by copy
                HIP_TRY(e, hipMemcpyAsync(e->d_eig + (size_t)p * e->Kmax * kEigStride,
                                          e->d_eig_save + (size_t)e->h_eig_save[p] * e->Kmax * kEigStride,
                                          (size_t)e->h_nclass[p] * kEigStride * sizeof(double),
                                          hipMemcpyDeviceToDevice, e->stream));
                e->eig_save_free.push_back(e->h_eig_save[p]);
                e->h_eig_save[p] = -1;
                e->pend_full.push_back(p);
            } else {
                eig_list.push_back(p);
                p_list.push_back(p);
            }
        } else if (e->h_dirty[p] & 1) {
            for (int n : e->h_pnodes[p]) e->pend_nodes.push_back(Item{p, n});
        }
    }
    hipStream_t s = e->stream;
    // restore the committed parameters of rejected particles: one upload of the
    // span they cover (accepted particles in it are committed already)
    if (!rej.empty()) {
        const int rn = (int)rej.size();
        std::vector<int32_t> ch((size_t)rn * Ti * 2);
        std::vector<double> bl((size_t)rn * nn), base((size_t)rn * 190), pv(rn), sh(rn);
        for (int q = 0; q < rn; ++q) {
            const smc_engine::Committed &c = e->committed[rej[q]];
            std::copy(c.children.begin(), c.children.end(), ch.begin() + (size_t)q * Ti * 2);
            std::copy(c.blen.begin(), c.blen.end(), bl.begin() + (size_t)q * nn);
            std::copy(c.base.begin(), c.base.end(), base.begin() + (size_t)q * 190);
            pv[q] = c.pinv;
            sh[q] = c.shape;
        }
        int rc = upload_params(e, 0, rn, ch.data(), bl.data(), base.data(), pv.data(), sh.data(), rej.data(), true);
        if (rc != SMC_OK) return rc;
    }
    if (!p_list.empty() || !p_nodes.empty()) {
        int rc = set_single_wave(e, first, count);
        if (rc != SMC_OK) return rc;
        rc = recompute_models(e, first, eig_list, p_list, p_nodes);
        if (rc != SMC_OK) return rc;
    }
    for (int p : rej) {
        e->h_dirty[p] = 0;
        e->h_pnodes[p].clear();
    }
    int rc = h2d(e, e->d_xflag + (size_t)first * Ti, e->h_xflag.data() + (size_t)first * Ti, (size_t)count * Ti);
    if (rc == SMC_OK)
        rc = h2d(e, e->d_lnl + first, e->h_lnl.data() + first, (size_t)count * sizeof(double));
    if (rc != SMC_OK) return rc;
    // the restores run on while the caller draws its next proposals (the stream
    // orders them before any later work); timed launches need their events done
    if (e->timing) {
        rc = sync_stream(e);
        if (rc != SMC_OK) return rc;
        collect_timings(e);
    }
    return SMC_OK;
}

extern "C++" {
namespace {

// Gathers rows of a host mirror: row dst[k] <- old row src[k] (all sources read first).
template <typename V>
void gather_host(V &v, size_t row, const std::vector<int> &dst, const std::vector<int> &src)
{
    V tmp(src.size() * row);
    for (size_t k = 0; k < src.size(); ++k)
        std::copy_n(v.begin() + (size_t)src[k] * row, row, tmp.begin() + k * row);
    for (size_t k = 0; k < dst.size(); ++k) std::copy_n(tmp.begin() + k * row, row, v.begin() + (size_t)dst[k] * row);
}

}  // namespace
}

int smc_copy_particles(smc_engine *e, int n, const int32_t *parents)
{
    if (e) {
        int rc = ensure_programs(e);  // programs are copied with the other rows
        if (rc != SMC_OK) return rc;
    }
    if (!e) return SMC_ERR_ARG;
    if (n < 1 || n > e->Pmax || !parents) return fail(e, SMC_ERR_ARG, "invalid particle copy arguments");
    for (int i = 0; i < n; ++i) {
        if (parents[i] < 0 || parents[i] >= n)
            return fail(e, SMC_ERR_ARG, "copy source " + std::to_string(parents[i]) + " out of [0, n)");
        if (!e->h_valid[parents[i]]) return fail(e, SMC_ERR_STATE, "copy source slot not uploaded");
    }
    std::vector<int> dst, src;
    for (int i = 0; i < n; ++i)
        if (parents[i] != i) {
            dst.push_back(i);
            src.push_back(parents[i]);
        }
    if (dst.empty()) return SMC_OK;
    const int T = e->T, Ti = T - 1, nn = e->nn, S = e->S, Sp = e->S_pad, R = e->R, K = e->Kmax;
    const int m = (int)dst.size();
    // host mirrors
    gather_host(e->h_children, (size_t)Ti * 2, dst, src);
    gather_host(e->h_root, 1, dst, src);
    gather_host(e->h_height, (size_t)Ti, dst, src);
    gather_host(e->h_prog, (size_t)site_prog_stride(T), dst, src);
    gather_host(e->h_valid, 1, dst, src);
    gather_host(e->h_nclass, 1, dst, src);
    gather_host(e->h_blen, (size_t)nn, dst, src);
    gather_host(e->h_base, 190, dst, src);
    gather_host(e->h_logits, (size_t)K * 20, dst, src);
    gather_host(e->h_assign, (size_t)S, dst, src);
    gather_host(e->h_pinv, 1, dst, src);
    gather_host(e->h_shape, 1, dst, src);
    gather_host(e->h_rates, (size_t)R, dst, src);
    gather_host(e->h_nslots, 1, dst, src);
    gather_host(e->h_xflag_old, (size_t)Ti, dst, src);
    gather_host(e->h_lnl_old, 1, dst, src);
    {
        std::vector<smc_engine::Committed> tmp(m);
        for (int k = 0; k < m; ++k) tmp[k] = e->committed[src[k]];
        for (int k = 0; k < m; ++k) e->committed[dst[k]] = tmp[k];
    }
    for (int d : dst) {  // a copy is a committed state (copyFromSerialized ends in a refresh)
        std::copy_n(e->h_xflag_old.begin() + (size_t)d * Ti, Ti, e->h_xflag.begin() + (size_t)d * Ti);
        e->h_lnl[d] = e->h_lnl_old[d];
        e->h_dirty[d] = 0;
        e->h_pnodes[d].clear();
        e->h_touched[d] = 0;
    }
    for (int &o : e->h_owner)  // copied slots' resident partials are not their parents'
        if (o >= 0 && o < n && parents[o] != o) o = -1;
    // device rows: gather the sources into staging, then scatter to the destinations
    struct Buf {
        void *ptr;
        size_t row;
    };
    const Buf bufs[] = {{e->d_children, sizeof(int) * Ti * 2}, {e->d_blen, sizeof(double) * nn},
                        {e->d_base, sizeof(double) * 190},    {e->d_logits, sizeof(double) * K * 20},
                        {e->d_nclass, sizeof(int)},           {e->d_pinv, sizeof(double)},
                        {e->d_rates, sizeof(double) * R},     {e->d_root, sizeof(int)},
                        {e->d_perm, sizeof(int) * Sp},        {e->d_tiles, sizeof(Tile) * e->tile_stride},
                        {e->d_ntiles, sizeof(int)},           {e->d_prog, sizeof(int) * site_prog_stride(T)},
                        {e->d_eig, sizeof(double) * K * kEigStride}, {e->d_pcodes, (size_t)T * Sp},
                        {e->d_pweight, sizeof(double) * Sp}};
    size_t max_row = 0;
    for (const Buf &b : bufs) max_row = std::max(max_row, b.row);
    const int64_t need = (int64_t)m * (int64_t)max_row;
    if (need > e->gather_cap) {
        HIP_TRY(e, hipStreamSynchronize(e->stream));
        if (e->d_gather) (void)hipFree(e->d_gather);
        e->d_gather = nullptr;
        e->gather_cap = 0;
        const int64_t cap = std::max<int64_t>(need, (int64_t)1 << 20);
        HIP_TRY(e, hipMalloc((void **)&e->d_gather, (size_t)cap));
        e->gather_cap = cap;
    }
    // the row lists go through d_plist ([2 * Pmax]): sources, then destinations
    int *d_src = e->d_plist, *d_dst = e->d_plist + e->Pmax;
    int rc = h2d(e, d_src, src.data(), src.size() * sizeof(int));
    if (rc == SMC_OK) rc = h2d(e, d_dst, dst.data(), dst.size() * sizeof(int));
    if (rc != SMC_OK) return rc;
    for (const Buf &b : bufs) {
        launch_gather_rows(b.ptr, e->d_gather, b.row, d_src, nullptr, m, e->stream);
        launch_gather_rows(e->d_gather, b.ptr, b.row, nullptr, d_dst, m, e->stream);
    }
    HIP_TRY(e, hipGetLastError());
    // the committed workspace flags and lnL of the range
    rc = h2d(e, e->d_xflag, e->h_xflag.data(), (size_t)n * Ti);
    if (rc == SMC_OK) rc = h2d(e, e->d_lnl, e->h_lnl.data(), (size_t)n * sizeof(double));
    if (rc == SMC_OK) rc = sync_stream(e);
    if (rc != SMC_OK) return rc;
    e->class_version++;
    e->topo_version++;
    e->sched_first = -1;
    return SMC_OK;
}

extern "C++" {
int smc_refresh_stale(smc_engine *e, int first, int count)
{
    if (e->WS != 2 || count > e->Wmax) return smc_refresh_likelihood(e, first, count, nullptr, nullptr);
    const int T = e->T, nn = e->nn;
    std::vector<uint8_t> flags((size_t)count * nn, 0);
    int stale = 0;
    for (int i = 0; i < count; ++i)
        if (e->h_owner[i] != first + i) {
            stale++;
            std::fill(flags.begin() + (size_t)i * nn + T, flags.begin() + (size_t)(i + 1) * nn, 1);
        }
    if (stale == 0) return SMC_OK;
    if (stale == count) return smc_refresh_likelihood(e, first, count, nullptr, nullptr);
    for (int p = first; p < first + count; ++p)
        if (!e->h_valid[p]) return fail(e, SMC_ERR_STATE, "slot " + std::to_string(p) + " not uploaded");
    int rc = set_single_wave(e, first, count);
    if (rc != SMC_OK) return rc;
    // eigensystems and P of every slot, in this range's class layout; partials only of the stale
    std::vector<int> all(count);
    for (int i = 0; i < count; ++i) all[i] = first + i;
    rc = recompute_models(e, first, all, all);
    if (rc != SMC_OK) return rc;
    drop_pending(e, first, count);
    rc = smc_refresh_incremental(e, first, count, flags.data(), nullptr, nullptr, nullptr);
    if (rc != SMC_OK) return rc;
    smc_commit_range(e, first, count);
    for (int i = 0; i < count; ++i) e->h_owner[i] = first + i;
    return SMC_OK;
}

int smc_refresh_slots(smc_engine *e, int first, int count, const uint8_t *which)
{
    if (e->WS != 2 || count > e->Wmax) return fail(e, SMC_ERR_STATE, "slot refresh needs one resident wave");
    std::vector<uint8_t> flags((size_t)count * e->nn, 0);
    int rc = smc_refresh_incremental(e, first, count, flags.data(), which, nullptr, nullptr);
    if (rc != SMC_OK) return rc;
    for (int i = 0; i < count; ++i)
        if (which[i]) {
            smc_commit_range(e, first + i, 1);
            e->h_owner[i] = first + i;
        }
    return SMC_OK;
}
}  // extern "C++"

static int ensure_smc(smc_engine *e, int n)
{
    if (n <= e->smc_cap) return SMC_OK;
    void *bufs[] = {e->d_w, e->d_l, e->d_norm, e->d_out2, e->d_assign};
    for (void *b : bufs)
        if (b) (void)hipFree(b);
    e->d_w = e->d_l = e->d_norm = e->d_out2 = nullptr;
    e->d_assign = nullptr;
    e->smc_cap = 0;
    int rc = dev_alloc(e, &e->d_w, n);
    if (rc == SMC_OK) rc = dev_alloc(e, &e->d_l, n);
    if (rc == SMC_OK) rc = dev_alloc(e, &e->d_norm, n);
    if (rc == SMC_OK) rc = dev_alloc(e, &e->d_out2, 2);
    if (rc == SMC_OK) rc = dev_alloc(e, &e->d_assign, n);
    if (rc == SMC_OK) e->smc_cap = n;
    return rc;
}

int smc_compute_next_step(smc_engine *e, int n, double *raw_weights, const double *raw_lnl,
                          double *normalized, double *ess, int target_ess, double *current_temp,
                          double last_temp)
{
    if (!e) return SMC_ERR_ARG;
    if (n < 1 || !raw_weights || !raw_lnl || !normalized || !ess || !current_temp)
        return fail(e, SMC_ERR_ARG, "invalid computeNextStep arguments");
    int rc = ensure_smc(e, n);
    if (rc != SMC_OK) return rc;
    hipStream_t s = e->stream;
    HIP_TRY(e, hipMemcpyAsync(e->d_w, raw_weights, n * sizeof(double), hipMemcpyHostToDevice, s));
    HIP_TRY(e, hipMemcpyAsync(e->d_l, raw_lnl, n * sizeof(double), hipMemcpyHostToDevice, s));
    {
        TimedLaunch tl(e, SMC_K_SMC);
        launch_next_step(n, e->d_w, e->d_l, e->d_norm, e->d_out2, target_ess, last_temp, s);
    }
    HIP_TRY(e, hipGetLastError());
    double out2[2];
    HIP_TRY(e, hipMemcpyAsync(raw_weights, e->d_w, n * sizeof(double), hipMemcpyDeviceToHost, s));
    HIP_TRY(e, hipMemcpyAsync(normalized, e->d_norm, n * sizeof(double), hipMemcpyDeviceToHost, s));
    HIP_TRY(e, hipMemcpyAsync(out2, e->d_out2, 2 * sizeof(double), hipMemcpyDeviceToHost, s));
    HIP_TRY(e, hipStreamSynchronize(s));
    collect_timings(e);
    *current_temp = out2[0];
    *ess = out2[1];
    return SMC_OK;
}

int smc_systematic_resample(smc_engine *e, int n, const double *normalized, double u,
                            int32_t *assignments)
{
    if (!e) return SMC_ERR_ARG;
    if (n < 1 || !normalized || !assignments || !(u >= 0.0 && u < 1.0))
        return fail(e, SMC_ERR_ARG, "invalid resampling arguments");
    int rc = ensure_smc(e, n);
    if (rc != SMC_OK) return rc;
    hipStream_t s = e->stream;
    HIP_TRY(e, hipMemcpyAsync(e->d_norm, normalized, n * sizeof(double), hipMemcpyHostToDevice, s));
    {
        TimedLaunch tl(e, SMC_K_SMC);
        launch_resample(n, e->d_norm, u, e->d_assign, s);
    }
    HIP_TRY(e, hipGetLastError());
    HIP_TRY(e, hipMemcpyAsync(assignments, e->d_assign, n * sizeof(int), hipMemcpyDeviceToHost, s));
    HIP_TRY(e, hipStreamSynchronize(s));
    collect_timings(e);
    return SMC_OK;
}

int smc_site_class_lnl(smc_engine *e, int count, const int32_t *particles, const int32_t *sites,
                       int n_cand, const int32_t *ncand, const double *cand_logits,
                       double tempering, double *out)
{
    return smc_site_class_lnl_ex(e, count, particles, sites, n_cand, ncand, cand_logits, nullptr, tempering,
                                 out);
}

int smc_site_class_lnl_ex(smc_engine *e, int count, const int32_t *particles, const int32_t *sites,
                          int n_cand, const int32_t *ncand, const double *cand_logits,
                          const int32_t *cand_class, double tempering, double *out)
{
    if (!e) return SMC_ERR_ARG;
    PhaseTimer phase_timer(e, SMC_PHASE_SITE_CLASS);
    if (count < 1 || n_cand < 1 || !particles || !sites || !ncand || !cand_logits || !out)
        return fail(e, SMC_ERR_ARG, "invalid site-class arguments");
    if (!e->have_alignment) return fail(e, SMC_ERR_STATE, "no alignment uploaded");
    for (int i = 0; i < count; ++i) {
        if (particles[i] < 0 || particles[i] >= e->Pmax || !e->h_valid[particles[i]])
            return fail(e, SMC_ERR_ARG, "entry " + std::to_string(i) + ": particle not uploaded");
        if (sites[i] < 0 || sites[i] >= e->S)
            return fail(e, SMC_ERR_ARG, "entry " + std::to_string(i) + ": site out of range");
        if (ncand[i] < 0 || ncand[i] > n_cand)
            return fail(e, SMC_ERR_ARG, "entry " + std::to_string(i) + ": candidate count");
    }
    const int64_t need = (int64_t)count * n_cand;
    int rc = ensure_site_class_buffers(e, count, need);
    if (rc != SMC_OK) return rc;
    std::vector<int> ent((size_t)3 * count);
    for (int i = 0; i < count; ++i) {
        ent[i] = particles[i];
        ent[count + i] = sites[i];
        ent[2 * count + i] = ncand[i];
    }
    rc = h2d(e, e->d_sc_ent, ent.data(), ent.size() * sizeof(int));
    if (rc == SMC_OK) rc = h2d(e, e->d_sc_logits, cand_logits, (size_t)need * 20 * sizeof(double));
    if (rc == SMC_OK && cand_class) rc = h2d(e, e->d_sc_cls, cand_class, (size_t)need * sizeof(int));
    if (rc == SMC_OK) rc = issue_site_class(e, count, n_cand, cand_class != nullptr, true, tempering);
    if (rc == SMC_OK) rc = d2h(e, out, e->d_sc_out, (size_t)need * sizeof(double));
    if (rc == SMC_OK) rc = sync_stream(e);
    if (rc != SMC_OK) return rc;
    collect_timings(e);
    return SMC_OK;
}

int smc_discretize_gamma(double shape, int num, double *out)
{
    if (!(shape > 0.0) || num < 1 || !out) return SMC_ERR_ARG;
    discretize_gamma(shape, num, out);
    return SMC_OK;
}

void smc_lg_base_matrix(double out[190])
{
    if (out) std::memcpy(out, kLG190, sizeof(double) * 190);
}

int smc_engine_set_variant(smc_engine *e, int family, int variant)
{
    if (!e) return SMC_ERR_ARG;
    if (family != SMC_K_PRUNE ||
        (variant != SMC_PRUNE_MFMA && variant != SMC_PRUNE_VALU && variant != SMC_PRUNE_MFMA_LEVELS &&
         variant != SMC_PRUNE_MFMA_WALK))
        return fail(e, SMC_ERR_ARG, "unknown kernel variant");
    e->prune_variant = variant;
    return SMC_OK;
}

int smc_engine_set_patterns(smc_engine *e, int compress)
{
    if (!e) return SMC_ERR_ARG;
    if (compress != 0 && compress != 1) return fail(e, SMC_ERR_ARG, "compress must be 0 or 1");
    e->compress = compress != 0;
    return SMC_OK;
}

int64_t smc_pattern_count(const smc_engine *e, int first, int count)
{
    if (!e || first < 0 || count < 0 || first + count > e->Pmax) return -1;
    int64_t n = 0;
    for (int p = first; p < first + count; ++p) n += e->h_nslots[p];
    return n;
}

int smc_engine_set_mixture(smc_engine *e, int mode)
{
    if (!e) return SMC_ERR_ARG;
    if (mode != SMC_MIX_REFERENCE && mode != SMC_MIX_LOGSPACE)
        return fail(e, SMC_ERR_ARG, "unknown mixture mode");
    e->mix_mode = mode;
    return SMC_OK;
}

int smc_engine_set_lane(smc_engine *e, int enabled)
{
    if (!e) return SMC_ERR_ARG;
    bool done = true;
    int rc = lane_finish(e, true, &done);
    if (rc != SMC_OK) return rc;
    e->lane_enabled = enabled != 0;
    return SMC_OK;
}

int smc_engine_set_timing(smc_engine *e, int enabled)
{
    if (!e) return SMC_ERR_ARG;
    e->timing = enabled != 0;
    return SMC_OK;
}

int smc_engine_kernel_stats(smc_engine *e, int family, int64_t *launches, double *total_ms)
{
    if (!e) return SMC_ERR_ARG;
    if (family < 0 || family >= SMC_K_COUNT) return fail(e, SMC_ERR_ARG, "bad kernel family");
    if (launches) *launches = e->launches[family];
    if (total_ms) *total_ms = e->total_ms[family];
    return SMC_OK;
}

int smc_engine_reset_stats(smc_engine *e)
{
    if (!e) return SMC_ERR_ARG;
    for (int i = 0; i < SMC_K_COUNT; ++i) {
        e->launches[i] = 0;
        e->total_ms[i] = 0.0;
    }
    return SMC_OK;
}

}  // extern "C"

int smc_engine_phase_stats(smc_engine *e, double *ms, int64_t *calls)
{
    if (!e) return SMC_ERR_ARG;
    for (int k = 0; k < SMC_PHASE_COUNT; ++k) {
        if (ms) ms[k] = e->phase_ms[k];
        if (calls) calls[k] = e->phase_calls[k];
    }
    return SMC_OK;
}

int smc_engine_reset_phase_stats(smc_engine *e)
{
    if (!e) return SMC_ERR_ARG;
    for (int k = 0; k < SMC_PHASE_COUNT; ++k) {
        e->phase_ms[k] = 0.0;
        e->phase_calls[k] = 0;
    }
    return SMC_OK;
}

int smc_refresh_eigen(smc_engine *e, int count, const int32_t *slots)
{
    if (count < 1) return SMC_OK;
    if (count > e->Pmax) return fail(e, SMC_ERR_ARG, "too many slots");
    hipStream_t s = e->stream;
    HIP_TRY(e, hipMemcpyAsync(e->d_plist, slots, (size_t)count * sizeof(int), hipMemcpyHostToDevice, s));
    {
        TimedLaunch tl(e, SMC_K_EIGEN);
        EigenArgs ea{e->d_base, e->d_logits, e->d_nclass, e->d_eig, e->Kmax, 0, count, e->d_plist};
        launch_eigen(ea, s);
    }
    HIP_TRY(e, hipGetLastError());
    HIP_TRY(e, hipStreamSynchronize(s));
    collect_timings(e);
    return SMC_OK;
}

int smc_engine_set_graphs(smc_engine *e, int enable)
{
    if (!e) return SMC_ERR_ARG;
    e->use_graph = enable != 0;
    return SMC_OK;
}

int smc_engine_graph_stats(smc_engine *e, int64_t *captures, int64_t *launches)
{
    if (!e) return SMC_ERR_ARG;
    if (captures) *captures = e->graph_captures;
    if (launches) *launches = e->graph_launches;
    return SMC_OK;
}

int smc_engine_get_tuning(smc_engine *e, smc_tuning *t)
{
    if (!e || !t) return SMC_ERR_ARG;
    t->blocks_per_cu = e->tune.blocks_per_cu;
    for (int k = 0; k < 3; ++k) t->unroll[k] = e->tune.unroll[k];
    t->xcd_groups = e->tune.xcd_groups;
    t->unroll_runtime = e->tune.unroll_rt;
    t->chain_max_nodes = e->tune.chain_max;
    t->ms_per_sweep = 0.0;
    t->configs_tried = 0;
    return SMC_OK;
}

int smc_engine_set_tuning(smc_engine *e, const smc_tuning *t)
{
    if (!e || !t) return SMC_ERR_ARG;
    if (t->blocks_per_cu < 1 || t->blocks_per_cu > 64) return fail(e, SMC_ERR_ARG, "blocks_per_cu out of 1..64");
    for (int k = 0; k < 4; ++k) {
        const int u = k < 3 ? t->unroll[k] : t->unroll_runtime;
        if (u != 1 && u != 2 && u != 4) return fail(e, SMC_ERR_ARG, "unroll must be 1, 2 or 4");
    }
    if (t->xcd_groups != 1 && t->xcd_groups != 2 && t->xcd_groups != 4 && t->xcd_groups != 8)
        return fail(e, SMC_ERR_ARG, "xcd_groups must be 1, 2, 4 or 8");
    if (t->chain_max_nodes < 0) return fail(e, SMC_ERR_ARG, "chain_max_nodes must be >= 0");
    e->tune.blocks_per_cu = t->blocks_per_cu;
    for (int k = 0; k < 3; ++k) e->tune.unroll[k] = t->unroll[k];
    e->tune.xcd_groups = t->xcd_groups;
    e->tune.unroll_rt = t->unroll_runtime;
    e->tune.chain_max = t->chain_max_nodes;
    return SMC_OK;
}

// Measured search over the launch shape of the full sweep: the persistent grid
// (workgroups per CU) and the column-block unroll of each child kind, by
// coordinate descent from the current setting.  Each candidate is timed over
// `sweeps` replays of its captured sweep of slots [first, first+count) after one
// untimed replay; the fastest is kept.  The particles' lnL are recomputed
// (the results do not depend on the launch shape).
int smc_engine_autotune(smc_engine *e, int first, int count, int sweeps, smc_tuning *out)
{
    if (!e) return SMC_ERR_ARG;
    if (sweeps < 1 || first < 0 || count < 1 || first + count > e->Pmax)
        return fail(e, SMC_ERR_ARG, "invalid autotune arguments");
    if (e->prune_variant == SMC_PRUNE_VALU) return fail(e, SMC_ERR_STATE, "autotune needs the MFMA variant");
    for (int p = first; p < first + count; ++p)
        if (!e->h_valid[p]) return fail(e, SMC_ERR_STATE, "slot " + std::to_string(p) + " not uploaded");
    const bool timing = e->timing, graphs = e->use_graph;
    e->timing = false;
    e->use_graph = true;
    hipEvent_t t0 = nullptr, t1 = nullptr;
    HIP_TRY(e, hipEventCreate(&t0));
    HIP_TRY(e, hipEventCreate(&t1));
    int tried = 0;
    auto measure = [&](const smc_engine::Tune &t, double *ms) -> int {
        e->tune = t;
        int rc = smc_refresh_likelihood(e, first, count, nullptr, nullptr);  // captures, warms
        if (rc != SMC_OK) return rc;
        HIP_TRY(e, hipEventRecord(t0, e->stream));
        for (int k = 0; k < sweeps; ++k) {
            rc = run_sweep(e, first, count);
            if (rc != SMC_OK) return rc;
        }
        HIP_TRY(e, hipEventRecord(t1, e->stream));
        HIP_TRY(e, hipEventSynchronize(t1));
        float f = 0.f;
        HIP_TRY(e, hipEventElapsedTime(&f, t0, t1));
        *ms = (double)f / sweeps;
        tried++;
        return SMC_OK;
    };
    smc_engine::Tune best = e->tune;
    double best_ms = 0.0;
    int rc = measure(best, &best_ms);
    const int grids[] = {2, 4, 8, 16, 32, 64};
    const int unrolls[] = {1, 2, 4};
    if (rc == SMC_OK)
        for (int g : grids) {
            if (g == best.blocks_per_cu) continue;
            smc_engine::Tune t = best;
            t.blocks_per_cu = g;
            double ms;
            if ((rc = measure(t, &ms)) != SMC_OK) break;
            if (ms < best_ms) {
                best_ms = ms;
                best = t;
            }
        }
    for (int x : {1, 8}) {
        if (x == best.xcd_groups || rc != SMC_OK) continue;
        smc_engine::Tune t = best;
        t.xcd_groups = x;
        double ms;
        if ((rc = measure(t, &ms)) != SMC_OK) break;
        if (ms < best_ms) {
            best_ms = ms;
            best = t;
        }
    }
    for (int k = 0; k < 4 && rc == SMC_OK; ++k)
        for (int u : unrolls) {
            int &cur = k < 3 ? best.unroll[k] : best.unroll_rt;
            if (u == cur) continue;
            smc_engine::Tune t = best;
            (k < 3 ? t.unroll[k] : t.unroll_rt) = u;
            double ms;
            if ((rc = measure(t, &ms)) != SMC_OK) break;
            if (ms < best_ms) {
                best_ms = ms;
                best = t;
            }
        }
    (void)hipEventDestroy(t0);
    (void)hipEventDestroy(t1);
    e->tune = best;
    e->timing = timing;
    e->use_graph = graphs;
    if (rc != SMC_OK) return rc;
    // leave the particles as a sweep with the chosen shape computed them
    rc = smc_refresh_likelihood(e, first, count, nullptr, nullptr);
    if (rc != SMC_OK) return rc;
    if (out) {
        out->blocks_per_cu = best.blocks_per_cu;
        for (int k = 0; k < 3; ++k) out->unroll[k] = best.unroll[k];
        out->xcd_groups = best.xcd_groups;
        out->unroll_runtime = best.unroll_rt;
        out->chain_max_nodes = best.chain_max;
        out->ms_per_sweep = best_ms;
        out->configs_tried = tried;
    }
    return SMC_OK;
}

int smc_pool_refill(smc_engine *e, int count, const int32_t *slots, int q,
                    const std::function<void(int, double *)> &draw, int fresh)
{
    ProfilePool &pl = e->pool;
    hipStream_t s = e->stream;
    const int64_t n = (int64_t)count * q;
    auto grow = [&](auto *&ptr, int64_t &cap, int64_t need) -> int {
        if (need <= cap) return SMC_OK;
        if (ptr) (void)hipFree(ptr);
        ptr = nullptr;
        cap = 0;
        const int64_t c2 = std::max<int64_t>(need, 1024);
        HIP_TRY(e, hipMalloc((void **)&ptr, (size_t)c2 * sizeof(*ptr)));
        cap = c2;
        return SMC_OK;
    };
    int rc = grow(pl.d_eig, pl.cap, n * kEigStride);
    if (rc == SMC_OK) rc = grow(pl.d_logits, pl.lcap, n * 20);
    if (rc == SMC_OK) rc = grow(pl.d_ent, pl.ecap, 2 * (int64_t)count);
    if (rc != SMC_OK) return rc;
    // unconsumed profiles stay first (a slot uses its draws in order, whatever
    // the batch), then each slot draws as many new ones as it consumed
    const bool keep = fresh == 0 && pl.count == count && pl.q == q && (int)pl.next.size() == count;
    std::vector<double> lg((size_t)n * 20, 0.0);
    for (int i = 0; i < count; ++i) {
        int j = 0;
        if (keep)
            for (int k = pl.next[i]; k < q; ++k, ++j)
                std::copy_n(&pl.logits[((size_t)i * q + k) * 20], 20, &lg[((size_t)i * q + j) * 20]);
        for (; j < q; ++j) draw(i, &lg[((size_t)i * q + j) * 20]);
    }
    pl.logits.swap(lg);
    pl.count = count;
    pl.q = q;
    pl.next.assign(count, 0);
    std::vector<int> ent(2 * (size_t)count);
    for (int i = 0; i < count; ++i) {
        ent[i] = slots[i];
        ent[count + i] = q;
    }
    HIP_TRY(e, hipMemcpyAsync(pl.d_ent, ent.data(), ent.size() * sizeof(int), hipMemcpyHostToDevice, s));
    HIP_TRY(e, hipMemcpyAsync(pl.d_logits, pl.logits.data(), pl.logits.size() * sizeof(double),
                              hipMemcpyHostToDevice, s));
    {
        TimedLaunch tl(e, SMC_K_EIGEN);
        launch_eigen_cand(e->d_base, pl.d_ent, pl.d_ent + count, pl.d_logits, count, q, pl.d_eig, nullptr,
                          e->d_eig, e->Kmax, nullptr, s);
    }
    HIP_TRY(e, hipGetLastError());
    HIP_TRY(e, hipStreamSynchronize(s));
    collect_timings(e);
    return SMC_OK;
}

int smc_pool_to_class(smc_engine *e, int slot, int pool_index, int cls)
{
    if (cls < 0 || cls >= e->Kmax) return fail(e, SMC_ERR_STATE, "no free class slot");
    HIP_TRY(e, hipMemcpyAsync(e->d_eig + ((size_t)slot * e->Kmax + cls) * kEigStride,
                              e->pool.d_eig + (size_t)pool_index * kEigStride, kEigStride * sizeof(double),
                              hipMemcpyDeviceToDevice, e->stream));
    return SMC_OK;
}
