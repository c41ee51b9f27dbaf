// CAT site-allocation Gibbs move, batched over particles.
//
// Reference: Particle::gibbsPartitionMove(tempering) (Particle.cpp:785-995),
// Neal (2000) Algorithm 8 with m = 5 auxiliary classes and numChar/2 random
// site updates.  Per iteration and particle: draw a site, remove it from its
// class (erasing the class if it empties, :817-821), draw m auxiliary profiles
// from the logit prior (19 N(0,1) logits and a 0, TransitionProbabilityClass.cpp:26-35),
// score the site under every class and auxiliary (smc_site_class_lnl, one batched
// launch for all particles), weight tempering * log L + log n_k for existing
// classes and + log(alpha / m) for auxiliaries (:933-940), draw the class
// (:943-978).  The per-rate tip index and rescale stride of the reference
// (Q15) are the corrected ones.  Random numbers come from a counter-based
// Philox4x32-10 stream keyed by (seed, particle slot), so results do not depend
// on batching or thread schedules.
//
// Two differences make each update an exact Algorithm-8 step:
//  * a site that was alone in its class keeps that class's profile as the first
//    auxiliary (Neal 2000, Algorithm 8); the reference draws all five afresh;
//  * classes are capped at max_classes, so the target is the DP posterior
//    restricted to K <= max_classes: with K classes left after the removal the
//    update uses m' = min(5, max_classes - K) auxiliaries, each weighted
//    log(alpha / m') (a valid Algorithm 8 with m' auxiliaries), and with m' = 0
//    the site can only join an existing class (the exact conditional of the
//    restricted target).  The learned move targets the same restriction.
#include <algorithm>
#include <array>
#include <cmath>
#include <cstring>

#include "engine_state.hpp"

namespace {

// Philox4x32-10 (Salmon, Moraes, Dror & Shaw, "Parallel random numbers: as easy
// as 1, 2, 3", SC 2011).
struct Philox {
    uint32_t key[2];
    uint64_t counter = 0;

    Philox(uint64_t seed, uint64_t stream)
    {
        key[0] = (uint32_t)(seed ^ (stream * 0x9E3779B97F4A7C15ull));
        key[1] = (uint32_t)((seed >> 32) ^ stream);
    }
    void block(uint32_t out[4])
    {
        uint32_t c[4] = {(uint32_t)counter, (uint32_t)(counter >> 32), 0x5eed1e55u, 0xca7u};
        uint32_t k0 = key[0], k1 = key[1];
        for (int r = 0; r < 10; r++) {
            const uint64_t p0 = (uint64_t)0xD2511F53u * c[0];
            const uint64_t p1 = (uint64_t)0xCD9E8D57u * c[2];
            const uint32_t n0 = (uint32_t)(p1 >> 32) ^ c[1] ^ k0;
            const uint32_t n2 = (uint32_t)(p0 >> 32) ^ c[3] ^ k1;
            c[0] = n0;
            c[1] = (uint32_t)p1;
            c[2] = n2;
            c[3] = (uint32_t)p0;
            k0 += 0x9E3779B9u;
            k1 += 0xBB67AE85u;
        }
        counter++;
        std::memcpy(out, c, sizeof(c));
    }
    // uniform in [0, 1) with 53 random bits
    double uniform()
    {
        uint32_t b[4];
        block(b);
        const uint64_t v = ((uint64_t)b[0] << 32 | b[1]) >> 11;
        return (double)v * (1.0 / 9007199254740992.0);
    }
    double normal()  // Box-Muller
    {
        const double u1 = uniform(), u2 = uniform();
        return std::sqrt(-2.0 * std::log1p(-u1)) * std::cos(6.283185307179586 * u2);
    }
};

struct PState {
    int K;
    std::vector<std::array<double, 20>> logits;
    std::vector<int> orig;  // engine class slot holding the class's eigensystem
    std::vector<int> free;  // engine class slots no class uses
    std::vector<int> assign;
    std::vector<int> size;
};

}  // namespace

extern "C" int smc_gibbs_partition_move(smc_engine *e, int count, const int32_t *slots,
                                        double tempering, uint64_t seed, int num_iter,
                                        double dpp_alpha, int64_t *stats_out)
{
    return smc_gibbs_move(e, count, slots, tempering, seed, num_iter, dpp_alpha, stats_out, nullptr);
}

int smc_gibbs_move(smc_engine *e, int count, const int32_t *slots, double tempering, uint64_t seed,
                   int num_iter, double dpp_alpha, int64_t *stats_out, AllocationSink *sink)
{
    if (!e) return SMC_ERR_ARG;
    if (count < 1 || !slots || !(dpp_alpha >= 0.0) || !(tempering >= 0.0))
        return smc_fail(e, SMC_ERR_ARG, "invalid Gibbs move arguments");
    const int S = e->S, Kmax = e->Kmax;
    const int m = 5;  // auxiliary classes (Particle.cpp:790)
    for (int i = 0; i < count; ++i)
        if (slots[i] < 0 || slots[i] >= e->Pmax || !e->h_valid[slots[i]])
            return smc_fail(e, SMC_ERR_ARG, "Gibbs move: slot not uploaded");
    if (S < 2) return smc_fail(e, SMC_ERR_ARG, "Gibbs move needs at least two sites");
    const int iters = num_iter > 0 ? num_iter : (int)(S / 2.0);  // Particle.cpp:791

    std::vector<PState> st(count);
    std::vector<Philox> rng;
    rng.reserve(count);
    for (int i = 0; i < count; ++i) {
        const int p = slots[i];
        PState &ps = st[i];
        ps.K = e->h_nclass[p];
        ps.logits.resize(ps.K);
        for (int k = 0; k < ps.K; ++k)
            for (int j = 0; j < 20; ++j) ps.logits[k][j] = e->h_logits[((size_t)p * Kmax + k) * 20 + j];
        ps.assign.assign(e->h_assign.begin() + (size_t)p * S, e->h_assign.begin() + (size_t)(p + 1) * S);
        ps.size.assign(ps.K, 0);
        for (int a : ps.assign) ps.size[a]++;
        ps.orig.resize(ps.K);
        for (int k = 0; k < ps.K; ++k) ps.orig[k] = k;
        for (int k = Kmax - 1; k >= ps.K; --k) ps.free.push_back(k);
        rng.emplace_back(seed, (uint64_t)p);
    }
    // existing classes are scored with the engine's eigensystems: make them current
    int rc0 = smc_refresh_eigen(e, count, slots);
    if (rc0 != SMC_OK) return rc0;
    // auxiliary profiles come from a pool of prior draws decomposed in batches
    // (each slot's own stream, consumed in draw order)
    // (a stream of their own per slot, so refills do not shift the site draws)
    const int pool_q = m * 16;
    std::vector<Philox> prof_rng;
    prof_rng.reserve(count);
    for (int i = 0; i < count; ++i) prof_rng.emplace_back(seed ^ 0x70726f66696c65ull, (uint64_t)slots[i]);
    auto draw = [&](int i, double *lg) {
        for (int j = 0; j < 19; ++j) lg[j] = prof_rng[i].normal();
        lg[19] = 0.0;
    };
    rc0 = smc_pool_refill(e, count, slots, pool_q, draw, 1);
    if (rc0 != SMC_OK) return rc0;
    const int n_cand = Kmax + m;
    std::vector<int32_t> ent_p(count), ent_s(count), ent_n(count);
    std::vector<double> cand((size_t)count * n_cand * 20, 0.0), ll((size_t)count * n_cand);
    std::vector<int32_t> cand_cls((size_t)count * n_cand, -1);
    // per particle of the current iteration: auxiliaries, and the engine class slot and
    // profile of a class the removal emptied (-1: none), which is auxiliary 0
    std::vector<int> naux(count, 0), single_slot(count, -1);
    std::vector<std::array<double, 20>> single_logits(count);
    int64_t moved = 0, created = 0, erased = 0;

    for (int it = 0; it < iters; ++it) {
        for (int i = 0; i < count; ++i)
            if (e->pool.next[i] + m > pool_q) {  // a slot is running out: top every slot up
                int rc = smc_pool_refill(e, count, slots, pool_q, draw, 0);
                if (rc != SMC_OK) return rc;
                break;
            }
        for (int i = 0; i < count; ++i) {
            PState &ps = st[i];
            Philox &g = rng[i];
            const int site = std::min(S - 1, (int)(g.uniform() * S));
            const int k0 = ps.assign[site];
            ps.size[k0]--;
            ps.assign[site] = -1;
            int was = k0;
            single_slot[i] = -1;
            if (ps.size[k0] == 0) {  // erase the emptied class (Particle.cpp:819-821)
                was = -1;
                single_slot[i] = ps.orig[k0];
                single_logits[i] = ps.logits[k0];
                ps.logits.erase(ps.logits.begin() + k0);
                ps.free.push_back(ps.orig[k0]);
                ps.orig.erase(ps.orig.begin() + k0);
                ps.size.erase(ps.size.begin() + k0);
                for (int &a : ps.assign)
                    if (a > k0) a--;
                ps.K--;
            }
            // auxiliary classes cannot exceed the engine's class capacity (an emptied
            // class leaves at least one slot)
            const int n_aux = std::min(m, Kmax - ps.K);
            naux[i] = n_aux;
            ent_p[i] = slots[i];
            ent_s[i] = site;
            ent_n[i] = ps.K + n_aux;
            double *c = &cand[(size_t)i * n_cand * 20];
            int32_t *cc = &cand_cls[(size_t)i * n_cand];
            for (int k = 0; k < ps.K; ++k) {
                std::copy(ps.logits[k].begin(), ps.logits[k].end(), c + 20 * k);
                cc[k] = ps.orig[k];
            }
            int a0 = 0;
            if (single_slot[i] >= 0) {  // the emptied class's profile, scored from its eigensystem
                std::copy(single_logits[i].begin(), single_logits[i].end(), c + 20 * ps.K);
                cc[ps.K] = single_slot[i];
                a0 = 1;
            }
            for (int a = a0; a < n_aux; ++a) {
                const int slot = e->pool.next[i]++;
                cc[ps.K + a] = -2 - (i * pool_q + slot);
                std::copy_n(&e->pool.logits[((size_t)i * pool_q + slot) * 20], 20, c + 20 * (ps.K + a));
            }
            ps.assign[site] = -2 - was;  // remembers the previous class until the draw
        }
        int rc = smc_site_class_lnl_ex(e, count, ent_p.data(), ent_s.data(), n_cand, ent_n.data(),
                                       cand.data(), cand_cls.data(), tempering, ll.data());
        if (rc != SMC_OK) return rc;
        for (int i = 0; i < count; ++i) {
            PState &ps = st[i];
            Philox &g = rng[i];
            const int site = ent_s[i];
            const int nc = ent_n[i];
            const double *l = &ll[(size_t)i * n_cand];
            std::vector<double> w(nc);
            double mx = -INFINITY;
            const double log_aux = naux[i] > 0 ? std::log(dpp_alpha / naux[i]) : -INFINITY;
            for (int k = 0; k < nc; ++k) {
                w[k] = l[k] + (k < ps.K ? std::log((double)ps.size[k]) : log_aux);
                if (w[k] > mx) mx = w[k];
            }
            double tot = 0.0;
            for (int k = 0; k < nc; ++k) {
                w[k] = std::exp(w[k] - mx);
                tot += w[k];
            }
            const double draw = tot * g.uniform();
            int pick = nc - 1;
            double acc = 0.0;
            for (int k = 0; k < nc; ++k) {
                acc += w[k];
                if (acc > draw) {
                    pick = k;
                    break;
                }
            }
            const int was = -2 - ps.assign[site];
            if (sink) sink->add(site, ps.logits, ps.size, std::min(pick, ps.K), tempering, dpp_alpha);
            bool stayed = false;
            if (pick >= ps.K) {  // a new class from the chosen auxiliary
                std::array<double, 20> lg;
                std::copy(&cand[((size_t)i * n_cand + pick) * 20],
                          &cand[((size_t)i * n_cand + pick) * 20] + 20, lg.begin());
                ps.logits.push_back(lg);
                // the free slot pushed last is the emptied class's, if any: the site's
                // own profile keeps its eigensystem there, a pool profile's moves in
                const int f = ps.free.back();
                ps.free.pop_back();
                const int32_t code = cand_cls[(size_t)i * n_cand + pick];
                if (code >= 0) {
                    if (code != f) return smc_fail(e, SMC_ERR_STATE, "Gibbs move: class slot bookkeeping");
                    stayed = true;
                } else {
                    int rc = smc_pool_to_class(e, slots[i], -2 - code, f);
                    if (rc != SMC_OK) return rc;
                    created++;
                }
                ps.orig.push_back(f);
                ps.size.push_back(0);
                pick = ps.K++;
            }
            if (single_slot[i] >= 0 && !stayed) erased++;
            ps.assign[site] = pick;
            ps.size[pick]++;
            if (pick != was && !stayed) moved++;
        }
    }
    // write the new allocations back and re-upload the particles
    std::vector<double> lg((size_t)Kmax * 20);
    for (int i = 0; i < count; ++i) {
        std::fill(lg.begin(), lg.end(), 0.0);
        for (int k = 0; k < st[i].K; ++k) std::copy(st[i].logits[k].begin(), st[i].logits[k].end(), &lg[(size_t)k * 20]);
        int rc = smc_upload_partition(e, slots[i], st[i].K, lg.data(), st[i].assign.data());
        if (rc != SMC_OK) return rc;
    }
    if (stats_out) {
        stats_out[0] = moved;
        stats_out[1] = created;
        stats_out[2] = erased;
    }
    return SMC_OK;
}

int smc_upload_partition(smc_engine *e, int slot, int K, const double *logits, const int32_t *assign)
{
    const int T = e->T, nn = e->nn;
    std::vector<int32_t> ch(e->h_children.begin() + (size_t)slot * (T - 1) * 2,
                            e->h_children.begin() + (size_t)(slot + 1) * (T - 1) * 2);
    const double pinv = e->h_pinv[slot], shape = e->h_shape[slot];
    std::vector<double> bl(e->h_blen.begin() + (size_t)slot * nn, e->h_blen.begin() + (size_t)(slot + 1) * nn);
    std::vector<double> base(e->h_base.begin() + (size_t)slot * 190, e->h_base.begin() + (size_t)(slot + 1) * 190);
    return smc_set_particles(e, slot, 1, ch.data(), bl.data(), base.data(), &K, logits, assign, &pinv, &shape);
}

extern "C" int smc_get_particle_state(smc_engine *e, int slot, int32_t *num_classes,
                                      double *logits, int32_t *assignment)
{
    if (!e) return SMC_ERR_ARG;
    if (slot < 0 || slot >= e->Pmax || !e->h_valid[slot])
        return smc_fail(e, SMC_ERR_ARG, "slot not uploaded");
    const int S = e->S, Kmax = e->Kmax;
    if (num_classes) *num_classes = e->h_nclass[slot];
    if (logits)
        std::copy(e->h_logits.begin() + (size_t)slot * Kmax * 20,
                  e->h_logits.begin() + (size_t)(slot + 1) * Kmax * 20, logits);
    if (assignment)
        std::copy(e->h_assign.begin() + (size_t)slot * S, e->h_assign.begin() + (size_t)(slot + 1) * S,
                  assignment);
    return SMC_OK;
}

extern "C" int smc_get_particle_params(smc_engine *e, int slot, int32_t *children, double *branch_len,
                                       double *base, double *pinv, double *shape)
{
    if (!e) return SMC_ERR_ARG;
    if (slot < 0 || slot >= e->Pmax || !e->h_valid[slot])
        return smc_fail(e, SMC_ERR_ARG, "slot not uploaded");
    const int Ti = e->T - 1, nn = e->nn;
    if (children)
        std::copy(e->h_children.begin() + (size_t)slot * Ti * 2,
                  e->h_children.begin() + (size_t)(slot + 1) * Ti * 2, children);
    if (branch_len)
        std::copy(e->h_blen.begin() + (size_t)slot * nn, e->h_blen.begin() + (size_t)(slot + 1) * nn,
                  branch_len);
    if (base)
        std::copy(e->h_base.begin() + (size_t)slot * 190, e->h_base.begin() + (size_t)(slot + 1) * 190,
                  base);
    if (pinv) *pinv = e->h_pinv[slot];
    if (shape) *shape = e->h_shape[slot];
    return SMC_OK;
}
