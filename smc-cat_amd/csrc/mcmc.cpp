// Batched MCMC rejuvenation of particles at a tempering T.
//
// Reference: the rejuvenation loop of main.cpp:293-313 (per particle: MCMC of
// r iterations, then with probability alg8Probability the CAT Gibbs move and a
// forced refresh), Mcmc::operator() (Mcmc.cpp:25-73: move chosen by weight,
// repeated numGibbs times, each repetition a Metropolis-Hastings step with
// ratio (T lnL' + prior') - (T lnL + prior) + log Hastings, accepted iff
// log u <= ratio), the move table of main.cpp:54-105, the moves of
// Particle.cpp:677-783 / Tree.cpp:553-811 and Particle::lnPrior
// (Particle.cpp:484-517).
//
// All particles of the range step in lockstep: each MH step of every particle
// that still has repetitions left is one proposal; one incremental refresh
// (smc_refresh_incremental) evaluates all proposals at once, one
// smc_accept_reject commits or reverts them.  Differences from the reference
// (each stated where it happens): adaptive NNI proposes a plain NNI (the split
// posteriors it is biased by live outside this engine); SPR marks the old and
// the new path to the root dirty (the reference marks only the new one,
// Tree.cpp:795-800); rate-matrix moves recompute the eigensystem (the
// reference keeps stale eigenvectors, Q13); the stationary move, a no-op in
// the reference (Q14), is skipped.
#include <algorithm>
#include <array>
#include <cmath>
#include <cstring>
#include <optional>

#include "engine_state.hpp"
#include "host_rng.hpp"
#include "moves.hpp"

namespace {

double log_gamma_pdf(double k, double theta, double x)
{
    return (k - 1.0) * std::log(x) - x / theta - std::lgamma(k) - k * std::log(theta);
}
double log_std_normal(double z) { return -0.5 * z * z - 0.9189385332046727; }

enum Move { TREE = 0, SHAPE = 1, INVAR = 2, RATES = 3, N_MOVES = 4 };

// The move table of main.cpp:54-105: weights, repetitions per iteration and the
// weighted draw (stationary moves are no-ops and skipped: -1).  The draws come from
// a stream of their own (seed ^ kMoveStream), so a chain's step count is known
// before it runs.
constexpr uint64_t kMoveStream = 0x6d6f766573ull;
struct MoveTable {
    double w[N_MOVES];
    double wtot;
    int reps[N_MOVES];
    template <typename G>
    int draw(G &g) const
    {
        const double u = g.uniform() * wtot;
        double c = 0.0;
        for (int k = 0; k < N_MOVES; ++k) {
            c += w[k];
            if (u < c) return k;
        }
        return -1;
    }
};

MoveTable move_table(const smc_engine *e, const smc_rejuvenation_options &o)
{
    MoveTable m{};
    const double w[N_MOVES] = {3.0, e->R > 1 ? 1.0 : 0.0, o.invariant ? 1.0 : 0.0, o.using_lg ? 0.0 : 1.0};
    const int reps[N_MOVES] = {std::max(1, (int)(e->T * 0.25)), 2, 2, 5};
    m.wtot = 1.0;  // the stationary move's weight
    for (int k = 0; k < N_MOVES; ++k) {
        m.w[k] = w[k];
        m.wtot += w[k];
        m.reps[k] = reps[k];
    }
    return m;
}

using Params = MoveParams;

// Particle::lnPrior (Particle.cpp:484-517), split in two: the terms of the
// tree, shape and base matrix, which the moves change, and the CRP and class
// terms, which they do not (computed once per particle and added back so the
// value matches the reference's).
double ln_prior_params(const smc_engine *e, const Params &q, bool using_lg)
{
    const int nn = e->nn;
    double len = 0.0;
    for (int n = 0; n < nn; ++n) len += q.bl[n];  // the root's length is 0
    double lp = log_gamma_pdf((double)(nn - 1), 0.1, len);
    if (e->R > 1) lp += log_gamma_pdf(3.0, 2.0, q.shape);
    if (!using_lg)
        for (double b : q.base) lp += log_std_normal(b);
    return lp;
}

double ln_prior_classes(const smc_engine *e, int p, double dpp_alpha)
{
    const int S = e->S, Kmax = e->Kmax;
    const int K = e->h_nclass[p];
    double lp = std::log(dpp_alpha) * K;
    std::vector<int> sz(K, 0);
    for (int s = 0; s < S; ++s) sz[e->h_assign[(size_t)p * S + s]]++;
    for (int k = 0; k < K; ++k) {
        for (int j = 0; j < 20; ++j) lp += log_std_normal(e->h_logits[((size_t)p * Kmax + k) * 20 + j]);
        lp += std::lgamma((double)sz[k]);
    }
    lp += std::lgamma(dpp_alpha) - std::lgamma(dpp_alpha + S);
    return lp;
}

}  // namespace

extern "C" int smc_rejuvenate(smc_engine *e, int first, int count, double tempering,
                              const smc_rejuvenation_options *opt, double *lnl_out,
                              int64_t *stats_out)
{
    return smc_rejuvenate_ex(e, first, count, tempering, opt, lnl_out, stats_out, nullptr);
}

extern "C" int smc_rejuvenation_steps(smc_engine *e, const smc_rejuvenation_options *opt, int count,
                                      const int32_t *keys, int32_t *steps_out)
{
    if (!e) return SMC_ERR_ARG;
    if (!opt || count < 0 || (count > 0 && (!keys || !steps_out)) || opt->iterations < 0)
        return smc_fail(e, SMC_ERR_ARG, "invalid rejuvenation-steps arguments");
    const MoveTable mt = move_table(e, *opt);
    for (int i = 0; i < count; ++i) {
        Rng g(opt->seed ^ kMoveStream, (uint64_t)(uint32_t)keys[i]);
        int32_t n = 0;
        for (int it = 0; it < opt->iterations; ++it) {
            const int m = mt.draw(g);
            n += m >= 0 ? mt.reps[m] : 0;
        }
        steps_out[i] = n;
    }
    return SMC_OK;
}

int smc_cat_moves(smc_engine *e, const std::vector<int32_t> &slots, double tempering,
                  const smc_rejuvenation_options &o, int64_t *cat_stats)
{
    if (slots.empty()) return SMC_OK;
    PhaseTimer cat_timer(e, SMC_PHASE_CAT_MOVE);
    const int n = (int)slots.size();
    const uint64_t seed = o.seed ^ 0x61626273ull;
    int64_t st[5] = {};
    int rc;
    if (o.allocation_proposal && !o.record_allocations) {
        rc = smc_learned_partition_move(o.allocation_proposal, n, slots.data(), tempering, seed, 0, o.dpp_alpha,
                                        -1.0, st);
        if (cat_stats) {
            cat_stats[0] += st[0];
            cat_stats[1] += st[2];
        }
    } else {
        rc = o.allocation_proposal
                 ? smc_mlp_collect_gibbs(o.allocation_proposal, n, slots.data(), tempering, seed, 0, o.dpp_alpha, st)
                 : smc_gibbs_partition_move(e, n, slots.data(), tempering, seed, 0, o.dpp_alpha, st);
        if (cat_stats) {
            cat_stats[0] += (int64_t)n * (int64_t)(e->S / 2.0);
            cat_stats[1] += st[0];
        }
    }
    return rc;
}

int smc_rejuvenate_ex(smc_engine *e, int first, int count, double tempering,
                      const smc_rejuvenation_options *opt, double *lnl_out, int64_t *stats_out,
                      int64_t *cat_stats, std::vector<int32_t> *deferred_cat)
{
    if (!e) return SMC_ERR_ARG;
    if (!opt || first < 0 || count < 1 || first + count > e->Pmax || opt->iterations < 0)
        return smc_fail(e, SMC_ERR_ARG, "invalid rejuvenation arguments");
    if (e->WS != 2) return smc_fail(e, SMC_ERR_STATE, "rejuvenation needs workspaces = 2");
    const smc_rejuvenation_options &o = *opt;
    const int T = e->T, Ti = T - 1, nn = e->nn;
    // full refresh of the freshly copied particles (copyFromSerialized, Particle.cpp:276);
    // slots whose committed partials are still resident keep them
    int rc = smc_refresh_stale(e, first, count);
    if (rc != SMC_OK) return rc;

    const MoveTable mt = move_table(e, o);
    const int *reps = mt.reps;

    std::vector<Rng> rng, mrng;
    rng.reserve(count);
    mrng.reserve(count);
    for (int i = 0; i < count; ++i) {
        const uint64_t key = o.stream_keys ? (uint64_t)(uint32_t)o.stream_keys[i] : (uint64_t)(first + i);
        rng.emplace_back(o.seed, key);
        mrng.emplace_back(o.seed ^ kMoveStream, key);
    }
    std::vector<double> lnl(e->h_lnl.begin() + first, e->h_lnl.begin() + first + count);
    std::vector<double> prior(count);
    auto load = [&](int p) {
        const smc_engine::Committed &c = e->committed[p];
        return Params{c.children, c.blen, c.base, c.pinv, c.shape};
    };
    std::vector<double> prior_cls(count);
    for (int i = 0; i < count; ++i) {
        prior_cls[i] = ln_prior_classes(e, first + i, o.dpp_alpha);
        prior[i] = ln_prior_params(e, load(first + i), o.using_lg) + prior_cls[i];
    }

    std::vector<int32_t> ch((size_t)count * Ti * 2);
    std::vector<double> bl((size_t)count * nn), base((size_t)count * 190), pv(count), sh(count);
    std::vector<uint8_t> flags((size_t)count * nn), eig(count), acc(count);
    std::vector<double> hastings(count), new_prior(count), lnl_new(count);
    std::vector<int> move(count), left(count);
    int64_t proposed[N_MOVES] = {}, accepted[N_MOVES] = {};

    // Each particle walks its own sequence of iterations (move drawn by weight,
    // then its repetitions); a round advances every particle that still has
    // steps by one MH step, so a particle never waits for another's longer move
    // (the chains are the ones a per-particle loop would run).
    std::vector<int> started(count, 0);
    auto next_move = [&](int i) {
        while (left[i] <= 0 && started[i] < o.iterations) {
            move[i] = mt.draw(mrng[i]);
            left[i] = move[i] >= 0 ? reps[move[i]] : 0;
            started[i]++;
        }
    };
    for (int i = 0; i < count; ++i) {
        left[i] = 0;
        next_move(i);
    }
    // the rows handed to each round: a particle's committed state, or its proposal
    // while it moves (rows of particles that do not move are left as they are, and
    // the parameter upload sends only rows that changed)
    auto put_committed = [&](int i) {
        const smc_engine::Committed &c = e->committed[first + i];
        std::copy(c.children.begin(), c.children.end(), ch.begin() + (size_t)i * Ti * 2);
        std::copy(c.blen.begin(), c.blen.end(), bl.begin() + (size_t)i * nn);
        std::copy(c.base.begin(), c.base.end(), base.begin() + (size_t)i * 190);
        pv[i] = c.pinv;
        sh[i] = c.shape;
    };
    for (int i = 0; i < count; ++i) put_committed(i);
    Params q;  // one proposal at a time (its vectors keep their capacity)
    std::vector<uint8_t> f(nn);
    // Whole-tree steps (shape and rate-matrix moves: every P, and for the latter
    // every eigensystem, then every node) run on the engine's lane: their refresh is
    // issued on its own stream and collected when it is done, while the other
    // particles go on with their rounds.  A particle waits while the lane is busy
    // with other particles.  Each particle's chain is the same as without the lane
    // (its random numbers are drawn in the same order), so the results are too.  When
    // nothing but waiting whole-tree steps is left beside a busy lane, they run in a
    // round of their own.
    const bool lane_ok = count > 1 && e->lane_enabled && !e->timing;
    std::vector<uint8_t> role(count, 0), on_lane(count, 0), flags_l((size_t)count * nn), eig_l(count);
    enum { IDLE = 0, MAIN = 1, LANE = 2 };
    auto whole_tree = [&](int i) { return move[i] == SHAPE || move[i] == RATES; };
    // the MH decision of particle i on its new lnL (Mcmc.cpp:62)
    auto decide = [&](int i, double lnl_prop) {
        const double ratio = (tempering * lnl_prop + new_prior[i]) - (tempering * lnl[i] + prior[i]) + hastings[i];
        const double lu = std::log(rng[i].uniform());
        acc[i] = (lu <= ratio) ? 1 : 0;
        proposed[move[i]]++;
        if (acc[i]) {
            accepted[move[i]]++;
            lnl[i] = lnl_prop;
            prior[i] = new_prior[i];
        }
        left[i]--;
    };
    for (;;) {
        bool active = false;
        for (int i = 0; i < count; ++i) active |= left[i] > 0;
        if (!active) break;
        const bool lane_free = lane_ok && !e->lane.in_flight;
        int n_main = 0, n_lane = 0;
        int n_wait = 0;  // whole-tree steps waiting for the lane
        for (int i = 0; i < count; ++i) {
            role[i] = IDLE;
            if (left[i] <= 0 || on_lane[i]) continue;
            if (lane_ok && whole_tree(i)) {
                if (lane_free) {
                    role[i] = LANE;
                    n_lane++;
                } else {
                    n_wait++;
                }
            } else {
                role[i] = MAIN;
                n_main++;
            }
        }
        if (n_main == 0 && n_wait > 0)  // nothing else to run beside the lane: the waiting
            for (int i = 0; i < count; ++i)  // whole-tree steps take a round of their own
                if (role[i] == IDLE && left[i] > 0 && !on_lane[i]) {
                    role[i] = MAIN;
                    n_main++;
                }
        if (n_main + n_lane > 0) {
            std::optional<PhaseTimer> propose_timer(std::in_place, e, SMC_PHASE_PROPOSE);
            std::fill(flags.begin(), flags.end(), 0);
            std::fill(flags_l.begin(), flags_l.end(), 0);
            std::fill(eig.begin(), eig.end(), 0);
            std::fill(eig_l.begin(), eig_l.end(), 0);
            for (int i = 0; i < count; ++i) {
                if (role[i] == IDLE) continue;
                const int p = first + i;
                uint8_t &eg = role[i] == LANE ? eig_l[i] : eig[i];
                hastings[i] = 0.0;
                new_prior[i] = prior[i];
                const smc_engine::Committed &c = e->committed[p];
                q.ch.assign(c.children.begin(), c.children.end());
                q.bl.assign(c.blen.begin(), c.blen.end());
                q.base.assign(c.base.begin(), c.base.end());
                q.pinv = c.pinv;
                q.shape = c.shape;
                std::fill(f.begin(), f.end(), 0);
                Rng &g = rng[i];
                switch (move[i]) {
                    case TREE:
                        hastings[i] = tree_move(g, q, T, f, o);
                        break;
                    case SHAPE:
                        hastings[i] = shape_move(g, q, f, o.shape_delta);
                        break;
                    case INVAR:
                        hastings[i] = invar_move(g, q, o.invar_alpha);
                        break;
                    case RATES:  // eigensystems recomputed (Q13 fixed)
                        hastings[i] = rate_matrix_move(g, q, o);
                        eg = 1;
                        break;
                }
                std::copy(q.ch.begin(), q.ch.end(), ch.begin() + (size_t)i * Ti * 2);
                std::copy(q.bl.begin(), q.bl.end(), bl.begin() + (size_t)i * nn);
                std::copy(q.base.begin(), q.base.end(), base.begin() + (size_t)i * 190);
                pv[i] = q.pinv;
                sh[i] = q.shape;
                std::copy(f.begin(), f.end(), (role[i] == LANE ? flags_l : flags).begin() + (size_t)i * nn);
                new_prior[i] = ln_prior_params(e, q, o.using_lg) + prior_cls[i];
            }
            propose_timer.reset();
            rc = smc_update_params(e, first, count, ch.data(), bl.data(), base.data(), pv.data(), sh.data());
            if (rc != SMC_OK) return rc;
        }
        if (n_lane > 0) {
            rc = refresh_incremental(e, first, count, flags_l.data(), eig_l.data(), nullptr, nullptr, true);
            if (rc != SMC_OK) return rc;
            for (int i = 0; i < count; ++i)
                if (role[i] == LANE) on_lane[i] = 1;
        }
        if (n_main > 0) {
            rc = smc_refresh_incremental(e, first, count, flags.data(), eig.data(), lnl_new.data(), nullptr);
            if (rc != SMC_OK) return rc;
            for (int i = 0; i < count; ++i) {
                acc[i] = 1;  // a particle without a proposal: commit is a no-op
                if (role[i] == MAIN) decide(i, lnl_new[i]);
            }
            rc = smc_accept_reject(e, first, count, acc.data());  // the lane's particles are left alone
            if (rc != SMC_OK) return rc;
            for (int i = 0; i < count; ++i) {
                if (role[i] != MAIN) continue;
                if (!acc[i]) put_committed(i);  // the rejected proposal's rows go back
                if (left[i] <= 0) next_move(i);
            }
        }
        if (e->lane.in_flight) {  // collected when done, or waited for when nothing else runs
            bool done = false;
            rc = lane_finish(e, n_main == 0, &done);
            if (rc != SMC_OK) return rc;
            if (done) {
                for (int i = 0; i < count; ++i) {
                    acc[i] = 1;
                    if (on_lane[i]) decide(i, e->h_lnl[first + i]);
                }
                rc = smc_accept_reject(e, first, count, acc.data());
                if (rc != SMC_OK) return rc;
                for (int i = 0; i < count; ++i) {
                    if (!on_lane[i]) continue;
                    on_lane[i] = 0;
                    if (!acc[i]) put_committed(i);
                    if (left[i] <= 0) next_move(i);
                }
            }
        } else if (n_main + n_lane == 0) {
            return smc_fail(e, SMC_ERR_STATE, "rejuvenation made no progress");
        }
    }
    // with probability gibbs_probability: the CAT Gibbs move and a forced refresh
    // (main.cpp:303-305)
    if (o.gibbs_probability > 0.0) {
        std::vector<int32_t> slots;
        for (int i = 0; i < count; ++i)
            if (rng[i].uniform() < o.gibbs_probability) slots.push_back(first + i);
        if (deferred_cat) {
            deferred_cat->insert(deferred_cat->end(), slots.begin(), slots.end());
        } else if (!slots.empty()) {
            rc = smc_cat_moves(e, slots, tempering, o, cat_stats);
            if (rc != SMC_OK) return rc;
            rc = smc_refresh_likelihood(e, first, count, nullptr, nullptr);
            if (rc != SMC_OK) return rc;
            std::copy(e->h_lnl.begin() + first, e->h_lnl.begin() + first + count, lnl.begin());
        }
    }
    if (lnl_out) std::copy(lnl.begin(), lnl.end(), lnl_out);
    if (stats_out)
        for (int k = 0; k < N_MOVES; ++k) {
            stats_out[2 * k] = proposed[k];
            stats_out[2 * k + 1] = accepted[k];
        }
    return SMC_OK;
}

extern "C" int smc_ln_prior(smc_engine *e, int slot, int using_lg, double dpp_alpha, double *out)
{
    if (!e || !out) return SMC_ERR_ARG;
    if (slot < 0 || slot >= e->Pmax || !e->h_valid[slot])
        return smc_fail(e, SMC_ERR_ARG, "slot not uploaded");
    if (!(dpp_alpha > 0.0)) return smc_fail(e, SMC_ERR_ARG, "dpp_alpha must be > 0");
    const int Ti = e->T - 1, nn = e->nn;
    Params q{std::vector<int32_t>(e->h_children.begin() + (size_t)slot * Ti * 2,
                                  e->h_children.begin() + (size_t)(slot + 1) * Ti * 2),
             std::vector<double>(e->h_blen.begin() + (size_t)slot * nn, e->h_blen.begin() + (size_t)(slot + 1) * nn),
             std::vector<double>(e->h_base.begin() + (size_t)slot * 190,
                                 e->h_base.begin() + (size_t)(slot + 1) * 190),
             e->h_pinv[slot], e->h_shape[slot]};
    *out = ln_prior_params(e, q, using_lg != 0) + ln_prior_classes(e, slot, dpp_alpha);
    return SMC_OK;
}
