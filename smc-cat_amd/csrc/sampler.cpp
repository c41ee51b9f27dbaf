// The SMC sampler loop: adaptive tempering, resampling, proposal tuning and
// batched rejuvenation of the whole population on the engine.
//
// Reference: main.cpp:143-336 (loop, ESS check, systematic resampling,
// tuning, rejuvenation), computeNextStep (Miscellaneous.hpp:42-109).  The
// per-thread Particle objects of the reference become slot ranges of one
// engine; the serialised copy of a parent into its child slot
// (copyFromSerialized, Particle.cpp:218-276) becomes a device-side gather of the
// parents' states (smc_copy_particles), and only slots whose partials are not
// resident are recomputed.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <vector>

#include "engine_state.hpp"
#include "host_rng.hpp"


extern "C" int smc_run_sampler(smc_engine *e, const smc_sampler_options *opt, smc_sampler_result *res,
                               double *lnl_out, double *weights_out)
{
    if (!e) return SMC_ERR_ARG;
    if (!opt || !res) return smc_fail(e, SMC_ERR_ARG, "null sampler options or result");
    if (opt->allocation_proposal && (opt->proposal_warmup < 0 || opt->train_epochs < 0 ||
                                     (opt->train_epochs > 0 && !(opt->train_lr > 0.0))))
        return smc_fail(e, SMC_ERR_ARG, "invalid proposal training options");
    const smc_sampler_options &o = *opt;
    const int N = o.num_particles;
    if (N < 1 || N > e->Pmax || o.rejuvenation_iterations < 0 || o.batch < 0 || o.reserved != 0 ||
        !(o.rejuvenation_threshold >= 0.0))
        return smc_fail(e, SMC_ERR_ARG, "invalid sampler options");
    if (e->WS != 2) return smc_fail(e, SMC_ERR_STATE, "the sampler needs workspaces = 2");
    for (int p = 0; p < N; ++p)
        if (!e->h_valid[p]) return smc_fail(e, SMC_ERR_STATE, "slot without an initial particle");
    const int batch = std::min(o.batch > 0 ? o.batch : e->Wmax, e->Wmax);
    *res = smc_sampler_result{};

    using clk = std::chrono::steady_clock;
    const auto t0 = clk::now();
    double rejuv_s = 0.0;
    smccat::Rng g(o.seed, 0xFFFFFFFFull);

    std::vector<double> lnl(N), w(N, -std::log((double)N)), norm(N);
    int rc = smc_refresh_likelihood(e, 0, N, lnl.data(), nullptr);
    if (rc != SMC_OK) return rc;

    std::vector<int32_t> assign(N);
    std::vector<double> chol;
    const double rej_threshold = o.rejuvenation_threshold * N;
    const int target = (int)rej_threshold;  // bound to const int& in the reference (Q17)
    double last = 0.0, ess = 0.0;
    int it = 0;
    for (it = 1; last < 1.0; ++it) {
        if (o.max_iterations > 0 && it > o.max_iterations) break;
        double cur = last;
        rc = smc_compute_next_step(e, N, w.data(), lnl.data(), norm.data(), &ess, target, &cur, last);
        if (rc != SMC_OK) return rc;
        res->iterations = it;
        res->ess_sum += ess;
        if (std::isnan(ess)) return smc_fail(e, SMC_ERR_STATE, "NaN ESS (a particle likelihood failed)");
        if (o.on_iteration) o.on_iteration(o.user, it, cur, ess);
        if (ess <= rej_threshold && cur != 1.0) {
            rc = smc_systematic_resample(e, N, norm.data(), g.uniform(), assign.data());
            if (rc != SMC_OK) return rc;
            // copy every parent into its new slot, on the device
            rc = smc_copy_particles(e, N, assign.data());
            if (rc != SMC_OK) return rc;
            // tuning from the resampled population (main.cpp:213-289), on the GPU
            double tune[4];
            chol.assign(190 * 190, 0.0);
            rc = smc_proposal_tuning(e, N, o.using_lg, tune, o.using_lg ? nullptr : chol.data());
            if (rc != SMC_OK) return rc;
            smc_rejuvenation_options ro{};
            ro.iterations = o.rejuvenation_iterations;
            ro.using_lg = o.using_lg;
            ro.invariant = o.invariant;
            ro.seed = o.seed * 0x9E3779B97F4A7C15ull + (uint64_t)it;
            ro.shape_delta = std::max(0.5 * tune[0], 0.01);
            ro.scale_delta = std::max(tune[1], 0.01);
            ro.subtree_scale_delta = ro.scale_delta;
            ro.invar_alpha = tune[2] * 0.5;
            ro.gibbs_probability = o.cat ? o.alg8_probability : 0.0;
            ro.dpp_alpha = o.dpp_alpha;
            ro.rate_full_scale = 0.0;
            // the learned CAT move: Gibbs with recording during the warm-up steps
            const bool learned = o.cat && o.allocation_proposal;
            ro.allocation_proposal = learned ? o.allocation_proposal : nullptr;
            ro.record_allocations = learned && res->resamples < o.proposal_warmup;
            if (!o.using_lg) {
                // the uncentred second moment of the base matrices (Q21), scaled, shrunk
                // towards its mean diagonal and factored (smc_proposal_tuning)
                ro.rate_cholesky = chol.data();
                ro.rate_delta = std::max(tune[3], 0.01);
            }
            const auto r0 = clk::now();
            // Batched (batch > 1): the CAT moves of every batch run together after the
            // last batch, one lockstep move over all chosen particles of the population
            // (its per-site steps are latency-bound, so many particles per step cost
            // little more than a few).  A slot's move depends only on the slot and the
            // seed, so the particles end as they would with the move inside each batch.
            // batch = 1 keeps the reference's shape: each particle's move right after
            // its MCMC.
            std::vector<int32_t> cat_slots;
            // Stream keys sorted by their chains' MH step counts: a lockstep batch runs
            // as many rounds as its longest chain has steps, so chains of similar length
            // share a batch.  The keys of a slot do not depend on the batch size.
            std::vector<int32_t> keys(N), steps(N);
            for (int n = 0; n < N; ++n) keys[n] = n;
            rc = smc_rejuvenation_steps(e, &ro, N, keys.data(), steps.data());
            if (rc != SMC_OK) return rc;
            std::stable_sort(keys.begin(), keys.end(), [&](int32_t a, int32_t b) { return steps[a] > steps[b]; });
            for (int first = 0; first < N; first += batch) {
                const int cnt = std::min(batch, N - first);
                int64_t st[8];
                ro.stream_keys = keys.data() + first;
                rc = smc_rejuvenate_ex(e, first, cnt, cur, &ro, lnl.data() + first, st, res->cat_stats,
                                       batch > 1 ? &cat_slots : nullptr);
                if (rc != SMC_OK) return rc;
                for (int k = 0; k < 8; ++k) res->move_stats[k] += st[k];
            }
            if (!cat_slots.empty()) {
                rc = smc_cat_moves(e, cat_slots, cur, ro, res->cat_stats);
                if (rc != SMC_OK) return rc;
                // the moved slots' likelihoods, batch by batch (each batch resident in turn)
                std::vector<uint8_t> moved(N, 0);
                for (int32_t sl : cat_slots) moved[sl] = 1;
                for (int first = 0; first < N; first += batch) {
                    const int cnt = std::min(batch, N - first);
                    if (std::none_of(moved.begin() + first, moved.begin() + first + cnt, [](uint8_t v) { return v; }))
                        continue;
                    rc = smc_refresh_slots(e, first, cnt, moved.data() + first);
                    if (rc != SMC_OK) return rc;
                }
                std::copy(e->h_lnl.begin(), e->h_lnl.begin() + N, lnl.begin());
            }
            rejuv_s += std::chrono::duration<double>(clk::now() - r0).count();
            if (ro.record_allocations && res->resamples + 1 == o.proposal_warmup) {
                const auto t1 = clk::now();
                res->cat_stats[2] = smc_mlp_num_examples(o.allocation_proposal);
                if (res->cat_stats[2] > 0 && o.train_epochs > 0) {
                    rc = smc_mlp_train(o.allocation_proposal, o.train_epochs, 4096, o.train_lr, o.seed ^ 0x6c6561726eull,
                                       nullptr);
                    if (rc != SMC_OK) return rc;
                }
                res->training_seconds += std::chrono::duration<double>(clk::now() - t1).count();
            }
            std::fill(w.begin(), w.end(), -std::log((double)N));
            res->resamples++;
        }
        last = cur;
    }
    res->final_temp = last;
    res->final_ess = ess;
    res->seconds = std::chrono::duration<double>(clk::now() - t0).count();
    res->rejuvenation_seconds = rejuv_s;
    if (lnl_out) std::copy(lnl.begin(), lnl.end(), lnl_out);
    if (weights_out) std::copy(w.begin(), w.end(), weights_out);
    return SMC_OK;
}
