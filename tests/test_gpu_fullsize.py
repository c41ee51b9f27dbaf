"""The CAT configurations at their full sizes on the GPU.

c4: 200 taxa x 10,000 sites, GTR+G4+I, 4,096 particles with K ~ U{5..120}, FP32
pruning, partials resident for about 240 particles at a time (multi-wave), the
launch shape autotuned, the sweep captured and replayed as a graph.  Every lnL is
finite or flagged non-finite; replays equal the eager sweep bit for bit; particles
from the first, a middle and the last wave match the FP64 oracle within the fast
path's bound; the P kernel of the large waves (256 node-rate pairs per workgroup)
gives the same likelihoods, bit for bit, as the one of small waves (32 pairs).

c5: the learned CAT-allocation proposal on full c3 data (100 taxa x 5,000 sites,
LG+G4, DP-CAT): Gibbs draws recorded on the GPU, the network trained on them by
the GPU kernels, held-out loss below the CRP-only proposal's, acceptance above
it, and the exact site scores of the move's candidates equal to the oracle's."""
import sys

import numpy as np
import pytest

import oracle_binding as orc
from smccat import Engine, ProposalNetwork, lg_base_matrix
from smccat import simulate as sim

pytestmark = pytest.mark.gpu

RTOL_FP32 = 1e-5


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _c4_sizes(N, T, S, R, kr):
    per_particle = (T - 1) * R * S * (20 * 4 + 4)
    slots = int(min(N, max(1, 160e9 // per_particle)))
    return slots, int(slots * (kr[0] + kr[1]) // 2 * 1.25)


def test_c4_full_population_waves_graph_autotune(gpu):
    codes, cfg = sim.simulate_config("c4")
    T, S, R, kr = cfg["taxa"], cfg["sites"], cfg["rates"], cfg["k_range"]
    N = cfg["particles"]
    batch = sim.prior_particles(np.random.default_rng(404), N, T, S, num_rates=R, invar=True, k_range=kr)
    slots, pcls = _c4_sizes(N, T, S, R, kr)
    _say("case c4 full size: %d particles, K %s, waves of %d particles" % (N, kr, slots))
    with Engine(T, S, R, N, kr[1], "fp32", partial_slots=slots, pmat_classes=pcls) as eng:
        eng.set_mixture("logspace")
        eng.set_alignment(codes)
        eng.set_particles(batch)
        tune = eng.autotune(0, slots, sweeps=1)
        _say("  autotuned on one wave: %s" % tune)
        eng.set_graphs(False)
        eager, nf = eng.refresh_likelihood()
        eng.set_graphs(True)
        for rep in range(2):
            got, nf2 = eng.refresh_likelihood()
            assert np.array_equal(got, eager), rep
            assert np.array_equal(nf2, nf), rep
        st = eng.graph_stats()
        assert st["launches"] >= 2
    bad = np.flatnonzero(~np.isfinite(eager))
    _say("  %d non-finite particles %s" % (len(bad), bad[:8].tolist()))
    assert np.array_equal(nf.astype(bool), ~np.isfinite(eager))
    assert len(bad) <= N // 100
    for i in bad[:1]:  # flagged exactly where the oracle's value is not finite either
        assert not np.isfinite(orc.particle_lnl_mode(codes, batch.particle(i), R, 1)[0])
    # one particle (the fewest classes) of the first, a middle and the last wave
    checks = []
    for lo in (0, (N // slots // 2) * slots, (N - 1) // slots * slots):
        ids = [i for i in range(lo, min(N, lo + slots)) if np.isfinite(eager[i])]
        checks.append(min(ids, key=lambda i: batch.num_classes[i]))
    for i in checks:
        want = orc.particle_lnl_mode(codes, batch.particle(i), R, 1)[0]
        _say("  particle %d (K %d): fp32 %.6f oracle %.6f rel %.2e" %
             (i, batch.num_classes[i], eager[i], want, abs(eager[i] - want) / abs(want)))
        assert eager[i] == pytest.approx(want, rel=RTOL_FP32)
    # the first 64 particles in waves of 16: the P kernel's 32-pair form (small waves)
    # against the 256-pair form the waves above ran
    sub = batch.take(np.arange(64))
    with Engine(T, S, R, 64, kr[1], "fp32", partial_slots=16,
                pmat_classes=int(16 * kr[1])) as small:
        small.set_mixture("logspace")
        small.set_alignment(codes)
        small.set_particles(sub)
        small.set_tuning(tune["blocks_per_cu"], tune["unroll"], tune["xcd_groups"], tune["unroll_runtime"])
        small.set_graphs(False)
        part, _ = small.refresh_likelihood()
    assert np.array_equal(part, eager[:64])


def test_c5_learned_proposal_on_full_c3_data(gpu):
    codes, cfg = sim.simulate_config("c3")
    T, S, R, kr = cfg["taxa"], cfg["sites"], cfg["rates"], cfg["k_range"]
    N = 96
    kmax = kr[1] + 8
    batch = sim.prior_particles(np.random.default_rng(505), N, T, S, num_rates=R, lg=True,
                                lg_table=lg_base_matrix(), k_range=kr, max_classes=kmax)
    _say("case c5: c3 data %d taxa x %d sites, %d particles, K %s" % (T, S, N, kr))
    with Engine(T, S, R, N, kmax, "fp32", partial_slots=8) as eng:
        eng.set_alignment(codes)
        eng.set_particles(batch)
        # exact site scores (the learned move's two-candidate entries use this kernel)
        p, sites = 3, np.array([0, 1234, 4999])
        K0 = int(batch.num_classes[p])
        cand = np.stack([batch.logits[p, [0, K0 - 1]] for _ in sites])
        got = eng.site_class_lnl(np.full(len(sites), p), sites, cand, tempering=0.7)
        for j, s in enumerate(sites):
            want = orc.site_class_lnl(codes, batch.particle(p), R, int(s), cand[j], 0.7)
            np.testing.assert_allclose(got[j], want, rtol=1e-10)
        with ProposalNetwork(eng, seed=1) as net, ProposalNetwork(eng, seed=2) as crp, \
                ProposalNetwork(eng, seed=3) as ev:
            for k in range(2):
                net.collect_gibbs(range(N), 1.0, seed=100 + k, dpp_alpha=1.0)
            n = net.num_examples
            assert n == 2 * N * (S // 2)
            # held out: the draws of a later Gibbs move, recorded by a network never trained
            ev.collect_gibbs(range(N), 1.0, seed=300, num_iter=200, dpp_alpha=1.0)
            hold = ev.num_examples
            _say("  %d recorded Gibbs draws for training, %d held out" % (n, hold))
            ev.params = np.zeros(net.num_params, dtype=np.float32)  # CRP weights only
            crp_loss, _ = ev.loss_grad(0, hold)
            losses = net.train(4, batch=4096, lr=2e-3, seed=1)
            ev.params = net.params
            after, _ = ev.loss_grad(0, hold)
            _say("  held-out loss: CRP-only %.3f trained %.3f; epochs %s" % (crp_loss, after, np.round(losses, 3)))
            assert after < 0.8 * crp_loss
            crp.params = np.zeros(net.num_params, dtype=np.float32)
            st_t = net.move(range(N), 1.0, seed=7, num_iter=150, dpp_alpha=1.0)
            st_c = crp.move(range(N), 1.0, seed=7, num_iter=150, dpp_alpha=1.0)
    acc_t = st_t[2] / max(st_t[1], 1)
    acc_c = st_c[2] / max(st_c[1], 1)
    _say("  acceptance: trained %.3f (%d moves), CRP-only %.3f (%d moves)" % (acc_t, st_t[2], acc_c, st_c[2]))
    assert acc_t > 1.3 * acc_c and st_t[2] > st_c[2]


def test_c2_smc_full_size(gpu):
    """The SMC loop (main.cpp:143-336) at the c2 workload's full size: 64 taxa x 2,000
    sites, GTR+G4+I, 1,024 particles resident in two FP64 workspaces, 8 MCMC
    iterations per rejuvenation, three reweighting steps (each lands at the ESS
    target, 0.6 N, and so resamples).  Every lnL is finite, the temperature rises, every
    move kind is proposed and accepted, and sampled final states match the oracle
    within the plain path's 1e-10."""
    codes, cfg = sim.simulate_config("c2")
    T, S, R, N = cfg["taxa"], cfg["sites"], cfg["rates"], cfg["particles"]
    _say("case c2 smc full size: %d taxa x %d sites, %d particles, 8 iterations, 3 steps" % (T, S, N))
    batch = sim.prior_particles(np.random.default_rng(606), N, T, S, num_rates=R, invar=True)
    with Engine(T, S, R, N, 1, "fp64", workspaces=2) as eng:
        eng.set_alignment(codes)
        eng.set_particles(batch)
        res, lnl, w = eng.run_sampler(N, 8, seed=3, max_iterations=3)
        _say("  %s" % {k: v for k, v in res.items() if k not in ("move_stats", "cat_stats")})
        assert res["iterations"] == 3 and res["resamples"] >= 2  # ESS lands at the target
        assert 0.0 < res["final_temp"] < 1.0 and np.isfinite(res["final_ess"])
        assert np.all(res["move_stats"][:, 0] > 0) and np.all(res["move_stats"][:, 1] > 0)
        assert np.all(np.isfinite(lnl)) and np.all(np.isfinite(w))
        check = [0, 511, 1023]
        final = eng.snapshot(check, batch)
    for j, i in enumerate(check):
        want = orc.particle_lnl(codes, final.particle(j), R)
        assert lnl[i] == pytest.approx(want, rel=1e-10)
