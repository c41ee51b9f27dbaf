"""The SMC sampler loop (main.cpp:143-336) on the engine: it reaches T = 1,
resamples, and leaves a population whose likelihoods equal the oracle's on the
final states; the GTR run exercises the tuned full rate-matrix move (Cholesky
of the population's base-matrix moments, main.cpp:249-283)."""
import sys

import numpy as np
import pytest

import oracle_binding as orc
from smccat import Engine, lg_base_matrix
from smccat import simulate as sim

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _check_final(eng, batch, codes, R, lnl, n_check=8):
    final = eng.snapshot(range(n_check), batch)
    want = np.array([orc.particle_lnl(codes, final.particle(i), R) for i in range(n_check)])
    np.testing.assert_allclose(lnl[:n_check], want, rtol=1e-10)


def test_sampler_lg_c1_shape(gpu):
    _say("case sampler c1: 16 taxa x 200 sites LG+G4, 64 particles, 4 iterations")
    codes, cfg = sim.simulate_config("c1")
    T, S, R, N = cfg["taxa"], cfg["sites"], cfg["rates"], 64
    batch = sim.prior_particles(np.random.default_rng(5), N, T, S, num_rates=R, lg=True,
                                lg_table=lg_base_matrix())
    with Engine(T, S, R, N, 1, workspaces=2) as eng:
        eng.set_alignment(codes)
        eng.set_particles(batch)
        l0, _ = eng.refresh_likelihood()
        res, lnl, w = eng.run_sampler(N, 4, seed=1, using_lg=True, invariant=False)
        _say("  %s" % {k: v for k, v in res.items() if k != "move_stats"})
        assert res["final_temp"] == 1.0 and res["resamples"] >= 1
        assert 0 < res["final_ess"] <= N
        assert res["move_stats"][0, 0] > 0 and res["move_stats"][0, 1] > 0
        assert np.all(np.isfinite(lnl)) and np.median(lnl) > np.median(l0)
        np.testing.assert_allclose(w, -np.log(N)) if res["final_ess"] == N else None
        _check_final(eng, batch, codes, R, lnl)


def test_sampler_gtr_invariant_batched_vs_single(gpu):
    _say("case sampler gtr: 10 taxa x 80 sites GTR+G4+I, 24 particles, batch 24 and batch 5")
    rng = np.random.default_rng(6)
    T, S, R, N = 10, 80, 4, 24
    codes = sim.simulate_config("c1")[0][:T, :S].copy()
    batch = sim.prior_particles(rng, N, T, S, num_rates=R, invar=True)
    out = {}
    for b in (0, 5):
        with Engine(T, S, R, N, 1, workspaces=2) as eng:
            eng.set_alignment(codes)
            eng.set_particles(batch)
            res, lnl, w = eng.run_sampler(N, 3, seed=2, batch=b, max_iterations=200)
            _say("  batch %d: %s" % (b, {k: v for k, v in res.items() if k != "move_stats"}))
            assert res["final_temp"] == 1.0
            assert res["move_stats"][3, 0] > 0 and res["move_stats"][2, 0] > 0
            _check_final(eng, batch, codes, R, lnl)
            out[b] = (res["final_ess"], lnl)
    # per-slot random streams: lockstep batches of any size run the same chains
    assert out[0][0] == out[5][0]
    np.testing.assert_allclose(out[0][1], out[5][1], rtol=1e-12)


def test_sampler_cat_with_gibbs(gpu):
    _say("case sampler cat: 12 taxa x 60 sites LG+G4 CAT K<=4, 16 particles, Gibbs p=0.5")
    rng = np.random.default_rng(7)
    T, S, R, N = 12, 60, 4, 16
    codes = sim.simulate_config("c1")[0][:T, :S].copy()
    batch = sim.prior_particles(rng, N, T, S, num_rates=R, lg=True, lg_table=lg_base_matrix(),
                                k_range=(1, 4))
    with Engine(T, S, R, N, 8, workspaces=2) as eng:
        eng.set_alignment(codes)
        eng.set_particles(batch)
        res, lnl, w = eng.run_sampler(N, 2, seed=3, using_lg=True, invariant=False, cat=True,
                                      alg8_probability=0.5, dpp_alpha=1.0)
        assert res["final_temp"] == 1.0
        _check_final(eng, batch, codes, R, lnl, n_check=N)


def test_sampler_cat_learned_proposal(gpu):
    _say("case sampler learned: 12 taxa x 120 sites LG+G4 CAT K<=6, 24 particles, warm-up 1")
    from smccat import ProposalNetwork
    rng = np.random.default_rng(8)
    T, S, R, N = 12, 120, 4, 24
    codes, _ = sim.simulate_config("c3", taxa=T, sites=S)
    batch = sim.prior_particles(rng, N, T, S, num_rates=R, lg=True, lg_table=lg_base_matrix(),
                                k_range=(2, 6), max_classes=16)
    with Engine(T, S, R, N, 16, workspaces=2) as eng:
        eng.set_alignment(codes)
        eng.set_particles(batch)
        with ProposalNetwork(eng, seed=1) as net:
            res, lnl, w = eng.run_sampler(N, 2, seed=4, using_lg=True, invariant=False, cat=True,
                                          alg8_probability=0.5, dpp_alpha=1.0, proposal=net,
                                          proposal_warmup=2, train_epochs=3)
            _say("  %s" % res)
            assert res["final_temp"] == 1.0 and res["resamples"] > 2
            assert res["cat_stats"][2] > 0 and net.num_examples == res["cat_stats"][2]
            assert res["cat_stats"][0] > res["cat_stats"][2] and res["cat_stats"][1] > 0
            assert res["training_seconds"] > 0
        _check_final(eng, batch, codes, R, lnl, n_check=N)


def test_cat_moves_do_not_depend_on_their_lockstep_companions(gpu):
    """The sampler runs the CAT moves of all its lockstep batches together after the
    last batch; that gives the same particles because a slot's move depends only on
    the slot and the seed.  Checked here directly: rejuvenation with the CAT move of
    all 12 slots in one lockstep equals four locksteps of 3 slots, for the Gibbs and the
    learned move; and the sampler gives the same population with batch 1 (moves inline),
    3 and 12 (moves deferred)."""
    _say("case cat lockstep independence: 12 taxa x 60 sites LG+G4 CAT K<=4, 12 particles")
    from smccat import ProposalNetwork
    rng = np.random.default_rng(11)
    T, S, R, N = 12, 60, 4, 12
    codes = sim.simulate_config("c1")[0][:T, :S].copy()
    batch = sim.prior_particles(rng, N, T, S, num_rates=R, lg=True, lg_table=lg_base_matrix(),
                                k_range=(1, 4))

    def state(eng):
        out = []
        for i in range(N):
            k, lg, a = eng.particle_state(i)
            out.append((k, lg[:k].copy(), a.copy()))
        return out

    with Engine(T, S, R, N, 8, workspaces=2) as eng:
        eng.set_alignment(codes)
        net = ProposalNetwork(eng, seed=4)
        for learned in (False, True):
            runs = []
            for groups in ((0, N),), ((0, 3), (3, 3), (6, 3), (9, 3)):
                eng.set_particles(batch)
                lnl = np.concatenate([eng.rejuvenate(0.7, 2, seed=9, first=f, count=c, using_lg=True,
                                                     invariant=False, gibbs_probability=0.6, dpp_alpha=1.0,
                                                     proposal=net if learned else None)[0]
                                      for f, c in groups])
                runs.append((lnl, state(eng)))
            np.testing.assert_array_equal(runs[0][0], runs[1][0])
            for (k0, l0, a0), (k1, l1, a1) in zip(runs[0][1], runs[1][1]):
                assert k0 == k1
                np.testing.assert_array_equal(l0, l1)
                np.testing.assert_array_equal(a0, a1)
        net.close()
        outs = []
        for b in (1, 3, N):  # 1: each move inline after its particle's MCMC; 3, N: deferred
            eng.set_particles(batch)
            res, lnl, w = eng.run_sampler(N, 2, seed=3, using_lg=True, invariant=False, cat=True,
                                          alg8_probability=0.5, dpp_alpha=1.0, batch=b, max_iterations=6)
            outs.append((lnl, state(eng)))
        for other in outs[1:]:
            np.testing.assert_array_equal(outs[0][0], other[0])
            for (k0, l0, a0), (k1, l1, a1) in zip(outs[0][1], other[1]):
                assert k0 == k1
                np.testing.assert_array_equal(a0, a1)
